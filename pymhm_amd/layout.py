"""Band ordering, CSR pattern and assembly maps of one local shape (host side).

A ``LocalLayout`` is built once per local mesh shape and uploaded with
``mhm_layout_create``.  It fixes:

* the band permutation (breadth-first level sets from one macro side, ordered
  along that side, or reverse Cuthill-McKee when narrower), which makes the
  pinned operator ``A + sigma e_p e_p^T`` a band matrix of half-bandwidth b;
* the symmetric CSR pattern in band numbering, row-sorted;
* for every CSR slot the ordered list of element contributions (element index
  ascending, as the COO -> CSC summation of ``fem/scalar/triangle.py:250-254``
  visits them), and likewise for every load row (``np.bincount`` order);
* the template coupling B, constraint C and retained mode Z in band rows.
"""

from __future__ import annotations

import ctypes
from collections import deque

import numpy as np
from scipy import sparse
from scipy.sparse.csgraph import reverse_cuthill_mckee

from . import _lib


def element_pattern(n: int, edofs):
    """Symmetric CSR pattern (rows sorted) of all DOF pairs sharing an element."""
    edofs = np.asarray(edofs, dtype=np.int64)
    nb = edofs.shape[1]
    rows = np.repeat(edofs, nb, axis=1).ravel()
    cols = np.tile(edofs, (1, nb)).ravel()
    mat = sparse.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n)).tocsr()
    mat.sum_duplicates()
    mat.sort_indices()
    return mat.indptr.astype(np.int64), mat.indices.astype(np.int64)


def summation_order(n: int, edofs):
    """Order in which SciPy's COO -> CSC conversion sums duplicate element entries.

    The reference assembles ``coo_matrix((blocks.ravel(), (repeat(dofs), tile(dofs)))).tocsc()``
    (``fem/scalar/triangle.py:250-254``): entries are placed column by column in
    stream order, each column's row indices are sorted (an unstable sort whose
    result depends only on the index sequence) and equal neighbours are summed
    left to right.  Running SciPy's own conversion on entry ids recovers that
    order, so the GPU gather adds every slot's contributions in the same order.
    Returns the COO stream positions (``e*nb*nb + i*nb + j``) in summation order.
    """
    edofs = np.asarray(edofs, dtype=np.int64)
    nb = edofs.shape[1]
    rows = np.repeat(edofs, nb, axis=1).ravel()
    cols = np.tile(edofs, (1, nb)).ravel()
    first = np.argsort(cols, kind="stable")          # column placement keeps stream order
    counts = np.bincount(cols, minlength=n)
    indptr = np.r_[0, np.cumsum(counts)].astype(np.int64)
    ids = first.astype(np.float64)                   # exact below 2**53
    csc = sparse.csc_matrix((ids, rows[first].copy(), indptr), shape=(n, n))
    csc.has_sorted_indices = False
    csc.sort_indices()
    return csc.data.astype(np.int64)


def bandwidth(perm, indptr, indices) -> int:
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    return int(np.max(np.abs(inv[rows] - inv[indices]), initial=0))


def level_order(indptr, indices, start_nodes, coords=None, axis=None):
    """Breadth-first level sets from ``start_nodes``.

    Within a level, nodes are sorted by their distance from the start side,
    then along ``axis`` (the start side's direction), which keeps lattice rows
    contiguous on structured local meshes.
    """
    n = len(indptr) - 1
    level = np.full(n, -1, dtype=np.int64)
    q = deque()
    for s in start_nodes:
        if level[s] < 0:
            level[s] = 0
            q.append(int(s))
    while q:
        v = q.popleft()
        for w in indices[indptr[v]:indptr[v + 1]]:
            if level[w] < 0:
                level[w] = level[v] + 1
                q.append(int(w))
    if np.any(level < 0):
        raise ValueError("local pattern is disconnected")
    if coords is None:
        along = across = np.zeros(n)
    else:
        xy = np.asarray(coords, dtype=float)
        t = np.asarray(axis, dtype=float)
        t = t / np.sqrt(t @ t)
        along = np.round(xy @ t, 12)
        across = np.round(xy[:, 0] * -t[1] + xy[:, 1] * t[0], 12)
    return np.lexsort((np.arange(n), along, across, level)).astype(np.int64)


def choose_permutation(indptr, indices, start_nodes=None, coords=None, axis=None):
    """Narrowest of the side-level order and reverse Cuthill-McKee."""
    n = len(indptr) - 1
    cands = []
    if start_nodes is not None and len(start_nodes):
        cands.append(level_order(indptr, indices, start_nodes, coords, axis))
    g = sparse.csr_matrix((np.ones(len(indices)), indices, indptr), shape=(n, n))
    cands.append(np.asarray(reverse_cuthill_mckee(g, symmetric_mode=True), dtype=np.int64))
    widths = [bandwidth(p, indptr, indices) for p in cands]
    return cands[int(np.argmin(widths))]


class LocalLayout:
    """Device-resident description of one local shape (see module docstring)."""

    def __init__(self, n, perm, indptr, indices, *, m, k=1, coupling=None, constraint=None,
                 kernel=None, assembly=None, device=0):
        self.n, self.m, self.k = int(n), int(m), int(k)
        perm = np.asarray(perm, dtype=np.int64)
        if sorted(perm.tolist()) != list(range(self.n)):
            raise ValueError("perm must be a permutation")
        self.perm = perm
        inv = np.empty_like(perm)
        inv[perm] = np.arange(self.n)
        self.inv = inv
        # pattern in band numbering
        rows = np.repeat(np.arange(self.n), np.diff(indptr))
        pat = sparse.coo_matrix((np.ones(len(indices)), (inv[rows], inv[np.asarray(indices)])),
                                shape=(self.n, self.n)).tocsr()
        pat.sort_indices()
        self.rowptr = pat.indptr.astype(np.int32)
        self.cols = pat.indices.astype(np.int32)
        self.nnz = len(self.cols)
        brow = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        self.bandwidth = int(np.max(np.abs(brow - self.cols), initial=0))
        self.pin = self.n - 1 if self.k else -1
        self.slot_of = None
        keep = []
        self._host = {}

        def arr(name, a, dtype):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            self._host[name] = a
            keep.append(a)
            return a.ctypes.data

        d = _lib.LayoutDesc()
        d.n, d.k, d.m, d.nnz = self.n, self.k, self.m, self.nnz
        d.bandwidth, d.pin = self.bandwidth, self.pin
        d.perm = arr("perm", perm, np.int32)
        d.rowptr = arr("rowptr", self.rowptr, np.int32)
        d.cols = arr("cols", self.cols, np.int32)
        if coupling is not None:
            bt = np.asarray(coupling, dtype=float)[perm, :]  # band rows
            d.coupling = arr("coupling", bt.T, np.float64)
            crow = np.flatnonzero(np.any(bt != 0.0, axis=1)).astype(np.int32)
            d.coupling_rows = arr("coupling_rows", crow, np.int32)
            d.ncoupling_rows = len(crow)
        if kernel is not None:
            d.kernel = arr("kernel", np.asarray(kernel, dtype=float)[perm], np.float64)
        if constraint is not None:
            d.constraint = arr("constraint", np.asarray(constraint, dtype=float)[perm], np.float64)
        if assembly is not None:
            self._assembly_maps(assembly, d, arr)
        self._desc = d
        self._keep = keep
        self.device = int(device)
        self._handle = None

    @property
    def handle(self):
        """Device layout, created on first use (host-side maps need no GPU)."""
        if self._handle is None:
            handle = ctypes.c_void_p()
            _lib.call("mhm_layout_create", self.device, ctypes.byref(self._desc), ctypes.byref(handle))
            self._handle = handle
            if getattr(self, "_skel_desc", None) is not None:
                _lib.call("mhm_layout_set_skeleton", handle, ctypes.byref(self._skel_desc))
        return self._handle

    def set_skeleton(self, tables):
        """Attach the macroface skeleton (``meshing.skeleton_tables`` of the
        template macro, local DOF ids) so batches can compute their couplings on
        the device (``LocalBatch.skeleton_coupling``)."""
        t = dict(tables)
        if t["nsides"] * t["nseg"] * (t["degree"] + 1) != self.m:
            raise ValueError("skeleton size does not match the layout's multiplier count")
        rows = self.inv[np.asarray(t["entry_node"], dtype=np.int64)]
        keep = {}

        def arr(name, a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep[name] = a
            return a.ctypes.data

        d = _lib.SkeletonDesc()
        d.nsides, d.nseg, d.degree, d.edeg = t["nsides"], t["nseg"], t["degree"], t["edeg"]
        d.ng = len(t["gx"])
        d.breaks = arr("breaks", t["breaks"], np.float64)
        d.gx = arr("gx", t["gx"], np.float64)
        d.gw = arr("gw", t["gw"], np.float64)
        d.npieces = len(t["piece_f"])
        d.piece_f = arr("piece_f", t["piece_f"], np.float64)
        d.nentries = len(rows)
        d.entry_row = arr("entry_row", rows, np.int32)
        d.entry_side = arr("entry_side", t["entry_side"], np.int32)
        d.entry_seg = arr("entry_seg", t["entry_seg"], np.int32)
        d.entry_ptr = arr("entry_ptr", t["entry_ptr"], np.int32)
        d.contrib = arr("contrib", t["contrib"], np.int32)
        self._skel_keep = keep
        self._skel_desc = d
        if self._handle is not None:
            _lib.call("mhm_layout_set_skeleton", self._handle, ctypes.byref(d))

    def _assembly_maps(self, asm, d, arr):
        edofs = np.asarray(asm["dofs"], dtype=np.int64)
        ne, nb = edofs.shape
        bd = self.inv[edofs]
        # slot index of every (e, i, j)
        rows = np.repeat(bd, nb, axis=1)          # (ne, nb*nb): row of (i, j) = bd[e, i]
        cols = np.tile(bd, (1, nb))               # col of (i, j) = bd[e, j]
        key = rows.astype(np.int64) * self.n + cols
        rowkey = np.repeat(np.arange(self.n), np.diff(self.rowptr)).astype(np.int64) * self.n + self.cols
        slot = np.searchsorted(rowkey, key)
        if np.any(rowkey[slot] != key):
            raise ValueError("element pair outside the pattern")
        ij = np.arange(nb * nb)
        code = (ij[None, :] * ne + np.arange(ne)[:, None]).ravel()   # (i*nb+j)*ne + e
        flat = slot.ravel()
        order = summation_order(self.n, edofs)
        flat_sorted = flat[order]
        # contributions of one slot must be consecutive in the summation order
        starts = np.r_[True, flat_sorted[1:] != flat_sorted[:-1]]
        if len(np.unique(flat_sorted[starts])) != starts.sum():
            raise ValueError("summation order does not group the slots")
        order = order[np.argsort(flat_sorted, kind="stable")]
        slot_ptr = np.zeros(self.nnz + 1, dtype=np.int64)
        np.add.at(slot_ptr, flat + 1, 1)
        slot_ptr = np.cumsum(slot_ptr)
        lrow = bd.ravel()
        lcode = (np.arange(nb)[None, :] * ne + np.arange(ne)[:, None]).ravel()  # b*ne + e
        lorder = np.lexsort((np.repeat(np.arange(ne), nb), lrow))
        load_ptr = np.zeros(self.n + 1, dtype=np.int64)
        np.add.at(load_ptr, lrow + 1, 1)
        load_ptr = np.cumsum(load_ptr)
        d.ne, d.nb, d.dim, d.nq = ne, nb, int(asm["dim"]), len(asm["qweights"])
        d.slot_ptr = arr("slot_ptr", slot_ptr, np.int32)
        d.slot_code = arr("slot_code", code[order], np.int32)
        d.load_ptr = arr("load_ptr", load_ptr, np.int32)
        d.load_code = arr("load_code", lcode[lorder], np.int32)
        d.grad_lambda = arr("grad_lambda", asm["grad_lambda"], np.float64)
        d.measure = arr("measure", asm["measure"], np.float64)
        d.qweights = arr("qweights", asm["qweights"], np.float64)
        d.dtable = arr("dtable", asm["dtable"], np.float64)
        d.vtable = arr("vtable", asm["vtable"], np.float64)
        self.ne, self.nb, self.nq, self.dim = ne, nb, d.nq, d.dim

    def slot_map(self, indptr, indices):
        """Slot of every stored entry of an original-numbering CSR/CSC with this pattern."""
        rows = np.repeat(np.arange(self.n), np.diff(indptr))
        key = self.inv[rows].astype(np.int64) * self.n + self.inv[np.asarray(indices)]
        rowkey = np.repeat(np.arange(self.n), np.diff(self.rowptr)).astype(np.int64) * self.n + self.cols
        slot = np.searchsorted(rowkey, key)
        if np.any(slot >= self.nnz) or np.any(rowkey[np.minimum(slot, self.nnz - 1)] != key):
            raise ValueError("matrix entry outside the layout pattern")
        return slot

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            _lib.load().mhm_layout_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
