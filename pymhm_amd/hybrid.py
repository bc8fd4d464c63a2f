"""Ordered global scatter of condensed blocks (replaces ``core/contributions.py:77-195``).

The reference expands each cell's ``(m+k)^2`` block into row-major COO
coordinates in cell order (``core/_assembly.py:9-25``), sums duplicates in the
COO -> CSC conversion, drops explicit zeros, and accumulates the reduced loads
and their absolute values serially in cell order (``core/_assembly.py:28-42``).

Here the integer work is fixed once per macro mesh: the COO stream, the CSC
pattern it produces and, for every global nonzero and every global load row,
the ordered list of stream positions that sum into it.  Values are then
gathered on the GPU in that fixed order (``mhm_ordered_gather``), so results
are deterministic and independent of batching.  Shared macrofaces give at most
two contributions per entry, for which the order of an IEEE sum is immaterial,
so the values equal the reference's bit for bit given equal blocks.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch
from scipy import sparse

from . import _lib


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _groups(keys):
    """Stable grouping of stream positions by key -> (unique keys, ptr, positions)."""
    order = np.argsort(keys, kind="stable")
    uniq, counts = np.unique(keys[order], return_counts=True)
    ptr = np.zeros(len(uniq) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(counts)
    return uniq, ptr, order.astype(np.int64)


class HybridScatter:
    """Fixed scatter plan of one macro mesh.

    ``indices`` is ``(M, w)`` per cell in cell order, or a list of per-cell
    index arrays of different widths (uneven macros).  The values gathered are
    a flat device buffer of the cells' row-major blocks and one of their
    loads; ``positions`` / ``load_positions`` give where each cell's block and
    load start in those buffers (default: cell order, back to back), so
    blocks produced bucket by bucket need no reordering.
    """

    def __init__(self, indices, trace_size: int, kernel_offsets, device=0, positions=None, load_positions=None):
        fixed = isinstance(indices, np.ndarray) and indices.ndim == 2
        rows_of = [np.asarray(r, dtype=np.int64) for r in (indices if not fixed else list(indices))]
        if not fixed and not all(r.ndim == 1 for r in rows_of):
            raise ValueError("indices must be (cells, width) or a list of 1-D index arrays")
        offs = np.asarray(kernel_offsets, dtype=np.int64)
        self.size = int(offs[-1])
        self.trace_size = int(trace_size)
        M = len(rows_of)
        widths = np.array([len(r) for r in rows_of], dtype=np.int64)
        if len(offs) != M + 1 or offs[0] != trace_size or np.any(np.diff(offs) < 0):
            raise ValueError("kernel_offsets must be an ordered partition starting at trace_size")
        allidx = np.concatenate(rows_of) if M else np.zeros(0, dtype=np.int64)
        if np.any(allidx < 0) or np.any(allidx >= self.size):
            raise ValueError("contribution indices must be valid global integers")
        for c, row in enumerate(rows_of):
            if len(np.unique(row)) != len(row):
                raise ValueError("contribution indices must be distinct")
            if not np.array_equal(np.sort(row[row >= trace_size]), np.arange(offs[c], offs[c + 1])):
                raise ValueError("contribution coarse indices must match its ordered cell partition")
        if not np.array_equal(np.unique(allidx), np.arange(self.size)):
            raise ValueError("contributions must cover the contiguous global numbering")
        bpos = (np.concatenate(([0], np.cumsum(widths * widths)[:-1])) if positions is None
                else np.asarray(positions, dtype=np.int64))
        lpos = (np.concatenate(([0], np.cumsum(widths)[:-1])) if load_positions is None
                else np.asarray(load_positions, dtype=np.int64))
        self.M = M
        self.w = int(widths[0]) if fixed else None
        self.indices = np.stack(rows_of) if fixed else rows_of
        self.device = torch.device("cuda", device) if torch.cuda.is_available() else None
        # the COO stream in the reference's order: cells in order, each block row-major
        rows = np.concatenate([np.repeat(r, len(r)) for r in rows_of])
        cols = np.concatenate([np.tile(r, len(r)) for r in rows_of])
        codes = np.concatenate([bpos[c] + np.arange(widths[c] ** 2) for c in range(M)])
        keys, vptr, order = _groups(cols * self.size + rows)
        vcode = codes[order]
        self.csc_indices = (keys % self.size).astype(np.int32)
        colcount = np.bincount((keys // self.size).astype(np.int64), minlength=self.size)
        self.csc_indptr = np.zeros(self.size + 1, dtype=np.int32)
        self.csc_indptr[1:] = np.cumsum(colcount)
        lcodes = np.concatenate([lpos[c] + np.arange(widths[c]) for c in range(M)])
        _, lptr, lorder = _groups(allidx)
        lcode = lcodes[lorder]
        self._host = dict(vptr=vptr, vcode=vcode, lptr=lptr, lcode=lcode)
        if self.device is not None:
            self.vptr = torch.as_tensor(vptr, device=self.device)
            self.vcode = torch.as_tensor(vcode, device=self.device)
            self.lptr = torch.as_tensor(lptr, device=self.device)
            self.lcode = torch.as_tensor(lcode, device=self.device)
            if fixed:
                self.d_indices = torch.as_tensor(self.indices, device=self.device)

    def coordinates(self):
        """Row-major COO ``(rows, cols)`` of the whole stream on the GPU (fixed width)."""
        if self.w is None:
            raise ValueError("device coordinates need one contribution width")
        rows = torch.empty(self.M * self.w * self.w, dtype=torch.int64, device=self.device)
        cols = torch.empty_like(rows)
        _lib.call("mhm_scatter_coordinates", _ptr(self.d_indices), self.w, self.M, _ptr(rows), _ptr(cols),
                  _stream())
        return rows, cols

    def gather(self, blocks, loads, boundary=None):
        """Global CSC values (pattern before zero elimination), rhs and load scale on the GPU.

        ``boundary`` = ``(DirichletBoundary, g)`` applies the weak Dirichlet
        load of the data ``g`` on the device (``rhs -= load``, ``scale += |load|``)."""
        blocks = blocks.reshape(-1).contiguous()
        loads = loads.reshape(-1).contiguous()
        vals = torch.empty(len(self.csc_indices), dtype=torch.float64, device=self.device)
        rhs = torch.empty(self.size, dtype=torch.float64, device=self.device)
        scale = torch.empty_like(rhs)
        _lib.call("mhm_ordered_gather", _ptr(self.vptr), _ptr(self.vcode), _ptr(blocks),
                  int(len(self.csc_indices)), _ptr(vals), None, _stream())
        _lib.call("mhm_ordered_gather", _ptr(self.lptr), _ptr(self.lcode), _ptr(loads), int(self.size),
                  _ptr(rhs), _ptr(scale), _stream())
        if boundary is not None:
            bnd, g = boundary
            if bnd.trace_size != self.trace_size:
                raise ValueError("the boundary belongs to another skeleton")
            bnd.load(g, rhs=rhs, scale=scale)
        return vals, rhs, scale

    def gather_loads(self, loads, boundary=None):
        """Global rhs and load scale of new loads on the GPU, the matrix unchanged
        (the online assembly of ``core/offline.py:272-286``); ``boundary`` as in
        ``gather``."""
        loads = loads.reshape(-1).contiguous()
        rhs = torch.empty(self.size, dtype=torch.float64, device=self.device)
        scale = torch.empty_like(rhs)
        _lib.call("mhm_ordered_gather", _ptr(self.lptr), _ptr(self.lcode), _ptr(loads), int(self.size),
                  _ptr(rhs), _ptr(scale), _stream())
        if boundary is not None:
            bnd, g = boundary
            if bnd.trace_size != self.trace_size:
                raise ValueError("the boundary belongs to another skeleton")
            bnd.load(g, rhs=rhs, scale=scale)
        return rhs, scale

    def assemble(self, blocks, loads, boundary_load=None, boundary=None):
        """``(matrix, rhs, load_scale)`` as ``assemble_hybrid_contributions`` returns them
        (``boundary_load`` a host vector, or ``boundary`` applied on the device)."""
        vals, rhs, scale = self.gather(blocks, loads, boundary)
        mat = sparse.csc_matrix((vals.cpu().numpy(), self.csc_indices.copy(), self.csc_indptr.copy()),
                                shape=(self.size, self.size))
        mat.eliminate_zeros()
        rhs = rhs.cpu().numpy()
        scale = scale.cpu().numpy()
        if boundary_load is not None:
            b = np.asarray(boundary_load, dtype=float)
            if b.shape != (self.trace_size,) or not np.all(np.isfinite(b)):
                raise ValueError("boundary_load must be finite with shape (trace_size,)")
            rhs[:self.trace_size] -= b
            scale[:self.trace_size] += np.abs(b)
        return mat, rhs, scale


class DirichletBoundary:
    """Weak Dirichlet data on the boundary macrofaces of a 2D macro mesh, on the
    device (``boundary_data``, ``src/pymhm/fem/scalar/operators.py:247-321``,
    scalar Dirichlet data, no Neumann faces).

    Every boundary face uses the face space's segment quadrature of order
    ``max(order, degree + 2)`` from the face's stored start to its end;
    ``points`` ``(faces, q, 2)`` are where the caller evaluates the data ``g``.
    ``load(g)`` returns the trace load ``|f| sum_q mu_a(t_q) w_q g_q`` per
    face DOF (zero on interior faces) and can apply it to a global rhs.
    """

    def __init__(self, macro, face_space, order: int = 5, device: int = 0):
        self.space = face_space
        fs = face_space.size
        faces = np.asarray(macro.boundary_faces, dtype=np.int64)
        par, w = face_space.quadrature(max(int(order), max(face_space.degrees) + 2))
        basis = face_space.evaluate(par)
        start = macro.points[macro.faces[faces, 0]]
        end = macro.points[macro.faces[faces, 1]]
        self.faces = faces
        self.points = start[:, None, :] + par[None, :, None] * (end - start)[:, None, :]
        self.length = np.linalg.norm(end - start, axis=1)
        self.trace_size = len(macro.faces) * fs
        self.nq, self.fs = len(par), fs
        self.device = torch.device("cuda", device)
        self.d_basis = torch.as_tensor(np.ascontiguousarray(basis), device=self.device)
        self.d_weights = torch.as_tensor(w, device=self.device)
        self.d_length = torch.as_tensor(self.length, device=self.device)
        self.d_dof0 = torch.as_tensor(faces * fs, device=self.device)

    def load(self, g, rhs=None, scale=None):
        """The trace load of data ``g`` ``(faces, q)`` at ``points``; with
        ``rhs``/``scale`` (device vectors of the global system) also
        ``rhs -= load`` and ``scale += |load|`` on the boundary rows."""
        g = torch.as_tensor(g, dtype=torch.float64, device=self.device).contiguous()
        if tuple(g.shape) != (len(self.faces), self.nq):
            raise ValueError(f"g must have shape ({len(self.faces)}, {self.nq})")
        out = torch.zeros(self.trace_size, dtype=torch.float64, device=self.device)
        _lib.call("mhm_dirichlet_load", len(self.faces), self.nq, self.fs, _ptr(self.d_basis), _ptr(self.d_weights),
                  _ptr(g), _ptr(self.d_length), _ptr(self.d_dof0), _ptr(out), _ptr(rhs), _ptr(scale), _stream())
        return out


class HybridSolver:
    """The assembled global hybrid system solved on the device.

    Mirrors ``SaddleBlockSolver`` (``src/pymhm/linalg/block.py:99-237``):
    prepared once per matrix (``set_values``), then ``solve`` for any number of
    right-hand sides against the resident preconditioner, and ``close``.  The
    Krylov method is restarted GMRES, right-preconditioned by weighted additive
    Schwarz over the plan's contribution cells (each cell's assembled block
    inverted on the device); a solve is accepted when the double-double
    residual of the original equations meets ``||b - S x|| <= rtol ||b||``
    (``linalg/linear.py:132-152``), else it raises ``LinearSolveError``.
    ``restart`` defaults to ``maxiter``: the whole Krylov basis stays in HBM.
    """

    def __init__(self, plan: HybridScatter, rtol: float = 1e-10, maxiter: int = 3000, restart: int | None = None,
                 device: int = 0):
        if not (0.0 < rtol < 1.0):
            raise ValueError("rtol must lie in (0, 1)")
        self.plan = plan
        self.rtol, self.maxiter = float(rtol), int(maxiter)
        self.restart = self.maxiter if restart is None else int(restart)
        if self.maxiter < 1 or self.restart < 1:
            raise ValueError("maxiter and restart must be positive")
        self.device = torch.device("cuda", device)
        cells = [np.asarray(r, dtype=np.int64) for r in (plan.indices if plan.w is None else list(plan.indices))]
        cell_ptr = np.zeros(len(cells) + 1, dtype=np.int64)
        cell_ptr[1:] = np.cumsum([len(r) for r in cells])
        cell_idx = np.ascontiguousarray(np.concatenate(cells))
        indptr = np.ascontiguousarray(plan.csc_indptr, dtype=np.int32)
        indices = np.ascontiguousarray(plan.csc_indices, dtype=np.int32)
        h = ctypes.c_void_p()
        _lib.call("mhm_global_create", device, int(plan.size), int(len(indices)), indptr.ctypes.data,
                  indices.ctypes.data, len(cells), cell_ptr.ctypes.data, cell_idx.ctypes.data, ctypes.byref(h))
        self._handle = h.value
        self.iterations: list[int] = []
        self.residuals: list[float] = []
        self.bad_cell = -1

    @property
    def handle(self):
        if self._handle is None:
            raise RuntimeError("hybrid solver is closed")
        return ctypes.c_void_p(self._handle)

    def set_values(self, vals):
        """Device CSC values of the plan's pattern (``HybridScatter.gather``)."""
        vals = torch.as_tensor(vals, dtype=torch.float64, device=self.device).contiguous()
        if vals.numel() != len(self.plan.csc_indices):
            raise ValueError("values must match the plan's CSC pattern")
        bad = ctypes.c_int32(-1)
        try:
            _lib.call("mhm_global_set_values", self.handle, _ptr(vals), ctypes.byref(bad), _stream())
        finally:
            self.bad_cell = bad.value

    def solve(self, rhs):
        """``x`` (device) with ``S x = rhs`` for one right-hand side."""
        rhs = torch.as_tensor(rhs, dtype=torch.float64, device=self.device).contiguous()
        if tuple(rhs.shape) != (self.plan.size,):
            raise ValueError(f"rhs must have shape ({self.plan.size},)")
        x = torch.empty_like(rhs)
        it, res = ctypes.c_int32(0), ctypes.c_double(0.0)
        try:
            _lib.call("mhm_global_solve", self.handle, _ptr(rhs), _ptr(x), self.rtol, self.maxiter, self.restart,
                      ctypes.byref(it), ctypes.byref(res), _stream())
        finally:
            self.iterations.append(it.value)
            self.residuals.append(res.value)
        return x

    def close(self):
        if self._handle is not None:
            _lib.call("mhm_global_destroy", ctypes.c_void_p(self._handle))
            self._handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_hybrid(plan: HybridScatter, blocks, loads, boundary=None, rtol: float = 1e-10, maxiter: int = 3000):
    """Gather, prepare and solve on the device: ``(trace, coarse, solver)``
    (device tensors; the solver keeps the preconditioner for more right-hand
    sides).  The device counterpart of ``solve_hybrid_system`` without gauges
    (``src/pymhm/core/system.py:355-528``)."""
    vals, rhs, _ = plan.gather(blocks, loads, boundary)
    solver = HybridSolver(plan, rtol=rtol, maxiter=maxiter)
    solver.set_values(vals)
    x = solver.solve(rhs)
    return x[:plan.trace_size], x[plan.trace_size:], solver
