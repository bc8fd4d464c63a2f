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
