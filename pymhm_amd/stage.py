"""Resident batched local stage on one GPU (thin host wrapper of the C ABI).

``LocalBatch`` owns one ``mhm_batch``: the band factors, the binary64
responses and the work buffers of ``macros`` local problems of one
``LocalLayout``.  Device arrays are PyTorch tensors used only as buffers;
every numerical operation runs in ``libmhm_hip.so``.

Mirrors ``BatchedFactorization`` (``src/pymhm/execution/cuda.py:165-288``):
inputs are copied, the object stays on its construction device, ``close()``
is idempotent and runs on exceptions, failures raise ``ValueError`` or
``LinearSolveError`` and never fall back to another solver.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import LinearSolveError

PRECISIONS = {"fp64": _lib.MHM_FP64, "mixed": _lib.MHM_MIXED}
DEFAULT_REFINE = {"fp64": 2, "mixed": 8}


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class LocalBatch:
    """``macros`` local problems of one layout with resident factors."""

    def __init__(self, layout, macros: int, *, precision: str = "fp64", max_columns: int = 0):
        if precision not in PRECISIONS:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if not torch.cuda.is_available():
            raise _lib.SolverUnavailableError("the HIP local stage requires a visible GPU")
        self.layout = layout
        self.macros = int(macros)
        self.precision = precision
        self.max_columns = int(max_columns)
        self.device = torch.device("cuda", layout.device)
        self.n, self.m, self.k = layout.n, layout.m, layout.k
        self.N = self.n + self.k
        self.R = 1 + self.m + self.k
        self.width = self.m + self.k
        handle = ctypes.c_void_p()
        _lib.call("mhm_batch_create", layout.handle, self.macros, PRECISIONS[precision],
                  self.max_columns, ctypes.byref(handle))
        self.handle = handle
        self.solve_variants = (_lib.SOLVE_AUTO, _lib.SOLVE_AUTO)
        self._keep = []
        self.blocks = None
        self.loads = None

    # -------------------------------------------------------------- inputs
    def _dev(self, a, dtype):
        t = torch.as_tensor(a, dtype=dtype, device=self.device).contiguous()
        return t

    def assemble(self, kq, fq=None, f_const: float = 0.0, check: bool = False):
        """Assemble from diffusion (and source) values at quadrature points ``(M, ne, nq)``.

        The device checks that the diffusion is finite and positive and the
        source finite at every point (``materials/evaluation.py:16-29``); with
        ``check`` a violation raises ``ValueError`` here (one synchronisation),
        otherwise it is reported by ``status``/``raise_on_failure`` later."""
        kq = self._dev(kq, torch.float64)
        shape = (self.macros, self.layout.ne, self.layout.nq)
        if tuple(kq.shape) != shape:
            raise ValueError(f"kq must have shape {shape}")
        if fq is not None:
            fq = self._dev(fq, torch.float64)
            if tuple(fq.shape) != shape:
                raise ValueError(f"fq must have shape {shape}")
        if not np.isfinite(f_const):
            raise ValueError("f_const must be finite")
        self._keep = [kq, fq]
        _lib.call("mhm_batch_assemble", self.handle, _ptr(kq), _ptr(fq), float(f_const), _stream())
        if check:
            self._raise_coefficients(self.status())

    def set_operator(self, values, loads):
        """Assembled operators: CSR values in layout slot order ``(M, nnz)``, loads ``(M, n)``."""
        values = self._dev(values, torch.float64)
        loads = self._dev(loads, torch.float64)
        if tuple(values.shape) != (self.macros, self.layout.nnz) or tuple(loads.shape) != (self.macros, self.n):
            raise ValueError("values/loads have the wrong shape")
        self._keep = [values, loads]
        _lib.call("mhm_batch_set_operator", self.handle, _ptr(values), _ptr(loads), _stream())

    def operator(self):
        """Assembled ``(values (M, nnz) slot order, loads (M, n), A Z (M, n))``."""
        vals = torch.empty((self.macros, self.layout.nnz), dtype=torch.float64, device=self.device)
        loads = torch.empty((self.macros, self.n), dtype=torch.float64, device=self.device)
        az = torch.empty((self.macros, self.n), dtype=torch.float64, device=self.device)
        _lib.call("mhm_batch_operator", self.handle, _ptr(vals), _ptr(loads), _ptr(az), _stream())
        return vals, loads, az

    def _shaped(self, name, a, dtype, shape):
        """Device copy of ``a``, which must have exactly ``shape`` (ValueError otherwise)."""
        if a is None:
            return None
        t = self._dev(a, dtype)
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
        return t

    def set_coupling(self, colmap=None, colscale=None, couplings=None, constraints=None, kernels=None):
        """Per-macro coupling column maps/scales or full couplings, constraints, kernels.

        Changing any of them invalidates the factors and responses: factor and
        condense again before solving."""
        M = self.macros
        cm = self._shaped("colmap", colmap, torch.int32, (M, self.m))
        cs = self._shaped("colscale", colscale, torch.float64, (M, self.m))
        cp = self._shaped("couplings", couplings, torch.float64, (M, self.m, self.n))
        cn = self._shaped("constraints", constraints, torch.float64, (M, self.n))
        kn = self._shaped("kernels", kernels, torch.float64, (M, self.n))
        if (constraints is not None or kernels is not None) and not self.k:
            raise ValueError("constraints and kernels need a layout with a retained mode")
        if cm is not None and self.m and (int(cm.min()) < 0 or int(cm.max()) >= self.m):
            raise ValueError("colmap entries outside the coupling columns")
        _lib.call("mhm_batch_set_coupling", self.handle, _ptr(cm), _ptr(cs), _ptr(cp), _ptr(cn),
                  _ptr(kn), _stream())
        torch.cuda.current_stream().synchronize()

    def set_geometry(self, grad_lambda, measure):
        """Per-macro element geometry ``(M, elements, dim+1, dim)`` barycentric
        gradients and ``(M, elements)`` measures, for macros that share this
        layout's topology but not its shape."""
        lay = self.layout
        if getattr(lay, "ne", 0) <= 0:
            raise ValueError("the layout has no assembly description")
        g = self._shaped("grad_lambda", grad_lambda, torch.float64, (self.macros, lay.ne, lay.dim + 1, lay.dim))
        ms = self._shaped("measure", measure, torch.float64, (self.macros, lay.ne))
        _lib.call("mhm_batch_set_geometry", self.handle, _ptr(g), _ptr(ms), _stream())
        torch.cuda.current_stream().synchronize()

    def skeleton_coupling(self, side_len, side_sign):
        """Couplings computed on the device from the layout's skeleton:
        ``side_len`` and ``side_sign`` are ``(M, sides)`` (sign +1 for the first
        owner of the macroface, -1 for the second)."""
        ns = int(self.layout._skel_desc.nsides) if getattr(self.layout, "_skel_desc", None) else 0
        if not ns:
            raise ValueError("the layout has no skeleton (LocalLayout.set_skeleton)")
        ln = self._shaped("side_len", side_len, torch.float64, (self.macros, ns))
        sg = self._shaped("side_sign", side_sign, torch.float64, (self.macros, ns))
        _lib.call("mhm_batch_skeleton_coupling", self.handle, _ptr(ln), _ptr(sg), _stream())
        torch.cuda.current_stream().synchronize()

    def mass_constraints(self):
        """Constraints ``M 1 / (1^T M 1)`` computed on the device from the
        batch's element geometry (set it first for uneven macros)."""
        _lib.call("mhm_batch_mass_constraints", self.handle, _stream())

    def coupling(self):
        """The effective ``(coupling (M, m, n), constraints (M, n))`` in original rows."""
        cp = torch.empty((self.macros, self.m, self.n), dtype=torch.float64, device=self.device)
        cn = torch.empty((self.macros, self.n), dtype=torch.float64, device=self.device) if self.k else None
        _lib.call("mhm_batch_coupling", self.handle, _ptr(cp), _ptr(cn), _stream())
        return cp, cn

    # -------------------------------------------------------------- numerics
    def factor(self, check: bool = True):
        _lib.call("mhm_batch_factor", self.handle, _stream())
        if check:
            self.raise_on_failure("factor")

    def condense(self, max_refine: int | None = None, rtol: float = 1e-10, check: bool = True):
        """Responses to [f, B, A Z] and condensed blocks ``(M, w, w)`` / loads ``(M, w)``."""
        if max_refine is None:
            max_refine = DEFAULT_REFINE[self.precision]
        if self.blocks is None:
            self.blocks = torch.empty((self.macros, self.width, self.width), dtype=torch.float64,
                                      device=self.device)
            self.loads = torch.empty((self.macros, self.width), dtype=torch.float64, device=self.device)
        _lib.call("mhm_batch_condense", self.handle, int(max_refine), float(rtol), _ptr(self.blocks),
                  _ptr(self.loads), _stream())
        if check:
            self.raise_on_failure("condense")
        return self.blocks, self.loads

    def set_solve_variants(self, first: int = _lib.SOLVE_AUTO, refine: int = _lib.SOLVE_AUTO):
        """Band solve kernel family (``_lib.SOLVE_*``) of the first and of the
        refinement passes; ``SOLVE_AUTO`` picks by band width and column count.
        A family that cannot serve the batch fails the next solving call."""
        _lib.call("mhm_batch_set_solve_variants", self.handle, int(first), int(refine))
        self.solve_variants = (int(first), int(refine))

    def set_side_stream(self, enable: bool):
        """Condense the macros that converge early on the batch's side stream
        (default) or everything on the caller's stream; batches that run
        concurrently on streams of their own turn it off."""
        _lib.call("mhm_batch_set_side_stream", self.handle, int(bool(enable)))

    def responses(self):
        """(source ``(M,n)``, lifts ``(M,n,m)``, retained ``(M,n)``) in original rows."""
        src = torch.empty((self.macros, self.n), dtype=torch.float64, device=self.device)
        lifts = torch.empty((self.macros, self.n, self.m), dtype=torch.float64, device=self.device)
        ret = torch.empty((self.macros, self.n), dtype=torch.float64, device=self.device) if self.k else None
        _lib.call("mhm_batch_responses", self.handle, _ptr(src), _ptr(lifts), _ptr(ret), _stream())
        return src, lifts, ret

    def solve(self, rhs, max_refine: int | None = None, rtol: float = 1e-10, check: bool = True):
        """Solve ``(M, ncols, N)`` right-hand sides of the augmented system (original rows)."""
        if max_refine is None:
            max_refine = DEFAULT_REFINE[self.precision]
        rhs = self._dev(rhs, torch.float64)
        if rhs.dim() != 3 or rhs.shape[0] != self.macros or rhs.shape[2] != self.N:
            raise ValueError(f"rhs must have shape (macros, ncols, {self.N})")
        if not 1 <= rhs.shape[1] <= self.max_columns:
            raise ValueError(f"ncols must be in [1, {self.max_columns}] (max_columns)")
        x = torch.empty_like(rhs)
        _lib.call("mhm_batch_solve", self.handle, _ptr(rhs), _ptr(x), int(rhs.shape[1]), int(max_refine),
                  float(rtol), _stream())
        if check:
            self.raise_on_failure("solve")
        return x

    def sources(self, fq=None, f_const: float = 0.0, *, ncols: int | None = None, max_refine: int | None = None,
                rtol: float = 1e-10, source: bool = False, check: bool = True):
        """Condensed loads of new source fields against the resident factors.

        ``fq`` holds ``(ncols, M, elements, q)`` source values at the quadrature
        points (``None``: ``f_const`` everywhere, ``ncols`` fields).  Returns the
        loads ``(ncols, M, m+k)`` and, with ``source=True``, the source responses
        ``(ncols, M, n)`` in original rows (``OfflineLocalProblem.response``,
        ``core/offline.py:121-127``; ``local_condensed_load``,
        ``core/reconstruction.py:91-127``).
        """
        if max_refine is None:
            max_refine = DEFAULT_REFINE[self.precision]
        lay = self.layout
        if getattr(lay, "ne", 0) <= 0:
            raise ValueError("the layout has no assembly description")
        if fq is not None:
            fq = self._dev(fq, torch.float64)
            if fq.dim() != 4 or tuple(fq.shape[1:]) != (self.macros, lay.ne, lay.nq):
                raise ValueError(f"fq must have shape (fields, {self.macros}, {lay.ne}, {lay.nq}), "
                                 f"got {tuple(fq.shape)}")
            ncols = int(fq.shape[0])
        elif ncols is None:
            raise ValueError("ncols is required without fq")
        if not 1 <= int(ncols) <= self.max_columns:
            raise ValueError(f"ncols must be in [1, {self.max_columns}] (max_columns)")
        loads = torch.empty((ncols, self.macros, self.width), dtype=torch.float64, device=self.device)
        src = (torch.empty((ncols, self.macros, self.n), dtype=torch.float64, device=self.device)
               if source else None)
        _lib.call("mhm_batch_sources", self.handle, _ptr(fq), float(f_const), int(ncols), int(max_refine),
                  float(rtol), _ptr(loads), _ptr(src), _stream())
        if check:
            self.raise_on_failure("sources")
        return (loads, src) if source else loads

    def reconstruct(self, trace, coarse=None, source=None):
        """Fields ``(M, cases, n)`` from local traces ``(M, cases, m)`` and amplitudes ``(M, cases, k)``;
        ``source`` ``(cases, M, n)`` (from ``sources(..., source=True)``) replaces the standard
        source response (the online reconstruction of ``core/offline.py:234-307``)."""
        trace = self._dev(trace, torch.float64)
        if trace.dim() == 2:
            trace = trace[:, None, :]
        if trace.dim() != 3 or trace.shape[0] != self.macros or trace.shape[2] != self.m or trace.shape[1] < 1:
            raise ValueError(f"trace must have shape ({self.macros}, cases, {self.m}), got {tuple(trace.shape)}")
        cases = trace.shape[1]
        co = None
        if self.k:
            if coarse is None:
                raise ValueError("coarse amplitudes are required with a retained mode")
            co = self._dev(coarse, torch.float64)
            if co.dim() == 2 and self.k == 1:  # (M, cases) amplitudes of the one retained mode
                co = co[:, :, None]
            if tuple(co.shape) != (self.macros, cases, self.k):
                raise ValueError(f"coarse must have shape ({self.macros}, {cases}, {self.k}), got {tuple(co.shape)}")
        fields = torch.empty((self.macros, cases, self.n), dtype=torch.float64, device=self.device)
        if source is None:
            _lib.call("mhm_batch_reconstruct", self.handle, _ptr(trace.contiguous()), _ptr(co), _ptr(fields),
                      int(cases), _stream())
            return fields
        src = self._dev(source, torch.float64).contiguous()
        if tuple(src.shape) != (cases, self.macros, self.n):
            raise ValueError(f"source must have shape ({cases}, {self.macros}, {self.n}), got {tuple(src.shape)}")
        _lib.call("mhm_batch_reconstruct_sources", self.handle, _ptr(src), _ptr(trace.contiguous()), _ptr(co),
                  _ptr(fields), int(cases), _stream())
        return fields

    # -------------------------------------------------------------- status
    def status(self):
        out = np.zeros(self.macros, dtype=np.int32)
        anyv = ctypes.c_int32()
        _lib.call("mhm_batch_status", self.handle, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(anyv),
                  _stream())
        return out

    def iterations(self):
        out = np.zeros(self.macros, dtype=np.int32)
        _lib.call("mhm_batch_iterations", self.handle, out.ctypes.data_as(ctypes.c_void_p), _stream())
        return out

    def _raise_coefficients(self, st):
        bad = np.flatnonzero(st & _lib.STATUS_COEFFICIENT)
        if len(bad):
            raise ValueError("diffusion must be finite and positive and the source finite at every quadrature "
                             f"point: macros {bad[:8].tolist()}" + ("..." if len(bad) > 8 else ""))

    def raise_on_failure(self, what: str):
        st = self.status()
        self._raise_coefficients(st)
        if st.any():
            bad = np.flatnonzero(st)
            kinds = []
            if np.any(st & _lib.STATUS_PIVOT):
                kinds.append("singular or rank-deficient local operator")
            if np.any(st & _lib.STATUS_NONFINITE):
                kinds.append("nonfinite values")
            if np.any(st & _lib.STATUS_RESIDUAL):
                kinds.append("residual criterion failed")
            raise LinearSolveError(f"{what}: {', '.join(kinds)} in macros {bad[:8].tolist()}"
                                   + ("..." if len(bad) > 8 else ""))

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            torch.cuda.synchronize(self.device)
            _lib.load().mhm_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()
            self._keep = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
