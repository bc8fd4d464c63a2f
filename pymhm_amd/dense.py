"""Dense batched LU drop-ins (host side of ``mhm_dense.hip``).

* ``BatchedFactorization(matrices, *, rtol=1e-10, atol=0.0)`` mirrors the
  reference's resident factor object ``src/pymhm/execution/cuda.py:165-288``:
  the matrices ``(batch, n, n)`` are copied to the device, equilibrated by rows
  and columns and factored once; ``solve(rhs, host=False)`` takes
  ``(batch, n, nrhs)`` right-hand sides, refines at most twice and checks every
  matrix and column against ``rtol``/``atol``; ``close()`` is idempotent and
  the object is a context manager.
* ``HipLocalSolver`` is a ``LocalLinearSolver`` (``core/assembly.py:32-47``):
  ``solver(matrix, rhs)`` solves one augmented local system on the GPU and
  returns a host array, for ``SolverConfig(local_solver=HipLocalSolver())``.
* ``factor_solve(matrices, rhs)`` is the one-shot form of both.

Errors follow the reference: ``ValueError`` for malformed or nonfinite input,
``LinearSolveError`` for a zero row or column, a singular pivot, a numerically
rank-deficient matrix, nonfinite solutions or a failed residual criterion.
No other solver is ever substituted.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import LinearSolveError

_BITS = (
    (_lib.DENSE_NONFINITE_MATRIX, ValueError, "matrices must be finite real arrays of shape (batch,n,n)"),
    (_lib.DENSE_ZERO_ROW, LinearSolveError, "batched matrix contains a zero row"),
    (_lib.DENSE_ZERO_COLUMN, LinearSolveError, "batched matrix contains a zero column"),
    (_lib.DENSE_SINGULAR, LinearSolveError, "batched LU failed: a local matrix has a singular pivot"),
    (_lib.DENSE_RANK, LinearSolveError, "batched local matrix is numerically rank deficient"),
    (_lib.DENSE_NONFINITE_RHS, ValueError, "rhs must be finite real arrays of shape (batch,n,nrhs)"),
    (_lib.DENSE_NONFINITE_RESULT, LinearSolveError, "batched solver returned nonfinite coefficients"),
    (_lib.DENSE_RESIDUAL, LinearSolveError, "batched residual criterion failed"),
)


def _tolerances(rtol: float, atol: float) -> None:
    """The reference's tolerance contract (``linalg/linear.py:76-81``)."""
    if not np.isfinite(rtol) or not np.isfinite(atol) or min(rtol, atol) < 0:
        raise ValueError("rtol and atol must be finite and nonnegative")
    if rtol == 0 and atol == 0:
        raise ValueError("at least one tolerance must be positive")


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _raise_status(status: np.ndarray, what: str) -> None:
    """Raise for the first failure kind present, in the reference's check order."""
    if not status.any():
        return
    for bit, kind, message in _BITS:
        bad = np.flatnonzero(status & bit)
        if len(bad):
            raise kind(f"{message} ({what}: matrices {bad[:8].tolist()}{'...' if len(bad) > 8 else ''})")


def _real_array(value, name: str, device):
    """A float64 device tensor of ``value``; complex or non-numeric input is a ValueError."""
    if isinstance(value, torch.Tensor):
        if value.is_complex():
            raise ValueError(f"{name} must be finite real arrays")
        if value.is_cuda and value.device != device:
            raise ValueError("resident arrays must belong to the factorization's device")
        return value.to(device=device, dtype=torch.float64).contiguous()
    arr = np.asarray(value)
    if arr.dtype.kind not in "biuf":
        raise ValueError(f"{name} must be finite real arrays")
    return torch.as_tensor(np.ascontiguousarray(arr, dtype=np.float64), device=device)


def _device(device):
    if not torch.cuda.is_available():
        raise _lib.SolverUnavailableError("the HIP dense solver requires a visible GPU")
    if device is None:
        device = torch.cuda.current_device()
    return torch.device("cuda", int(device.index if isinstance(device, torch.device) else device))


class BatchedFactorization:
    """Reusable resident LU of real ``(batch, n, n)`` matrices on one GPU.

    The object belongs to the device current at construction (or ``device``);
    resident tensors of another device are rejected.  It must not be shared
    between threads.  ``solve`` returns device tensors unless ``host=True``.
    """

    def __init__(self, matrices, *, rtol: float = 1e-10, atol: float = 0.0, device=None) -> None:
        _tolerances(rtol, atol)
        self._closed = True
        self._handle = ctypes.c_void_p()
        self.device = _device(device)
        a = _real_array(matrices, "matrices", self.device)
        if a.dim() != 3 or a.shape[0] == 0 or a.shape[1] == 0 or a.shape[1] != a.shape[2]:
            raise ValueError("matrices must be finite real arrays of shape (batch,n,n)")
        self.batch, self.n = int(a.shape[0]), int(a.shape[1])
        self.rtol, self.atol = float(rtol), float(atol)
        self._status = torch.zeros(self.batch, dtype=torch.int32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.call("mhm_dense_create", self.device.index, self.batch, self.n, _ptr(a), _ptr(self._status),
                      _stream(self.device), ctypes.byref(self._handle))
            self._closed = False
            try:
                _raise_status(self._status.cpu().numpy(), "factor")
            except BaseException:
                self.close()
                raise

    def solve(self, rhs, *, host: bool = False):
        """Solve ``(batch, n, nrhs)`` loads; device tensor unless ``host=True``."""
        if self._closed:
            raise RuntimeError("batched factorization is closed")
        b = _real_array(rhs, "rhs", self.device)
        if b.dim() != 3 or tuple(b.shape[:2]) != (self.batch, self.n) or b.shape[2] == 0:
            raise ValueError("rhs must be finite real arrays of shape (batch,n,nrhs)")
        x = torch.zeros_like(b)
        with torch.cuda.device(self.device):
            _lib.call("mhm_dense_solve", self._handle, _ptr(b), int(b.shape[2]), self.rtol, self.atol, _ptr(x),
                      _ptr(self._status), _stream(self.device))
            _raise_status(self._status.cpu().numpy(), "solve")
        return x.cpu().numpy() if host else x

    def close(self) -> None:
        if not self._closed:
            self._closed = True
            with torch.cuda.device(self.device):
                torch.cuda.current_stream(self.device).synchronize()
                _lib.load().mhm_dense_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __enter__(self):
        if self._closed:
            raise RuntimeError("batched factorization is closed")
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def factor_solve(matrices, rhs, *, rtol: float = 1e-10, atol: float = 0.0, device=None, host: bool = False):
    """Factor ``(batch, n, n)`` and solve ``(batch, n, nrhs)`` in one call (synchronous)."""
    _tolerances(rtol, atol)
    dev = _device(device)
    a = _real_array(matrices, "matrices", dev)
    if a.dim() != 3 or a.shape[0] == 0 or a.shape[1] == 0 or a.shape[1] != a.shape[2]:
        raise ValueError("matrices must be finite real arrays of shape (batch,n,n)")
    b = _real_array(rhs, "rhs", dev)
    if b.dim() != 3 or tuple(b.shape[:2]) != tuple(a.shape[:2]) or b.shape[2] == 0:
        raise ValueError("rhs must be finite real arrays of shape (batch,n,nrhs)")
    x = torch.zeros_like(b)
    status = torch.zeros(a.shape[0], dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        _lib.call("mhm_dense_factor_solve", dev.index, int(a.shape[0]), int(a.shape[1]), int(b.shape[2]), _ptr(a),
                  _ptr(b), _ptr(x), float(rtol), float(atol), _ptr(status), _stream(dev))
        _raise_status(status.cpu().numpy(), "factor_solve")
    return x.cpu().numpy() if host else x


class HipLocalSolver:
    """``LocalLinearSolver`` on the GPU (``core/assembly.py:32-47``).

    ``solver(matrix, rhs)`` receives the augmented local operator (SciPy
    sparse or dense, ``(N, N)``) and the right-hand sides ``(N, R)`` and returns
    the ``(N, R)`` solution as a host array, checked per column at ``rtol``.
    The caller's ``check_linear_solution`` still applies its own acceptance.
    The object holds no device state between calls, so it is picklable for
    the reference's spawn backend; ``close()`` is a no-op kept for the
    provider protocol (``core/assembly.py:233-242``).
    """

    def __init__(self, *, rtol: float = 1e-10, atol: float = 0.0, device=None) -> None:
        _tolerances(rtol, atol)
        self.rtol, self.atol, self.device = float(rtol), float(atol), device

    def __call__(self, matrix, rhs):
        dense = matrix.toarray() if hasattr(matrix, "toarray") else np.asarray(matrix)
        b = np.asarray(rhs, dtype=np.float64)
        squeeze = b.ndim == 1
        if squeeze:
            b = b[:, None]
        x = factor_solve(dense[None], b[None], rtol=self.rtol, atol=self.atol, device=self.device, host=True)[0]
        return x[:, 0] if squeeze else x

    def close(self) -> None:
        return None
