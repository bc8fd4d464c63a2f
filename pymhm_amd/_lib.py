"""ctypes binding of ``libmhm_hip.so`` (declarations in ``include/mhm_hip.h``).

The library is built in-tree by ``__graft_entry__.build()`` (``hipcc
--offload-arch=gfx950``).  Loading fails loudly when it is missing: there is
no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIBRARY = _HERE / "libmhm_hip.so"

MHM_OK, MHM_EINVAL, MHM_ESINGULAR, MHM_ERESIDUAL, MHM_ENOMEM, MHM_EDEVICE = range(6)
MHM_FP64, MHM_MIXED = 0, 1
SOLVE_AUTO, SOLVE_MFMA, SOLVE_MFMA_TILES, SOLVE_WAVE, SOLVE_BLOCKED = range(5)
STATUS_PIVOT, STATUS_RESIDUAL, STATUS_NONFINITE, STATUS_COEFFICIENT = 1, 2, 4, 8
(DENSE_NONFINITE_MATRIX, DENSE_ZERO_ROW, DENSE_ZERO_COLUMN, DENSE_SINGULAR, DENSE_RANK, DENSE_NONFINITE_RHS,
 DENSE_NONFINITE_RESULT, DENSE_RESIDUAL) = (1, 2, 4, 8, 16, 32, 64, 128)

EXPORTS = (
    "mhm_last_error", "mhm_device_count", "mhm_version",
    "mhm_layout_create", "mhm_layout_destroy", "mhm_layout_info",
    "mhm_batch_create", "mhm_batch_destroy", "mhm_batch_assemble", "mhm_batch_set_operator",
    "mhm_batch_set_side_stream", "mhm_batch_set_solve_variants",
    "mhm_batch_operator", "mhm_batch_set_coupling", "mhm_batch_set_geometry", "mhm_batch_factor", "mhm_batch_condense", "mhm_batch_responses",
    "mhm_batch_solve", "mhm_batch_sources", "mhm_batch_reconstruct", "mhm_batch_reconstruct_sources",
    "mhm_batch_status",
    "mhm_batch_iterations", "mhm_scatter_coordinates", "mhm_ordered_gather",
    "mhm_timing_enable", "mhm_timing_reset", "mhm_timing_read",
    "mhm_layout_set_skeleton", "mhm_batch_skeleton_coupling", "mhm_batch_mass_constraints",
    "mhm_batch_coupling", "mhm_dirichlet_load",
    "mhm_global_create", "mhm_global_set_values", "mhm_global_solve", "mhm_global_destroy",
    "mhm_dense_create", "mhm_dense_solve", "mhm_dense_destroy", "mhm_dense_factor_solve",
    "mhm_dense_factor_solve_noalloc", "mhm_dense_scratch_bytes",
)
TIMERS = {"assemble": 0, "factor": 1, "solve": 2, "residual": 3, "condense": 4}


class LinearSolveError(RuntimeError):
    """A factorisation, refinement or residual check failed (``linalg/linear.py:32-33``)."""


class SolverUnavailableError(ImportError):
    """The HIP library or a GPU is unavailable (``linalg/linear.py:36-37``)."""


class SkeletonDesc(ctypes.Structure):
    _fields_ = [
        ("nsides", ctypes.c_int32), ("nseg", ctypes.c_int32), ("degree", ctypes.c_int32),
        ("edeg", ctypes.c_int32), ("ng", ctypes.c_int32),
        ("breaks", ctypes.c_void_p), ("gx", ctypes.c_void_p), ("gw", ctypes.c_void_p),
        ("npieces", ctypes.c_int32), ("piece_f", ctypes.c_void_p),
        ("nentries", ctypes.c_int32), ("entry_row", ctypes.c_void_p), ("entry_side", ctypes.c_void_p),
        ("entry_seg", ctypes.c_void_p), ("entry_ptr", ctypes.c_void_p), ("contrib", ctypes.c_void_p),
    ]


class LayoutDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("k", ctypes.c_int32), ("m", ctypes.c_int32),
        ("nnz", ctypes.c_int32), ("bandwidth", ctypes.c_int32), ("pin", ctypes.c_int32),
        ("perm", ctypes.c_void_p), ("rowptr", ctypes.c_void_p), ("cols", ctypes.c_void_p),
        ("coupling", ctypes.c_void_p), ("coupling_rows", ctypes.c_void_p),
        ("ncoupling_rows", ctypes.c_int32),
        ("kernel", ctypes.c_void_p), ("constraint", ctypes.c_void_p),
        ("ne", ctypes.c_int32), ("nb", ctypes.c_int32), ("dim", ctypes.c_int32),
        ("nq", ctypes.c_int32),
        ("slot_ptr", ctypes.c_void_p), ("slot_code", ctypes.c_void_p),
        ("load_ptr", ctypes.c_void_p), ("load_code", ctypes.c_void_p),
        ("grad_lambda", ctypes.c_void_p), ("measure", ctypes.c_void_p),
        ("qweights", ctypes.c_void_p), ("dtable", ctypes.c_void_p), ("vtable", ctypes.c_void_p),
    ]


_lib = None


def load(path: str | os.PathLike | None = None):
    """Load (once) and return the ctypes library, raising if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    target = Path(path) if path is not None else LIBRARY
    if not target.exists():
        raise SolverUnavailableError(
            f"{target} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(str(target))
    P, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "mhm_last_error": (ctypes.c_char_p, ctypes.c_size_t),
        "mhm_device_count": (P,),
        "mhm_version": (P, P),
        "mhm_layout_create": (I32, P, P),
        "mhm_layout_destroy": (P,),
        "mhm_layout_info": (P, P, P, P),
        "mhm_batch_create": (P, I32, I32, I32, P),
        "mhm_batch_destroy": (P,),
        "mhm_batch_assemble": (P, P, P, D, P),
        "mhm_batch_set_operator": (P, P, P, P),
        "mhm_batch_operator": (P, P, P, P, P),
        "mhm_batch_set_coupling": (P, P, P, P, P, P, P),
        "mhm_batch_set_geometry": (P, P, P, P),
        "mhm_batch_factor": (P, P),
        "mhm_batch_condense": (P, I32, D, P, P, P),
        "mhm_batch_set_side_stream": (P, I32),
        "mhm_batch_set_solve_variants": (P, I32, I32),
        "mhm_batch_responses": (P, P, P, P, P),
        "mhm_batch_solve": (P, P, P, I32, I32, D, P),
        "mhm_batch_sources": (P, P, D, I32, I32, D, P, P, P),
        "mhm_batch_reconstruct": (P, P, P, P, I32, P),
        "mhm_batch_reconstruct_sources": (P, P, P, P, P, I32, P),
        "mhm_batch_status": (P, P, P, P),
        "mhm_batch_iterations": (P, P, P),
        "mhm_scatter_coordinates": (P, I32, I32, P, P, P),
        "mhm_ordered_gather": (P, P, P, I64, P, P, P),
        "mhm_timing_enable": (I32,),
        "mhm_timing_reset": (),
        "mhm_timing_read": (I32, P, P),
        "mhm_layout_set_skeleton": (P, P),
        "mhm_batch_skeleton_coupling": (P, P, P, P),
        "mhm_batch_mass_constraints": (P, P),
        "mhm_batch_coupling": (P, P, P, P),
        "mhm_dirichlet_load": (I32, I32, I32, P, P, P, P, P, P, P, P, P),
        "mhm_global_create": (I32, I32, I64, P, P, I32, P, P, P),
        "mhm_global_set_values": (P, P, P, P),
        "mhm_global_solve": (P, P, P, D, I32, I32, P, P, P),
        "mhm_global_destroy": (P,),
        "mhm_dense_create": (I32, I32, I32, P, P, P, P),
        "mhm_dense_solve": (P, P, I32, D, D, P, P, P),
        "mhm_dense_destroy": (P,),
        "mhm_dense_factor_solve": (I32, I32, I32, I32, P, P, P, D, D, P, P),
        "mhm_dense_factor_solve_noalloc": (I32, I32, I32, I32, P, P, P, D, D, P, P),
        "mhm_dense_scratch_bytes": (I32, I32, I32, P, P),
    }
    for name in EXPORTS:
        fn = getattr(lib, name)
        fn.argtypes = list(sig[name])
        fn.restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load().mhm_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    """Map a status code to the reference's exception types."""
    if rc == MHM_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == MHM_EINVAL:
        raise ValueError(msg)
    if rc in (MHM_ESINGULAR, MHM_ERESIDUAL):
        raise LinearSolveError(msg)
    if rc == MHM_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def kernel_times():
    """``{name: (total_ms, launches)}`` of the event-timed kernels since the last reset."""
    out = {}
    for name, kid in TIMERS.items():
        ms, n = ctypes.c_double(), ctypes.c_int64()
        call("mhm_timing_read", kid, ctypes.byref(ms), ctypes.byref(n))
        out[name] = (ms.value, n.value)
    return out
