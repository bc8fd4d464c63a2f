"""Autotuner of the band solve kernels of a batch.

The library has several band-solve kernel families (``_lib.SOLVE_*``): the
binary64 matrix-core solve with one wavefront per macro or one per 16-column
tile, the lane-per-column wavefront solve and the blocked solve over column
groups.  Which one is fastest depends on the band width and column count and
on how many macros the batch holds: a bucket of a few dozen uneven macros
leaves most of the chip idle, and splitting each macro over more wavefronts
shortens its latency, while a batch of thousands of macros is better served
by reading each band once.  The first pass of a solve (every column, every
macro) and its refinement passes (few macros pending) are tuned separately.

``tune_batch`` times the families that can serve the batch on the batch
itself (``run`` is one full use of it, for instance factor and condense),
first for the first pass with the automatic refinement choice, then for the
refinement passes with the chosen first pass, and keeps the fastest.
Choices are cached by batch shape in memory and, when the environment names
a file in ``PYMHM_AMD_TUNE_CACHE``, in that JSON file across processes.
"""

from __future__ import annotations

import json
import os
import time

import numpy as np

from . import _lib
from ._lib import LinearSolveError

FAMILIES = {
    _lib.SOLVE_AUTO: "auto",
    _lib.SOLVE_MFMA: "mfma",
    _lib.SOLVE_MFMA_TILES: "mfma_tiles",
    _lib.SOLVE_WAVE: "wave",
    _lib.SOLVE_BLOCKED: "blocked",
}
# both precisions solve on the matrix cores (binary64 arithmetic on the
# factor's storage type) or blocked; binary32 factors also lane per column
CANDIDATES = {
    "fp64": (_lib.SOLVE_AUTO, _lib.SOLVE_MFMA, _lib.SOLVE_MFMA_TILES, _lib.SOLVE_BLOCKED),
    "mixed": (_lib.SOLVE_AUTO, _lib.SOLVE_MFMA, _lib.SOLVE_MFMA_TILES, _lib.SOLVE_WAVE, _lib.SOLVE_BLOCKED),
}
_CACHE: dict[str, list[int]] = {}


def shape_key(batch) -> str:
    """What the choice depends on: device, precision, local pattern, macros."""
    import torch

    lay = batch.layout
    return (f"{torch.cuda.get_device_name(batch.device)}|{batch.precision}|n={lay.n}|b={lay.bandwidth}|"
            f"nnz={lay.nnz}|m={lay.m}|k={lay.k}|macros={batch.macros}|cols={batch.max_columns}")


def _file():
    return os.environ.get("PYMHM_AMD_TUNE_CACHE") or None


def _load_file():
    path = _file()
    if path and os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
        _CACHE.update({k: [int(v[0]), int(v[1])] for k, v in data.items()})


def _store_file():
    path = _file()
    if path:
        tmp = f"{path}.tmp{os.getpid()}"
        with open(tmp, "w") as fh:
            json.dump(_CACHE, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _time(batch, run, first, refine, repeats):
    """Median wall time of ``run`` under the given families, or None when a
    family cannot serve the batch or the results fail their criterion."""
    import torch

    batch.set_solve_variants(first, refine)
    try:
        run()  # warm-up: kernel loading and LDS limits
        torch.cuda.synchronize()
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        batch.raise_on_failure("tuning")
    except (ValueError, LinearSolveError):
        return None
    return float(np.median(times))


def tune_batch(batch, run, repeats: int = 3, use_cache: bool = True):
    """Choose and set the solve families of ``batch``; returns
    ``(first, refine, timings)`` with timings ``{(first, refine): seconds}``
    (empty when the choice came from the cache)."""
    key = shape_key(batch)
    if use_cache:
        if not _CACHE:
            _load_file()
        if key in _CACHE:
            first, refine = _CACHE[key]
            batch.set_solve_variants(first, refine)
            return first, refine, {}
    timings = {}
    fams = CANDIDATES[batch.precision]
    for f in fams:
        timings[(f, _lib.SOLVE_AUTO)] = _time(batch, run, f, _lib.SOLVE_AUTO, repeats)
    ok = {k: v for k, v in timings.items() if v is not None}
    if not ok:
        raise RuntimeError("no band solve family can serve this batch")
    first = min(ok, key=ok.get)[0]
    for r in fams[1:]:
        timings[(first, r)] = _time(batch, run, first, r, repeats)
    ok = {k: v for k, v in timings.items() if v is not None}
    first, refine = min(ok, key=ok.get)
    batch.set_solve_variants(first, refine)
    _CACHE[key] = [first, refine]
    if use_cache:
        _store_file()
    return first, refine, timings


def describe(first: int, refine: int) -> str:
    return f"{FAMILIES[first]}/{FAMILIES[refine]}"
