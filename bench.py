#!/usr/bin/env python3
"""Benchmark of the MHM local-problem stage on one MI355X.

Workload: 64x64 square macros, each 16x16x2 P2 triangles (n = 1089 local
DOFs), discontinuous P1 multipliers on 4 segments per macroface (m = 32, so
R = 34 response columns), log K a seeded Gaussian random field (exponential
covariance, correlation length 0.05, variance 4) at the 16 quadrature points
of every triangle, f = 1.  Inputs (K at quadrature points) are resident in
HBM before timing.

A step is one pass of the local stage over all 4096 macros: assembly, band
factorisation, solves of [f, B, A Z] with binary64 refinement, condensation
into the 33x33 blocks and the ordered gather into the global hybrid matrix
and right-hand side.  ``value`` is macros/s of the mixed-precision path
(binary32 factor/solves + binary64 refinement); ``fp64_value`` is the plain
binary64 path on the same inputs.  Both are checked against the FP64 oracle
on the CPU-baseline sample and the maximum relative errors are reported.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "macro solves/s at <=1e-10 rel err vs FP64 oracle (vs plain FP64); p50/p99 latency"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


_LIMITS = None


def _cpu_worker_init():
    """One BLAS/OpenMP thread per worker, as the reference's process policy
    (``execution/cpu.py:25-82``); imports happen outside the timed region."""
    global _LIMITS
    from threadpoolctl import threadpool_limits

    _LIMITS = threadpool_limits(limits=1)
    import oracle.workloads  # noqa: F401


def cpu_baseline(case, sample_macros, cores):
    """Oracle (FP64 NumPy/SciPy restatement) on a bounded sample, process pool."""
    from oracle import workloads as ow

    chunks = [c for c in np.array_split(np.asarray(sample_macros), cores) if len(c)]
    specs = [ow.grid_spec(case, c.tolist()) for c in chunks]
    with ProcessPoolExecutor(max_workers=cores, initializer=_cpu_worker_init) as pool:
        # warm-up: each worker builds its macro mesh once (mesh-only, untimed)
        list(pool.map(ow.condense_spec, [ow.grid_spec(case, [int(c[0])]) for c in chunks]))
        t0 = time.perf_counter()
        results = list(pool.map(ow.condense_spec, specs))
        wall = time.perf_counter() - t0
    blocks = np.zeros((len(sample_macros),) + results[0][1].shape[1:])
    loads = np.zeros((len(sample_macros),) + results[0][2].shape[1:])
    pos = {m: i for i, m in enumerate(sample_macros)}
    for macros, blk, ld in results:
        for r, m in enumerate(macros):
            blocks[pos[m]], loads[pos[m]] = blk[r], ld[r]
    return wall, blocks, loads


def rel_err(a, b):
    num = np.sqrt(np.sum((a - b) ** 2, axis=tuple(range(1, a.ndim))))
    den = np.sqrt(np.sum(b ** 2, axis=tuple(range(1, b.ndim))))
    return float(np.max(num / den))


def per_macro_traffic(prob, precision):
    """Algorithmic bytes/flops per macro of the factor and of one solve of R columns."""
    lay = prob.layout
    n, b = lay.n, lay.bandwidth
    W = b + 1
    ts = 8 if precision == "fp64" else 4
    R = 1 + lay.m + lay.k
    return dict(
        factor_bytes=2 * n * W * ts, factor_flops=n * b * (b + 1),
        solve_bytes=2 * n * W * ts + 4 * n * R * ts, solve_flops=4 * n * b * R,
    )


def run_gpu(case, precision, steps, warmup, plan):
    import torch
    from pymhm_amd import _lib

    prob = case.problem
    kq = torch.as_tensor(case.kq, device="cuda")
    batch = prob.batch(precision)

    def step():
        batch.assemble(kq, None, case.f_const)
        batch.factor(check=False)
        blocks, loads = batch.condense(check=False)
        plan.gather(blocks, loads)
        return blocks, loads

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    _lib.call("mhm_timing_reset")
    _lib.call("mhm_timing_enable", 1)
    lat = []
    solves = 0
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        solves += int(batch.iterations().sum())
    elapsed = time.perf_counter() - t_start
    _lib.call("mhm_timing_enable", 0)
    times = _lib.kernel_times()
    status = batch.status()
    if status.any():
        raise RuntimeError(f"{precision}: {np.count_nonzero(status)} macros failed (status bits)")
    blocks, loads = (t.cpu().numpy() for t in batch.condense(check=True))
    iters = batch.iterations()
    batch.close()
    return dict(elapsed=elapsed, lat=np.array(lat), times=times, blocks=blocks, loads=loads,
                iters=iters, macro_solves=solves)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=192)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        return 2
    from pymhm_amd import hybrid
    from pymhm_amd import workloads as pw

    t0 = time.perf_counter()
    case = pw.config2(nx=args.nx)
    prob = case.problem
    M = prob.M
    print(f"[bench] workload built in {time.perf_counter() - t0:.1f}s: {M} macros, n={prob.n}, "
          f"m={prob.m}, bandwidth={prob.layout.bandwidth}", file=sys.stderr, flush=True)

    cpu = None
    sample = np.unique(np.linspace(0, M - 1, min(args.cpu_sample, M)).astype(np.int64)).tolist()
    if not args.no_cpu:
        cores = max(1, min(16, len(os.sched_getaffinity(0))))
        wall, oblocks, oloads = cpu_baseline(case, sample, cores)
        cpu = dict(value=len(sample) / wall, unit="macro solves/s", cores=cores, kind="port",
                   sample=f"{len(sample)} of {M} macros of the same workload, assembly+SuperLU "
                          f"condensation, {cores}-process pool, wall {wall:.2f}s")
        print(f"[bench] cpu baseline {cpu['value']:.1f} macros/s on {cores} cores", file=sys.stderr, flush=True)

    plan = hybrid.HybridScatter(prob.global_indices(), prob.trace_size, prob.kernel_offsets())
    res = {}
    for precision in ("mixed", "fp64"):
        res[precision] = run_gpu(case, precision, args.steps, args.warmup, plan)
        r = res[precision]
        print(f"[bench] {precision}: {M * args.steps / r['elapsed']:.0f} macros/s, "
              f"p50 {np.percentile(r['lat'], 50) * 1e3:.2f} ms", file=sys.stderr, flush=True)

    mixed = res["mixed"]
    value = M * args.steps / mixed["elapsed"]
    traffic = per_macro_traffic(prob, "mixed")
    times = mixed["times"]
    dominant = max(("factor", "solve"), key=lambda k: times[k][0])
    tot_ms, launches = times[dominant]
    avg_s = tot_ms / launches / 1e3
    if dominant == "factor":
        units = M  # every launch factors every macro
        per_launch_bytes = units * traffic["factor_bytes"]
    else:
        units = mixed["macro_solves"] / launches  # macros actually solved per launch
        per_launch_bytes = units * traffic["solve_bytes"]
    achieved = per_launch_bytes / avg_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "macro solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "64x64 square macros x (16x16x2 P2 triangles, n=1089), m=32 P1 multipliers, "
                        "log K Gaussian field (ell 0.05, var 4), f=1",
            "macros": M, "precision": "mixed (f32 band Cholesky/solves, f64 refinement)",
        },
        "p50_ms": float(np.percentile(mixed["lat"], 50) * 1e3),
        "p99_ms": float(np.percentile(mixed["lat"], 99) * 1e3),
        "fp64_value": M * args.steps / res["fp64"]["elapsed"],
        "fp64_p50_ms": float(np.percentile(res["fp64"]["lat"], 50) * 1e3),
        "rhs_solves_per_s": value * (1 + prob.m + 1),
        "refinement_solves_per_macro": {p: float(res[p]["iters"].mean()) for p in res},
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in times.items()},
        "fp64_kernel_ms_per_step": {k: v[0] / args.steps for k, v in res["fp64"]["times"].items()},
        "roofline": {
            "kernel": "k_factor<float>" if dominant == "factor" else "k_solve<float>",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "units_per_launch": units, "avg_launch_ms": avg_s * 1e3,
        },
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        idx = np.asarray(sample)
        out["max_rel_err_vs_oracle"] = {
            p: {"blocks": rel_err(res[p]["blocks"][idx], oblocks), "loads": rel_err(res[p]["loads"][idx], oloads)}
            for p in res}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
