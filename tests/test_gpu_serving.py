"""Re-solves of new sources against resident factors and the serving loop (GPU).

New-source loads are compared with the FP64 oracle at 1e-10 relative; a
request's result must be bit-for-bit the same alone and inside a batch, and
the served results bit-for-bit the same as each request computed alone from
the coefficient state it was scheduled against.
"""

import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from oracle import workloads as ow  # noqa: E402
from pymhm_amd import serving  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402

TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    axes = tuple(range(1, a.ndim))
    num = np.sqrt(np.sum((a - b) ** 2, axis=axes))
    den = np.sqrt(np.sum(b ** 2, axis=axes))
    return float(np.max(num / np.where(den > 0, den, 1.0)))


@pytest.mark.parametrize("which", ["config1", "config3"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sources_against_oracle_and_batch_independent(which, precision):
    case = pw.config1(nx=4) if which == "config1" else pw.config3(nx=2, refinement=2)
    prob = case.problem
    print(f"sources {which} {precision}", flush=True)
    pts = torch.as_tensor(prob.quadrature_points(), dtype=torch.float64, device="cuda")
    fq = serving.bump_sources(pts, [11, 22, 33])
    with prob.batch(precision, max_columns=4) as b:
        b.assemble(case.kq, case.fq, case.f_const)
        b.factor()
        b.condense()
        loads, src = b.sources(fq, source=True)
        alone = b.sources(fq[1:2])
        loads, src, alone = loads.cpu().numpy(), src.cpu().numpy(), alone.cpu().numpy()
    assert np.array_equal(alone[0], loads[1])
    for c in range(3):
        ref = ow.oracle_config(dataclasses.replace(case, fq=fq[c].cpu().numpy()))
        assert rel(loads[c], ref["loads"]) < TOL, c
        stacked = np.concatenate((src[c][..., None], ref["lifts"]), axis=-1)
        ref_stacked = np.concatenate((ref["source"][..., None], ref["lifts"]), axis=-1)
        assert rel(stacked, ref_stacked) < TOL, c


def test_served_results_equal_alone():
    cases = [pw.config1(nx=4), pw.config3(nx=2, refinement=2)]
    print("serving: 14 requests over two problems", flush=True)
    reqs = serving.poisson_requests(14, rate=1000.0, nproblems=2, coefficient_fraction=0.3, seed=5)
    srv = serving.Server(cases, "fp64", max_columns=4, source_pool=5)
    try:
        done = srv.run(reqs, realtime=False)
        assert all(np.isfinite(r.done) for r in done)
        # replay each problem's requests one at a time on fresh batches
        for p, case in enumerate(cases):
            mine = [r for r in done if r.problem == p]
            kq = torch.as_tensor(case.kq, dtype=torch.float64, device="cuda")
            with case.problem.batch("fp64", max_columns=4) as b:
                b.assemble(kq, case.fq, case.f_const)
                b.factor()
                b.condense()
                for r in mine:
                    if r.kind == serving.COEFFICIENT:
                        kq = serving.coefficient_field(torch.as_tensor(case.kq, dtype=torch.float64, device="cuda"),
                                                       r.seed)
                        b.assemble(kq, case.fq, case.f_const)
                        b.factor()
                        blocks, loads = b.condense()
                        assert torch.equal(blocks, r.result[0]) and torch.equal(loads, r.result[1]), r.rid
                    else:
                        fq = srv.source_fields(p, [r.seed])
                        assert torch.equal(b.sources(fq)[0], r.result), r.rid
        stats = serving.latency_stats(done)
        assert stats["requests"] == 14 and stats["p99_ms"] >= stats["p50_ms"]
    finally:
        srv.close()
