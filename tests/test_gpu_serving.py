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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sources_config2_mesh(precision):
    """New sources on the config-2 local shape (P2, 16x16x2 triangles, P1
    multipliers) against the oracle, and a column's bits alone and in a batch."""
    case = pw.config2(nx=2)
    prob = case.problem
    print(f"sources config2 2x2 {precision}", flush=True)
    pts = torch.as_tensor(prob.quadrature_points(), dtype=torch.float64, device="cuda")
    fq = serving.bump_sources(pts, [5, 6])
    with prob.batch(precision, max_columns=2) as b:
        b.assemble(case.kq, None, case.f_const)
        b.factor()
        b.condense()
        loads = b.sources(fq).cpu().numpy()
        alone = b.sources(fq[:1]).cpu().numpy()
    assert np.array_equal(alone[0], loads[0])
    for c in range(2):
        ref = ow.oracle_config(dataclasses.replace(case, fq=fq[c].cpu().numpy(), f_const=0.0))
        assert rel(loads[c], ref["loads"]) < TOL, c


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sources_uneven_macros(precision):
    """New sources on Voronoi buckets (per-macro geometry and element measures
    in the source load) against the oracle."""
    case = pw.config4(npts=12, h=1 / 64, dof_range=(30, 600))
    prob = case.problem
    print(f"sources Voronoi ({len(prob.buckets)} buckets) {precision}", flush=True)
    per_macro = [None] * prob.M
    for bk in prob.buckets:
        pts = torch.as_tensor(bk.qpts, dtype=torch.float64, device="cuda")
        fq = serving.bump_sources(pts, [3])
        with bk.batch(precision, max_columns=1) as b:
            b.assemble(np.stack([case.kq[c] for c in bk.cells]), None, case.f_const)
            b.factor()
            b.condense()
            loads = b.sources(fq).cpu().numpy()[0]
        fqh = fq[0].cpu().numpy()
        for r, c in enumerate(bk.cells):
            per_macro[c] = (loads[r], fqh[r])
    ref = ow.oracle_config(dataclasses.replace(case, fq=[f for _, f in per_macro], f_const=0.0))
    for c in range(prob.M):
        assert rel(per_macro[c][0][None], np.asarray(ref["loads"][c])[None]) < TOL, c


def replay_alone(srv, cases, done, precision):
    """Each problem's requests one at a time on fresh batches, bit for bit."""
    for p, case in enumerate(cases):
        mine = [r for r in done if r.problem == p]
        kq = torch.as_tensor(case.kq, dtype=torch.float64, device="cuda")
        with case.problem.batch(precision, max_columns=srv.max_columns) as b:
            b.assemble(kq, case.fq, case.f_const)
            b.factor()
            b.condense()
            for r in mine:
                assert r.error is None, (r.rid, r.error)
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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_results_equal_alone(precision):
    """Three problems (the config-1 and config-2 square meshes, the config-3
    cubes) served concurrently on their own streams: every result equals the
    same request computed alone."""
    cases = [pw.config1(nx=4), pw.config2(nx=6, refinement=16), pw.config3(nx=2, refinement=2)]
    print(f"serving {precision}: 30 requests over three problems", flush=True)
    reqs = serving.poisson_requests(30, rate=1000.0, nproblems=3, coefficient_fraction=0.3, seed=5)
    srv = serving.Server(cases, precision, max_columns=4, source_pool=5)
    try:
        done = srv.run(reqs, realtime=False)
        assert all(np.isfinite(r.done) for r in done)
        assert srv.peak_bytes <= srv.hbm_budget
        replay_alone(srv, cases, done, precision)
        stats = serving.latency_stats(done)
        assert stats["requests"] == 30 and stats["p99_ms"] >= stats["p50_ms"]
    finally:
        srv.close()


def test_admission_by_footprint():
    """A budget of one and a half source columns: every source batch is cut to
    one column, the in-flight bytes never pass the budget, and the results are
    unchanged; a budget below one batch is refused."""
    cases = [pw.config1(nx=4), pw.config2(nx=4, refinement=8)]
    reqs = [serving.Request(i, 0.0, i % 2, serving.SOURCE, 100 + i) for i in range(8)]
    probe = serving.Server(cases, "fp64", max_columns=4, source_pool=0)
    try:
        per = max(probe.request_bytes(p, serving.SOURCE, 1) for p in range(2))
    finally:
        probe.close()
    srv = serving.Server(cases, "fp64", max_columns=4, source_pool=0, hbm_budget=int(1.5 * per))
    try:
        done = srv.run(reqs, realtime=False)
        assert srv.peak_bytes <= srv.hbm_budget
        replay_alone(srv, cases, done, "fp64")
    finally:
        srv.close()
    tiny = serving.Server(cases, "fp64", max_columns=4, source_pool=0, hbm_budget=16)
    try:
        with pytest.raises(MemoryError):
            tiny.run(reqs[:1], realtime=False)
    finally:
        tiny.close()
