"""Uneven polygonal (Voronoi) macros on the GPU against the oracle (marked ``gpu``).

Buckets of equal local topology run as one batch each with per-macro
geometry; assembly must equal the oracle's bit for bit, blocks, loads and
responses 1e-10 relative per macro.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from oracle import workloads as ow  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402

TOL = 1e-10


def rel1(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def voronoi():
    case = pw.config4(npts=24, h=1 / 80, dof_range=(30, 900))
    print(f"config4 (24 Voronoi macros, {len(case.problem.buckets)} buckets): oracle", flush=True)
    return case, ow.oracle_config(case)


def test_polygon_assembly_bit_exact(voronoi):
    case, ref = voronoi
    prob = case.problem
    for bk in prob.buckets:
        print(f"assembly bucket k={bk.k} r={bk.r} macros={len(bk.cells)}", flush=True)
        with bk.batch("fp64") as b:
            b.assemble(np.stack([case.kq[c] for c in bk.cells]), None, case.f_const)
            vals, loads, _ = (t.cpu().numpy() for t in b.operator())
        for r, c in enumerate(bk.cells):
            _, p, _ = ref["problems"][c]
            a = p.matrix.tocsr()
            ref_vals = np.zeros(bk.layout.nnz)
            ref_vals[bk.layout.slot_map(a.indptr, a.indices)] = a.data
            assert np.array_equal(vals[r], ref_vals), c
            assert np.array_equal(loads[r], p.load), c


@pytest.fixture(scope="module")
def voronoi_wide():
    """24 Voronoi macros refined to h = 1/1024 within 1000-3000 P1 DOFs: six
    buckets, n up to 2835, one of half-bandwidth 131 (past the LDS-window
    factor: the wide-band factor and the blocked solve)."""
    case = pw.config4(npts=24, h=1 / 1024, dof_range=(1000, 3000))
    shapes = [(bk.n, bk.layout.bandwidth, len(bk.cells)) for bk in case.problem.buckets]
    print(f"config4 wide (24 Voronoi macros, buckets (n, b, macros) {shapes}): oracle", flush=True)
    return case, ow.oracle_config(case)


def check_stage(case, ref, precision):
    out = pw.run_polygon_stage(case, precision)
    for c in range(case.problem.M):
        for key in ("blocks", "loads", "retained"):
            assert rel1(out[key][c], ref[key][c]) < TOL, (key, c)
        got = np.concatenate((out["source"][c][:, None], out["lifts"][c]), axis=1)
        want = np.concatenate((ref["source"][c][:, None], ref["lifts"][c]), axis=1)
        assert rel1(got, want) < TOL, c


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_polygon_stage(voronoi, precision):
    case, ref = voronoi
    print(f"config4 {precision}", flush=True)
    check_stage(case, ref, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_polygon_stage_wide_buckets(voronoi_wide, precision):
    """The bench's local sizes: n up to 2835 and a bucket past the LDS window."""
    case, ref = voronoi_wide
    prob = case.problem
    assert max(bk.n for bk in prob.buckets) == 2835
    assert max(bk.layout.bandwidth for bk in prob.buckets) > 111
    print(f"config4 wide {precision}", flush=True)
    check_stage(case, ref, precision)


def test_bucketed_stage_concurrent_and_global_scatter(voronoi):
    """Buckets on 16 streams give the sequential results bit for bit, and the
    ragged global scatter matches the reference assembly of the oracle's
    contributions (indices exact, values 1e-10)."""
    from oracle.local import assemble_hybrid_contributions
    from pymhm_amd import hybrid
    from pymhm_amd.problems import BucketedStage

    case, ref = voronoi
    prob = case.problem
    seq = pw.run_polygon_stage(case, "fp64")
    stage = BucketedStage(prob, "fp64", streams=16)
    try:
        bpos, lpos = stage.positions()
        plan = hybrid.HybridScatter(prob.global_indices(), prob.trace_size, prob.kernel_offsets(),
                                    positions=bpos, load_positions=lpos)
        kqs = [torch.as_tensor(np.stack([case.kq[c] for c in bk.cells]), device="cuda") for bk in prob.buckets]
        blocks, loads = stage.run(kqs, None, case.f_const)
        torch.cuda.synchronize()
        flat_b, flat_l = blocks.cpu().numpy(), loads.cpu().numpy()
        for c in range(prob.M):
            w = len(prob.global_indices()[c])
            assert np.array_equal(flat_b[bpos[c]:bpos[c] + w * w].reshape(w, w), seq["blocks"][c]), c
            assert np.array_equal(flat_l[lpos[c]:lpos[c] + w], seq["loads"][c]), c
        mat, rhs, _ = plan.assemble(blocks, loads)
    finally:
        stage.close()
    idx = prob.global_indices()
    omat, orhs, _ = assemble_hybrid_contributions([(idx[c], ref["blocks"][c], ref["loads"][c])
                                                    for c in range(prob.M)], prob.trace_size, prob.kernel_offsets())
    dense, odense = mat.toarray(), omat.toarray()
    floor = 1e-13 * np.abs(odense).max()
    assert np.array_equal(np.abs(dense) > floor, np.abs(odense) > floor)
    assert np.linalg.norm(dense - odense) / np.linalg.norm(odense) < TOL
    assert np.linalg.norm(rhs - orhs) / np.linalg.norm(orhs) < TOL
