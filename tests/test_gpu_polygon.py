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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_polygon_stage(voronoi, precision):
    case, ref = voronoi
    print(f"config4 {precision}", flush=True)
    out = pw.run_polygon_stage(case, precision)
    for c in range(case.problem.M):
        for key in ("blocks", "loads", "retained"):
            assert rel1(out[key][c], ref[key][c]) < TOL, (key, c)
        got = np.concatenate((out["source"][c][:, None], out["lifts"][c]), axis=1)
        want = np.concatenate((ref["source"][c][:, None], ref["lifts"][c]), axis=1)
        assert rel1(got, want) < TOL, c
