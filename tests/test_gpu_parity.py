"""GPU parity of the HIP local stage against the FP64 oracle (marked ``gpu``).

Tolerances: integer maps and the global CSC pattern are compared bit for bit;
condensed blocks, loads, responses and reconstructed fields must agree with
the oracle to 1e-10 relative (per macro, Euclidean norm), the bar set for the
plain FP64 GPU path and required of the mixed-precision path as well.
"""

from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # collected on CPU-only hosts, skipped there
    pytest.skip("no GPU", allow_module_level=True)

from oracle import workloads as ow  # noqa: E402
from pymhm_amd import batched, hybrid  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402

TOL = 1e-10


def responses(d):
    """Stacked [s, L] per macro: normwise error of the solved columns.

    The source response alone can vanish exactly (f = 1 makes the load
    proportional to the constraint, so s = 0), so it is compared inside the
    stacked solution block rather than relative to its own norm."""
    return np.concatenate((d["source"][..., None], d["lifts"]), axis=-1)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    axes = tuple(range(1, a.ndim))
    num = np.sqrt(np.sum((a - b) ** 2, axis=axes))
    den = np.sqrt(np.sum(b ** 2, axis=axes))
    return float(np.max(num / np.where(den > 0, den, 1.0)))


@pytest.fixture(scope="module")
def config1():
    case = pw.config1(nx=8)
    print("config1: oracle", flush=True)
    return case, ow.oracle_config(case)


@pytest.fixture(scope="module")
def config2_small():
    case = pw.config2(nx=4)
    sel = [0, 3, 6, 9, 12, 15]
    print("config2 (4x4 macros): oracle", flush=True)
    return case, sel, ow.oracle_config(case, macros=sel)


@pytest.fixture(scope="module")
def config3_small():
    case = pw.config3(nx=2, refinement=2)
    print("config3 (2x2x2 cube macros of P2 tetrahedra): oracle", flush=True)
    return case, ow.oracle_config(case)


@pytest.mark.parametrize("which", ["config1", "config2", "config3"])
def test_assembly_bit_exact(request, which):
    """Stiffness values and loads equal the oracle's assembly bit for bit."""
    if which == "config2":
        case, sel, ref = request.getfixturevalue("config2_small")
    else:
        case, ref = request.getfixturevalue(which if which == "config1" else "config3_small")
        sel = list(range(case.problem.M))
    prob = case.problem
    print(f"assembly {which}", flush=True)
    with prob.batch("fp64") as b:
        b.assemble(case.kq, case.fq, case.f_const)
        vals, loads, az = (t.cpu().numpy() for t in b.operator())
    for r, (c, p, _) in enumerate(ref["problems"]):
        a = p.matrix.tocsr()
        ref_vals = np.zeros(prob.layout.nnz)
        ref_vals[prob.layout.slot_map(a.indptr, a.indices)] = a.data
        assert np.array_equal(vals[c], ref_vals), c
        assert np.array_equal(loads[c], p.load), c
        scale = np.abs(a.data).max()
        assert np.abs(az[c] - p.retained_action[:, 0]).max() <= 1e-15 * scale


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config1_stage(config1, precision):
    case, ref = config1
    print(f"config1 {precision}", flush=True)
    out = pw.run_stage(case, precision)
    for key in ("blocks", "loads", "retained"):
        assert rel(out[key], ref[key]) < TOL, key
    assert rel(responses(out), responses(ref)) < TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config2_stage(config2_small, precision):
    case, sel, ref = config2_small
    print(f"config2 {precision}", flush=True)
    out = pw.run_stage(case, precision)
    for key in ("blocks", "loads", "retained"):
        assert rel(out[key][sel], ref[key]) < TOL, key
    assert rel(responses(out)[sel], responses(ref)) < TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_stage(config3_small, precision):
    """Tetrahedral P2 cube macros: blocks, loads, responses against the oracle."""
    case, ref = config3_small
    print(f"config3 {precision}", flush=True)
    out = pw.run_stage(case, precision)
    for key in ("blocks", "loads", "retained"):
        assert rel(out[key], ref[key]) < TOL, key
    assert rel(responses(out), responses(ref)) < TOL


@pytest.fixture(scope="module")
def config3_wide():
    case = pw.config3(nx=2, refinement=4)
    print(f"config3 at 4^3 x 6 tetrahedra per cube (bandwidth {case.problem.layout.bandwidth}): oracle",
          flush=True)
    return case, ow.oracle_config(case)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_wide_band_stage(config3_wide, precision):
    """Cube macros of 4^3 x 6 P2 tetrahedra (n = 729, bandwidth 189): the band
    is too wide for the LDS window, so the in-place wide-band factor and the
    block solve run; blocks, loads and responses against the oracle."""
    case, ref = config3_wide
    assert case.problem.layout.bandwidth + 16 > 127
    print(f"config3 wide {precision}", flush=True)
    out = pw.run_stage(case, precision)
    for key in ("blocks", "loads", "retained"):
        assert rel(out[key], ref[key]) < TOL, key
    assert rel(responses(out), responses(ref)) < TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_refinement_below_the_rounding_floor(config2_small, precision):
    """rtol below the rounding floor keeps every column pending through every
    pass: the binary64 residual stage hands the columns near the floor to the
    double-double stage, every pass solves every macro (in binary64 one
    wavefront per column tile), no macro is condensed early.  The results
    still meet the oracle bar, and the missed criterion is reported."""
    from pymhm_amd import _lib

    case, sel, ref = config2_small
    print(f"config2 {precision}, rtol 1e-17", flush=True)
    with case.problem.batch(precision) as b:
        b.assemble(case.kq, None, case.f_const)
        b.factor()
        blocks, loads = b.condense(max_refine=3, rtol=1e-17, check=False)
        solves, status = b.iterations(), b.status()
        blocks, loads = blocks.cpu().numpy(), loads.cpu().numpy()
    assert (solves == 4).all()  # the first solve and three refinement passes for every macro
    assert (status & _lib.STATUS_RESIDUAL).any()
    assert rel(blocks[sel], ref["blocks"]) < TOL
    assert rel(loads[sel], ref["loads"]) < TOL


def test_results_independent_of_batch_composition(config2_small):
    """A macro's result is bitwise the same alone and inside a larger batch."""
    case, _, _ = config2_small
    full = pw.run_stage(case, "mixed")
    alone = pw.run_stage(case, "mixed", macros=slice(5, 6))
    pair = pw.run_stage(case, "mixed", macros=slice(4, 7))
    assert np.array_equal(full["blocks"][5], alone["blocks"][0])
    assert np.array_equal(full["blocks"][5], pair["blocks"][1])


def test_global_scatter_and_reconstruction(config1):
    case, ref = config1
    prob = case.problem
    plan = hybrid.HybridScatter(prob.global_indices(), prob.trace_size, prob.kernel_offsets())
    with prob.batch("fp64") as b:
        b.assemble(case.kq, case.fq, case.f_const)
        b.factor()
        blocks, loads = b.condense()
        mat, rhs, scale = plan.assemble(blocks, loads)
        rows, cols = plan.coordinates()
        idx = prob.global_indices()
        w = idx.shape[1]
        assert np.array_equal(rows.cpu().numpy(), np.repeat(idx, w, axis=1).ravel())
        assert np.array_equal(cols.cpu().numpy(), np.tile(idx, (1, w)).ravel())
        from oracle.local import assemble_hybrid_contributions
        contribs = [(idx[c], ref["blocks"][c], ref["loads"][c]) for c in range(prob.M)]
        omat, orhs, oscale = assemble_hybrid_contributions(contribs, prob.trace_size, prob.kernel_offsets())
        # Coarse-coarse entries are -(A Z)^T E: exactly zero (and eliminated) when A Z
        # rounds to zero, ~1e-17 otherwise - the value-dependent branch of
        # core/contracts.py:207-211.  Compare the pattern of entries above roundoff.
        dense, odense = mat.toarray(), omat.toarray()
        floor = 1e-13 * np.abs(odense).max()
        assert np.array_equal(np.abs(dense) > floor, np.abs(odense) > floor)
        assert np.linalg.norm(dense - odense) / np.linalg.norm(odense) < TOL
        assert np.linalg.norm(rhs - orhs) / np.linalg.norm(orhs) < TOL
        # the global solve is outside the stage: gauge-free Dirichlet system, direct solve
        from scipy.sparse import linalg as sla
        sol = sla.spsolve(mat.tocsc(), rhs)
        trace, coarse = sol[:prob.trace_size], sol[prob.trace_size:]
        fields = b.reconstruct(trace[prob.trace_dofs][:, None, :], coarse[:, None, None])
        fields = fields[:, 0].cpu().numpy()
    rs = [ref["problems"][c][1] for c in range(prob.M)]
    expected = np.stack([ref["source"][c] - ref["lifts"][c] @ trace[prob.trace_dofs[c]]
                         + ref["retained"][c] * coarse[c] for c in range(prob.M)])
    assert rel(fields, expected) < TOL
    assert len(rs) == prob.M


def test_golden_npz_through_condense_batched():
    """The reference's archived P2 case through the drop-in condense_batched."""
    from tests.test_oracle_golden import golden_2d_case  # oracle builds the LocalProblems
    from oracle.local import assemble_hybrid_contributions
    golden = np.load(Path(__file__).parent / "golden" / "original-2d-2d-n4-r8-p2-q6.npz")
    mesh, skel, fines, oresp, _, _ = golden_2d_case()
    problems = [r.problem for r in oresp]
    for precision in ("fp64", "mixed"):
        print(f"golden npz {precision}", flush=True)
        resp = batched.condense_batched(problems, precision=precision)
        contribs = [(np.r_[p.trace_dofs, skel.size + c], r.block, r.load)
                    for c, (p, r) in enumerate(zip(problems, resp))]
        mat, rhs, _ = assemble_hybrid_contributions(
            contribs, skel.size, skel.size + np.arange(len(problems) + 1))
        assert np.array_equal(mat.indptr, golden["global_indptr"])
        assert np.array_equal(mat.indices, golden["global_indices"])
        gv, gl = golden["global_values"], golden["global_load"]
        assert np.linalg.norm(mat.data - gv) / np.linalg.norm(gv) < TOL
        assert np.linalg.norm(rhs - gl) / np.linalg.norm(gl) < TOL
        from scipy.sparse import linalg as sla
        sol = sla.spsolve(mat.tocsc(), rhs)
        tr, co = sol[:skel.size], sol[skel.size:]
        pres = np.concatenate([r.reconstruct(tr[p.trace_dofs], co[c:c + 1])
                               for c, (p, r) in enumerate(zip(problems, resp))])
        ref = golden["pressure"]
        assert np.linalg.norm(pres - ref) / np.linalg.norm(ref) < TOL


def test_local_factorization_resident_solves(config1):
    """Resident band factors: more columns than max_columns (solved in chunks)
    and a nonzero constraint-row right-hand side (the bordered row's beta)."""
    case, ref = config1
    problems = [p for _, p, _ in ref["problems"]][:6]
    with batched.LocalFactorization(problems, max_columns=4) as fac:
        rng = np.random.default_rng(3)
        rhs = rng.normal(size=(6, problems[0].n + 1, 7))
        x = fac.solve(rhs, host=True)
    for c, p in enumerate(problems):
        s = p.condensation_matrix().toarray()
        expected = np.linalg.solve(s, rhs[c])
        assert np.linalg.norm(x[c] - expected) / np.linalg.norm(expected) < TOL
    with pytest.raises(RuntimeError):
        fac.solve(rhs)


def test_singular_local_operator_is_rejected(config1):
    case, _ = config1
    bad = case.kq.copy()
    bad[3] = 0.0                           # zero diffusion: singular macro 3
    with case.problem.batch("fp64") as b:
        b.assemble(bad, case.fq, case.f_const)
        b.factor(check=False)
        status = b.status()
    assert status[3] != 0
    assert np.count_nonzero(status) == 1
