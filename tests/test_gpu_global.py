"""The global hybrid solve on the GPU (marked ``gpu``).

* the reference's archived P2 case (golden NPZ): local condensation and the
  global solve both on the device reproduce the archived trace and pressure
  to 1e-10, with no host sparse solve;
* square (P2, four-segment P1 multipliers) and Voronoi (uneven widths)
  workloads: the device solution of the assembled system matches a host
  sparse direct solve (the checker) to 1e-10 and the device reconstruction
  matches the host one;
* the affine patch test fully on the device: stage, Dirichlet load, scatter,
  global solve and reconstruction;
* reuse (a second right-hand side against the resident preconditioner equals
  a fresh solve bit for bit) and the error paths (singular blocks, too few
  iterations) end as ``LinearSolveError``, an unprepared solve as ``ValueError``.
"""

from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from scipy.sparse import linalg as sla  # noqa: E402

from pymhm_amd import batched, hybrid, meshing  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402
from pymhm_amd._lib import LinearSolveError  # noqa: E402
from pymhm_amd.problems import PolygonProblem  # noqa: E402

TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def device_flat(arrays):
    return torch.as_tensor(np.concatenate([np.ravel(a) for a in arrays]), device="cuda")


def test_golden_npz_fully_on_device():
    from tests.test_oracle_golden import golden_2d_case  # oracle builds the LocalProblems

    golden = np.load(Path(__file__).parent / "golden" / "original-2d-2d-n4-r8-p2-q6.npz")
    _, skel, _, oresp, _, _ = golden_2d_case()
    problems = [r.problem for r in oresp]
    M = len(problems)
    for precision in ("fp64", "mixed"):
        print(f"golden npz global solve, {precision} local stage", flush=True)
        resp = batched.condense_batched(problems, precision=precision)
        idx = [np.r_[p.trace_dofs, skel.size + c].astype(np.int64) for c, p in enumerate(problems)]
        plan = hybrid.HybridScatter(idx, skel.size, skel.size + np.arange(M + 1))
        trace, coarse, solver = hybrid.solve_hybrid(plan, device_flat([r.block for r in resp]),
                                                    device_flat([r.load for r in resp]), rtol=1e-13)
        with solver:
            tr, co = trace.cpu().numpy(), coarse.cpu().numpy()
            assert rel(tr, golden["trace"]) < TOL
            pres = np.concatenate([r.reconstruct(tr[p.trace_dofs], co[c:c + 1])
                                   for c, (p, r) in enumerate(zip(problems, resp))])
            assert rel(pres, golden["pressure"]) < TOL
            assert solver.residuals[-1] <= 1e-13


def stage_square(nx, r):
    case = pw.config2(nx=nx, refinement=r)
    prob = case.problem
    b = prob.batch("fp64")
    b.assemble(case.kq, None, case.f_const)
    b.factor()
    blocks, loads = b.condense()
    return prob, b, blocks, loads


@pytest.mark.parametrize("nx,r", [(8, 16), (16, 8)])
def test_square_global_solve_matches_direct(nx, r):
    print(f"square global solve nx={nx} r={r}", flush=True)
    prob, b, blocks, loads = stage_square(nx, r)
    with b:
        plan = hybrid.HybridScatter(prob.global_indices(), prob.trace_size, prob.kernel_offsets())
        vals, rhs, _ = plan.gather(blocks, loads)
        with hybrid.HybridSolver(plan, rtol=1e-12) as solver:
            solver.set_values(vals)
            x = solver.solve(rhs)
            print(f"  {plan.size} unknowns, {solver.iterations[-1]} GMRES steps, residual {solver.residuals[-1]:.2e}",
                  flush=True)
            mat, hrhs, _ = plan.assemble(blocks, loads)
            exact = sla.spsolve(mat.tocsc(), hrhs)
            got = x.cpu().numpy()
            assert rel(got, exact) < TOL
            # reconstruction on the device from the device solution
            ts = prob.trace_size
            tr = x[:ts][torch.as_tensor(prob.trace_dofs, device="cuda")]
            fields = b.reconstruct(tr[:, None, :], x[ts:][:, None, None])[:, 0].cpu().numpy()
            src, lifts, ret = (t.cpu().numpy() for t in b.responses())
            want = np.stack([src[c] - lifts[c] @ exact[prob.trace_dofs[c]] + ret[c] * exact[ts + c]
                             for c in range(prob.M)])
            assert rel(fields, want) < TOL
            # a second right-hand side against the resident preconditioner,
            # bit for bit a fresh solve of the same system
            rhs2 = rhs * 0.5 + 1.0
            x2 = solver.solve(rhs2)
            with hybrid.HybridSolver(plan, rtol=1e-12) as fresh:
                fresh.set_values(vals)
                assert torch.equal(fresh.solve(rhs2), x2)
            assert rel(x2.cpu().numpy(), sla.spsolve(mat.tocsc(), hrhs * 0.5 + 1.0)) < TOL


def test_voronoi_global_solve_matches_direct():
    case = pw.config4(npts=24, h=1 / 80, dof_range=(30, 900))
    prob = case.problem
    print(f"voronoi global solve: {prob.M} macros, {len(prob.buckets)} buckets", flush=True)
    out = pw.run_polygon_stage(case, "fp64")
    idx = prob.global_indices()
    plan = hybrid.HybridScatter(idx, prob.trace_size, prob.kernel_offsets())
    blocks = device_flat([out["blocks"][c] for c in range(prob.M)])
    loads = device_flat([out["loads"][c] for c in range(prob.M)])
    trace, coarse, solver = hybrid.solve_hybrid(plan, blocks, loads, rtol=1e-12)
    with solver:
        mat, hrhs, _ = plan.assemble(blocks, loads)
        exact = sla.spsolve(mat.tocsc(), hrhs)
        got = torch.cat((trace, coarse)).cpu().numpy()
        assert rel(got, exact) < TOL


def test_affine_patch_fully_on_device():
    macro = meshing.voronoi_square(12, seed=5)
    space = meshing.FaceSpace.uniform(1, 1)
    prob = PolygonProblem(macro, [1 + c % 3 for c in range(len(macro.cells))], 1, space, order=4)
    print(f"device patch test: {prob.M} macros", flush=True)
    affine = lambda x: 0.7 + x[..., 0] - 2 * x[..., 1]  # noqa: E731
    nq = len(prob.qw)
    kq = [None] * prob.M
    for bk in prob.buckets:
        for c in bk.cells:
            kq[c] = np.full((len(bk.mesh.cells), nq), 2.0)
    out = pw.run_polygon_stage(pw.Case("patch", prob, kq, None, 0.0), "fp64")
    idx = prob.global_indices()
    plan = hybrid.HybridScatter(idx, prob.trace_size, prob.kernel_offsets())
    bnd = hybrid.DirichletBoundary(macro, space)
    trace, coarse, solver = hybrid.solve_hybrid(plan, device_flat([out["blocks"][c] for c in range(prob.M)]),
                                                device_flat([out["loads"][c] for c in range(prob.M)]),
                                                boundary=(bnd, affine(bnd.points)), rtol=1e-13)
    solver.close()
    tr, co = trace.cpu().numpy(), coarse.cpu().numpy()
    err = scale = 0.0
    for c in range(prob.M):
        u = out["source"][c] - out["lifts"][c] @ tr[idx[c][:-1]] + out["retained"][c] * co[c]
        verts = macro.points[macro.cells[c]]
        pts, _ = meshing.refine_triangles(verts, meshing.ear_split(verts), int(prob.refinements[c]))
        err = max(err, float(np.abs(u - affine(pts)).max()))
        scale = max(scale, float(np.abs(affine(pts)).max()))
    assert err / scale < 1e-10


def test_error_paths():
    prob, b, blocks, loads = stage_square(4, 8)
    with b:
        plan = hybrid.HybridScatter(prob.global_indices(), prob.trace_size, prob.kernel_offsets())
        vals, rhs, _ = plan.gather(blocks, loads)
        with hybrid.HybridSolver(plan, rtol=1e-13, maxiter=2) as solver:
            with pytest.raises(ValueError):
                solver.solve(rhs)  # not prepared: refused before any work
        with hybrid.HybridSolver(plan, rtol=1e-13) as solver:
            with pytest.raises(LinearSolveError):
                solver.set_values(torch.zeros_like(vals))
            assert solver.bad_cell == 0
        with hybrid.HybridSolver(plan, rtol=1e-13, maxiter=2) as solver:
            solver.set_values(vals)
            with pytest.raises(LinearSolveError):
                solver.solve(rhs)
            assert solver.iterations[-1] == 2
        with hybrid.HybridSolver(plan) as solver:
            solver.set_values(vals)
            assert not torch.any(solver.solve(torch.zeros_like(rhs)))
