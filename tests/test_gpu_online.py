"""Offline/online hybrid solves fully on the device (marked ``gpu``).

``OnlineHybridSystem`` answers source queries against resident local factors
and a resident global preconditioner.  Its global solutions match a host
sparse direct solve of the same system (the checker) to 1e-10, its fields
match the host reconstruction from the device responses, a query's bits are
the same alone and among others, and an online answer equals the offline
solve of the same source to 1e-10.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from scipy.sparse import linalg as sla  # noqa: E402

from pymhm_amd import hybrid, serving  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402
from pymhm_amd.online import OnlineHybridSystem  # noqa: E402

TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_online_queries(precision):
    case = pw.config2(nx=6, refinement=8)
    prob = case.problem
    print(f"online {precision}: {prob.M} macros", flush=True)
    pts = torch.as_tensor(prob.quadrature_points(), dtype=torch.float64, device="cuda")
    fq = serving.bump_sources(pts, [7, 8, 9])
    with OnlineHybridSystem(prob, case.kq, None, case.f_const, precision=precision, max_columns=4,
                            rtol=1e-13) as online:
        sols, fields = online.solve(fq)
        sol1, field1 = online.solve(fq[1:2])
        assert torch.equal(sol1[0], sols[1]) and torch.equal(field1[0], fields[1])
        loads, src = online.batch.sources(fq, source=True)
        src = src.cpu().numpy()
        _, lifts, ret = (t.cpu().numpy() for t in online.batch.responses())
        blocks = online.batch.blocks
        for c in range(3):
            mat, rhs, _ = online.plan.assemble(blocks, loads[c])
            exact = sla.spsolve(mat.tocsc(), rhs)
            assert rel(sols[c].cpu().numpy(), exact) < TOL, c
            ts = prob.trace_size
            want = np.stack([src[c, m] - lifts[m] @ exact[prob.trace_dofs[m]] + ret[m] * exact[ts + m]
                             for m in range(prob.M)])
            assert rel(fields[c].cpu().numpy(), want) < TOL, c
    # the online answer of source 0 against the offline solve with that source
    with OnlineHybridSystem(prob, case.kq, fq[0], 0.0, precision=precision, rtol=1e-13) as offline:
        assert rel(sols[0].cpu().numpy(), offline.solver.solve(offline.offline_rhs).cpu().numpy()) < TOL


def test_online_dirichlet_data():
    """New Dirichlet data per query on the device (the boundary load moves to
    the right-hand side only)."""
    case = pw.config2(nx=4, refinement=8)
    prob = case.problem
    bnd = hybrid.DirichletBoundary(prob.macro, prob.space)
    g0 = 1.0 + bnd.points[..., 0]
    with OnlineHybridSystem(prob, case.kq, None, case.f_const, boundary=(bnd, g0), rtol=1e-13) as online:
        g1 = np.sin(bnd.points[..., 1])
        sols, _ = online.solve(None, case.f_const, ncols=1, boundary=(bnd, g1))
        loads = online.batch.sources(None, case.f_const, ncols=1)
        mat, rhs, _ = online.plan.assemble(online.batch.blocks, loads[0],
                                           boundary_load=bnd.load(g1).cpu().numpy())
        assert rel(sols[0].cpu().numpy(), sla.spsolve(mat.tocsc(), rhs)) < TOL
