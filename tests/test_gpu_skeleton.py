"""Skeleton and boundary forms on the GPU against the oracle (marked ``gpu``).

* per-macro trace couplings (``mhm_batch_skeleton_coupling``) and
  constraints ``M 1 / 1^T M 1`` (``mhm_batch_mass_constraints``) equal the
  oracle's ``trace_coupling`` and ``mz / (Z^T mz)`` to 1e-13 relative;
* the weak Dirichlet trace load (``mhm_dirichlet_load``) equals the oracle's
  ``boundary_data`` to 1e-14 relative, and its application to the rhs and
  load scale matches the host path bit for bit;
* the affine patch test (``tests/test_darcy.py:16-32``) end to end: device
  stage, device scatter with the device Dirichlet load, a host sparse solve
  and the device responses reproduce ``0.7 + x - 2 y`` to 1e-10.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from scipy.sparse import linalg as sla  # noqa: E402

from oracle.fem import FaceSpace, SkeletonSpace, boundary_data, scalar_operators, trace_coupling  # noqa: E402
from oracle.meshes import PolygonMesh  # noqa: E402
from pymhm_amd import hybrid, meshing  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402
from pymhm_amd.problems import PolygonProblem  # noqa: E402

CASES = [
    ("squares", lambda: meshing.square_grid(3), (1, 4), 1, lambda c: 4),
    ("squares-p2", lambda: meshing.square_grid(3), (1, 4), 2, lambda c: 2),
    ("voronoi", lambda: meshing.voronoi_square(12, seed=5), (1, 1), 1, lambda c: 1 + c % 3),
    ("voronoi-p2", lambda: meshing.voronoi_square(10, seed=7), (2, 2), 2, lambda c: 1 + c % 2),
]
IDS = [c[0] for c in CASES]


def affine(x):
    return 0.7 + x[..., 0] - 2 * x[..., 1]


def build(make, space_args, degree, refine):
    macro = make()
    space = meshing.FaceSpace.uniform(*space_args)
    prob = PolygonProblem(macro, [refine(c) for c in range(len(macro.cells))], degree, space, order=4)
    pm = PolygonMesh(macro.points, [list(cell) for cell in macro.cells])
    skel = SkeletonSpace(pm, [FaceSpace(space.breaks, space.degrees)] * len(pm.faces))
    return macro, space, prob, pm, skel


@pytest.mark.parametrize("name,make,space_args,degree,refine", CASES, ids=IDS)
def test_device_couplings_and_constraints(name, make, space_args, degree, refine):
    macro, space, prob, pm, skel = build(make, space_args, degree, refine)
    print(f"skeleton {name}: {prob.M} macros, {len(prob.buckets)} buckets", flush=True)
    worst_b = worst_c = 0.0
    for bk in prob.buckets:
        with bk.batch("fp64") as b:
            cp, cn = (t.cpu().numpy() for t in b.coupling())
        for r, c in enumerate(bk.cells):
            fine = pm.submesh(c, int(prob.refinements[c]))
            want = trace_coupling(pm, c, fine, skel, degree)
            worst_b = max(worst_b, np.abs(cp[r].T - want).max() / np.abs(want).max())
            _, mass, _ = scalar_operators(fine, degree, order=4)
            mz = mass @ np.ones(mass.shape[0])
            worst_c = max(worst_c, np.abs(cn[r] - mz / mz.sum()).max() / np.abs(mz / mz.sum()).max())
    assert worst_b < 1e-13
    assert worst_c < 1e-13


@pytest.mark.parametrize("name,make,space_args,degree,refine", CASES[::2], ids=IDS[::2])
def test_device_dirichlet_load(name, make, space_args, degree, refine):
    macro, space, prob, pm, skel = build(make, space_args, degree, refine)
    print(f"dirichlet {name}", flush=True)
    bnd = hybrid.DirichletBoundary(macro, space)
    g = np.sin(3 * bnd.points[..., 0]) + bnd.points[..., 0] * bnd.points[..., 1]
    got = bnd.load(g).cpu().numpy()
    want = boundary_data(skel, lambda x: np.sin(3 * x[:, 0]) + x[:, 0] * x[:, 1])
    assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max()
    # applied to a global rhs and load scale exactly as the host path applies a load vector
    size = prob.trace_size + prob.M
    rhs0 = np.random.default_rng(1).standard_normal(size)
    rhs = torch.as_tensor(rhs0, device="cuda")
    scale = torch.as_tensor(np.abs(rhs0), device="cuda")
    bnd.load(g, rhs=rhs, scale=scale)
    want_rhs, want_scale = rhs0.copy(), np.abs(rhs0)
    want_rhs[:prob.trace_size] -= got
    want_scale[:prob.trace_size] += np.abs(got)
    assert np.array_equal(rhs.cpu().numpy(), want_rhs)
    assert np.array_equal(scale.cpu().numpy(), want_scale)
    with pytest.raises(ValueError):
        bnd.load(g[:, :-1])


@pytest.mark.parametrize("name,make,space_args,degree,refine", CASES, ids=IDS)
def test_affine_patch_on_device(name, make, space_args, degree, refine):
    macro, space, prob, pm, skel = build(make, space_args, degree, refine)
    print(f"patch {name}", flush=True)
    nq = len(prob.qw)
    kq = [None] * prob.M
    for bk in prob.buckets:
        for c in bk.cells:
            kq[c] = np.full((len(bk.mesh.cells), nq), 2.0)
    out = pw.run_polygon_stage(pw.Case("patch", prob, kq, None, 0.0), "fp64")
    idx = prob.global_indices()
    plan = hybrid.HybridScatter(idx, prob.trace_size, prob.kernel_offsets())
    blocks = torch.as_tensor(np.concatenate([out["blocks"][c].ravel() for c in range(prob.M)]), device="cuda")
    loads = torch.as_tensor(np.concatenate([out["loads"][c] for c in range(prob.M)]), device="cuda")
    bnd = hybrid.DirichletBoundary(macro, space)
    mat, rhs, _ = plan.assemble(blocks, loads, boundary=(bnd, affine(bnd.points)))
    sol = sla.spsolve(mat.tocsc(), rhs)
    err = scale = 0.0
    for c in range(prob.M):
        lam, coarse = sol[idx[c][:-1]], sol[idx[c][-1]]
        u = out["source"][c] - out["lifts"][c] @ lam + out["retained"][c] * coarse
        verts = macro.points[macro.cells[c]]
        pts, cells = meshing.refine_triangles(verts, meshing.ear_split(verts), int(prob.refinements[c]))
        exact = affine(meshing.local_mesh(pts, cells, degree).nodes)
        err = max(err, float(np.abs(u - exact).max()))
        scale = max(scale, float(np.abs(exact).max()))
    assert err / scale < 1e-10
