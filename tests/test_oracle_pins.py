"""Pins of the oracle pieces the golden vectors leave open.

The reference's golden NPZ uses one P0 multiplier per face of triangular
macros, so it fixes neither the discontinuous-P1 Legendre face modes (their
values, their orientation along the face, the degree-1 trace coupling) nor
the polygon ear split and its refinement.  They are pinned here by:

* the reference's own polynomial contract: conventional Legendre values equal
  ``numpy.polynomial.legendre.legvander`` (``tests/test_basix_polynomial_contracts.py:80-91``)
  and the documented segment mass ``L / (2j + 1)`` (``fem/traces/interval.py:54-62``);
* closed forms of the trace coupling ``int phi_i mu_j`` for P1 and P2 nodal
  functions against the modes L0 and L1 on every fine edge of every macroface,
  derived by hand below (not from the oracle's quadrature), with the sign of
  the face's first or second owner;
* invariants of the ear split on the reference's L-shaped macro
  (``tests/test_polygon_local_meshes.py:17-20``: area 3, every vertex kept,
  counterclockwise, conforming) and on Voronoi macros;
* the affine patch test (``tests/test_darcy.py:16-32``): with weak Dirichlet
  data ``0.7 + x - 2y`` the condensed global solve reproduces the affine
  solution to 1e-12 on square macros with ``FaceSpace.uniform(1, 4)`` and on
  Voronoi macros with ``FaceSpace.uniform(1, 1)``;
* a planted mistake in the oracle (odd modes not reversed on the second owner
  of a face) that the closed forms and the patch test must both catch.
"""

import numpy as np
import pytest
from numpy.polynomial.legendre import legvander
from scipy.sparse import linalg as sla

from oracle.fem import FaceSpace, SkeletonSpace, boundary_data, legendre, nodal_space, scalar_operators, trace_coupling
from oracle.local import LocalProblem, assemble_hybrid_contributions
from oracle.meshes import PolygonMesh, ear_triangulation


def affine(x):
    return 0.7 + x[:, 0] - 2 * x[:, 1]


def cross(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


# ---------------------------------------------------------------- Legendre modes
@pytest.mark.parametrize("degree", [0, 1, 9])
def test_legendre_equals_legvander(degree):
    points = np.linspace(-1.07, 1.13, 9)
    np.testing.assert_allclose(legendre(points, degree), legvander(points, degree), atol=8e-14)


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_face_modes_mass_and_orthogonality(degree):
    space = FaceSpace.uniform(degree, 3)
    par, w = space.quadrature(degree + 2)
    vals = space.evaluate(par)
    gram = vals.T @ (w[:, None] * vals)
    seg = 1.0 / 3
    expect = np.kron(np.eye(3), np.diag([seg / (2 * j + 1) for j in range(degree + 1)]))
    np.testing.assert_allclose(gram, expect, atol=1e-15)
    assert abs(w.sum() - 1.0) < 1e-15


# ---------------------------------------------------------------- trace coupling closed forms
# On a straight fine edge with local parameter s in [0, 1] from the node at face
# parameter t0 to the node at t1 (t0 < t1), the segment coordinate of mode L1 is
# x(s) = x0 + (x1 - x0) s, and with h the edge length:
#   P1: int phi_0 = h/2,  int phi_1 = h/2
#       int phi_0 x = h (x0/3 + x1/6),  int phi_1 x = h (x0/6 + x1/3)
#   P2 (end, middle, end): int phi = h (1/6, 2/3, 1/6)
#       int phi x = h (x0/6, (x0 + x1)/3, x1/6)
P1_L0 = np.array([1 / 2, 1 / 2])
P1_L1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])          # rows: phi_0, phi_1; cols: x0, x1
P2_L0 = np.array([1 / 6, 2 / 3, 1 / 6])
P2_L1 = np.array([[1 / 6, 0.0], [1 / 3, 1 / 3], [0.0, 1 / 6]])


def closed_form_coupling(pm, cell, fine, space, degree):
    """B[i, (side, segment, mode)] from the closed forms above (modes L0, L1)."""
    _, pts = nodal_space(fine, degree)
    S = len(space.degrees)
    assert all(d == 1 for d in space.degrees)
    out = np.zeros((len(pts), len(pm.cell_faces[cell]) * 2 * S))
    for side, face in enumerate(pm.cell_faces[cell]):
        sign = pm.signs[cell][side]
        start, end = pm.points[pm.faces[face]]
        tan = end - start
        length = np.linalg.norm(tan)
        t = (pts - start) @ tan / length ** 2
        dist = np.abs(cross(tan, pts - start)) / length
        on = np.flatnonzero((dist < 1e-12) & (t > -1e-12) & (t < 1 + 1e-12))
        on = on[np.argsort(t[on])]
        step = degree
        for e in range(0, len(on) - 1, step):
            nodes = on[e:e + step + 1]
            t0, t1 = t[nodes[0]], t[nodes[-1]]
            k = min(int(0.5 * (t0 + t1) * S), S - 1)
            lo, hi = space.breaks[k], space.breaks[k + 1]
            assert lo - 1e-12 <= t0 and t1 <= hi + 1e-12, "fine edge crosses a segment break"
            x0, x1 = 2 * (t0 - lo) / (hi - lo) - 1, 2 * (t1 - lo) / (hi - lo) - 1
            h = (t1 - t0) * length
            m0, m1 = (P1_L0, P1_L1) if degree == 1 else (P2_L0, P2_L1)
            col = side * 2 * S + 2 * k
            out[nodes, col] += sign * h * m0
            out[nodes, col + 1] += sign * h * (m1 @ [x0, x1])
    return out


def planted_no_reversal(pm, cell, b, space):
    """The oracle coupling with the odd modes of second owners not reversed."""
    bad = b.copy()
    S = len(space.degrees)
    for side in range(len(pm.cell_faces[cell])):
        if pm.signs[cell][side] < 0:
            bad[:, side * 2 * S + 1: (side + 1) * 2 * S: 2] *= -1
    return bad


@pytest.mark.parametrize("degree,refinement", [(1, 4), (1, 8), (2, 4)])
def test_trace_coupling_closed_form_square_macros(degree, refinement):
    pm = PolygonMesh.unit_square_grid(3)
    space = FaceSpace.uniform(1, 4)
    skel = SkeletonSpace(pm, [space] * len(pm.faces))
    reversed_sides = 0
    for c in range(len(pm.cells)):
        fine = pm.submesh(c, refinement)
        b = trace_coupling(pm, c, fine, skel, degree)
        expect = closed_form_coupling(pm, c, fine, space, degree)
        np.testing.assert_allclose(b, expect, atol=1e-15)
        if np.any(pm.signs[c] < 0):
            reversed_sides += 1
            assert np.abs(planted_no_reversal(pm, c, b, space) - expect).max() > 1e-3
    assert reversed_sides > 0


def test_trace_coupling_closed_form_voronoi_macros():
    from pymhm_amd import meshing

    macro = meshing.voronoi_square(10, seed=3)
    pm = PolygonMesh(macro.points, [list(cell) for cell in macro.cells])
    space = FaceSpace.uniform(1, 1)
    skel = SkeletonSpace(pm, [space] * len(pm.faces))
    for c in range(len(pm.cells)):
        fine = pm.submesh(c, 3)
        b = trace_coupling(pm, c, fine, skel, 1)
        np.testing.assert_allclose(b, closed_form_coupling(pm, c, fine, space, 1), atol=1e-15)


# ---------------------------------------------------------------- ear split
def l_macro():
    points = np.array([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], dtype=float)
    return PolygonMesh(points, (np.arange(6),))


def check_local_mesh(pm, cell, r):
    """Area, orientation, kept vertices and conformity of submesh(cell, r)."""
    fine = pm.submesh(cell, r)
    poly = pm.points[pm.cells[cell]]
    k = len(poly)
    area = 0.5 * np.sum(cross(poly, np.roll(poly, -1, axis=0)))
    assert abs(fine.areas.sum() - area) < 1e-12 * max(1.0, area)
    assert np.all(fine.areas > 0)
    assert len(fine.cells) == (k - 2) * r * r
    for v in poly:  # every polygon vertex is a node of the local mesh
        assert np.min(np.linalg.norm(fine.points - v, axis=1)) < 1e-14
    # conforming: interior edges have two cells, boundary edges lie on the
    # polygon's sides and cover each side exactly once
    inner = fine.face_cells[:, 1] >= 0
    bnd = ~inner
    cover = np.zeros(k)
    for f in np.flatnonzero(bnd):
        a, b = fine.points[fine.faces[f]]
        hit = []
        for s in range(k):
            p, q = poly[s], poly[(s + 1) % k]
            tan = q - p
            L = np.linalg.norm(tan)
            ok = all(abs(cross(tan, x - p)) / L < 1e-12 and -1e-12 <= (x - p) @ tan / L ** 2 <= 1 + 1e-12
                     for x in (a, b))
            if ok:
                hit.append(s)
        assert len(hit) == 1, "boundary edge not on exactly one polygon side"
        cover[hit[0]] += np.linalg.norm(b - a)
    sides = np.linalg.norm(np.roll(poly, -1, axis=0) - poly, axis=1)
    np.testing.assert_allclose(cover, sides, rtol=1e-12)
    # each ear contains no other polygon vertex (the ears are a triangulation)
    tris = ear_triangulation(poly)
    assert len(tris) == k - 2
    for t in tris:
        tri = poly[t]
        assert cross(tri[1] - tri[0], tri[2] - tri[0]) > 0  # counterclockwise ears
        for j in set(range(k)) - set(t.tolist()):
            bc = np.linalg.solve(np.vstack([tri.T, np.ones(3)]), np.r_[poly[j], 1.0])
            assert np.any(bc < 1e-12)
    return fine


@pytest.mark.parametrize("r", [1, 3])
def test_ear_split_of_l_macro(r):
    fine = check_local_mesh(l_macro(), 0, r)
    assert abs(fine.areas.sum() - 3.0) < 1e-12


def test_ear_split_of_voronoi_macros():
    from pymhm_amd import meshing

    macro = meshing.voronoi_square(24, seed=7)
    pm = PolygonMesh(macro.points, [list(cell) for cell in macro.cells])
    total = sum(check_local_mesh(pm, c, 2).areas.sum() for c in range(len(pm.cells)))
    assert abs(total - 1.0) < 1e-12


def test_ear_split_of_translated_l_macro():
    """The local mesh of a translated macro is the translated local mesh."""
    base = l_macro().submesh(0, 3)
    pm = PolygonMesh(l_macro().points + [500.0, 2100.0], (np.arange(6),))
    moved = pm.submesh(0, 3)
    np.testing.assert_array_equal(moved.cells, base.cells)
    np.testing.assert_allclose(moved.points, base.points + [500.0, 2100.0], atol=1e-12)


# ---------------------------------------------------------------- affine patch test
def patch_error(pm, space, degree, refinements, diffusion=2.0, planted=False):
    """Max nodal error of the condensed global solve against the affine data,
    relative to the data's largest value."""
    skel = SkeletonSpace(pm, [space] * len(pm.faces))
    responses, contributions, fines = [], [], []
    for c in range(len(pm.cells)):
        fine = pm.submesh(c, refinements[c])
        a, m, f = scalar_operators(fine, degree, diffusion=diffusion, source=0.0, order=4)
        b = trace_coupling(pm, c, fine, skel, degree)
        if planted:
            b = planted_no_reversal(pm, c, b, space)
        z = np.ones((a.shape[0], 1))
        mz = m @ z
        p = LocalProblem(a, b, f, skel.cell_dofs(c), z, mz / (z.T @ mz).item())
        r = p.condense()
        responses.append(r)
        fines.append(fine)
        contributions.append(r.global_contribution([skel.size + c]))
    load = boundary_data(skel, affine)
    mat, rhs, _ = assemble_hybrid_contributions(
        contributions, skel.size, skel.size + np.arange(len(pm.cells) + 1), boundary_load=load)
    sol = sla.spsolve(mat.tocsc(), rhs)
    trace, coarse = sol[:skel.size], sol[skel.size:]
    err, scale = 0.0, 0.0
    for c, (r, fine) in enumerate(zip(responses, fines)):
        u = r.reconstruct(trace[r.problem.trace_dofs], coarse[c:c + 1])
        _, pts = nodal_space(fine, degree)
        err = max(err, float(np.max(np.abs(u - affine(pts)))))
        scale = max(scale, float(np.max(np.abs(affine(pts)))))
    return err / scale


@pytest.mark.parametrize("degree,refinement", [(1, 4), (2, 2)])
def test_affine_patch_square_macros(degree, refinement):
    pm = PolygonMesh.unit_square_grid(3)
    assert patch_error(pm, FaceSpace.uniform(1, 4), degree, [refinement] * len(pm.cells)) < 1e-12


def test_affine_patch_voronoi_macros():
    from pymhm_amd import meshing

    macro = meshing.voronoi_square(12, seed=5)
    pm = PolygonMesh(macro.points, [list(cell) for cell in macro.cells])
    refine = [1 + c % 3 for c in range(len(pm.cells))]
    assert patch_error(pm, FaceSpace.uniform(1, 1), 1, refine) < 1e-12


def test_patch_test_catches_unreversed_odd_modes():
    pm = PolygonMesh.unit_square_grid(3)
    assert patch_error(pm, FaceSpace.uniform(1, 4), 1, [4] * len(pm.cells), planted=True) > 1e-4
