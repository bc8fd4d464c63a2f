"""Skeleton tables of the device trace-coupling kernel, checked on the CPU.

``meshing.skeleton_tables`` describes a macro's skeleton coupling by fine
boundary-edge pieces at fixed fractions of the sides; the device kernel turns
them into ``B`` with each macro's side lengths and orientations.  Evaluating
the tables here with plain numpy (Gauss rule per piece, nodal edge functions,
Legendre modes on the segment, ``t -> 1 - t`` on a second owner) and
comparing with the oracle's ``trace_coupling`` pins the tables' pieces,
entry ids and segment assignment for square and Voronoi macros, P1 and P2,
one and several segments; a planted orientation slip must fail.
"""

import numpy as np
import pytest
from numpy.polynomial.legendre import legvander

from oracle.fem import FaceSpace, SkeletonSpace, trace_coupling
from oracle.meshes import PolygonMesh
from pymhm_amd import meshing


def eval_tables(t, side_len, side_sign, n, flip=True):
    """``B (n, m)`` of one macro from the tables (``flip=False`` plants the
    mistake of not mirroring the second owner's parameter)."""
    deg, nseg, breaks = t["degree"], t["nseg"], t["breaks"]
    fs = nseg * (deg + 1)
    out = np.zeros((n, t["nsides"] * fs))
    for e in range(len(t["entry_node"])):
        s, node = t["entry_side"][e], t["entry_node"][e]
        fwd = side_sign[s] > 0 or not flip
        seg = t["entry_seg"][e] if fwd else nseg - 1 - t["entry_seg"][e]
        lo, hi = breaks[seg], breaks[seg + 1]
        acc = np.zeros(deg + 1)
        for code in t["contrib"][t["entry_ptr"][e]:t["entry_ptr"][e + 1]]:
            f0, f1, e0, e1 = t["piece_f"][code // 3]
            f = f0 + (t["gx"] + 1) * (f1 - f0) / 2
            u = (f - e0) / (e1 - e0)
            phi = {1: [1 - u, u], 2: [(1 - u) * (1 - 2 * u), u * (2 * u - 1), 4 * u * (1 - u)]}[t["edeg"]][code % 3]
            tq = f if fwd else 1 - f
            modes = legvander(2 * (tq - lo) / (hi - lo) - 1, deg)
            acc += (t["gw"] * (f1 - f0) / 2 * phi) @ modes
        out[node, s * fs + seg * (deg + 1):s * fs + (seg + 1) * (deg + 1)] += acc * side_len[s] * side_sign[s]
    return out


CASES = [
    ("squares", lambda: meshing.square_grid(3), FaceSpace.uniform(1, 4), 1, lambda c: 3),
    ("squares-p2", lambda: meshing.square_grid(2), FaceSpace.uniform(2, 2), 2, lambda c: 2),
    ("voronoi", lambda: meshing.voronoi_square(12, seed=5), FaceSpace.uniform(1, 1), 1, lambda c: 1 + c % 3),
    ("voronoi-p2", lambda: meshing.voronoi_square(8, seed=3), FaceSpace.uniform(3, 2), 2, lambda c: 1 + c % 2),
]


@pytest.mark.parametrize("name,make,space,degree,refine", CASES, ids=[c[0] for c in CASES])
def test_tables_reproduce_oracle_coupling(name, make, space, degree, refine):
    macro = make()
    pm = PolygonMesh(macro.points, [list(cell) for cell in macro.cells])
    skel = SkeletonSpace(pm, [space] * len(pm.faces))
    ours = meshing.FaceSpace(space.breaks, space.degrees)
    worst = 0.0
    for c in range(len(macro.cells)):
        verts = macro.points[macro.cells[c]]
        pts, cells = meshing.refine_triangles(verts, meshing.ear_split(verts), refine(c))
        mesh = meshing.local_mesh(pts, cells, degree)
        t = meshing.skeleton_tables(mesh, verts, ours)
        lens = np.linalg.norm(np.roll(verts, -1, axis=0) - verts, axis=1)
        got = eval_tables(t, lens, np.asarray(macro.signs[c], dtype=float), mesh.n)
        want = trace_coupling(pm, c, pm.submesh(c, refine(c)), skel, degree)
        worst = max(worst, np.abs(got - want).max() / np.abs(want).max())
    assert worst < 1e-14


def test_planted_orientation_slip_fails():
    macro = meshing.voronoi_square(12, seed=5)
    pm = PolygonMesh(macro.points, [list(cell) for cell in macro.cells])
    space = FaceSpace.uniform(1, 1)
    skel = SkeletonSpace(pm, [space] * len(pm.faces))
    worst = 0.0
    for c in range(len(macro.cells)):
        if np.all(np.asarray(macro.signs[c]) > 0):
            continue
        verts = macro.points[macro.cells[c]]
        pts, cells = meshing.refine_triangles(verts, meshing.ear_split(verts), 2)
        mesh = meshing.local_mesh(pts, cells, 1)
        t = meshing.skeleton_tables(mesh, verts, meshing.FaceSpace(space.breaks, space.degrees))
        lens = np.linalg.norm(np.roll(verts, -1, axis=0) - verts, axis=1)
        got = eval_tables(t, lens, np.asarray(macro.signs[c], dtype=float), mesh.n, flip=False)
        want = trace_coupling(pm, c, pm.submesh(c, 2), skel, 1)
        worst = max(worst, np.abs(got - want).max() / np.abs(want).max())
    assert worst > 1e-3


def test_tables_reject_mixed_degrees():
    verts = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    pts, cells = meshing.refine_triangles(verts, meshing.ear_split(verts), 2)
    mesh = meshing.local_mesh(pts, cells, 1)
    with pytest.raises(ValueError):
        meshing.skeleton_tables(mesh, verts, meshing.FaceSpace((0.0, 0.5, 1.0), (1, 2)))
