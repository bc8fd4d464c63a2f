"""Voronoi macro meshes and the bucketed polygon problem (CPU)."""

import numpy as np

from oracle.meshes import PolygonMesh
from pymhm_amd import meshing
from pymhm_amd import workloads as pw


def _area(xy):
    return 0.5 * float(np.sum(meshing._cross(xy - xy[0], np.roll(xy, -1, axis=0) - xy[0])))


def test_voronoi_square_is_a_convex_partition():
    npts = 64
    m = meshing.voronoi_square(npts, 11)
    assert len(m.cells) == npts
    assert abs(sum(_area(m.points[c]) for c in m.cells) - 1.0) < 1e-12
    for c in m.cells:
        xy = m.points[c]
        e = np.roll(xy, -1, axis=0) - xy
        assert np.all(meshing._cross(e, np.roll(e, -1, axis=0)) > 0)  # counterclockwise, convex
    lengths = np.linalg.norm(m.points[m.faces[:, 1]] - m.points[m.faces[:, 0]], axis=1)
    assert lengths.min() >= 0.05 / np.sqrt(npts) * (1 - 1e-12)
    owners = np.zeros(len(m.faces), dtype=int)
    for cf in m.cell_faces:
        np.add.at(owners, cf, 1)
    assert owners.max() == 2
    # boundary faces lie on the unit square's sides
    b = m.points[m.faces[owners == 1]]
    assert np.all(np.any((b == 0.0) | (b == 1.0), axis=2).all(axis=1))


def test_oracle_polygon_mesh_has_the_same_faces_and_signs():
    m = meshing.voronoi_square(48, 5)
    pm = PolygonMesh(m.points, [list(c) for c in m.cells])
    assert np.array_equal(pm.faces, m.faces)
    for c in range(len(m.cells)):
        assert np.array_equal(pm.cell_faces[c], m.cell_faces[c])
        assert np.array_equal(pm.signs[c], m.signs[c])


def test_buckets_share_topology_and_cover_every_macro():
    case = pw.config4(npts=40, h=1 / 96, dof_range=(30, 800))
    prob = case.problem
    seen = np.zeros(prob.M, dtype=int)
    for bk in prob.buckets:
        seen[bk.cells] += 1
        for c in bk.cells:  # every macro's own refinement has the bucket's cells
            verts = prob.macro.points[prob.macro.cells[c]]
            pts, cells = meshing.refine_triangles(verts, meshing.ear_split(verts), bk.r)
            assert np.array_equal(cells, bk.mesh.cells) and len(pts) == bk.n
    assert np.all(seen == 1)
    assert all(case.kq[c].shape == prob.buckets[prob.bucket_of[c]].qpts.shape[1:3] for c in range(prob.M))
