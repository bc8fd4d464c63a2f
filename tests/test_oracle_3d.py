"""Pins of the tetrahedral oracle and of the 3D host tables (CPU only).

* local tetrahedral meshes: the reference's golden fixture
  ``original-3d-3d-n2-r4-p2-q6.npz`` (48 red-refined tetrahedral macros of
  ``TetraMesh.unit_cube(2)``), points and cells bit for bit;
* the Duffy tetrahedron rule: exact monomial integrals
  ``int lambda^alpha = alpha! 3! / (|alpha| + 3)!`` (unit-volume weights) up to
  the degree the rule integrates exactly;
* P2 tabulation: the Lagrange property at the lattice nodes;
* operators: closed forms (zero row sums of the stiffness, the energy of a
  linear function, integrals of the P2 vertex/edge functions -V/20 and V/5,
  the P1 stiffness ``V g_i . g_j``);
* cube-face coupling: every column integrates the constant over its face cell,
  P2 vertex rows vanish;
* the GPU host's tables equal the oracle's bit for bit (assembly parity).
"""

import math
from itertools import product
from pathlib import Path

import numpy as np
import pytest

from oracle import fem3d
from oracle.fem import simplex_lagrange
from oracle.meshes import TetraMesh

GOLDEN = Path(__file__).parent / "golden" / "original-3d-3d-n2-r4-p2-q6.npz"


def test_tetra_meshes_match_reference_golden():
    g = np.load(GOLDEN)
    macro = TetraMesh.unit_cube(2)
    assert len(macro.cells) == 48
    for c in range(48):
        fine = macro.submesh(c, 4)
        assert np.array_equal(fine.points, g["local_points"][c]), c
        assert np.array_equal(fine.cells, g["local_cells"][c]), c


def test_tetra_quadrature_monomials():
    bary, w = fem3d.tetrahedron_quadrature(4)
    for alpha in product(range(6), repeat=4):
        if sum(alpha) > 5:
            continue
        exact = math.prod(math.factorial(a) for a in alpha) * 6 / math.factorial(sum(alpha) + 3)
        assert abs(np.sum(w * np.prod(bary ** np.array(alpha), axis=1)) - exact) < 1e-15


def test_p2_lagrange_property():
    alphas = fem3d.tetra_multiindices(2)
    vals, ders = simplex_lagrange(2, alphas / 2.0, alphas)
    assert np.array_equal(vals, np.eye(10))


@pytest.mark.parametrize("degree", [1, 2])
def test_tetra_operators_closed_forms(degree):
    mesh = TetraMesh.unit_cube(2)
    A, M, f = fem3d.tetra_operators(mesh, degree, source=1.0)
    dofs, nodes = fem3d.tetra_nodal_space(mesh, degree)
    one = np.ones(A.shape[0])
    assert np.abs(A @ one).max() < 1e-13
    for axis in range(3):  # int |grad x_a|^2 over the unit cube = 1
        x = nodes[:, axis]
        assert abs(x @ A @ x - 1.0) < 1e-13
    vol = mesh.volumes
    expect = np.zeros(len(nodes))
    if degree == 1:
        np.add.at(expect, dofs.ravel(), np.repeat(vol / 4, 4))
        g = fem3d.tetra_gradients(mesh)
        blocks = np.einsum("t,tia,tja->tij", vol, g, g)
        dense = np.zeros((len(nodes), len(nodes)))
        for t in range(len(mesh.cells)):
            dense[np.ix_(dofs[t], dofs[t])] += blocks[t]
        assert np.abs(A.toarray() - dense).max() < 1e-13
    else:
        np.add.at(expect, dofs[:, :4].ravel(), np.repeat(-vol / 20, 4))
        np.add.at(expect, dofs[:, 4:].ravel(), np.repeat(vol / 5, 6))
    assert np.abs(M @ one - expect).max() < 1e-15
    assert np.abs(f - expect).max() < 1e-15


def test_cube_face_coupling_closed_form():
    mesh = TetraMesh(np.array([0.25, 0.5, 0.0]) + TetraMesh.unit_cube(2).points * 0.25, TetraMesh.unit_cube(2).cells)
    B = fem3d.cube_face_coupling(mesh, 2, [0.25, 0.5, 0.0], [0.5, 0.75, 0.25], split=2)
    assert B.shape == (125, 24)
    assert np.abs(B.sum(axis=0) - 0.125 ** 2).max() < 1e-15
    dofs, _ = fem3d.tetra_nodal_space(mesh, 2)
    vertices = np.unique(dofs[:, :4])
    assert np.abs(B[vertices]).max() < 1e-15


def test_host_tables_equal_oracle():
    from pymhm_amd import elements, meshing

    pts, cells = meshing.freudenthal_cube(2, (0.5, 0.25, 0.75), 0.25)
    oracle_mesh = TetraMesh(np.array([0.5, 0.25, 0.75]) + TetraMesh.unit_cube(2).points * 0.25,
                            TetraMesh.unit_cube(2).cells)
    assert np.array_equal(pts, oracle_mesh.points) and np.array_equal(cells, oracle_mesh.cells)
    host = meshing.tet_local_mesh(pts, cells, 2)
    dofs, nodes = fem3d.tetra_nodal_space(oracle_mesh, 2)
    assert np.array_equal(host.dofs, dofs) and np.array_equal(host.nodes, nodes)
    bary, w = elements.tetra_rule(4)
    ob, ow = fem3d.tetrahedron_quadrature(4)
    assert np.array_equal(bary, ob) and np.array_equal(w, ow)
    vals, der = elements.tetra_tables(2, bary)
    ov, od = simplex_lagrange(2, ob, fem3d.tetra_multiindices(2))
    assert np.array_equal(vals, ov) and np.array_equal(der, od)
    g, vol = elements.tetra_gradients(pts, cells)
    assert np.array_equal(g, fem3d.tetra_gradients(oracle_mesh)) and np.array_equal(vol, oracle_mesh.volumes)
    B = meshing.cube_coupling(host, (0.5, 0.25, 0.75), 0.25, 2)
    OB = fem3d.cube_face_coupling(oracle_mesh, 2, [0.5, 0.25, 0.75], [0.75, 0.5, 1.0], 2)
    assert np.abs(B - OB).max() < 1e-16


def test_cube_spec_worker_matches_oracle_config():
    """The picklable per-worker path of the CPU baseline gives the same blocks
    and loads as the direct oracle evaluation."""
    from oracle import workloads as ow
    from pymhm_amd import workloads as pw

    case = pw.config3(nx=2, refinement=2)
    sample = [1, 6]
    macros, blocks, loads = ow.condense_spec(ow.cube_spec(case, sample))
    ref = ow.oracle_config(case, sample)
    assert macros == sample
    assert np.array_equal(blocks, ref["blocks"]) and np.array_equal(loads, ref["loads"])
