"""Pin the oracle against the reference's own golden vectors.

``tests/golden/original-2d-2d-n4-r8-p2-q6.npz`` is the archived output of the
reference's P2 Darcy case (``examples/numba_performance.py:49-183``, records
``benchmarks/results/numba-20261009/``): 32 macro triangles refined 8 times,
P2 pressure, one P0 multiplier per macroface, K = exp(0.25 prod sin(2 pi x_i /
0.137)), manufactured source, quadrature order 6, weak homogeneous Dirichlet
data, constraints M 1.  The reference's original and compiled runs differ by
up to 9.6e-15, which bounds the agreement expected here.
"""

from pathlib import Path

import numpy as np
import pytest
from scipy.sparse import linalg as sla

from oracle.fem import SkeletonSpace, scalar_operators, trace_coupling
from oracle.local import LocalProblem, assemble_hybrid_contributions
from oracle.meshes import TriangleMesh

GOLDEN = Path(__file__).parent / "golden"
PERIOD, AMP = 0.137, 0.25


def permeability(x):
    return np.exp(AMP * np.prod(np.sin(2 * np.pi * x / PERIOD), axis=-1))


def exact(x):
    return np.prod(np.sin(np.pi * x), axis=-1)


def source(x):
    fr = 2 * np.pi / PERIOD
    s, c = np.sin(fr * x), np.cos(fr * x)
    k = permeability(x)
    grad_k = fr * AMP * k[..., None] * np.stack([c[..., 0] * s[..., 1], c[..., 1] * s[..., 0]], -1)
    sp, cp = np.sin(np.pi * x), np.cos(np.pi * x)
    grad_u = np.pi * np.stack([cp[..., 0] * sp[..., 1], cp[..., 1] * sp[..., 0]], -1)
    return 2 * np.pi ** 2 * k * exact(x) - np.sum(grad_k * grad_u, axis=-1)


def golden_2d_case():
    mesh = TriangleMesh.unit_square(4)
    skel = SkeletonSpace(mesh)
    responses, contributions, fines = [], [], []
    for c in range(len(mesh.cells)):
        fine = mesh.submesh(c, 8)
        a, m, f = scalar_operators(fine, 2, diffusion=permeability, source=source, order=6)
        b = trace_coupling(mesh, c, fine, skel, 2)
        z = np.ones((a.shape[0], 1))
        p = LocalProblem(a, b, f, skel.cell_dofs(c), z, m @ z, test_coupling=b)
        r = p.condense()
        responses.append(r)
        fines.append(fine)
        contributions.append(r.global_contribution([skel.size + c]))
    mat, rhs, _ = assemble_hybrid_contributions(
        contributions, skel.size, skel.size + np.arange(len(mesh.cells) + 1))
    return mesh, skel, fines, responses, mat, rhs


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN / "original-2d-2d-n4-r8-p2-q6.npz")


@pytest.fixture(scope="module")
def case():
    return golden_2d_case()


def test_local_meshes_match(golden, case):
    _, _, fines, _, _, _ = case
    for c, fine in enumerate(fines):
        assert np.array_equal(fine.points, golden["local_points"][c])
        assert np.array_equal(fine.cells, golden["local_cells"][c])


def test_global_pattern_bit_exact(golden, case):
    mat = case[4]
    assert np.array_equal(mat.indptr, golden["global_indptr"])
    assert np.array_equal(mat.indices, golden["global_indices"])


def test_condensed_values_and_load(golden, case):
    mat, rhs = case[4], case[5]
    gv, gl = golden["global_values"], golden["global_load"]
    assert np.linalg.norm(mat.data - gv) / np.linalg.norm(gv) < 3e-14
    assert np.linalg.norm(rhs - gl) / np.linalg.norm(gl) < 3e-15


def test_trace_and_reconstructed_pressure(golden, case):
    mesh, skel, _, responses, mat, rhs = case
    sol = sla.spsolve(mat.tocsc(), rhs)
    trace, coarse = sol[:skel.size], sol[skel.size:]
    assert np.linalg.norm(trace - golden["trace"]) / np.linalg.norm(golden["trace"]) < 3e-14
    pressure = np.concatenate([r.reconstruct(trace[r.problem.trace_dofs], coarse[i:i + 1])
                               for i, r in enumerate(responses)])
    ref = golden["pressure"]
    assert np.linalg.norm(pressure - ref) / np.linalg.norm(ref) < 3e-14


def test_planted_mistake_is_caught(golden):
    """A dropped sign in the coupling (a plausible slip) must break the pin."""
    mesh = TriangleMesh.unit_square(4)
    skel = SkeletonSpace(mesh)
    contributions = []
    for c in range(len(mesh.cells)):
        fine = mesh.submesh(c, 8)
        a, m, f = scalar_operators(fine, 2, diffusion=permeability, source=source, order=6)
        b = np.abs(trace_coupling(mesh, c, fine, skel, 2))  # orientation signs lost
        z = np.ones((a.shape[0], 1))
        p = LocalProblem(a, b, f, skel.cell_dofs(c), z, m @ z, test_coupling=b)
        contributions.append(p.condense().global_contribution([skel.size + c]))
    mat, _, _ = assemble_hybrid_contributions(
        contributions, skel.size, skel.size + np.arange(len(mesh.cells) + 1))
    gv = golden["global_values"]
    same = np.array_equal(mat.indices, golden["global_indices"]) and len(mat.data) == len(gv)
    assert not same or np.linalg.norm(mat.data - gv) / np.linalg.norm(gv) > 1e-3
