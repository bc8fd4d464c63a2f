"""Known answers and closed forms that pin the oracle's algebra and elements.

* ``2u + lambda = 3`` gives block [[0.5]] and load [1.5] exactly
  (reference ``tests/test_hybrid_contributions.py:12-28``).
* Condensation equals the monolithic saddle solve (reference
  ``tests/test_hybrid.py:27-41``, seed 12).
* Two-cell interval system with a Neumann kernel: condensed matrix and load
  equal a hand elimination.
* P1/P2 triangle stiffness and mass reproduce exact monomial integrals
  (``int x^p y^q = p! q! / (p+q+2)!`` on the unit right triangle), which pins
  quadrature, node order and tabulation without reusing the oracle's formula.
* Equilibrated SuperLU agrees with a dense LAPACK solve.
"""

from math import factorial

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal
from scipy import sparse

from oracle.fem import scalar_operators, nodal_space
from oracle.local import Factorization, LocalProblem, assemble_hybrid_contributions
from oracle.meshes import TriangleMesh


def test_scalar_closed_form_block_and_load():
    r = LocalProblem([[2.0]], [[1.0]], [3.0], [4]).condense()
    idx, block, load = r.global_contribution(np.array([], dtype=int))
    assert_array_equal(idx, [4])
    assert_array_equal(block, [[0.5]])
    assert_array_equal(load, [1.5])
    full = np.linalg.solve([[2.0, 1.0], [-1.0, 0.0]], [3.0, 0.0])
    lam = np.linalg.solve(block, load)
    assert_allclose(lam, full[1:], atol=1e-14)
    assert_allclose(r.reconstruct(lam, np.empty(0)), full[:1], atol=1e-14)


def test_condensation_matches_monolithic_solve():
    rng = np.random.default_rng(12)
    x = rng.normal(size=(5, 5))
    a = x.T @ x + np.eye(5)
    b = rng.normal(size=(5, 3))
    f = rng.normal(size=5)
    g = rng.normal(size=3)
    p = LocalProblem(a, b, f, np.arange(3))
    r = p.condense()
    _, block, load = r.global_contribution(np.array([], dtype=int))
    lam = np.linalg.solve(block, load - g)
    expected = np.linalg.solve(np.block([[a, b], [b.T, np.zeros((3, 3))]]), np.r_[f, g])
    assert_allclose(lam, expected[5:], atol=1e-13)
    assert_allclose(r.reconstruct(lam, np.empty(0)), expected[:5], atol=1e-13)
    assert_allclose(r.source, np.linalg.solve(a, f), atol=1e-14)


def _interval_cells():
    a = [[1.0, -1.0], [-1.0, 1.0]]
    first = LocalProblem(a, np.eye(2), [0.0, 0.0], [0, 1], [[1.0], [1.0]], [[0.5], [0.5]])
    second = LocalProblem(a, np.diag([-1.0, 1.0]), [0.0, 0.0], [1, 2], [[1.0], [1.0]], [[0.5], [0.5]])
    return first, second


def test_neumann_interval_pair_hand_elimination():
    first, second = _interval_cells()
    assert not first.correct_kernel  # A 1 == 0 exactly: kernel shortcut
    contribs = [first.condense().global_contribution([3]), second.condense().global_contribution([4])]
    mat, rhs, scale = assemble_hybrid_contributions(contribs, 3, [3, 4, 5])
    # A^+ restricted to mean-zero: L = R B with R the constrained inverse
    r = np.linalg.solve(np.block([[np.array([[1.0, -1], [-1, 1]]), np.full((2, 1), 0.5)],
                                  [np.full((1, 2), 0.5), np.zeros((1, 1))]]),
                        np.r_[np.eye(2), np.zeros((1, 2))])[:2]
    blk1 = np.block([[r, -np.ones((2, 1))],
                     [-np.ones((1, 2)), np.zeros((1, 1))]])
    d = np.diag([-1.0, 1.0])
    blk2 = np.block([[d.T @ (r @ d), -(d.T @ np.ones((2, 1)))],
                     [-(np.ones((1, 2)) @ d), np.zeros((1, 1))]])
    dense = np.zeros((5, 5))
    dense[np.ix_([0, 1, 3], [0, 1, 3])] += blk1
    dense[np.ix_([1, 2, 4], [1, 2, 4])] += blk2
    assert_allclose(mat.toarray(), dense, atol=1e-15)
    assert_array_equal(rhs, np.zeros(5))
    assert_array_equal(scale, np.zeros(5))


def _monomials(xy):
    x, y = xy[:, 0], xy[:, 1]
    return np.stack([np.ones_like(x), x, y, x * x, x * y, y * y], axis=1)


def _grad_monomials():
    # gradient of each monomial as coefficients over the same monomial basis
    g = np.zeros((6, 2, 6))
    g[1, 0, 0] = 1
    g[2, 1, 0] = 1
    g[3, 0, 1] = 2
    g[4, 0, 2] = 1
    g[4, 1, 1] = 1
    g[5, 1, 2] = 2
    return g


def _moment(p, q):
    return factorial(p) * factorial(q) / factorial(p + q + 2)


def _product_integral(c1, c2):
    powers = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2)]
    return sum(c1[i] * c2[j] * _moment(powers[i][0] + powers[j][0], powers[i][1] + powers[j][1])
               for i in range(6) for j in range(6) if c1[i] and c2[j])


@pytest.mark.parametrize("degree", [1, 2])
def test_element_matrices_reproduce_monomial_integrals(degree):
    mesh = TriangleMesh(np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]]), np.array([[0, 1, 2]]))
    a, m, f = scalar_operators(mesh, degree, diffusion=1.0, source=1.0, order=4)
    _, nodes = nodal_space(mesh, degree)
    nm = 3 if degree == 1 else 6
    u = _monomials(nodes)[:, :nm]            # nodal interpolants of the monomials
    grads = _grad_monomials()
    eye = np.eye(6)
    stiff = np.array([[sum(_product_integral(grads[i, d], grads[j, d]) for d in range(2))
                       for j in range(nm)] for i in range(nm)])
    mass = np.array([[_product_integral(eye[i], eye[j]) for j in range(nm)] for i in range(nm)])
    assert_allclose(u.T @ a.toarray() @ u, stiff, atol=2e-15)
    assert_allclose(u.T @ m.toarray() @ u, mass, atol=2e-15)
    assert_allclose(f @ u, mass[0], atol=2e-15)  # int 1 * monomial


def test_superlu_oracle_matches_dense_lapack():
    rng = np.random.default_rng(5)
    x = sparse.random(40, 40, density=0.2, random_state=6) + sparse.eye(40) * 3
    rhs = rng.normal(size=(40, 4))
    fac = Factorization(x)
    assert_allclose(fac.solve(rhs), np.linalg.solve(x.toarray(), rhs), rtol=1e-12, atol=1e-12)
