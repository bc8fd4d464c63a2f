"""Reference-element tables consumed by the assembly kernel.

Quadrature points/weights and Lagrange values with barycentric derivatives in
the reference node order (vertices, then oriented edges; for P2 simplices
there are no interior nodes):

* triangle Duffy-Gauss rule: ``src/pymhm/fem/scalar/operators.py:63-69``
* triangle node order: ``src/pymhm/fem/scalar/triangle.py:14-24``
* tetrahedron rule and node order: ``src/pymhm/fem/scalar/tetrahedron.py``

The device kernel forms physical gradients as
``grad phi_b = sum_i dphi_b/dlambda_i grad lambda_i`` per element.
"""

from __future__ import annotations

import numpy as np
from numpy.polynomial.legendre import leggauss

TRIANGLE_EDGES = ((0, 1), (1, 2), (2, 0))


def triangle_rule(order: int):
    """Collapsed (Duffy) product Gauss rule: barycentric points ``(q,3)``, unit-sum weights."""
    x, w = leggauss(int(order))
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    a = np.repeat(x, len(x))
    b = np.tile(x, len(x))
    bary = np.stack((1.0 - a - (1.0 - a) * b, a, (1.0 - a) * b), axis=1)
    weights = 2.0 * np.repeat(w, len(w)) * np.tile(w, len(w)) * (1.0 - a)
    return bary, weights


def triangle_tables(degree: int, bary):
    """P1/P2 values ``(q, nb)`` and ``dphi/dlambda`` ``(q, nb, 3)``."""
    lam = np.asarray(bary, dtype=float)
    q = len(lam)
    if degree == 1:
        return lam.copy(), np.broadcast_to(np.eye(3), (q, 3, 3)).copy()
    if degree != 2:
        raise ValueError("triangle tables support degree 1 or 2")
    vals = np.empty((q, 6))
    der = np.zeros((q, 6, 3))
    for i in range(3):
        vals[:, i] = lam[:, i] * (2.0 * lam[:, i] - 1.0)
        der[:, i, i] = 4.0 * lam[:, i] - 1.0
    for e, (a, b) in enumerate(TRIANGLE_EDGES):
        vals[:, 3 + e] = 4.0 * lam[:, a] * lam[:, b]
        der[:, 3 + e, a] = 4.0 * lam[:, b]
        der[:, 3 + e, b] = 4.0 * lam[:, a]
    return vals, der


def interval_tables(degree: int, s):
    """1D Lagrange values at ``s`` on nodes ``[0, 1, interior...]``."""
    s = np.asarray(s, dtype=float)
    if degree == 1:
        return np.stack((1.0 - s, s), axis=1)
    if degree == 2:
        return np.stack((2.0 * (s - 1.0) * (s - 0.5), 2.0 * s * (s - 0.5), 4.0 * s * (1.0 - s)), axis=1)
    raise ValueError("interval tables support degree 1 or 2")


def legendre_modes(x, degree: int):
    """Unnormalised Legendre polynomials ``P_0..P_degree`` at ``x`` in [-1, 1]."""
    x = np.asarray(x, dtype=float)
    out = [np.ones_like(x)]
    if degree >= 1:
        out.append(x.copy())
    for k in range(1, degree):
        out.append(((2 * k + 1) * x * out[k] - k * out[k - 1]) / (k + 1))
    return np.stack(out[:degree + 1], axis=1)


def barycentric_gradients(points, cells):
    """Physical gradients of the barycentric coordinates ``(ne, d+1, d)`` and measures."""
    v = np.asarray(points, dtype=float)[cells]
    d = v.shape[2]
    jac = (v[:, 1:, :] - v[:, :1, :]).transpose(0, 2, 1)  # columns are edge vectors
    inv = np.linalg.inv(jac)  # rows: gradients of lambda_1..lambda_d
    grads = np.empty((len(v), d + 1, d))
    grads[:, 1:, :] = inv
    grads[:, 0, :] = -inv.sum(axis=1)
    det = np.linalg.det(jac)
    measure = np.abs(det) / (2.0 if d == 2 else 6.0)
    return grads, measure
