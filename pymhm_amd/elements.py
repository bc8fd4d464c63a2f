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
TETRA_EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))  # combinations(range(4), 2)


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


def lattice_nodes(degree: int, dim: int = 2):
    """Barycentric multi-indices: vertices, then edge nodes along each oriented edge.

    Triangle edges (0,1), (1,2), (2,0) as ``src/pymhm/fem/scalar/triangle.py:14-24``;
    tetrahedron edges in ``combinations(range(4), 2)`` order as
    ``src/pymhm/fem/scalar/tetrahedron_topology.py:27-45``; interior nodes
    (degree >= 3) are not needed for P1/P2.
    """
    k = int(degree)
    edges = TRIANGLE_EDGES if dim == 2 else TETRA_EDGES
    nodes = []
    for v in range(dim + 1):
        a = [0] * (dim + 1)
        a[v] = k
        nodes.append(tuple(a))
    for a_, b_ in edges:
        for j in range(1, k):
            a = [0] * (dim + 1)
            a[a_], a[b_] = k - j, j
            nodes.append(tuple(a))
    if k > 2:
        raise ValueError("tables support degree 1 or 2")
    return np.asarray(nodes, dtype=np.int64)


def lagrange_tables(degree: int, bary, nodes):
    """Equispaced Lagrange values ``(q, nb)`` and ``dphi/dlambda`` ``(q, nb, d+1)``.

    Textbook product form ``phi_a = prod_i prod_{m < a_i} (k lambda_i - m) / (m + 1)``
    with the product rule, evaluated factor by factor (the same sequence of
    operations as the oracle's tabulation, so the tables agree bit for bit).
    """
    k = int(degree)
    lam = np.asarray(bary, dtype=float)
    q, d1 = lam.shape
    nb = len(nodes)
    parts = np.empty((d1, q, nb))
    dparts = np.empty((d1, q, nb))
    for i in range(d1):
        kl = k * lam[:, i]
        for b, a in enumerate(nodes):
            val = np.ones(q)
            der = np.zeros(q)
            for m in range(int(a[i])):
                f = (kl - m) / (m + 1)
                der = der * f + val / (m + 1)
                val = val * f
            parts[i, :, b] = val
            dparts[i, :, b] = der * k
    vals = np.ones((q, nb))
    for i in range(d1):
        vals = vals * parts[i]
    ders = np.empty((q, nb, d1))
    for i in range(d1):
        prod = dparts[i].copy()
        for j in range(d1):
            if j != i:
                prod = prod * parts[j]
        ders[:, :, i] = prod
    return vals, ders


def triangle_tables(degree: int, bary):
    """P1/P2 triangle values ``(q, nb)`` and ``dphi/dlambda`` ``(q, nb, 3)``."""
    if degree not in (1, 2):
        raise ValueError("triangle tables support degree 1 or 2")
    return lagrange_tables(degree, bary, lattice_nodes(degree, 2))


def tetra_rule(order: int):
    """Duffy product Gauss rule on the tetrahedron (``fem/scalar/tetrahedron.py:41-51``):
    barycentric points ``(q, 4)`` and unit-sum weights."""
    x, w = leggauss(int(order))
    x, w = (x + 1) / 2, w / 2
    a, b, c = np.meshgrid(x, x, x, indexing="ij")
    wa, wb, wc = np.meshgrid(w, w, w, indexing="ij")
    xyz = np.array([a, (1 - a) * b, (1 - a) * (1 - b) * c]).reshape(3, -1).T
    bary = np.column_stack((1 - xyz.sum(axis=1), xyz))
    weights = (6 * wa * wb * wc * (1 - a) ** 2 * (1 - b)).ravel()
    return bary, weights


def tetra_tables(degree: int, bary):
    """P1/P2 tetrahedron values ``(q, nb)`` and ``dphi/dlambda`` ``(q, nb, 4)``."""
    if degree not in (1, 2):
        raise ValueError("tetrahedron tables support degree 1 or 2")
    return lagrange_tables(degree, bary, lattice_nodes(degree, 3))


def tetra_gradients(points, cells):
    """Gradients of lambda_0..lambda_3 ``(ne, 4, 3)`` and volumes: rows of the
    inverse vertex Jacobian with lambda_0 the negative sum, volume |det|/6
    (``meshes/tetrahedron.py:156-166,44-52``), so affine geometry enters
    assembly with the reference's rounding."""
    v = np.asarray(points, dtype=float)[cells]
    edges = v[:, 1:] - v[:, :1]
    inv = np.linalg.inv(edges.transpose(0, 2, 1))
    grads = np.concatenate((-inv.sum(axis=1)[:, None], inv), axis=1)
    return grads, np.abs(np.linalg.det(edges)) / 6


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
    """Physical gradients of the barycentric coordinates ``(ne, d+1, d)`` and measures.

    Gradients are rows 1.. of ``inv([1 x])`` per simplex and triangle areas the
    halved cross product, the operations of ``fem/scalar/operators.py:72-77`` and
    ``meshes/triangle.py:115-120``, so affine geometry enters assembly with the
    reference's rounding.
    """
    v = np.asarray(points, dtype=float)[cells]
    d = v.shape[2]
    m = np.concatenate((np.ones(v.shape[:2] + (1,)), v), axis=2)
    grads = np.linalg.inv(m)[:, 1:, :].swapaxes(1, 2)
    if d == 2:
        a, b = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
        measure = (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]) / 2
    else:
        measure = np.abs(np.linalg.det(v[:, 1:] - v[:, :1])) / 6.0
    return grads, measure
