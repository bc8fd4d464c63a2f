"""Tetrahedral P1/P2 operators and cube-face P0 coupling (oracle only).

Restates, with closed-form Lagrange polynomials instead of Basix:

* ``tetrahedron_quadrature`` ``src/pymhm/fem/scalar/tetrahedron.py:41-51`` (Duffy
  product of Gauss rules, weights normalised to unit volume)
* ``tetra_multiindices``     ``src/pymhm/fem/scalar/tetrahedron_topology.py:27-45``
  (vertices, then edge interiors over ``combinations(range(4), 2)``)
* ``tetra_nodal_space``      ``src/pymhm/fem/scalar/tetrahedron.py:54-66``
* ``tetra_gradients``        ``src/pymhm/meshes/tetrahedron.py:156-166``
* ``tetra_operators``        ``src/pymhm/fem/scalar/tetrahedron.py:157-189``
  (compensated diffusion blocks, consistent mass, ``np.add.at`` load)

``cube_face_coupling`` is the coupling ``B[i, a] = int_face phi_i mu_a`` of
``trace_coupling`` (``fem/scalar/triangle.py:160-205``) for the workload with
cube macros and piecewise-constant multipliers on a ``s x s`` split of every
cube face; the reference's three-dimensional skeleton
(``fem/traces/triangle_3d.py``) uses tetrahedral macros, so this face layout
is the workload's own.

Pinned by: the reference's golden tetrahedral local meshes (bit-exact, through
``oracle.meshes.TetraMesh``), exact monomial integrals of the quadrature, the
Lagrange property and closed-form P1/P2 integrals (``tests/test_oracle_3d.py``).
"""

from __future__ import annotations

from itertools import combinations

import numpy as np
from numpy.polynomial.legendre import leggauss
from scipy import sparse

from .fem import diffusion_blocks, simplex_lagrange, triangle_quadrature

TET_EDGES = tuple(combinations(range(4), 2))


def tetrahedron_quadrature(order: int = 5):
    """Barycentric points ``(q, 4)`` and unit-sum weights of the Duffy rule."""
    x, w = leggauss(int(order))
    x, w = (x + 1) / 2, w / 2
    a, b, c = np.meshgrid(x, x, x, indexing="ij")
    wa, wb, wc = np.meshgrid(w, w, w, indexing="ij")
    xyz = np.array([a, (1 - a) * b, (1 - a) * (1 - b) * c]).reshape(3, -1).T
    bary = np.column_stack((1 - xyz.sum(axis=1), xyz))
    weights = (6 * wa * wb * wc * (1 - a) ** 2 * (1 - b)).ravel()
    return bary, weights


def tetra_multiindices(degree: int):
    """Barycentric node multi-indices for degree 1 or 2 in the reference order."""
    k = int(degree)
    idx = []
    for v in range(4):
        e = [0] * 4
        e[v] = k
        idx.append(e)
    if k == 2:
        for i, j in TET_EDGES:
            e = [0] * 4
            e[i] = e[j] = 1
            idx.append(e)
    elif k != 1:
        raise ValueError("degree 1 or 2")
    return np.asarray(idx, dtype=np.int64)


def tetra_nodal_space(mesh, degree: int):
    """Continuous P1/P2 numbering: mesh vertices, then edges in sorted-pair order."""
    if degree == 1:
        return mesh.cells.copy(), mesh.points.copy()
    pairs = np.sort(mesh.cells[:, TET_EDGES].reshape(-1, 2), axis=1)
    edges, inverse = np.unique(pairs, axis=0, return_inverse=True)
    dofs = np.column_stack((mesh.cells, len(mesh.points) + inverse.reshape(-1, 6)))
    points = np.vstack((mesh.points, mesh.points[edges].mean(axis=1)))
    return dofs, points


def tetra_gradients(mesh):
    """Physical gradients of lambda_0..lambda_3 ``(cells, 4, 3)``: rows of the
    inverse vertex Jacobian, lambda_0 the negative sum."""
    inv = np.linalg.inv((mesh.points[mesh.cells[:, 1:]] - mesh.points[mesh.cells[:, :1]]).transpose(0, 2, 1))
    return np.concatenate((-inv.sum(axis=1)[:, None], inv), axis=1)


def tetra_element_tabulate(mesh, degree: int, bary):
    """Values ``(q, nb)`` and physical gradients ``(cells, q, nb, 3)``."""
    vals, dl = simplex_lagrange(degree, bary, tetra_multiindices(degree))
    glam = tetra_gradients(mesh)
    grads = np.zeros((len(mesh.cells), len(bary), vals.shape[1], 3))
    for i in range(4):  # sum over barycentric directions in order 0..3
        grads = grads + dl[None, :, :, i, None] * glam[:, None, None, i, :]
    return vals, grads


def tetra_operators(mesh, degree: int, *, diffusion_q=None, source_q=None, source=0.0, order: int = 4):
    """Continuous Pk stiffness (compensated Gram), consistent mass and load.

    ``diffusion_q``/``source_q`` give scalar values at the ``(cells, q)``
    quadrature points of ``tetrahedron_quadrature(max(order, degree + 2))``.
    The load of element t, basis i is ``sum_q ((w_q phi_i(q)) f_tq) |T|`` in
    quadrature order.  Returns CSC stiffness, CSC mass and the load.
    """
    bary, w = tetrahedron_quadrature(max(order, degree + 2))
    vals, grads = tetra_element_tabulate(mesh, degree, bary)
    dofs, nodes = tetra_nodal_space(mesh, degree)
    ncell, nq = len(mesh.cells), len(w)
    kq = np.ones((ncell, nq)) if diffusion_q is None else np.asarray(diffusion_q, dtype=float)
    fq = np.full((ncell, nq), float(source)) if source_q is None else np.asarray(source_q, dtype=float)
    tensors = kq[:, :, None, None] * np.eye(3)[None, None]
    stiff = diffusion_blocks(np.broadcast_to(w, (ncell, nq)), grads, tensors, mesh.volumes)
    mass = np.einsum("t,q,qi,qj->tij", mesh.volumes, w, vals, vals)
    load_e = np.zeros((ncell, vals.shape[1]))
    for q in range(nq):
        load_e = load_e + ((w[q] * vals[q])[None, :] * fq[:, q, None]) * mesh.volumes[:, None]
    size = len(nodes)
    rows = np.broadcast_to(dofs[:, :, None], stiff.shape).ravel()
    cols = np.broadcast_to(dofs[:, None, :], stiff.shape).ravel()
    A = sparse.coo_matrix((stiff.ravel(), (rows, cols)), shape=(size, size)).tocsc()
    Mm = sparse.coo_matrix((mass.ravel(), (rows, cols)), shape=(size, size)).tocsc()
    load = np.zeros(size)
    np.add.at(load, dofs, load_e)
    return A, Mm, load


def tetra_quadrature_points(mesh, degree: int = 2, order: int = 4):
    bary, _ = tetrahedron_quadrature(max(order, degree + 2))
    return np.einsum("qi,tij->tqj", bary, mesh.points[mesh.cells])


def cube_face_coupling(mesh, degree: int, lo, hi, split: int = 2, order: int = 4):
    """``B (n, 6 split^2)``: integrals of the tetrahedral basis over the cube's
    boundary triangles, one column per face cell.

    Faces in the order -x, +x, -y, +y, -z, +z; on a face the ``split x split``
    cells are numbered along the lower remaining axis first, then the higher
    (e.g. on an x face: index = iy + split * iz).  Boundary triangles are
    integrated with the Duffy triangle rule, the tetrahedral basis evaluated
    at the face points (the opposite barycentric coordinate is zero).
    """
    lo, hi = np.asarray(lo, dtype=float), np.asarray(hi, dtype=float)
    dofs, _ = tetra_nodal_space(mesh, degree)
    alphas = tetra_multiindices(degree)
    tb, tw = triangle_quadrature(order)
    n = dofs.max() + 1
    B = np.zeros((n, 6 * split * split))
    for f in mesh.boundary_faces:
        cell = int(mesh.face_cells[f, 0])
        side = int(np.flatnonzero(mesh.cell_faces[cell] == f)[0])
        pts = mesh.points[mesh.faces[f]]
        centre = pts.mean(axis=0)
        axis = None
        for ax in range(3):
            for s, bound in ((0, lo[ax]), (1, hi[ax])):
                if np.all(np.abs(pts[:, ax] - bound) <= 1e-12 * (hi[ax] - lo[ax])):
                    axis, sidx = ax, s
        if axis is None:
            raise ValueError("boundary triangle not on a cube face")
        others = [a for a in range(3) if a != axis]
        cell_ij = [min(split - 1, int((centre[a] - lo[a]) / (hi[a] - lo[a]) * split)) for a in others]
        col = (2 * axis + sidx) * split * split + cell_ij[0] + split * cell_ij[1]
        # barycentric coordinates of the face points in the tetrahedron
        bary4 = np.zeros((len(tb), 4))
        loc = [k for k in range(4) if k != side]
        for r, kk in enumerate(loc):
            # triangle vertex r of the face in the order of the cell's local vertices
            bary4[:, kk] = tb[:, r]
        vals, _ = simplex_lagrange(degree, bary4, alphas)
        area = mesh.areas[f]
        contrib = (vals * tw[:, None]).sum(axis=0) * area
        np.add.at(B[:, col], dofs[cell], contrib)
    return B
