"""Triangle P1/P2 operators, skeleton spaces and trace coupling (oracle only).

Restates, with closed-form Lagrange polynomials instead of Basix:

* ``triangle_quadrature``     ``src/pymhm/fem/scalar/operators.py:63-69``
* ``p1_geometry``             ``src/pymhm/fem/scalar/operators.py:72-77``
* ``multiindices``            ``src/pymhm/fem/scalar/triangle.py:14-24``
* ``nodal_space``             ``src/pymhm/fem/scalar/triangle.py:65-92``
* ``diffusion_blocks``        ``src/pymhm/fem/scalar/_integration.py:32-88`` (Neumaier)
* ``scalar_operators``        ``src/pymhm/fem/scalar/triangle.py:208-257``
* ``trace_coupling``          ``src/pymhm/fem/scalar/triangle.py:160-205``
* ``FaceSpace``/``SkeletonSpace`` ``src/pymhm/fem/traces/interval.py:13-161``
* ``boundary_data``           ``src/pymhm/fem/scalar/operators.py:247-321`` (scalar
  Dirichlet data), with ``boundary_moments`` ``fem/scalar/_integration.py:92-108``

Basix is not installed; Lagrange values are the textbook product formula on
the equispaced barycentric lattice, evaluated in the reference node order.
Pinned by the reference golden vectors (see ``oracle/__init__.py``).
"""

from __future__ import annotations

import numpy as np
from numpy.polynomial.legendre import leggauss
from scipy import sparse


# ---------------------------------------------------------------- quadrature
def triangle_quadrature(order: int = 4):
    """Duffy product Gauss rule, barycentric points and unit-sum weights."""
    t, w = leggauss(int(order))
    t, w = (t + 1) / 2, w / 2
    bary = np.array([(1 - a - (1 - a) * b, a, (1 - a) * b) for a in t for b in t])
    weights = np.array([2 * wa * wb * (1 - a) for a, wa in zip(t, w) for wb in w])
    return bary, weights


def p1_geometry(mesh):
    """Physical gradients of the barycentric coordinates (cells,3,2) and areas."""
    v = mesh.points[mesh.cells]
    m = np.concatenate((np.ones((len(v), 3, 1)), v), axis=2)
    grads = np.linalg.inv(m)[:, 1:, :].swapaxes(1, 2)
    return grads, mesh.areas


# ---------------------------------------------------------------- Lagrange
def multiindices(degree: int):
    """Vertices, oriented edges (0-1, 1-2, 2-0), then interior lattice nodes."""
    k = int(degree)
    idx = [(k, 0, 0), (0, k, 0), (0, 0, k)]
    for a, b in ((0, 1), (1, 2), (2, 0)):
        for j in range(1, k):
            node = [0, 0, 0]
            node[a], node[b] = k - j, j
            idx.append(tuple(node))
    idx.extend((i, j, k - i - j) for i in range(1, k) for j in range(1, k - i))
    return np.asarray(idx, dtype=np.int64)


def _factor(kl, m_count):
    """Product prod_{m<m_count} (kl - m)/(m+1) and its derivative w.r.t. kl."""
    val = np.ones_like(kl)
    der = np.zeros_like(kl)
    for m in range(m_count):
        f = (kl - m) / (m + 1)
        der = der * f + val / (m + 1)
        val = val * f
    return val, der


def simplex_lagrange(degree: int, bary, alphas):
    """Equispaced Lagrange values and barycentric derivatives.

    ``phi_alpha(lambda) = prod_i prod_{m<alpha_i} (k lambda_i - m)/(m+1)``;
    returns values ``(q, nb)`` and ``dphi/dlambda_i`` ``(q, nb, d+1)``.
    """
    k = int(degree)
    bary = np.asarray(bary, dtype=float)
    q, d1 = bary.shape
    nb = len(alphas)
    vals = np.ones((q, nb))
    ders = np.zeros((q, nb, d1))
    parts = np.empty((d1, q, nb))
    dparts = np.empty((d1, q, nb))
    for i in range(d1):
        for b, alpha in enumerate(alphas):
            v, dv = _factor(k * bary[:, i], int(alpha[i]))
            parts[i, :, b] = v
            dparts[i, :, b] = dv * k
    for i in range(d1):
        vals = vals * parts[i]
    for i in range(d1):
        prod = dparts[i].copy()
        for j in range(d1):
            if j != i:
                prod = prod * parts[j]
        ders[:, :, i] = prod
    return vals, ders


def interval_lagrange(degree: int, s):
    """1D Lagrange on nodes ``[0, 1, 1/k, ..., (k-1)/k]`` at parameters ``s``."""
    k = int(degree)
    nodes = np.r_[0.0, 1.0, np.arange(1, k) / k]
    s = np.asarray(s, dtype=float)
    out = np.ones((len(s), k + 1))
    for j in range(k + 1):
        for m in range(k + 1):
            if m != j:
                out[:, j] *= (s - nodes[m]) / (nodes[j] - nodes[m])
    return out


def nodal_space(mesh, degree: int):
    """Continuous nodal numbering: mesh vertices, then face nodes, then interior."""
    k = int(degree)
    idx = multiindices(k)
    pts = [mesh.points]
    dofs = np.empty((len(mesh.cells), len(idx)), dtype=np.int64)
    dofs[:, :3] = mesh.cells
    offset = len(mesh.points)
    if k > 1:
        par = np.arange(1, k) / k
        ep = (mesh.points[mesh.faces[:, 0], None] * (1 - par[None, :, None])
              + mesh.points[mesh.faces[:, 1], None] * par[None, :, None])
        pts.append(ep.reshape(-1, 2))
        for side, (a, _b) in enumerate(((0, 1), (1, 2), (2, 0))):
            fwd = mesh.cells[:, a] == mesh.faces[mesh.cell_faces[:, side], 0]
            loc = np.where(fwd[:, None], np.arange(k - 1), np.arange(k - 2, -1, -1))
            dofs[:, 3 + side * (k - 1):3 + (side + 1) * (k - 1)] = (
                offset + mesh.cell_faces[:, side, None] * (k - 1) + loc)
        offset += len(mesh.faces) * (k - 1)
    interior = idx[3 * k:] / k
    if len(interior):
        pts.append(np.einsum("qi,tij->tqj", interior, mesh.points[mesh.cells]).reshape(-1, 2))
        dofs[:, 3 * k:] = offset + np.arange(len(mesh.cells) * len(interior)).reshape(
            len(mesh.cells), -1)
    return dofs, np.concatenate(pts)


def element_tabulate(mesh, degree: int, bary):
    """Values ``(q,nb)`` and physical gradients ``(cells,q,nb,2)`` at shared points."""
    vals, dl = simplex_lagrange(degree, bary, multiindices(degree))
    glam, _ = p1_geometry(mesh)
    grads = np.einsum("qbi,tia->tqba", dl, glam)
    return vals, grads


# ---------------------------------------------------------------- compensated Gram
def _neumaier(total, corr, value):
    combined = total + value
    big = np.abs(total) >= np.abs(value)
    corr = corr + np.where(big, (total - combined) + value, (value - combined) + total)
    return combined, corr


def diffusion_blocks(weights, gradients, tensors, measures):
    """Cellwise ``sum_q w_q |T| grad(phi_i).K_q grad(phi_j)``, Neumaier-compensated.

    Shapes ``(cells,q)``, ``(cells,q,nb,d)``, ``(cells,q,d,d)``, ``(cells,)``;
    the same three compensated sums as ``_integration.py:43-88``.
    """
    cells, nq, nb, d = gradients.shape
    # transformed[c,q,j,a] = sum_b K[c,q,a,b] g[c,q,j,b]
    tr = np.empty((cells, nq, nb, d))
    for a in range(d):
        tot = np.zeros((cells, nq, nb))
        cor = np.zeros_like(tot)
        for b in range(d):
            tot, cor = _neumaier(tot, cor, tensors[:, :, a, b][:, :, None] * gradients[:, :, :, b])
        tr[:, :, :, a] = tot + cor
    total = np.zeros((cells, nb, nb))
    corr = np.zeros_like(total)
    for q in range(nq):
        ct = np.zeros((cells, nb, nb))
        cc = np.zeros_like(ct)
        for a in range(d):
            ct, cc = _neumaier(ct, cc, gradients[:, q, :, a][:, :, None] * tr[:, q, :, a][:, None, :])
        val = (ct + cc) * weights[:, q][:, None, None] * measures[:, None, None]
        total, corr = _neumaier(total, corr, val)
    return total + corr


def _eval(field, points):
    return np.broadcast_to(np.asarray(field(points) if callable(field) else field, dtype=float),
                           (len(points),)).copy()


def scalar_operators(mesh, degree: int, *, diffusion=1.0, source=0.0, order: int = 6,
                     diffusion_q=None, source_q=None):
    """Continuous Pk stiffness, mass and load (``fem/scalar/triangle.py:208-257``).

    ``diffusion``/``source`` are scalars or callables of ``(n,2)`` points;
    ``diffusion_q``/``source_q`` optionally give values directly at the
    ``(cells, q)`` quadrature points (the synthetic configurations do this).
    Returns CSC stiffness, CSC mass and the load vector.
    """
    bary, w = triangle_quadrature(max(order, degree + 2))
    vals, grads = element_tabulate(mesh, degree, bary)
    dofs, nodes = nodal_space(mesh, degree)
    ncell, nq = len(mesh.cells), len(w)
    weights = np.broadcast_to(w, (ncell, nq))
    phys = np.einsum("qi,tij->tqj", bary, mesh.points[mesh.cells])
    flat = phys.reshape(-1, 2)
    kq = (np.asarray(diffusion_q, dtype=float).reshape(ncell, nq) if diffusion_q is not None
          else _eval(diffusion, flat).reshape(ncell, nq))
    fq = (np.asarray(source_q, dtype=float).reshape(ncell, nq) if source_q is not None
          else _eval(source, flat).reshape(ncell, nq))
    if not np.all(kq > 0):
        raise ValueError("diffusion must be positive")
    tensors = kq[:, :, None, None] * np.eye(2)
    areas = mesh.areas
    basis = np.broadcast_to(vals, (ncell, nq, vals.shape[1]))
    mass_blocks = np.einsum("tq,tqi,tqj,t->tij", weights, basis, basis, areas)
    blocks = diffusion_blocks(weights, grads, tensors, areas)
    nb = vals.shape[1]
    row = np.repeat(dofs, nb, axis=1).ravel()
    col = np.tile(dofs, (1, nb)).ravel()
    n = len(nodes)
    matrix = sparse.coo_matrix((blocks.ravel(), (row, col)), shape=(n, n)).tocsc()
    mass = sparse.coo_matrix((mass_blocks.ravel(), (row, col)), shape=(n, n)).tocsc()
    load = np.einsum("tq,tqi,tq,t->ti", weights, basis, fq, areas)
    return matrix, mass, np.bincount(dofs.ravel(), weights=load.ravel(), minlength=n)


def quadrature_points(mesh, degree: int, order: int):
    """Physical quadrature points ``(cells, q, 2)`` used by ``scalar_operators``."""
    bary, _ = triangle_quadrature(max(order, degree + 2))
    return np.einsum("qi,tij->tqj", bary, mesh.points[mesh.cells])


# ---------------------------------------------------------------- skeleton
class FaceSpace:
    """Discontinuous Legendre segments on one oriented macroface (``interval.py:13-118``)."""

    def __init__(self, breaks=(0.0, 1.0), degrees=(0,)):
        self.breaks = tuple(float(x) for x in breaks)
        self.degrees = tuple(int(p) for p in degrees)
        if len(self.breaks) != len(self.degrees) + 1:
            raise ValueError("one degree per segment")

    @classmethod
    def uniform(cls, degree: int = 0, subdivisions: int = 1) -> "FaceSpace":
        return cls(tuple(np.linspace(0, 1, subdivisions + 1)), (degree,) * subdivisions)

    @property
    def size(self) -> int:
        return sum(p + 1 for p in self.degrees)

    def evaluate(self, parameter):
        """Unnormalised Legendre modes per segment; breaks belong to the right segment."""
        t = np.atleast_1d(np.asarray(parameter, dtype=float))
        out = np.zeros((len(t), self.size))
        off = 0
        for i, p in enumerate(self.degrees):
            lo, hi = self.breaks[i:i + 2]
            last = i == len(self.degrees) - 1
            mask = (t >= lo) & ((t < hi) | (last & (t <= hi)))
            x = 2 * (t[mask] - lo) / (hi - lo) - 1
            out[mask, off:off + p + 1] = legendre(x, p)
            off += p + 1
        return out


    def quadrature(self, order: int = 4):
        """Gauss points and weights per segment on [0, 1] (``interval.py:92-118``):
        the reference rule is renormalised to mass two in extended precision,
        then mapped onto every segment."""
        x, w = leggauss(int(order))
        w = np.asarray(w, dtype=np.longdouble)
        w *= np.longdouble(2) / np.sum(w)
        pts, wts = [], []
        for lo, hi in zip(self.breaks[:-1], self.breaks[1:]):
            pts.extend(lo + (x + 1) * (hi - lo) / 2)
            wts.extend(w * (hi - lo) / 2)
        return np.asarray(pts), np.asarray(wts, dtype=float)


def legendre(x, degree: int):
    """Conventional Legendre ``L_0..L_degree`` by Bonnet's recurrence."""
    x = np.asarray(x, dtype=float)
    out = np.ones((len(x), degree + 1))
    if degree >= 1:
        out[:, 1] = x
    for n in range(1, degree):
        out[:, n + 1] = ((2 * n + 1) * x * out[:, n] - n * out[:, n - 1]) / (n + 1)
    return out


class SkeletonSpace:
    """Face offsets and per-cell trace DOF concatenation (``interval.py:119-161``)."""

    def __init__(self, mesh, faces=None):
        self.mesh = mesh
        self.faces = tuple(FaceSpace() for _ in mesh.faces) if faces is None else tuple(faces)
        self.offsets = np.r_[0, np.cumsum([f.size for f in self.faces])].astype(np.int64)

    @property
    def size(self) -> int:
        return int(self.offsets[-1])

    def dofs(self, face: int):
        return np.arange(self.offsets[face], self.offsets[face + 1])

    def cell_dofs(self, cell: int):
        return np.concatenate([self.dofs(int(f)) for f in self.mesh.cell_faces[cell]])


def trace_coupling(coarse, cell: int, fine, skeleton, degree: int):
    """Signed nodal traces against skeleton modes (``fem/scalar/triangle.py:160-205``)."""
    _, points = nodal_space(fine, degree)
    width = sum(skeleton.faces[f].size for f in coarse.cell_faces[cell])
    mat = np.zeros((len(points), width))
    off = 0
    fine_bfaces = fine.boundary_faces
    for side, face in enumerate(coarse.cell_faces[cell]):
        space = skeleton.faces[face]
        start, end = coarse.points[coarse.faces[face]]
        tangent = end - start
        length = np.linalg.norm(tangent)
        gauss, gw = leggauss(max(space.degrees) + degree + 1)
        for ff in fine_bfaces:
            nodes = fine.faces[ff]
            coords = fine.points[nodes]
            t = (coords - start) @ tangent / length ** 2
            if (not np.allclose(coords, start + t[:, None] * tangent, atol=1e-12, rtol=1e-12)
                    or min(t) < -1e-12 or max(t) > 1 + 1e-12):
                continue
            lo, hi = np.clip(np.sort(t), 0, 1)
            cuts = sorted({lo, hi, *(b for b in space.breaks if lo < b < hi)})
            ids = np.r_[nodes, len(fine.points) + ff * (degree - 1) + np.arange(degree - 1)]
            for left, right in zip(cuts[:-1], cuts[1:]):
                par = left + (gauss + 1) * (right - left) / 2
                wts = gw * (right - left) / 2 * length * coarse.signs[cell][side]
                s = (par - t[0]) / (t[1] - t[0])
                basis = interval_lagrange(degree, s)
                mat[ids, off:off + space.size] += basis.T @ (wts[:, None] * space.evaluate(par))
        off += space.size
    return mat


def boundary_moments(basis, weights, values):
    """Neumaier-compensated ``sum_q basis[q, i] weights[q] values[q]``
    (``fem/scalar/_integration.py:92-108``, one scalar component)."""
    out = np.zeros(basis.shape[1])
    for i in range(basis.shape[1]):
        total, corr = 0.0, 0.0
        for q in range(len(weights)):
            total, corr = _neumaier(total, corr, basis[q, i] * weights[q] * values[q])
        out[i] = total + corr
    return out


def boundary_data(skeleton, dirichlet, order: int = 5):
    """Weak Dirichlet trace load of a scalar skeleton (``operators.py:247-321``
    without Neumann faces): on every boundary face, ``load[dofs] = |face| *
    sum_q mu(t_q) w_q g(x_q)`` with the face's own segment quadrature of order
    ``max(order, max degree + 2)``, ``x_q`` on the face from its stored start
    to its stored end.  Returns the load over the whole skeleton."""
    mesh = skeleton.mesh
    load = np.zeros(skeleton.size)
    for face in mesh.boundary_faces:
        space = skeleton.faces[face]
        par, wts = space.quadrature(max(order, max(space.degrees) + 2))
        start, end = mesh.points[mesh.faces[face]]
        pts = start + par[:, None] * (end - start)
        vals = np.broadcast_to(np.asarray(dirichlet(pts) if callable(dirichlet) else dirichlet, dtype=float),
                               (len(pts),))
        moments = boundary_moments(space.evaluate(par), wts, vals)
        load[skeleton.dofs(int(face))] = moments * mesh.lengths[face]
    return load
