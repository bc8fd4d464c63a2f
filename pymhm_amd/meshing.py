"""Macro/local meshes, nodal numbering and trace coupling for the local stage.

Mesh-only preparation done once per local shape on the host (the data the
reference providers rebuild per cell in Python):

* polygon macro topology and ear splitting: ``src/pymhm/meshes/polygonal.py:33-57,83-190``
* conforming lattice refinement of the ear triangles: ``polygonal.py:217-265``
* local triangle topology (face numbering by first encounter, orientation):
  ``src/pymhm/meshes/triangle.py:28-84``
* continuous P1/P2 numbering (vertices, then one node per face):
  ``src/pymhm/fem/scalar/triangle.py:65-92``
* trace coupling B = int_{dK} phi_i mu_j with quadrature split at fine
  vertices and skeleton breaks: ``src/pymhm/fem/scalar/triangle.py:160-205``
* skeleton numbering (face offsets, side order): ``src/pymhm/fem/traces/interval.py:119-161``

Node order and numbering must match the reference exactly (integer maps are
compared bit-for-bit in the tests).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from numpy.polynomial.legendre import leggauss

from . import elements


def _cross(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


# ----------------------------------------------------------------- macro topology
@dataclass
class MacroMesh:
    """Polygonal macro cells with signed faces (first-encounter face numbering)."""

    points: np.ndarray
    cells: list            # CCW vertex indices per cell
    faces: np.ndarray      # (F, 2) as stored by the first adjacent cell
    cell_faces: list       # per cell, face index per side
    signs: list            # per cell, +1 first owner / -1 second

    @property
    def boundary_faces(self):
        owners = np.zeros(len(self.faces), dtype=np.int64)
        for cf in self.cell_faces:
            np.add.at(owners, cf, 1)
        return np.flatnonzero(owners == 1)


def macro_topology(points, polygons) -> MacroMesh:
    points = np.asarray(points, dtype=float)
    cells, faces, cell_faces, signs = [], [], [], []
    owner = {}
    for poly in polygons:
        idx = np.asarray(poly, dtype=np.int64)
        xy = points[idx]
        if np.sum(_cross(xy - xy[0], np.roll(xy, -1, axis=0) - xy[0])) < 0:
            idx = idx[::-1].copy()
        cf, sg = [], []
        for a, b in zip(idx.tolist(), np.roll(idx, -1).tolist()):
            key = (a, b) if a < b else (b, a)
            face = owner.get(key)
            if face is None:
                owner[key] = face = len(faces)
                faces.append((a, b))
                sg.append(1)
            else:
                if faces[face] == (a, b):
                    raise ValueError("inconsistent macro orientation")
                sg.append(-1)
            cf.append(face)
        cells.append(idx)
        cell_faces.append(np.asarray(cf, dtype=np.int64))
        signs.append(np.asarray(sg, dtype=np.int64))
    return MacroMesh(points, cells, np.asarray(faces, dtype=np.int64), cell_faces, signs)


def square_grid(nx: int, ny: int | None = None) -> MacroMesh:
    """Unit square split into ``nx * ny`` counterclockwise square macros (row-major)."""
    ny = nx if ny is None else ny
    xs, ys = np.meshgrid(np.arange(nx + 1) / nx, np.arange(ny + 1) / ny)
    pts = np.stack((xs.ravel(), ys.ravel()), axis=1)
    a = (np.arange(ny)[:, None] * (nx + 1) + np.arange(nx)[None, :]).ravel()
    polys = np.stack((a, a + 1, a + nx + 2, a + nx + 1), axis=1)
    return macro_topology(pts, polys)


def voronoi_square(npts: int, seed: int, min_edge: float | None = None) -> MacroMesh:
    """Voronoi cells of ``npts`` seeded uniform points, bounded by the unit
    square: the seeds are mirrored across the four sides, so the bisectors
    with the mirrors are the sides themselves.  Vertices within 1e-12 of each
    other are merged and snapped onto the sides; cells are counterclockwise.

    Edges shorter than ``min_edge`` (default 0.05 / sqrt(npts), a twentieth of
    the mean seed spacing) are collapsed - a Voronoi diagram of uniform points
    has edges down to 1e-6 whose fan triangles have angles of 1e-3 degrees.
    The endpoints merge into a corner, else a side vertex, else the midpoint."""
    from scipy.spatial import Voronoi

    rng = np.random.default_rng(seed)
    pts = rng.random((npts, 2))
    mirrors = [pts, pts * (-1, 1), pts * (-1, 1) + (2, 0), pts * (1, -1), pts * (1, -1) + (0, 2)]
    vor = Voronoi(np.concatenate(mirrors))
    verts = vor.vertices.copy()
    verts[np.abs(verts) < 1e-12] = 0.0
    verts[np.abs(verts - 1.0) < 1e-12] = 1.0
    key = np.round(verts / 1e-12).astype(np.int64)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    polys = []
    for i in range(npts):
        region = vor.regions[vor.point_region[i]]
        if -1 in region or not region:
            raise ValueError("unbounded Voronoi region inside the mirrored set")
        idx = [int(inverse[v]) for v in region]
        idx = [v for k, v in enumerate(idx) if v != idx[k - 1]]
        polys.append(idx)
    points = verts[first]
    tol = 0.05 / np.sqrt(npts) if min_edge is None else float(min_edge)
    while True:  # collapse short edges until none is left
        parent = np.arange(len(points))

        def find(v):
            while parent[v] != v:
                parent[v] = parent[parent[v]]
                v = parent[v]
            return v

        short = False
        for poly in polys:
            for u, v in zip(poly, poly[1:] + poly[:1]):
                ru, rv = find(u), find(v)
                if ru != rv and np.linalg.norm(points[ru] - points[rv]) < tol:
                    on = [np.sum((points[w] == 0.0) | (points[w] == 1.0)) for w in (ru, rv)]
                    keep, drop = (ru, rv) if on[0] >= on[1] else (rv, ru)
                    if on[0] == on[1] == 0:
                        points[keep] = 0.5 * (points[ru] + points[rv])
                    parent[drop] = keep
                    short = True
        if not short:
            break
        merged = []
        for poly in polys:
            idx = [find(v) for v in poly]
            idx = [v for k, v in enumerate(idx) if v != idx[k - 1]]
            if len(idx) < 3:
                raise ValueError("edge collapse removed a cell")
            merged.append(idx)
        polys = merged
    used = np.unique(np.concatenate(polys))
    renum = -np.ones(len(points), dtype=np.int64)
    renum[used] = np.arange(len(used))
    return macro_topology(points[used], [renum[p] for p in polys])


# ----------------------------------------------------------------- local meshes
def ear_split(xy):
    """Ear clipping of a simple counterclockwise polygon (first valid ear, repeatedly)."""
    rest = list(range(len(xy)))
    out = []
    span = float(np.max(np.ptp(xy, axis=0)))
    tol = 64 * np.finfo(float).eps * span * span
    while len(rest) > 3:
        for j in range(len(rest)):
            a, b, c = rest[j - 1], rest[j], rest[(j + 1) % len(rest)]
            if _cross(xy[b] - xy[a], xy[c] - xy[a]) <= tol:
                continue
            tri = xy[[a, b, c]]
            others = np.asarray([v for v in rest if v not in (a, b, c)], dtype=np.int64)
            if len(others):
                edges = np.roll(tri, -1, axis=0) - tri
                rel = xy[others][:, None, :] - tri[None, :, :]
                inside = np.all(_cross(edges[None, :, :], rel) >= -tol, axis=1)
                if inside.any():
                    continue
            out.append((a, b, c))
            del rest[j]
            break
        else:
            raise ValueError("polygon has no resolvable ear")
    out.append(tuple(rest))
    return np.asarray(out, dtype=np.int64)


def refine_triangles(vertices, triangles, r: int):
    """Conforming ``r``-lattice refinement, nodes shared through topological keys."""
    keymap = {}
    pts = []
    cells = []
    for tri in triangles:
        tri = [int(v) for v in tri]
        ids = {}
        for i in range(r + 1):
            for j in range(r + 1 - i):
                w = (r - i - j, i, j)
                key = tuple(sorted((tri[t], w[t]) for t in range(3) if w[t]))
                node = keymap.get(key)
                if node is None:
                    keymap[key] = node = len(pts)
                    pts.append(np.dot(np.asarray(w, dtype=float), vertices[tri]) / r)
                ids[i, j] = node
        for i in range(r):
            for j in range(r - i):
                cells.append((ids[i, j], ids[i + 1, j], ids[i, j + 1]))
                if i + j < r - 1:
                    cells.append((ids[i + 1, j], ids[i + 1, j + 1], ids[i, j + 1]))
    return np.asarray(pts), np.asarray(cells, dtype=np.int64)


def triangle_topology(cells):
    """Faces by first encounter over (cell, side), side s = (v_s, v_{s+1})."""
    faces, owner = [], {}
    cell_faces = np.empty_like(cells)
    count = []
    for c, tri in enumerate(cells.tolist()):
        for s in range(3):
            a, b = tri[s], tri[(s + 1) % 3]
            key = (a, b) if a < b else (b, a)
            f = owner.get(key)
            if f is None:
                owner[key] = f = len(faces)
                faces.append((a, b))
                count.append(0)
            count[f] += 1
            cell_faces[c, s] = f
    return np.asarray(faces, dtype=np.int64), cell_faces, np.asarray(count)


@dataclass
class LocalMesh:
    """Local triangulation of one macro with its continuous nodal numbering."""

    points: np.ndarray       # vertex coordinates (np, 2)
    cells: np.ndarray        # (ne, 3)
    faces: np.ndarray        # (nf, 2)
    cell_faces: np.ndarray   # (ne, 3)
    boundary: np.ndarray     # boundary face indices
    degree: int
    dofs: np.ndarray         # (ne, nb) local DOF ids
    nodes: np.ndarray        # (n, 2) DOF coordinates

    @property
    def n(self) -> int:
        return len(self.nodes)


def local_mesh(points, cells, degree: int) -> LocalMesh:
    pts = np.asarray(points, dtype=float)
    cells = np.asarray(cells, dtype=np.int64).copy()
    v = pts[cells]
    det = _cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    flip = det < 0
    cells[flip] = cells[flip][:, [0, 2, 1]]
    faces, cf, count = triangle_topology(cells)
    nodes = [pts]
    if degree == 1:
        dofs = cells.copy()
    elif degree == 2:
        dofs = np.empty((len(cells), 6), dtype=np.int64)
        dofs[:, :3] = cells
        # one midpoint node per face, numbered after the vertices in face order
        dofs[:, 3:] = len(pts) + cf
        nodes.append(0.5 * (pts[faces[:, 0]] + pts[faces[:, 1]]))
    else:
        raise ValueError("degree must be 1 or 2")
    return LocalMesh(pts, cells, faces, cf, np.flatnonzero(count == 1), degree, dofs,
                     np.concatenate(nodes))


# ----------------------------------------------------------------- trace coupling
@dataclass(frozen=True)
class FaceSpace:
    """Discontinuous Legendre segments on a macroface parameterised by [0, 1]."""

    breaks: tuple = (0.0, 1.0)
    degrees: tuple = (0,)

    @classmethod
    def uniform(cls, degree: int = 0, segments: int = 1) -> "FaceSpace":
        return cls(tuple(np.linspace(0.0, 1.0, segments + 1).tolist()), (int(degree),) * segments)

    @property
    def size(self) -> int:
        return sum(p + 1 for p in self.degrees)

    def evaluate(self, t):
        t = np.asarray(t, dtype=float)
        out = np.zeros((len(t), self.size))
        col = 0
        last = len(self.degrees) - 1
        for s, p in enumerate(self.degrees):
            lo, hi = self.breaks[s], self.breaks[s + 1]
            sel = (t >= lo) & ((t < hi) | ((s == last) & (t <= hi)))
            out[sel, col:col + p + 1] = elements.legendre_modes(2.0 * (t[sel] - lo) / (hi - lo) - 1.0, p)
            col += p + 1
        return out

    def quadrature(self, order: int = 4):
        """Gauss points and weights per segment on [0, 1] (``fem/traces/interval.py:92-118``):
        the reference rule renormalised to mass two in extended precision, then
        mapped onto every segment; the weights sum to 1 over the face."""
        x, w = leggauss(int(order))
        w = np.asarray(w, dtype=np.longdouble)
        w *= np.longdouble(2) / np.sum(w)
        pts, wts = [], []
        for lo, hi in zip(self.breaks[:-1], self.breaks[1:]):
            pts.extend(lo + (x + 1) * (hi - lo) / 2)
            wts.extend(w * (hi - lo) / 2)
        return np.asarray(pts), np.asarray(wts, dtype=float)

    def reversal(self):
        """Signed column map of the same space under t -> 1 - t (uniform breaks only)."""
        b = np.asarray(self.breaks)
        if not np.allclose(b, 1.0 - b[::-1], rtol=0, atol=1e-15) or len(set(self.degrees)) != 1:
            raise ValueError("orientation reversal needs symmetric breaks and equal degrees")
        p = self.degrees[0]
        nseg = len(self.degrees)
        cmap = np.empty(self.size, dtype=np.int64)
        scale = np.empty(self.size)
        for s in range(nseg):
            for mode in range(p + 1):
                cmap[s * (p + 1) + mode] = (nseg - 1 - s) * (p + 1) + mode
                scale[s * (p + 1) + mode] = -1.0 if mode % 2 else 1.0
        return cmap, scale


def side_coupling(mesh: LocalMesh, start, end, space: FaceSpace, sign: float = 1.0):
    """``B[i, j] = sign * int_side phi_i mu_j`` with the side parameterised start -> end."""
    start = np.asarray(start, dtype=float)
    tangent = np.asarray(end, dtype=float) - start
    length = float(np.sqrt(tangent @ tangent))
    k = mesh.degree
    gx, gw = leggauss(max(space.degrees) + k + 1)
    out = np.zeros((mesh.n, space.size))
    for f in mesh.boundary:
        a, b = mesh.faces[f]
        xy = mesh.points[[a, b]]
        t = (xy - start) @ tangent / (length * length)
        if (not np.allclose(xy, start + t[:, None] * tangent, atol=1e-12, rtol=1e-12)
                or t.min() < -1e-12 or t.max() > 1 + 1e-12):
            continue
        lo, hi = np.clip(np.sort(t), 0.0, 1.0)
        cuts = sorted({lo, hi, *(x for x in space.breaks if lo < x < hi)})
        rows = [a, b] + ([len(mesh.points) + f] if k == 2 else [])
        for left, right in zip(cuts[:-1], cuts[1:]):
            par = left + (gx + 1.0) * (right - left) / 2.0
            w = gw * (right - left) / 2.0 * length * sign
            local = elements.interval_tables(k, (par - t[0]) / (t[1] - t[0]))
            out[rows, :] += local.T @ (w[:, None] * space.evaluate(par))
    return out


def skeleton_tables(mesh: LocalMesh, verts, space: FaceSpace):
    """Mesh-only description of a macro's skeleton coupling for the device
    kernel (``mhm_layout_set_skeleton``).

    Side ``s`` runs counterclockwise from ``verts[s]`` to ``verts[s+1]`` with
    fraction ``f`` in [0, 1].  Every fine boundary edge on it (endpoints at
    fractions ``e0 < e1``) is cut at the face segment breaks into pieces
    ``[f0, f1]``; a piece contributes its edge's nodal functions (0: the node at
    ``e0``, 1: the node at ``e1``, 2: the P2 midpoint) to the coupling entry
    (node, side, segment containing the piece).  Returns the piece table, the
    entries in sorted order with their contribution lists (local DOF ids), the
    face space's breaks and a Gauss-Legendre rule exact for the products.
    """
    verts = np.asarray(verts, dtype=float)
    k = len(verts)
    b = np.asarray(space.breaks, dtype=float)
    if len(set(space.degrees)) != 1:
        raise ValueError("the device skeleton needs one Legendre degree for every segment")
    pieces, contribs = [], {}
    for s in range(k):
        start, end = verts[s], verts[(s + 1) % k]
        tangent = end - start
        length = float(np.sqrt(tangent @ tangent))
        for f in mesh.boundary:
            a, c = mesh.faces[f]
            xy = mesh.points[[a, c]]
            t = (xy - start) @ tangent / (length * length)
            if (not np.allclose(xy, start + t[:, None] * tangent, atol=1e-12, rtol=1e-12)
                    or t.min() < -1e-12 or t.max() > 1 + 1e-12):
                continue
            nodes = [a, c] if t[0] < t[1] else [c, a]
            if mesh.degree == 2:
                nodes.append(len(mesh.points) + f)
            e0, e1 = float(np.clip(t.min(), 0.0, 1.0)), float(np.clip(t.max(), 0.0, 1.0))
            cuts = sorted({e0, e1, *(x for x in b if e0 < x < e1)})
            for f0, f1 in zip(cuts[:-1], cuts[1:]):
                seg = int(np.searchsorted(b, 0.5 * (f0 + f1), side="right") - 1)
                seg = min(max(seg, 0), len(b) - 2)
                pid = len(pieces)
                pieces.append((f0, f1, e0, e1))
                for basis, node in enumerate(nodes):
                    contribs.setdefault((int(node), s, seg), []).append(pid * 3 + basis)
    keys = sorted(contribs)
    ptr = np.zeros(len(keys) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(contribs[key]) for key in keys])
    ng = space.degrees[0] + mesh.degree + 1
    gx, gw = leggauss(ng)
    return dict(
        nsides=k, nseg=len(b) - 1, degree=int(space.degrees[0]), edeg=int(mesh.degree), breaks=b, gx=gx, gw=gw,
        piece_f=np.asarray(pieces, dtype=float).reshape(-1, 4),
        entry_node=np.array([key[0] for key in keys], dtype=np.int64),
        entry_side=np.array([key[1] for key in keys], dtype=np.int64),
        entry_seg=np.array([key[2] for key in keys], dtype=np.int64),
        entry_ptr=ptr, contrib=np.concatenate([contribs[key] for key in keys]).astype(np.int64) if keys
        else np.zeros(0, dtype=np.int64))


def local_mass_ones(mesh: LocalMesh):
    """``M 1`` (integral of each basis function), exact for P1/P2 on triangles."""
    _, meas = elements.barycentric_gradients(mesh.points, mesh.cells)
    out = np.zeros(mesh.n)
    if mesh.degree == 1:
        np.add.at(out, mesh.dofs.ravel(), np.repeat(meas / 3.0, 3))
    else:
        # P2 vertex functions integrate to 0, edge functions to |T|/3
        np.add.at(out, mesh.dofs[:, 3:].ravel(), np.repeat(meas / 3.0, 3))
    return out


def quadrature_points(mesh: LocalMesh, order: int, offset=(0.0, 0.0)):
    """Physical quadrature points ``(ne, q, 2)`` of the element rule used in assembly."""
    bary, _ = elements.triangle_rule(order)
    return np.einsum("qi,tij->tqj", bary, mesh.points[mesh.cells]) + np.asarray(offset)


# ----------------------------------------------------------------- tetrahedra
_PERMS3 = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
_TET_PAIRS = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


def freudenthal_cube(r: int, lo, size: float):
    """Points and positively oriented cells of ``r^3`` lattice cubes, six
    Freudenthal tetrahedra each (``meshes/tetrahedron.py:95-108``: corners in C
    order, axis permutations in lexicographic order, cells with a negative
    determinant swap local vertices 1 and 2)."""
    r = int(r)
    shape = (r + 1,) * 3
    pts = np.asarray(lo, dtype=float) + np.indices(shape).reshape(3, -1).T / r * float(size)
    cells = []
    for corner in np.ndindex((r,) * 3):
        for perm in _PERMS3:
            node = list(corner)
            path = [np.ravel_multi_index(tuple(node), shape)]
            for axis in perm:
                node[axis] += 1
                path.append(np.ravel_multi_index(tuple(node), shape))
            cells.append(path)
    cells = np.asarray(cells, dtype=np.int64)
    det = np.linalg.det(pts[cells[:, 1:]] - pts[cells[:, :1]])
    neg = det < 0
    cells[neg, 1], cells[neg, 2] = cells[neg, 2].copy(), cells[neg, 1].copy()
    return pts, cells


@dataclass(frozen=True)
class TetMesh:
    points: np.ndarray
    cells: np.ndarray
    degree: int
    dofs: np.ndarray
    nodes: np.ndarray

    @property
    def n(self) -> int:
        return len(self.nodes)


def tet_local_mesh(points, cells, degree: int) -> TetMesh:
    """P1/P2 numbering: vertices, then edges in sorted-pair order
    (``fem/scalar/tetrahedron.py:54-66``)."""
    pts = np.asarray(points, dtype=float)
    cells = np.asarray(cells, dtype=np.int64)
    if degree == 1:
        return TetMesh(pts, cells, 1, cells.copy(), pts.copy())
    if degree != 2:
        raise ValueError("degree must be 1 or 2")
    pairs = np.sort(cells[:, _TET_PAIRS].reshape(-1, 2), axis=1)
    edges, inverse = np.unique(pairs, axis=0, return_inverse=True)
    dofs = np.column_stack((cells, len(pts) + inverse.reshape(-1, 6)))
    return TetMesh(pts, cells, 2, dofs, np.vstack((pts, pts[edges].mean(axis=1))))


def tet_mass_ones(mesh: TetMesh):
    """``M 1`` (integral of every basis function): P1 |T|/4; P2 vertex
    functions -|T|/20, edge functions |T|/5."""
    _, vol = elements.tetra_gradients(mesh.points, mesh.cells)
    out = np.zeros(mesh.n)
    if mesh.degree == 1:
        np.add.at(out, mesh.dofs.ravel(), np.repeat(vol / 4.0, 4))
    else:
        np.add.at(out, mesh.dofs[:, :4].ravel(), np.repeat(-vol / 20.0, 4))
        np.add.at(out, mesh.dofs[:, 4:].ravel(), np.repeat(vol / 5.0, 6))
    return out


def cube_coupling(mesh: TetMesh, lo, size: float, split: int):
    """Coupling ``B (n, 6 split^2)`` of piecewise-constant multipliers on a
    ``split x split`` grid of every face of the cube ``[lo, lo + size]^3``.

    Faces -x, +x, -y, +y, -z, +z; a face cell's index is ``i + split * j``
    along the face's lower then higher remaining axis.  On a boundary triangle
    the integral of a P2 vertex function is 0 and of an edge function |T|/3
    (P1: |T|/3 per vertex).
    """
    lo = np.asarray(lo, dtype=float)
    size = float(size)
    nb = mesh.dofs.shape[1]
    B = np.zeros((mesh.n, 6 * split * split))
    tol = 1e-12 * size
    for c, (cell, dofs) in enumerate(zip(mesh.cells, mesh.dofs)):
        for side in range(4):
            loc = [k for k in range(4) if k != side]
            p = mesh.points[cell[loc]]
            for axis in range(3):
                for s, bound in ((0, lo[axis]), (1, lo[axis] + size)):
                    if not np.all(np.abs(p[:, axis] - bound) <= tol):
                        continue
                    others = [a for a in range(3) if a != axis]
                    centre = p.mean(axis=0)
                    ij = [min(split - 1, int((centre[a] - lo[a]) / size * split)) for a in others]
                    col = (2 * axis + s) * split * split + ij[0] + split * ij[1]
                    area = 0.5 * np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))
                    if nb == 4:
                        B[cell[loc], col] += area / 3.0
                    else:
                        for e, (i, j) in enumerate(_TET_PAIRS):
                            if i != side and j != side:
                                B[dofs[4 + e], col] += area / 3.0
    return B


def tet_quadrature_points(mesh: TetMesh, order: int, offset=(0.0, 0.0, 0.0)):
    bary, _ = elements.tetra_rule(order)
    return np.einsum("qi,tij->tqj", bary, mesh.points[mesh.cells]) + np.asarray(offset)
