"""Macro and local meshes of the oracle (test infrastructure only).

Restates the topology conventions of the reference meshes so that local DOF
numbering, face numbering and orientation signs match it exactly:

* ``TriangleMesh``: ``src/pymhm/meshes/triangle.py:12-147`` (face numbering by
  first encounter, sign +1 for the first adjacent cell, clockwise cells
  reoriented, ``unit_square`` southwest diagonals, ``submesh`` lattice order).
* ``PolygonMesh``: ``src/pymhm/meshes/polygonal.py:33-57`` (ear clipping) and
  ``:83-265`` (signed macroface incidence, conforming refinement of the ear
  triangulation keyed by topological barycentric weights).
* ``TetraMesh``: ``src/pymhm/meshes/tetrahedron.py`` (see that class).
"""

from __future__ import annotations

import numpy as np


def _cross(a, b):
    return a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]


class TriangleMesh:
    """Oriented triangles with signed cell-to-face incidence.

    Follows ``meshes/triangle.py:28-84``: faces are numbered in order of first
    encounter over ``(cell, side)``, side ``s`` joins local vertices ``s`` and
    ``s+1``; the first cell owning a face stores it as ``(i, j)`` with sign +1.
    """

    def __init__(self, points, cells):
        points = np.array(points, dtype=float, copy=True)
        cells = np.array(cells, dtype=np.int64, copy=True)
        if points.ndim != 2 or points.shape[1] != 2 or cells.ndim != 2 or cells.shape[1] != 3:
            raise ValueError("points (n,2) and cells (m,3) required")
        v = points[cells]
        a, b = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
        det = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
        if np.any(np.abs(det) <= np.finfo(float).eps * np.linalg.norm(a, axis=1)
                  * np.linalg.norm(b, axis=1)):
            raise ValueError("degenerate triangle")
        rev = det < 0
        cells[rev] = cells[rev][:, [0, 2, 1]]
        lookup = {}
        faces, neighbors = [], []
        cell_faces = np.empty_like(cells)
        signs = np.empty_like(cells)
        for c, nodes in enumerate(cells):
            for side in range(3):
                i, j = int(nodes[side]), int(nodes[(side + 1) % 3])
                key = (min(i, j), max(i, j))
                if key not in lookup:
                    lookup[key] = len(faces)
                    faces.append((i, j))
                    neighbors.append([c, -1])
                    s = 1
                else:
                    f = lookup[key]
                    if neighbors[f][1] != -1 or faces[f] == (i, j):
                        raise ValueError("nonmanifold face")
                    neighbors[f][1] = c
                    s = -1
                cell_faces[c, side] = lookup[key]
                signs[c, side] = s
        self.points = points
        self.cells = cells
        self.faces = np.array(faces, dtype=np.int64).reshape(-1, 2)
        self.cell_faces = cell_faces
        self.signs = signs
        self.face_cells = np.array(neighbors, dtype=np.int64).reshape(-1, 2)

    @classmethod
    def unit_square(cls, nx: int, ny: int | None = None) -> "TriangleMesh":
        """``meshes/triangle.py:86-97``: southwest-diagonal split of a Cartesian grid."""
        ny = nx if ny is None else ny
        pts = np.array([(x / nx, y / ny) for y in range(ny + 1) for x in range(nx + 1)])
        cells = []
        for y in range(ny):
            for x in range(nx):
                a = y * (nx + 1) + x
                cells.extend(((a, a + 1, a + nx + 2), (a, a + nx + 2, a + nx + 1)))
        return cls(pts, np.array(cells, dtype=np.int64))

    @property
    def boundary_faces(self):
        return np.flatnonzero(self.face_cells[:, 1] == -1)

    @property
    def lengths(self):
        return np.linalg.norm(self.points[self.faces[:, 1]] - self.points[self.faces[:, 0]], axis=1)

    @property
    def areas(self):
        v = self.points[self.cells]
        a, b = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
        return (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]) / 2

    def submesh(self, cell: int, subdivisions: int) -> "TriangleMesh":
        """``meshes/triangle.py:122-147``: uniform ``r*r`` split, lattice order (i, j)."""
        n = int(subdivisions)
        v = self.points[self.cells[cell]]
        keys = [(i, j) for i in range(n + 1) for j in range(n + 1 - i)]
        index = {k: t for t, k in enumerate(keys)}
        pts = np.array([v[0] + i / n * (v[1] - v[0]) + j / n * (v[2] - v[0]) for i, j in keys])
        cells = []
        for i in range(n):
            for j in range(n - i):
                cells.append((index[i, j], index[i + 1, j], index[i, j + 1]))
                if i + j < n - 1:
                    cells.append((index[i + 1, j], index[i + 1, j + 1], index[i, j + 1]))
        return TriangleMesh(pts, np.array(cells, dtype=np.int64))


def ear_triangulation(xy):
    """Ear clipping of one counterclockwise simple polygon (``polygonal.py:33-57``)."""
    remaining = list(range(len(xy)))
    tris = []
    scale = float(np.max(np.ptp(xy, axis=0)))
    tol = 64 * np.finfo(float).eps * scale * scale
    while len(remaining) > 3:
        for j, b in enumerate(remaining):
            a, c = remaining[j - 1], remaining[(j + 1) % len(remaining)]
            if _cross(xy[b] - xy[a], xy[c] - xy[a]) <= tol:
                continue
            other = [v for v in remaining if v not in (a, b, c)]
            tri = xy[[a, b, c]]
            diff = xy[other, None, :] - tri[None, :, :]
            sg = _cross(np.roll(tri, -1, axis=0) - tri, diff)
            if np.any(np.all(sg >= -tol, axis=1)):
                continue
            tris.append((a, b, c))
            remaining.pop(j)
            break
        else:
            raise ValueError("polygon has no resolvable ear")
    tris.append((remaining[0], remaining[1], remaining[2]))
    return np.asarray(tris, dtype=np.int64)


class PolygonMesh:
    """Polygonal macro cells with signed faces and ear-based local meshes.

    Topology follows ``meshes/polygonal.py:83-190``; ``submesh`` follows
    ``:217-265`` for the default ``'ears'`` local triangulation.
    """

    def __init__(self, points, cells):
        points = np.array(points, dtype=float, copy=True)
        out_cells, tris, faces, neighbors, cell_faces, signs = [], [], [], [], [], []
        lookup = {}
        for c, raw in enumerate(cells):
            idx = np.asarray(raw, dtype=np.int64).copy()
            xy = points[idx]
            area = float(np.sum(_cross(xy - xy[0], np.roll(xy, -1, axis=0) - xy[0])) / 2)
            if area < 0:
                idx = idx[::-1].copy()
                xy = points[idx]
            tris.append(ear_triangulation(xy))
            lf, ls = [], []
            for a, b in zip(idx, np.roll(idx, -1)):
                key = (min(int(a), int(b)), max(int(a), int(b)))
                if key not in lookup:
                    lookup[key] = len(faces)
                    faces.append((int(a), int(b)))
                    neighbors.append([c, -1])
                    s = 1
                else:
                    f = lookup[key]
                    if neighbors[f][1] != -1 or faces[f] == (a, b):
                        raise ValueError("nonmanifold macroface")
                    neighbors[f][1] = c
                    s = -1
                lf.append(lookup[key])
                ls.append(s)
            out_cells.append(idx)
            cell_faces.append(np.asarray(lf, dtype=np.int64))
            signs.append(np.asarray(ls, dtype=np.int64))
        self.points = points
        self.cells = tuple(out_cells)
        self.faces = np.asarray(faces, dtype=np.int64)
        self.cell_faces = tuple(cell_faces)
        self.signs = tuple(signs)
        self.face_cells = np.asarray(neighbors, dtype=np.int64)
        self._triangles = tuple(tris)

    @classmethod
    def unit_square_grid(cls, nx: int, ny: int | None = None) -> "PolygonMesh":
        """Cartesian grid of counterclockwise squares, row-major from the origin."""
        ny = nx if ny is None else ny
        pts = np.array([(x / nx, y / ny) for y in range(ny + 1) for x in range(nx + 1)])
        cells = []
        for y in range(ny):
            for x in range(nx):
                a = y * (nx + 1) + x
                cells.append((a, a + 1, a + nx + 2, a + nx + 1))
        return cls(pts, cells)

    @property
    def boundary_faces(self):
        return np.flatnonzero(self.face_cells[:, 1] < 0)

    @property
    def lengths(self):
        return np.linalg.norm(self.points[self.faces[:, 1]] - self.points[self.faces[:, 0]], axis=1)

    def submesh(self, cell: int, subdivisions: int) -> TriangleMesh:
        """Conforming refinement of the ear triangulation (``polygonal.py:217-265``)."""
        r = int(subdivisions)
        verts = self.points[self.cells[cell]]
        lookup, pts, tris = {}, [], []
        for tri in self._triangles[cell]:
            local = {}
            for i in range(r + 1):
                for j in range(r + 1 - i):
                    w = (r - i - j, i, j)
                    key = tuple(sorted((int(v), ww) for v, ww in zip(tri, w) if ww))
                    if key not in lookup:
                        lookup[key] = len(pts)
                        pts.append(np.asarray(w) @ verts[tri] / r)
                    local[i, j] = lookup[key]
            for i in range(r):
                for j in range(r - i):
                    tris.append((local[i, j], local[i + 1, j], local[i, j + 1]))
                    if i + j < r - 1:
                        tris.append((local[i + 1, j], local[i + 1, j + 1], local[i, j + 1]))
        return TriangleMesh(np.asarray(pts), np.asarray(tris, dtype=np.int64))


# ---------------------------------------------------------------- tetrahedra
_TET_PAIRS = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))
_PERMUTATIONS_3 = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


class TetraMesh:
    """Straight tetrahedra with signed triangular face incidence.

    Follows ``meshes/tetrahedron.py:31-92``: cells with a negative vertex
    determinant swap local vertices 1 and 2; face ``s`` of a cell is opposite
    local vertex ``s``, faces are numbered by first encounter over
    ``(cell, side)`` with the sorted vertex triple as key; a face normal points
    out of its first adjacent cell, which gets sign +1 (the second gets -1).
    """

    def __init__(self, points, cells):
        points = np.array(points, dtype=float, copy=True)
        cells = np.array(cells, dtype=np.int64, copy=True)
        if points.ndim != 2 or points.shape[1] != 3 or cells.ndim != 2 or cells.shape[1] != 4:
            raise ValueError("points (n,3) and cells (m,4) required")
        edges = points[cells[:, 1:]] - points[cells[:, :1]]
        det = np.linalg.det(edges)
        scale = np.prod(np.linalg.norm(edges, axis=2), axis=1)
        if np.any(np.abs(det) <= 32 * np.finfo(float).eps * scale):
            raise ValueError("degenerate tetrahedron")
        neg = det < 0
        cells[neg, 1], cells[neg, 2] = cells[neg, 2].copy(), cells[neg, 1].copy()
        self.points, self.cells = points, cells
        self.volumes = np.abs(det) / 6
        lookup, faces, owners = {}, [], []
        self.cell_faces = np.empty((len(cells), 4), dtype=np.int64)
        for c, ids in enumerate(cells):
            for side in range(4):
                key = tuple(sorted(int(v) for k, v in enumerate(ids) if k != side))
                f = lookup.get(key)
                if f is None:
                    f = lookup[key] = len(faces)
                    faces.append(key)
                    owners.append([])
                owners[f].append(c)
                self.cell_faces[c, side] = f
        self.faces = np.asarray(faces, dtype=np.int64)
        self.face_cells = np.array([o + [-1] * (2 - len(o)) for o in owners], dtype=np.int64)
        fp = points[self.faces]
        cross = np.cross(fp[:, 1] - fp[:, 0], fp[:, 2] - fp[:, 0])
        self.areas = np.linalg.norm(cross, axis=1) / 2
        self.signs = np.empty((len(cells), 4), dtype=np.int64)
        for f, adjacent in enumerate(owners):
            for index, c in enumerate(adjacent):
                self.signs[c, int(np.flatnonzero(self.cell_faces[c] == f)[0])] = 1 - 2 * index
        self.boundary_faces = np.flatnonzero(self.face_cells[:, 1] < 0)

    @classmethod
    def unit_cube(cls, n: int) -> "TetraMesh":
        """Six Freudenthal tetrahedra per lattice cube (``tetrahedron.py:95-108``):
        for every cube corner (C order) and every axis permutation, the path
        corner -> +e_p0 -> +e_p1 -> +e_p2."""
        n = int(n)
        shape = (n + 1,) * 3
        points = np.indices(shape).reshape(3, -1).T / n
        cells = []
        for corner in np.ndindex((n,) * 3):
            for perm in _PERMUTATIONS_3:
                node = list(corner)
                path = [np.ravel_multi_index(tuple(node), shape)]
                for axis in perm:
                    node[axis] += 1
                    path.append(np.ravel_multi_index(tuple(node), shape))
                cells.append(path)
        return cls(points, np.asarray(cells))

    def submesh(self, cell: int, refinement: int) -> "TetraMesh":
        """Red refinement of one cell into refinement**3 tetrahedra
        (``tetrahedron.py:111-153``), mapped affinely from barycentric points."""
        bary, cells = red_refinement(int(refinement))
        return TetraMesh(bary @ self.points[self.cells[cell]], cells)


def red_refinement(refinement: int):
    """Barycentric points and cells of ``log2(refinement)`` red-refinement
    rounds of the reference tetrahedron (``tetrahedron.py:121-153``): edge
    midpoints are numbered by first encounter over each cell's vertex pairs
    (0,1),(0,2),(0,3),(1,2),(1,3),(2,3); every cell becomes four corner cells
    and four octahedron cells split along the ab-cd diagonal."""
    if refinement < 1 or refinement & (refinement - 1):
        raise ValueError("tetrahedral refinement must be a power of two")
    points = [row for row in np.eye(4)]
    cells = [[0, 1, 2, 3]]
    for _ in range(refinement.bit_length() - 1):
        mid, out = {}, []
        for a, b, c, d in cells:
            ids = (a, b, c, d)
            m = []
            for i, j in _TET_PAIRS:
                key = tuple(sorted((ids[i], ids[j])))
                if key not in mid:
                    mid[key] = len(points)
                    points.append((points[ids[i]] + points[ids[j]]) / 2)
                m.append(mid[key])
            ab, ac, ad, bc, bd, cd = m
            out += [[a, ab, ac, ad], [ab, b, bc, bd], [ac, bc, c, cd], [ad, bd, cd, d],
                    [ab, ac, ad, cd], [ab, ac, bc, cd], [ab, ad, bd, cd], [ab, bc, bd, cd]]
        cells = out
    return np.asarray(points), np.asarray(cells, dtype=np.int64)
