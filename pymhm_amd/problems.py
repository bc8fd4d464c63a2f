"""Local-stage problem descriptions (the provider side of the drop-in).

``SquareGridProblem`` is primal MHM scalar diffusion ``-div(K grad u) = f``
on a Cartesian grid of square macros, each ear-split and refined into
``2 r^2`` P1/P2 triangles, with discontinuous Legendre multipliers on every
macroface.  It mirrors the primal Darcy provider
``src/pymhm/_legacy/models/darcy/primal.py:44-78`` (Z = 1, constraints
``M Z / Z^T M Z``, coupling from ``trace_coupling``) but prepares the
mesh-only data once per local shape: all macros of the grid are translates
of one template, and a macroface shared with a neighbour enters with its
orientation as a signed column map of the template coupling.
"""

from __future__ import annotations

import numpy as np

from . import elements, meshing
from .layout import LocalLayout, choose_permutation, element_pattern
from .stage import LocalBatch


class SquareGridProblem:
    def __init__(self, nx: int, refinement: int, degree: int, face_space: meshing.FaceSpace,
                 order: int = 4, device: int = 0):
        self.nx, self.refinement, self.degree = int(nx), int(refinement), int(degree)
        self.space = face_space
        self.order = max(int(order), self.degree + 2)
        self.macro = meshing.square_grid(self.nx)
        self.M = len(self.macro.cells)
        verts = self.macro.points[self.macro.cells[0]]
        self.origin = verts[0].copy()
        pts, cells = meshing.refine_triangles(verts, meshing.ear_split(verts), self.refinement)
        self.mesh = meshing.local_mesh(pts, cells, self.degree)
        mesh = self.mesh
        n = mesh.n
        nsides = 4
        fs = face_space.size
        self.m = nsides * fs
        coupling = np.concatenate(
            [meshing.side_coupling(mesh, verts[s], verts[(s + 1) % nsides], face_space) for s in range(nsides)],
            axis=1)
        mass1 = meshing.local_mass_ones(mesh)
        constraint = mass1 / mass1.sum()
        bary, w = elements.triangle_rule(self.order)
        vals, der = elements.triangle_tables(self.degree, bary)
        glam, meas = elements.barycentric_gradients(mesh.points, mesh.cells)
        indptr, indices = element_pattern(n, mesh.dofs)
        tangent = verts[1] - verts[0]
        dist = np.abs(meshing._cross(tangent, mesh.nodes - verts[0]))
        start = np.flatnonzero(dist <= 1e-12 * float(np.sqrt(tangent @ tangent)))
        perm = choose_permutation(indptr, indices, start, mesh.nodes - verts[0], tangent)
        self.layout = LocalLayout(
            n, perm, indptr, indices, m=self.m, k=1, coupling=coupling, constraint=constraint,
            assembly=dict(dofs=mesh.dofs, dim=2, qweights=w, dtable=der, vtable=vals,
                          grad_lambda=glam, measure=meas),
            device=device)
        self.coupling = coupling
        self.constraint = constraint
        self.qbary = bary
        # skeleton numbering and per-macro orientation maps
        nf = len(self.macro.faces)
        self.trace_size = nf * fs
        self.trace_dofs = np.stack(
            [np.concatenate([f * fs + np.arange(fs) for f in cf]) for cf in self.macro.cell_faces]
        ).astype(np.int64)
        rev_map, rev_scale = face_space.reversal()
        colmap = np.tile(np.arange(self.m), (self.M, 1))
        colscale = np.ones((self.M, self.m))
        for c, sg in enumerate(self.macro.signs):
            for s in range(nsides):
                if sg[s] < 0:
                    colmap[c, s * fs:(s + 1) * fs] = s * fs + rev_map
                    colscale[c, s * fs:(s + 1) * fs] = -rev_scale
        self.colmap, self.colscale = colmap.astype(np.int32), colscale
        self.offsets = np.stack([self.macro.points[c[0]] - self.origin for c in self.macro.cells])

    @property
    def n(self) -> int:
        return self.mesh.n

    def quadrature_points(self):
        """Physical quadrature points ``(M, ne, nq, 2)`` of every macro."""
        local = np.einsum("qi,tij->tqj", self.qbary, self.mesh.points[self.mesh.cells])
        return local[None] + self.offsets[:, None, None, :]

    def global_indices(self):
        """Per macro ``[trace_dofs, coarse_dof]`` (``core/contributions.py:70-74``)."""
        coarse = self.trace_size + np.arange(self.M, dtype=np.int64)
        return np.concatenate((self.trace_dofs, coarse[:, None]), axis=1)

    def kernel_offsets(self):
        return self.trace_size + np.arange(self.M + 1, dtype=np.int64)

    def batch(self, precision: str = "fp64", max_columns: int = 0, macros: slice | None = None):
        sl = slice(None) if macros is None else macros
        cm, cs = self.colmap[sl], self.colscale[sl]
        b = LocalBatch(self.layout, len(cm), precision=precision, max_columns=max_columns)
        b.set_coupling(colmap=cm, colscale=cs)
        return b


class CubeGridProblem:
    """Primal MHM scalar diffusion on an ``nx^3`` grid of cube macros.

    Each macro is ``r^3`` lattice cubes of six Freudenthal P1/P2 tetrahedra
    (``meshes/tetrahedron.py:95-108``); the multipliers are piecewise constant
    on a ``split x split`` grid of every macroface.  All macros are translates
    of one template (their binary-fraction coordinates make the geometry
    tables identical), a face shared by two macros enters the second one (in
    macro order) with sign -1, and global trace DOFs are numbered face by face
    in order of first encounter over (macro, face -x, +x, -y, +y, -z, +z).
    """

    def __init__(self, nx: int, refinement: int, degree: int = 2, split: int = 2, order: int = 4,
                 device: int = 0):
        self.nx, self.refinement, self.degree, self.split = int(nx), int(refinement), int(degree), int(split)
        self.order = max(int(order), self.degree + 2)
        self.M = self.nx ** 3
        self.h = 1.0 / self.nx
        pts, cells = meshing.freudenthal_cube(self.refinement, (0.0, 0.0, 0.0), self.h)
        self.mesh = mesh = meshing.tet_local_mesh(pts, cells, self.degree)
        n = mesh.n
        nfc = self.split * self.split
        self.m = 6 * nfc
        coupling = meshing.cube_coupling(mesh, (0.0, 0.0, 0.0), self.h, self.split)
        mass1 = meshing.tet_mass_ones(mesh)
        constraint = mass1 / mass1.sum()
        bary, w = elements.tetra_rule(self.order)
        vals, der = elements.tetra_tables(self.degree, bary)
        glam, vol = elements.tetra_gradients(mesh.points, mesh.cells)
        indptr, indices = element_pattern(n, mesh.dofs)
        start = np.flatnonzero(np.abs(mesh.nodes[:, 0]) <= 1e-12 * self.h)
        perm = choose_permutation(indptr, indices, start)
        self.layout = LocalLayout(
            n, perm, indptr, indices, m=self.m, k=1, coupling=coupling, constraint=constraint,
            assembly=dict(dofs=mesh.dofs, dim=3, qweights=w, dtable=der, vtable=vals,
                          grad_lambda=glam, measure=vol),
            device=device)
        self.coupling, self.constraint, self.qbary = coupling, constraint, bary
        # macro faces: global ids by first encounter, signs, trace numbering
        ijk = np.array(list(np.ndindex((self.nx,) * 3)))  # macro c = (i, j, k) in C order
        index = {tuple(t): c for c, t in enumerate(ijk)}
        face_id, signs = {}, np.ones((self.M, 6))
        self.cell_faces = np.empty((self.M, 6), dtype=np.int64)
        for c, t in enumerate(ijk):
            for axis in range(3):
                for s in (0, 1):
                    key = (axis, *(t + (np.arange(3) == axis) * s))  # the face plane and its cell
                    if key not in face_id:
                        face_id[key] = len(face_id)
                    else:
                        signs[c, 2 * axis + s] = -1.0
                    self.cell_faces[c, 2 * axis + s] = face_id[key]
        del index
        self.nfaces = len(face_id)
        self.trace_size = self.nfaces * nfc
        self.trace_dofs = (self.cell_faces[:, :, None] * nfc + np.arange(nfc)[None, None, :]).reshape(self.M, -1)
        self.colmap = np.tile(np.arange(self.m, dtype=np.int32), (self.M, 1))
        self.colscale = np.repeat(signs, nfc, axis=1)
        self.offsets = ijk * self.h

    @property
    def n(self) -> int:
        return self.mesh.n

    def quadrature_points(self):
        """Physical quadrature points ``(M, ne, nq, 3)`` of every macro."""
        local = np.einsum("qi,tij->tqj", self.qbary, self.mesh.points[self.mesh.cells])
        return local[None] + self.offsets[:, None, None, :]

    def global_indices(self):
        coarse = self.trace_size + np.arange(self.M, dtype=np.int64)
        return np.concatenate((self.trace_dofs, coarse[:, None]), axis=1)

    def kernel_offsets(self):
        return self.trace_size + np.arange(self.M + 1, dtype=np.int64)

    def batch(self, precision: str = "fp64", max_columns: int = 0, macros: slice | None = None):
        sl = slice(None) if macros is None else macros
        b = LocalBatch(self.layout, len(self.colmap[sl]), precision=precision, max_columns=max_columns)
        b.set_coupling(colmap=self.colmap[sl], colscale=self.colscale[sl])
        return b
