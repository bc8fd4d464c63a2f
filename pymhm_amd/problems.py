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
from . import _lib
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


class PolygonProblem:
    """Primal MHM scalar diffusion on a polygonal macro mesh with uneven macros.

    Every macro is ear-split (``meshes/polygonal.py:33-57``) and each ear
    refined into ``r^2`` triangles (``:217-265``), with ``r`` chosen per macro;
    the multipliers are ``face_space`` on every macroface.  Macros whose ear
    triangulation and refinement agree have the same local topology, so they
    form a **bucket** that shares one layout (pattern, band ordering, assembly
    plan) and one batch; the geometry (barycentric gradients, element
    measures), the constraints ``M 1 / 1^T M 1`` and the couplings are per
    macro and computed on the device: the constraints from the element
    measures, the couplings from the bucket's skeleton tables (its boundary
    edges at fixed fractions of the sides - the same in every macro of the
    bucket) with each macro's side lengths and macroface orientations
    (``fem/scalar/triangle.py:160-205``).
    """

    def __init__(self, macro: meshing.MacroMesh, refinements, degree: int = 1,
                 face_space: meshing.FaceSpace = meshing.FaceSpace.uniform(1, 1), order: int = 4, device: int = 0):
        self.macro = macro
        self.degree = int(degree)
        self.space = face_space
        self.order = max(int(order), self.degree + 2)
        self.M = len(macro.cells)
        self.refinements = np.asarray(refinements, dtype=np.int64)
        if self.refinements.shape != (self.M,) or self.refinements.min() < 1:
            raise ValueError("one refinement >= 1 per macro")
        fs = face_space.size
        self.trace_size = len(macro.faces) * fs
        groups = {}
        for c, cell in enumerate(macro.cells):
            tris = meshing.ear_split(macro.points[cell])
            key = (len(cell), int(self.refinements[c]), tuple(tris.ravel().tolist()))
            groups.setdefault(key, []).append(c)
        self.qbary, self.qw = elements.triangle_rule(self.order)
        self.vals, self.der = elements.triangle_tables(self.degree, self.qbary)
        self.buckets = [PolygonBucket(self, groups[key], key, device) for key in sorted(groups)]
        self.bucket_of = np.empty(self.M, dtype=np.int64)
        self.slot_of = np.empty(self.M, dtype=np.int64)
        for bi, bk in enumerate(self.buckets):
            self.bucket_of[bk.cells] = bi
            self.slot_of[bk.cells] = np.arange(len(bk.cells))

    def quadrature_points(self):
        """Per macro ``(ne_c, nq, 2)`` physical quadrature points (a list)."""
        out = [None] * self.M
        for bk in self.buckets:
            for r, c in enumerate(bk.cells):
                out[c] = bk.qpts[r]
        return out

    def global_indices(self):
        """Per macro ``[trace_dofs, coarse_dof]`` (a list: widths differ)."""
        out = [None] * self.M
        for bk in self.buckets:
            for r, c in enumerate(bk.cells):
                out[c] = np.concatenate((bk.trace_dofs[r], [self.trace_size + c])).astype(np.int64)
        return out

    def kernel_offsets(self):
        return self.trace_size + np.arange(self.M + 1, dtype=np.int64)


class PolygonBucket:
    """Macros of one local topology: a shared layout, per-macro geometry."""

    def __init__(self, prob: PolygonProblem, cells, key, device: int = 0):
        k, r, flat = key
        tris = np.asarray(flat, dtype=np.int64).reshape(-1, 3)
        macro, fs = prob.macro, prob.space.size
        self.cells = np.asarray(cells, dtype=np.int64)
        self.k, self.r = k, r
        Mb = len(cells)
        verts0 = macro.points[macro.cells[cells[0]]]
        pts0, tri_cells = meshing.refine_triangles(verts0, tris, r)
        mesh = meshing.local_mesh(pts0, tri_cells, prob.degree)
        self.mesh = mesh
        n, ne, m = mesh.n, len(mesh.cells), k * fs
        self.n, self.m = n, m
        lengths0 = np.linalg.norm(np.roll(verts0, -1, axis=0) - verts0, axis=1)
        template = np.concatenate(
            [meshing.side_coupling(mesh, verts0[s], verts0[(s + 1) % k], prob.space) / lengths0[s]
             for s in range(k)], axis=1)
        nq = len(prob.qw)
        self.glam = np.empty((Mb, ne, 3, 2))
        self.meas = np.empty((Mb, ne))
        self.qpts = np.empty((Mb, ne, nq, 2))
        self.side_len = np.empty((Mb, k))
        self.side_sign = np.empty((Mb, k))
        self.trace_dofs = np.empty((Mb, m), dtype=np.int64)
        for i, c in enumerate(cells):
            verts = macro.points[macro.cells[c]]
            pts = pts0 if i == 0 else meshing.refine_triangles(verts, tris, r)[0]
            g, ms = elements.barycentric_gradients(pts, mesh.cells)
            if np.any(ms <= 0):
                raise ValueError(f"macro {c}: inverted element in its ear refinement")
            self.glam[i], self.meas[i] = g, ms
            self.qpts[i] = np.einsum("qi,tij->tqj", prob.qbary, pts[mesh.cells])
            self.side_len[i] = np.linalg.norm(np.roll(verts, -1, axis=0) - verts, axis=1)
            self.side_sign[i] = macro.signs[c]
            for s in range(k):
                self.trace_dofs[i, s * fs:(s + 1) * fs] = macro.cell_faces[c][s] * fs + np.arange(fs)
        indptr, indices = element_pattern(n, mesh.dofs)
        tangent = verts0[1] - verts0[0]
        dist = np.abs(meshing._cross(tangent, mesh.nodes - verts0[0]))
        start = np.flatnonzero(dist <= 1e-12 * float(np.sqrt(tangent @ tangent)))
        perm = choose_permutation(indptr, indices, start, mesh.nodes - verts0[0], tangent)
        mass1 = meshing.local_mass_ones(mesh)
        self.layout = LocalLayout(
            n, perm, indptr, indices, m=m, k=1, coupling=template, constraint=mass1 / mass1.sum(),
            assembly=dict(dofs=mesh.dofs, dim=2, qweights=prob.qw, dtable=prob.der, vtable=prob.vals,
                          grad_lambda=self.glam[0], measure=self.meas[0]),
            device=device)
        self.layout.set_skeleton(meshing.skeleton_tables(mesh, verts0, prob.space))

    def batch(self, precision: str = "fp64", max_columns: int = 0):
        """A batch of the bucket's macros with their geometry, couplings and
        constraints set on the device."""
        b = LocalBatch(self.layout, len(self.cells), precision=precision, max_columns=max_columns)
        b.set_geometry(self.glam, self.meas)
        b.skeleton_coupling(self.side_len, self.side_sign)
        b.mass_constraints()
        return b


class BucketedStage:
    """The local stage over every bucket of a ``PolygonProblem``.

    Buckets are independent batches; a bucket of a few dozen macros leaves most
    of the GPU idle while its band solves walk their rows, so the buckets are
    issued largest first, round-robin over ``streams`` HIP streams (default
    priority), and joined before the global gather.  Streams beyond the HIP
    runtime's hardware queues share queues and run in order, so ``streams``
    defaults to ``GPU_MAX_HW_QUEUES`` (HIP's default 4); on 2048 uneven macros
    16 queues and streams ran the step 1.6x faster than 4, more than 20
    slower again.  The batches condense on their own bucket's stream only:
    a batch's side stream would share a queue with other buckets and hold up
    what is queued behind its wait.
    """

    def __init__(self, prob: PolygonProblem, precision: str = "fp64", streams: int | None = None,
                 max_columns: int = 0):
        import os

        import torch

        if streams is None:
            streams = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))

        self.prob = prob
        work = [len(bk.cells) * bk.n * (bk.layout.bandwidth + 1) for bk in prob.buckets]
        self.order = sorted(range(len(prob.buckets)), key=lambda i: -work[i])
        self.batches = {i: prob.buckets[i].batch(precision, max_columns=max_columns) for i in self.order}
        for b in self.batches.values():  # the buckets' own streams overlap them; no side streams
            b.set_side_stream(False)
        self.streams = [torch.cuda.Stream() for _ in range(max(1, int(streams)))]

    def tune(self, kqs, fqs=None, f_const: float = 0.0, repeats: int = 3):
        """Choose every bucket's band solve kernels with ``tuning.tune_batch``
        (assembled once, then factor and condense timed per candidate);
        returns ``{bucket: (first, refine, seconds saved alone)}``."""
        from . import tuning

        out = {}
        for j, i in enumerate(self.order):
            b = self.batches[i]
            b.assemble(kqs[i], None if fqs is None else fqs[i], f_const)

            def run(b=b):
                b.factor(check=False)
                b.condense(check=False)

            first, refine, timings = tuning.tune_batch(b, run, repeats=repeats)
            auto = timings.get((_lib.SOLVE_AUTO, _lib.SOLVE_AUTO))
            best = timings.get((first, refine))
            out[i] = (first, refine, (auto - best) if auto is not None and best is not None else 0.0)
        return out

    def positions(self):
        """Block and load start of every macro in the flat buffers of ``run``."""
        prob = self.prob
        bpos = np.zeros(prob.M, dtype=np.int64)
        lpos = np.zeros(prob.M, dtype=np.int64)
        ob = ol = 0
        for i in self.order:
            bk = prob.buckets[i]
            w = bk.m + 1
            bpos[bk.cells] = ob + np.arange(len(bk.cells)) * w * w
            lpos[bk.cells] = ol + np.arange(len(bk.cells)) * w
            ob += len(bk.cells) * w * w
            ol += len(bk.cells) * w
        return bpos, lpos

    def run(self, kqs, fqs=None, f_const: float = 0.0, check: bool = True):
        """Assemble, factor and condense every bucket (``kqs[i]``: the bucket's
        ``(macros, elements, q)`` coefficients); returns the flat blocks and
        loads in bucket order on the current stream."""
        import torch

        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        bl, ld = [], []
        for j, i in enumerate(self.order):
            b = self.batches[i]
            with torch.cuda.stream(self.streams[j % len(self.streams)]):
                b.assemble(kqs[i], None if fqs is None else fqs[i], f_const)
                b.factor(check=False)
                x, y = b.condense(check=False)
                bl.append(x.reshape(-1))
                ld.append(y.reshape(-1))
        for s in self.streams:
            main.wait_stream(s)
        if check:
            for i in self.order:
                self.batches[i].raise_on_failure(f"bucket {i}")
        return torch.cat(bl), torch.cat(ld)

    def close(self):
        for b in self.batches.values():
            b.close()
        self.batches = {}
