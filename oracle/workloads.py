"""Oracle evaluation of the synthetic workloads (test infrastructure only).

Builds every macro's ``LocalProblem`` the way the reference primal Darcy
provider does (``src/pymhm/_legacy/models/darcy/primal.py:55-78``: local
mesh from the macro polygon, ``scalar_operators``, ``trace_coupling``,
Z = 1, C = M Z / Z^T M Z) from the same coefficient values at quadrature
points as the GPU run, and condenses it with the FP64 SuperLU restatement.
"""

from __future__ import annotations

import numpy as np

from .fem import FaceSpace, SkeletonSpace, scalar_operators, trace_coupling
from .fem3d import cube_face_coupling, tetra_operators
from .local import LocalProblem
from .meshes import PolygonMesh, TetraMesh


def build_cube_problems(case, macros=None):
    """LocalProblems of a cube-macro workload (tetrahedral P1/P2, piecewise
    constant multipliers on ``split x split`` cells of every macroface).

    Macro c = (i, j, k) in C order spans ``[ijk h, ijk h + h]``, its local
    mesh is ``TetraMesh.unit_cube(r)`` scaled and translated; macrofaces are
    numbered by first encounter over (macro, -x, +x, -y, +y, -z, +z), the
    second macro meeting a face couples to it with sign -1.  Returns
    ``([(c, LocalProblem, mesh)], trace_size)``.
    """
    prob = case.problem
    nx, r, split, degree = prob.nx, prob.refinement, prob.split, prob.degree
    h = 1.0 / nx
    unit = TetraMesh.unit_cube(r)
    ijk = list(np.ndindex((nx,) * 3))
    faces, cell_faces, signs = {}, [], []
    for t in ijk:
        ids, sg = [], []
        for axis in range(3):
            for s in (0, 1):
                key = (axis,) + tuple(t[a] + (s if a == axis else 0) for a in range(3))
                sg.append(-1.0 if key in faces else 1.0)
                faces.setdefault(key, len(faces))
                ids.append(faces[key])
        cell_faces.append(ids)
        signs.append(sg)
    nfc = split * split
    trace_size = len(faces) * nfc
    out = []
    for c in (range(len(ijk)) if macros is None else macros):
        lo = np.asarray(ijk[c], dtype=float) * h
        mesh = TetraMesh(lo + unit.points * h, unit.cells)
        fq = None if case.fq is None else case.fq[c]
        A, M, f = tetra_operators(mesh, degree, diffusion_q=case.kq[c], source_q=fq,
                                  source=case.f_const, order=prob.order)
        B = cube_face_coupling(mesh, degree, lo, lo + h, split) * np.repeat(signs[c], nfc)[None, :]
        Z = np.ones((A.shape[0], 1))
        mz = M @ Z
        tdofs = np.concatenate([fid * nfc + np.arange(nfc) for fid in cell_faces[c]])
        out.append((c, LocalProblem(A, B, f, tdofs, Z, mz / (Z.T @ mz).item()), mesh))
    return out, trace_size


def build_problems(case, macros=None):
    prob = case.problem
    pm = PolygonMesh.unit_square_grid(prob.nx)
    sp = prob.space
    space = FaceSpace(sp.breaks, sp.degrees)
    skel = SkeletonSpace(pm, [space] * len(pm.faces))
    cells = range(len(pm.cells)) if macros is None else macros
    out = []
    for c in cells:
        fine = pm.submesh(c, prob.refinement)
        fq = None if case.fq is None else case.fq[c]
        A, M, f = scalar_operators(fine, prob.degree, diffusion_q=case.kq[c], source_q=fq,
                                   source=case.f_const, order=prob.order)
        B = trace_coupling(pm, c, fine, skel, prob.degree)
        Z = np.ones((A.shape[0], 1))
        mz = M @ Z
        C = mz / (Z.T @ mz).item()
        out.append((c, LocalProblem(A, B, f, skel.cell_dofs(c), Z, C), fine))
    return out, skel


def build_polygon_problems(case, macros=None):
    """LocalProblems of a polygonal macro workload with per-macro refinement:
    ``PolygonMesh`` of the macro cells (ear triangulation, first-encounter
    faces and signs, ``meshes/polygonal.py:83-190``), ``submesh(c, r_c)``
    (``:217-265``), ``scalar_operators`` and ``trace_coupling`` as the primal
    Darcy provider (``_legacy/models/darcy/primal.py:55-78``)."""
    prob = case.problem
    cached = getattr(prob, "_oracle_skeleton", None)
    if cached is None:  # mesh-only, built once per problem
        pm = PolygonMesh(prob.macro.points, [list(cell) for cell in prob.macro.cells])
        sp = prob.space
        cached = (pm, SkeletonSpace(pm, [FaceSpace(sp.breaks, sp.degrees)] * len(pm.faces)))
        prob._oracle_skeleton = cached
    pm, skel = cached
    out = []
    for c in (range(len(pm.cells)) if macros is None else macros):
        fine = pm.submesh(c, int(prob.refinements[c]))
        fq = None if case.fq is None else case.fq[c]
        A, M, f = scalar_operators(fine, prob.degree, diffusion_q=case.kq[c], source_q=fq,
                                   source=case.f_const, order=prob.order)
        B = trace_coupling(pm, c, fine, skel, prob.degree)
        Z = np.ones((A.shape[0], 1))
        mz = M @ Z
        out.append((c, LocalProblem(A, B, f, skel.cell_dofs(c), Z, mz / (Z.T @ mz).item()), fine))
    return out, skel


def _stack(rows):
    """An array when every row has one shape, else the list of rows."""
    return np.array(rows) if len({np.shape(r) for r in rows}) <= 1 else list(rows)


def oracle_config(case, macros=None):
    """Blocks, loads and responses of the selected macros (default: all)."""
    if hasattr(case.problem, "split"):
        problems, trace_size = build_cube_problems(case, macros)
    elif hasattr(case.problem, "buckets"):
        problems, skel = build_polygon_problems(case, macros)
        trace_size = skel.size
    else:
        problems, skel = build_problems(case, macros)
        trace_size = skel.size
    blocks, loads, src, lifts, ret, corr = [], [], [], [], [], []
    for c, p, _ in problems:
        r = p.condense()
        idx, blk, ld = r.global_contribution([trace_size + c])
        blocks.append(blk)
        loads.append(ld)
        src.append(r.source)
        lifts.append(r.lifts)
        ret.append(r.retained_basis[:, 0])
        corr.append(p.correct_kernel)
    return dict(blocks=_stack(blocks), loads=_stack(loads), source=_stack(src),
                lifts=_stack(lifts), retained=_stack(ret), correct_kernel=np.array(corr),
                trace_size=trace_size, problems=problems)


_MACRO_CACHE = {}


def grid_spec(case, macros):
    """Picklable description of selected macros of a square-grid case."""
    prob = case.problem
    macros = list(macros)
    return dict(nx=prob.nx, refinement=prob.refinement, degree=prob.degree, order=prob.order,
                breaks=tuple(prob.space.breaks), degrees=tuple(prob.space.degrees),
                macros=macros, kq=case.kq[macros],
                fq=None if case.fq is None else case.fq[macros], f_const=case.f_const)


def cube_spec(case, macros):
    """Picklable description of selected macros of a cube-grid case."""
    prob = case.problem
    macros = list(macros)
    return dict(kind="cube", nx=prob.nx, refinement=prob.refinement, degree=prob.degree,
                split=prob.split, order=prob.order, macros=macros, kq=case.kq[macros],
                fq=None if case.fq is None else case.fq[macros], f_const=case.f_const)


def polygon_spec(case, macros):
    """Picklable description of selected macros of a polygon (uneven) case."""
    prob = case.problem
    macros = [int(c) for c in macros]
    return dict(kind="polygon", points=prob.macro.points, cells=[list(map(int, c)) for c in prob.macro.cells],
                refinements=prob.refinements, degree=prob.degree, order=prob.order,
                breaks=tuple(prob.space.breaks), degrees=tuple(prob.space.degrees), macros=macros,
                kq=[case.kq[c] for c in macros], fq=None if case.fq is None else [case.fq[c] for c in macros],
                f_const=case.f_const)


def _polygon_case(spec):
    """A minimal case over the spec's macros (kq/fq indexed by macro id)."""
    key = ("polygon", len(spec["cells"]), float(spec["points"].sum()))
    if key not in _MACRO_CACHE:
        class P:
            pass
        prob = P()

        class Macro:
            pass
        prob.macro = Macro()
        prob.macro.points, prob.macro.cells = spec["points"], spec["cells"]
        prob.refinements, prob.degree, prob.order = spec["refinements"], spec["degree"], spec["order"]
        prob.space = FaceSpace(spec["breaks"], spec["degrees"])
        _MACRO_CACHE[key] = prob
    prob = _MACRO_CACHE[key]

    class C:
        pass
    case = C()
    case.problem = prob
    case.kq = {c: spec["kq"][r] for r, c in enumerate(spec["macros"])}
    case.fq = None if spec["fq"] is None else {c: spec["fq"][r] for r, c in enumerate(spec["macros"])}
    case.f_const = spec["f_const"]
    return case


class _SpecCase:
    """Minimal case object rebuilt from a spec inside a worker process."""

    def __init__(self, spec, kq, fq):
        class P:
            pass
        self.problem = P()
        for key in ("nx", "refinement", "degree", "split", "order"):
            setattr(self.problem, key, spec[key])
        self.kq, self.fq, self.f_const = kq, fq, spec["f_const"]


def condense_spec(spec):
    """Assemble and condense the macros of ``grid_spec`` (one worker's share).

    Same scope as the GPU stage: local mesh, operators, trace coupling,
    augmented SuperLU solve and condensed block/load per macro.
    """
    if spec.get("kind") == "polygon":
        problems, skel = build_polygon_problems(_polygon_case(spec), spec["macros"])
        blocks, loads = [], []
        for c, p, _ in problems:
            _, blk, ld = p.condense().global_contribution([skel.size + c])
            blocks.append(blk)
            loads.append(ld)
        return spec["macros"], blocks, loads
    if spec.get("kind") == "cube":
        # the macro list is re-indexed 0..len-1 against the spec's own kq/fq rows
        full_kq = {c: spec["kq"][r] for r, c in enumerate(spec["macros"])}
        full_fq = None if spec["fq"] is None else {c: spec["fq"][r] for r, c in enumerate(spec["macros"])}
        problems, trace_size = build_cube_problems(_SpecCase(spec, full_kq, full_fq), spec["macros"])
        blocks, loads = [], []
        for c, p, _ in problems:
            _, blk, ld = p.condense().global_contribution([trace_size + c])
            blocks.append(blk)
            loads.append(ld)
        return spec["macros"], np.array(blocks), np.array(loads)
    key = (spec["nx"], spec["breaks"], spec["degrees"])
    if key not in _MACRO_CACHE:   # macro mesh and skeleton are mesh-only: built once per worker
        pm = PolygonMesh.unit_square_grid(spec["nx"])
        space = FaceSpace(spec["breaks"], spec["degrees"])
        _MACRO_CACHE[key] = (pm, SkeletonSpace(pm, [space] * len(pm.faces)))
    pm, skel = _MACRO_CACHE[key]
    blocks, loads = [], []
    for r, c in enumerate(spec["macros"]):
        fine = pm.submesh(c, spec["refinement"])
        fq = None if spec["fq"] is None else spec["fq"][r]
        A, M, f = scalar_operators(fine, spec["degree"], diffusion_q=spec["kq"][r], source_q=fq,
                                   source=spec["f_const"], order=spec["order"])
        B = trace_coupling(pm, c, fine, skel, spec["degree"])
        Z = np.ones((A.shape[0], 1))
        mz = M @ Z
        p = LocalProblem(A, B, f, skel.cell_dofs(c), Z, mz / (Z.T @ mz).item())
        _, blk, ld = p.condense().global_contribution([skel.size + c])
        blocks.append(blk)
        loads.append(ld)
    return spec["macros"], np.array(blocks), np.array(loads)


def _macro_problems(spec):
    """(c, LocalProblem, mesh, fq row or None) of a spec's macros, assembled as in condense_spec."""
    if spec.get("kind") == "cube":
        full_kq = {c: spec["kq"][r] for r, c in enumerate(spec["macros"])}
        full_fq = None if spec["fq"] is None else {c: spec["fq"][r] for r, c in enumerate(spec["macros"])}
        problems, _ = build_cube_problems(_SpecCase(spec, full_kq, full_fq), spec["macros"])
        return [(c, p, mesh) for c, p, mesh in problems]
    key = (spec["nx"], spec["breaks"], spec["degrees"])
    if key not in _MACRO_CACHE:
        pm = PolygonMesh.unit_square_grid(spec["nx"])
        space = FaceSpace(spec["breaks"], spec["degrees"])
        _MACRO_CACHE[key] = (pm, SkeletonSpace(pm, [space] * len(pm.faces)))
    pm, skel = _MACRO_CACHE[key]
    out = []
    for r, c in enumerate(spec["macros"]):
        fine = pm.submesh(c, spec["refinement"])
        fq = None if spec["fq"] is None else spec["fq"][r]
        A, M, f = scalar_operators(fine, spec["degree"], diffusion_q=spec["kq"][r], source_q=fq,
                                   source=spec["f_const"], order=spec["order"])
        B = trace_coupling(pm, c, fine, skel, spec["degree"])
        Z = np.ones((A.shape[0], 1))
        mz = M @ Z
        out.append((c, LocalProblem(A, B, f, skel.cell_dofs(c), Z, mz / (Z.T @ mz).item()), fine))
    return out


def _load_only(spec, mesh, fq):
    """The load vector of source values ``fq (cells, q)`` - the load part of
    scalar_operators / tetra_operators alone (same formulas)."""
    from .fem import element_tabulate, nodal_space, triangle_quadrature
    from .fem3d import tetra_element_tabulate, tetra_nodal_space, tetrahedron_quadrature

    degree, order = spec["degree"], spec["order"]
    if spec.get("kind") == "cube":
        bary, w = tetrahedron_quadrature(max(order, degree + 2))
        vals, _ = tetra_element_tabulate(mesh, degree, bary)
        dofs, nodes = tetra_nodal_space(mesh, degree)
        load_e = np.zeros((len(mesh.cells), vals.shape[1]))
        for q in range(len(w)):
            load_e = load_e + ((w[q] * vals[q])[None, :] * fq[:, q, None]) * mesh.volumes[:, None]
        load = np.zeros(len(nodes))
        np.add.at(load, dofs, load_e)
        return load
    bary, w = triangle_quadrature(max(order, degree + 2))
    vals, _ = element_tabulate(mesh, degree, bary)
    dofs, nodes = nodal_space(mesh, degree)
    ncell, nq = len(mesh.cells), len(w)
    basis = np.broadcast_to(vals, (ncell, nq, vals.shape[1]))
    load = np.einsum("tq,tqi,tq,t->ti", np.broadcast_to(w, (ncell, nq)), basis, fq, mesh.areas)
    return np.bincount(dofs.ravel(), weights=load.ravel(), minlength=len(nodes))


def serving_cost_spec(spec):
    """CPU seconds, summed over the spec's macros, of the two request kinds of
    the serving workload: a coefficient request (local assembly and
    condensation, as condense_spec) and a source request against a resident
    factorization (the new load, one refined solve of [f; 0], the condensed
    load).  The source values are the spec's ``src`` rows."""
    import time

    from .local import Factorization

    t_coef = t_src = 0.0
    problems = _macro_problems(spec)  # warm the mesh caches
    t0 = time.perf_counter()
    problems = _macro_problems(spec)
    for _, p, _ in problems:
        p.condense()
    t_coef = time.perf_counter() - t0
    for r, (c, p, mesh) in enumerate(problems):
        fac = Factorization(p.condensation_matrix())
        t0 = time.perf_counter()
        load = _load_only(spec, mesh, spec["src"][r])
        rhs = np.zeros(p.n + p.k)
        rhs[:p.n] = load
        s = fac.solve(rhs)[:p.n]
        tr = p.test_coupling.T @ s
        co = -p.kernel.T @ load
        if p.correct_kernel:
            co = p.test_action.T @ s + co
        np.concatenate((tr, co))
        t_src += time.perf_counter() - t0
    return len(problems), t_coef, t_src
