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
from .local import LocalProblem
from .meshes import PolygonMesh


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


def oracle_config(case, macros=None):
    """Blocks, loads and responses of the selected macros (default: all)."""
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
    return dict(blocks=np.array(blocks), loads=np.array(loads), source=np.array(src),
                lifts=np.array(lifts), retained=np.array(ret), correct_kernel=np.array(corr),
                trace_size=trace_size, problems=problems)
