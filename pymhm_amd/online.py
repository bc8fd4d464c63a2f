"""Offline/online hybrid solves with every step on the device.

``OnlineHybridSystem`` is the device counterpart of ``OfflineHybridSystem``
(``src/pymhm/core/offline.py:234-307``): the offline preparation runs the
local stage once (assembly, band factors, condensation; the factors stay
resident), scatters the condensed blocks into the global system and prepares
the global solver's preconditioner (``HybridSolver.set_values``).  A query
brings new source fields (and optionally new Dirichlet data); it is answered
without refactoring anything:

1. loads and source responses of the new fields against the resident local
   factors (``LocalBatch.sources``, ``OfflineLocalProblem.response``,
   ``core/offline.py:121-127``);
2. the global right-hand side from those loads in the reference's order,
   minus the weak Dirichlet load (``HybridScatter.gather_loads``,
   ``core/offline.py:272-286``);
3. the global solve against the resident preconditioner (``HybridSolver``);
4. the fields ``u = s - L lambda + E c`` with the new source responses
   (``LocalBatch.reconstruct(source=...)``, ``core/reconstruction.py:130-141``).

Problems are those with one batch over all macros (square and cube macro
grids); fields come back per macro in local DOF order.
"""

from __future__ import annotations

import numpy as np
import torch

from . import hybrid


class OnlineHybridSystem:
    def __init__(self, problem, kq, fq=None, f_const: float = 0.0, *, precision: str = "fp64",
                 max_columns: int = 8, boundary=None, rtol: float = 1e-10, maxiter: int = 3000):
        """``boundary`` = ``(DirichletBoundary, g)`` of the offline data (the
        online default when a query brings none)."""
        self.problem = problem
        self.batch = problem.batch(precision, max_columns=max_columns)
        try:
            self.batch.assemble(kq, fq, f_const)
            self.batch.factor()
            blocks, loads = self.batch.condense()
            self.plan = hybrid.HybridScatter(problem.global_indices(), problem.trace_size, problem.kernel_offsets())
            vals, rhs, _ = self.plan.gather(blocks, loads, boundary)
            self.solver = hybrid.HybridSolver(self.plan, rtol=rtol, maxiter=maxiter)
            self.solver.set_values(vals)
            self.offline_rhs = rhs
        except BaseException:
            self.batch.close()
            raise
        self.boundary = boundary
        dev = self.batch.device
        self.trace_dofs = torch.as_tensor(np.asarray(problem.trace_dofs), dtype=torch.int64, device=dev)
        self.ts = int(problem.trace_size)
        self._closed = False

    def solve(self, fq=None, f_const: float = 0.0, *, ncols: int | None = None, boundary=None):
        """Answer ``ncols`` source queries (``fq`` ``(ncols, M, elements, q)``, or
        ``f_const`` everywhere): ``(solutions (ncols, size), fields (ncols, M, n))``
        on the device."""
        if self._closed:
            raise RuntimeError("online system is closed")
        loads, src = self.batch.sources(fq, f_const, ncols=ncols, source=True)
        bnd = self.boundary if boundary is None else boundary
        ncols = loads.shape[0]
        sols = torch.empty((ncols, self.plan.size), dtype=torch.float64, device=loads.device)
        for c in range(ncols):
            rhs, _ = self.plan.gather_loads(loads[c], bnd)
            sols[c] = self.solver.solve(rhs)
        trace = sols[:, :self.ts][:, self.trace_dofs].permute(1, 0, 2).contiguous()   # (M, ncols, m)
        coarse = sols[:, self.ts:].T.contiguous()[:, :, None]                           # (M, ncols, 1)
        fields = self.batch.reconstruct(trace, coarse, source=src)
        return sols, fields.permute(1, 0, 2)

    def close(self):
        if not self._closed:
            self.solver.close()
            self.batch.close()
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
