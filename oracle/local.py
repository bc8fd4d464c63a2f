"""FP64 local algebra of the MHM stage (oracle only; the checker, never the product).

Restates:

* ``accurate_residual``            ``src/pymhm/linalg/linear.py:98-126`` (long double)
* ``checked``                      ``src/pymhm/linalg/linear.py:132-152``
* ``equilibrated_lu``/``Factorization`` ``src/pymhm/linalg/linear.py:400-422,262-310,521-537``
* ``LocalProblem``                 ``src/pymhm/core/contracts.py:55-211`` (hot-path subset)
* ``condensation_matrix/system``   ``src/pymhm/core/condensation.py:125-174``
* ``condense``                     ``src/pymhm/core/condensation.py:118-122,177-191``
* ``global_contribution``          ``src/pymhm/core/contributions.py:14-74``
* ``condensed_load``               ``src/pymhm/core/reconstruction.py:91-127``
* ``assemble_hybrid_contributions`` ``src/pymhm/core/contributions.py:77-195``,
  ``src/pymhm/core/_assembly.py:9-42``
* ``reconstruct``                  ``src/pymhm/core/reconstruction.py:130-141``
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import sparse
from scipy.sparse import linalg as splinalg


class LinearSolveError(RuntimeError):
    """Singular pivot, failed residual criterion or non-finite result."""


# ---------------------------------------------------------------- residuals
def accurate_residual(matrix, rhs, solution):
    """``rhs - matrix @ solution`` accumulated in x87 long double."""
    ld = np.longdouble
    return np.asarray(rhs, dtype=ld) - sparse.csr_matrix(matrix).astype(ld) @ np.asarray(
        solution, dtype=ld)


def matrix_action(matrix, values):
    """``matrix @ values`` in long double, rounded to binary64 (``contracts.py:37-52``)."""
    return np.asarray(-accurate_residual(
        matrix, np.zeros((matrix.shape[0], *values.shape[1:])), values), dtype=float)


class _ResidualFailure(LinearSolveError):
    pass


def checked(matrix, rhs, solution, rtol=1e-10, atol=0.0):
    """Per-column acceptance ``||Ax-b||/s <= rtol ||b/s||``, ``s = max|b|``."""
    solution = np.asarray(solution)
    if solution.shape != rhs.shape or not np.all(np.isfinite(solution)):
        raise LinearSolveError("nonfinite or misshapen solution")
    with np.errstate(over="ignore", invalid="ignore"):
        diff = matrix @ solution - rhs
    if not np.all(np.isfinite(diff)):
        raise LinearSolveError("nonfinite residual")
    scale = np.maximum(atol, np.max(np.abs(rhs), axis=0))
    scale = np.where(scale > 0, scale, 1.0)
    res = np.linalg.norm(diff / scale, axis=0)
    thr = np.maximum(atol / scale, rtol * np.linalg.norm(rhs / scale, axis=0))
    if np.any(res > thr):
        res = np.linalg.norm(np.asarray(accurate_residual(matrix, rhs, solution), dtype=float)
                             / scale, axis=0)
    if np.any(res > thr):
        raise _ResidualFailure(f"residual criterion failed: {res} > {thr}")
    return solution


def residual_ok(matrix, rhs, solution, rtol=1e-10, atol=0.0):
    """Boolean per column of the acceptance criterion above (long-double residual)."""
    scale = np.maximum(atol, np.max(np.abs(rhs), axis=0))
    scale = np.where(scale > 0, scale, 1.0)
    res = np.linalg.norm(np.asarray(accurate_residual(matrix, rhs, solution), dtype=float)
                         / scale, axis=0)
    thr = np.maximum(atol / scale, rtol * np.linalg.norm(rhs / scale, axis=0))
    return res <= thr


# ---------------------------------------------------------------- factorization
def equilibrated_lu(matrix):
    """Row then column max-abs scaling, SuperLU, pivot-rank rejection."""
    matrix = sparse.csr_matrix(matrix, dtype=float, copy=True)
    matrix.sum_duplicates()
    row = np.asarray(abs(matrix).max(axis=1).toarray()).ravel()
    if np.any(row == 0):
        raise LinearSolveError("zero row")
    bal = matrix.copy()
    bal.data /= np.repeat(row, np.diff(bal.indptr))
    col = np.asarray(abs(bal).max(axis=0).toarray()).ravel()
    if np.any(col == 0):
        raise LinearSolveError("zero column")
    bal.data /= col[bal.indices]
    try:
        lu = splinalg.splu(bal.tocsc())
    except RuntimeError as exc:
        raise LinearSolveError(str(exc)) from exc
    piv = np.abs(lu.U.diagonal())
    if np.min(piv) <= np.finfo(float).eps * matrix.shape[0] * np.max(np.abs(lu.U.data)):
        raise LinearSolveError("insufficient numerical rank after equilibration")
    return lu, row, col


class Factorization:
    """Checked reusable SuperLU factorization with <=2 long-double refinements."""

    def __init__(self, matrix, rtol=1e-10, atol=0.0):
        self.matrix = sparse.csr_matrix(matrix, dtype=float, copy=True)
        self.matrix.sum_duplicates()
        self.lu, self.row, self.col = equilibrated_lu(self.matrix)
        self.rtol, self.atol = rtol, atol
        self.refinements = 0

    def _apply(self, rhs):
        r = self.row[:, None] if rhs.ndim == 2 else self.row
        c = self.col[:, None] if rhs.ndim == 2 else self.col
        return np.asarray(self.lu.solve(np.ascontiguousarray(rhs / r)) / c)

    def solve(self, rhs, refinement_steps: int = 2):
        rhs = np.asarray(rhs, dtype=float)
        x = self._apply(rhs)
        for _ in range(refinement_steps):
            try:
                return checked(self.matrix, rhs, x, self.rtol, self.atol)
            except _ResidualFailure:
                r = np.asarray(accurate_residual(self.matrix, rhs, x), dtype=float)
                x = x + self._apply(r)
                self.refinements += 1
        return checked(self.matrix, rhs, x, self.rtol, self.atol)


# ---------------------------------------------------------------- local contract
@dataclass
class LocalProblem:
    """``A u + B lambda = f`` with declared kernel Z and constraints C (``contracts.py:55-211``).

    Hot-path subset: ``kernel`` Z (right = left kernel), ``constraints`` C
    (default Z; test constraints equal C), ``test_coupling`` (default B).
    ``retained_action = A Z`` and ``test_action = A^T Z`` are accumulated in
    long double; ``correct_kernel`` is set when either is not exactly zero.
    """

    matrix: sparse.csc_matrix
    coupling: np.ndarray
    load: np.ndarray
    trace_dofs: np.ndarray
    kernel: np.ndarray
    constraints: np.ndarray
    test_coupling: np.ndarray

    def __init__(self, matrix, coupling, load, trace_dofs, kernel=None, constraints=None,
                 test_coupling=None):
        self.matrix = sparse.csc_matrix(matrix, dtype=float, copy=True)
        n = self.matrix.shape[0]
        self.coupling = np.array(coupling, dtype=float, copy=True).reshape(n, -1)
        self.trace_dofs = np.asarray(trace_dofs, dtype=np.int64).copy()
        if len(np.unique(self.trace_dofs)) != len(self.trace_dofs) or np.any(self.trace_dofs < 0):
            raise ValueError("trace_dofs must be distinct nonnegative integer indices")
        self.load = np.array(load, dtype=float, copy=True).reshape(n)
        self.kernel = np.empty((n, 0)) if kernel is None else np.array(kernel, dtype=float)
        self.constraints = (self.kernel.copy() if constraints is None
                            else np.array(constraints, dtype=float).reshape(self.kernel.shape))
        self.test_coupling = (self.coupling.copy() if test_coupling is None
                              else np.array(test_coupling, dtype=float))
        for name in ("coupling", "load", "kernel", "constraints", "test_coupling"):
            if not np.all(np.isfinite(getattr(self, name))):
                raise ValueError(f"{name} must be finite")
        if not np.all(np.isfinite(self.matrix.data)):
            raise ValueError("local matrix must be finite and square")
        k = self.kernel.shape[1]
        if k:
            self.retained_action = matrix_action(self.matrix, self.kernel)
            self.test_action = matrix_action(self.matrix.T.tocsc(), self.kernel)
        else:
            self.retained_action = np.empty_like(self.kernel)
            self.test_action = np.empty_like(self.kernel)
        self.correct_kernel = bool(k and (np.any(self.retained_action)
                                          or np.any(self.test_action)))

    @property
    def n(self) -> int:
        return self.matrix.shape[0]

    @property
    def k(self) -> int:
        return self.kernel.shape[1]

    @property
    def m(self) -> int:
        return self.coupling.shape[1]

    def condensation_matrix(self):
        """``[[A, C s],[(C s)^T, 0]]``, ``s = min(1, max|A| / max|C col|)``."""
        if not self.k:
            return self.matrix
        scale = float(np.max(np.abs(self.matrix.data), initial=0.0)) or 1.0
        left = self.constraints * np.minimum(1.0, scale / np.max(np.abs(self.constraints), axis=0))
        return sparse.bmat([[self.matrix, sparse.csc_matrix(left)],
                            [sparse.csc_matrix(left.T), None]], format="csc")

    def condensation_system(self):
        """Augmented operator and RHS columns ``[f, B, A Z (general only)]``."""
        n, k = self.n, self.k
        general = self.correct_kernel
        width = self.m + 1
        rhs = np.zeros((n + k, width + (k if general else 0)))
        rhs[:n, 0] = self.load
        rhs[:n, 1:width] = self.coupling
        if general:
            rhs[:n, width:] = self.retained_action
        return self.condensation_matrix(), rhs

    def condense(self) -> "LocalResponse":
        mat, rhs = self.condensation_system()
        fac = Factorization(mat)
        return self.response_from_solution(fac.solve(rhs))

    def response_from_solution(self, solution) -> "LocalResponse":
        n, width = self.n, self.m + 1
        sol = np.asarray(solution, dtype=float)[:n]
        coarse = self.kernel - sol[:, width:] if self.correct_kernel else None
        return LocalResponse(self, sol[:, 0].copy(), sol[:, 1:width].copy(), coarse)


@dataclass
class LocalResponse:
    """Source response s, lifts L and executed retained basis E (or None = Z)."""

    problem: LocalProblem
    source: np.ndarray
    lifts: np.ndarray
    coarse_vectors: np.ndarray | None = None

    @property
    def retained_basis(self):
        return self.problem.kernel if self.coarse_vectors is None else self.coarse_vectors

    def global_load(self):
        """Reduced load ``[B_t^T s ; (A^T W)^T s - W^T f]`` or ``[.. ; -Z^T f]``."""
        p = self.problem
        if self.coarse_vectors is not None:
            coarse = p.test_action.T @ self.source - p.kernel.T @ p.load
        else:
            coarse = -p.kernel.T @ p.load
        return np.concatenate((p.test_coupling.T @ self.source, coarse))

    def global_contribution(self, coarse_dofs):
        """Indices, condensed block and load of one cell (``contributions.py:14-74``)."""
        p = self.problem
        coarse_dofs = np.asarray(coarse_dofs, dtype=np.int64)
        e = self.retained_basis
        g = p.test_coupling.T @ e
        if self.coarse_vectors is None:
            ctrace = p.kernel.T @ p.coupling
            cmat = np.zeros((len(coarse_dofs), len(coarse_dofs)))
        else:
            ctrace = p.kernel.T @ p.coupling - p.test_action.T @ self.lifts
            cmat = p.test_action.T @ e
        block = np.block([[p.test_coupling.T @ self.lifts, -g], [-ctrace, -cmat]])
        return np.r_[p.trace_dofs, coarse_dofs].astype(np.int64), block, self.global_load()

    def reconstruct(self, trace, coarse):
        """``u = s - L lambda + E c`` (``reconstruction.py:130-141``)."""
        return self.source - self.lifts @ trace + self.retained_basis @ coarse


# ---------------------------------------------------------------- global scatter
def contribution_coordinates(indices):
    """Row-major ``(rows, cols)`` of one block (``_assembly.py:9-25``)."""
    idx = np.asarray(indices, dtype=np.int64)
    return np.repeat(idx, len(idx)), np.tile(idx, len(idx))


def assemble_hybrid_contributions(contributions, trace_size, kernel_offsets, boundary_load=None):
    """Ordered COO stream -> CSC (sum duplicates, eliminate zeros), rhs and load scale."""
    kernel_offsets = np.asarray(kernel_offsets, dtype=np.int64)
    size = int(kernel_offsets[-1])
    rhs = np.zeros(size)
    scale = np.zeros(size)
    rows, cols, vals = [], [], []
    for cell, (idx, block, load) in enumerate(contributions):
        idx = np.asarray(idx, dtype=np.int64)
        if not np.array_equal(np.sort(idx[idx >= trace_size]),
                              np.arange(kernel_offsets[cell], kernel_offsets[cell + 1])):
            raise ValueError("coarse indices must match the ordered cell partition")
        r, c = contribution_coordinates(idx)
        rows.append(r)
        cols.append(c)
        vals.append(np.asarray(block, dtype=float).ravel().copy())
        for i in range(len(idx)):  # ordered accumulation (_assembly.py:28-42)
            rhs[idx[i]] += load[i]
            scale[idx[i]] += abs(load[i])
    mat = sparse.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                            shape=(size, size)).tocsc()
    mat.eliminate_zeros()
    if boundary_load is not None:
        b = np.asarray(boundary_load, dtype=float)
        rhs[:trace_size] -= b
        scale[:trace_size] += np.abs(b)
    return mat, rhs, scale
