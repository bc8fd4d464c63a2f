"""Drop-ins for the reference's batch-level GPU seams.

* ``condense_batched(problems)`` mirrors ``src/pymhm/execution/cuda.py:291-312``:
  it takes ``LocalProblem``-shaped objects (attributes ``matrix``, ``coupling``,
  ``load``, ``trace_dofs``, ``kernel``, ``constraints``, ``test_coupling``,
  ``coarse_basis``, ``left_kernel``, ``test_constraints``), groups them by
  local sparsity pattern (the reference groups by ``(N, R)``), condenses every
  group on the GPU and returns responses in input order.
* ``BatchedFactorization`` is the reference's dense resident factor object
  (``cuda.py:165-288``: ``(batch, n, n)`` matrices, ``rtol``/``atol``),
  implemented by the dense batched LU of ``pymhm_amd.dense``.
* ``LocalFactorization`` keeps the band factors of the augmented operators of
  ``LocalProblem``-shaped inputs resident: factor once, ``solve`` new
  right-hand sides, ``close`` idempotently.  It is the fast path behind
  ``condense_batched`` for the primal scalar-diffusion family.

Scope of the HIP factorisation: symmetric (to rounding) local operators with
at most one retained mode, positive definite on the complement of the mode -
the primal scalar-diffusion family.  Other problems raise ``ValueError``; no
other solver is substituted.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
from scipy import sparse

from .dense import BatchedFactorization, HipLocalSolver, factor_solve  # noqa: F401  (re-exported)
from .layout import LocalLayout, choose_permutation
from .stage import LocalBatch


@dataclass(frozen=True)
class LocalResponse:
    """Source response, lifts and executed retained basis (``core/contracts.py:321-333``)."""

    problem: object
    source: np.ndarray
    lifts: np.ndarray
    coarse_vectors: np.ndarray | None = None
    block: np.ndarray | None = None
    load: np.ndarray | None = None

    @property
    def retained_basis(self):
        return self.problem.kernel if self.coarse_vectors is None else self.coarse_vectors

    def reconstruct(self, trace, coarse):
        return self.source - self.lifts @ trace + self.retained_basis @ coarse


def _check_supported(p):
    k = p.kernel.shape[1]
    if k > 1:
        raise ValueError("the HIP local stage supports at most one retained mode")
    if getattr(p, "coarse_basis", p.kernel).shape[1] != k:
        raise ValueError("a general coarse_basis is not supported by the HIP local stage")
    if not np.array_equal(p.test_coupling, p.coupling):
        raise ValueError("test_coupling must equal coupling for the HIP local stage")
    for name in ("left_kernel", "test_constraints"):
        other = getattr(p, name, None)
        ref = p.kernel if name == "left_kernel" else p.constraints
        if other is not None and not np.array_equal(other, ref):
            raise ValueError(f"{name} must equal the trial data for the HIP local stage")
    a = sparse.csr_matrix(p.matrix)
    scale = float(np.max(np.abs(a.data), initial=0.0))
    if abs(a - a.T).max() > 1e-12 * scale:
        raise ValueError("the HIP local stage needs a symmetric local operator")


def _pattern(p):
    a = sparse.csr_matrix(p.matrix, dtype=float, copy=True)
    a.sum_duplicates()
    a.sort_indices()
    ones = a.copy()
    ones.data[:] = 1.0            # stored zeros stay in the pattern
    sym = (ones + ones.T).tocsr()  # structurally symmetric pattern
    sym.sort_indices()
    return a, sym.indptr.astype(np.int64), sym.indices.astype(np.int64)


def _group_key(p, indptr, indices):
    return (p.matrix.shape[0], p.coupling.shape[1], p.kernel.shape[1],
            indptr.tobytes(), indices.tobytes())


class _Group:
    def __init__(self, members, indptr, indices, precision, max_columns, device):
        first = members[0][1]
        n, m, k = first.matrix.shape[0], first.coupling.shape[1], first.kernel.shape[1]
        perm = choose_permutation(indptr, indices)
        if k:
            # pin the last band row; the retained mode must not vanish there
            for _, p, _a in members:
                if p.kernel[perm[-1], 0] == 0.0:
                    raise ValueError("retained mode vanishes at the pinned row")
        self.layout = LocalLayout(n, perm, indptr, indices, m=m, k=k, device=device)
        self.members = members
        M = len(members)
        slot_values = np.zeros((M, self.layout.nnz))
        for r, (_, p, a) in enumerate(members):
            slots = self.layout.slot_map(a.indptr, a.indices)
            slot_values[r, slots] = a.data
        loads = np.stack([p.load for _, p, _ in members])
        couplings = np.stack([p.coupling.T for _, p, _ in members])         # (M, m, n)
        self.batch = LocalBatch(self.layout, M, precision=precision, max_columns=max_columns)
        if k:
            cons = np.stack([p.constraints[:, 0] for _, p, _ in members])
            kern = np.stack([p.kernel[:, 0] for _, p, _ in members])
            self.batch.set_coupling(couplings=couplings, constraints=cons, kernels=kern)
        else:
            self.batch.set_coupling(couplings=couplings)
        self.batch.set_operator(slot_values, loads)


def _prepare(problems, precision, max_columns, device):
    groups = {}
    for index, p in enumerate(problems):
        if not hasattr(p, "matrix") or not hasattr(p, "coupling"):
            raise TypeError("every batch entry must be a LocalProblem")
        _check_supported(p)
        a, indptr, indices = _pattern(p)
        groups.setdefault(_group_key(p, indptr, indices), (indptr, indices, []))[2].append((index, p, a))
    return [_Group(mem, ip, ix, precision, max_columns, device) for ip, ix, mem in groups.values()]


def condense_batched(problems, *, precision: str = "fp64", device: int = 0):
    """Condense on the GPU, preserving input order (``cuda.py:291-312``)."""
    problems = list(problems)
    responses = {}
    for g in _prepare(problems, precision, 0, device):
        with g.batch as b:
            b.factor()
            blocks, loads = b.condense()
            src, lifts, ret = b.responses()
            src, lifts = src.cpu().numpy(), lifts.cpu().numpy()
            ret = None if ret is None else ret.cpu().numpy()
            blocks, loads = blocks.cpu().numpy(), loads.cpu().numpy()
        for r, (index, p, _) in enumerate(g.members):
            corrected = bool(getattr(p, "correct_kernel", getattr(p, "_correct_kernel", True)))
            coarse = None if (ret is None or not corrected) else ret[r][:, None]
            responses[index] = LocalResponse(p, src[r], lifts[r], coarse, blocks[r], loads[r])
        g.layout.close()
    return tuple(responses[i] for i in range(len(problems)))


class LocalFactorization:
    """Resident band factors of the augmented operators of equal-pattern problems.

    ``solve(rhs)`` takes ``(batch, N, nrhs)`` like ``cuda.py:223-255`` (any
    ``nrhs``: columns go to the device ``max_columns`` at a time, each column
    computed independently) and returns device tensors unless ``host=True``.
    """

    def __init__(self, problems, *, precision: str = "fp64", max_columns: int = 8, device: int = 0):
        problems = list(problems)
        groups = _prepare(problems, precision, max_columns, device)
        if len(groups) != 1:
            raise ValueError("LocalFactorization needs problems of one local pattern")
        self._group = groups[0]
        self._group.batch.factor()
        self._closed = False

    def solve(self, rhs, *, host: bool = False):
        if self._closed:
            raise RuntimeError("batched factorization is closed")
        b = self._group.batch
        rhs = torch.as_tensor(rhs, dtype=torch.float64, device=b.device)
        if rhs.dim() != 3 or rhs.shape[0] != b.macros or rhs.shape[1] != b.N or rhs.shape[2] == 0:
            raise ValueError(f"rhs must have shape ({b.macros}, {b.N}, nrhs)")
        cols = rhs.transpose(1, 2).contiguous()
        parts = [b.solve(cols[:, c:c + b.max_columns]) for c in range(0, cols.shape[1], b.max_columns)]
        x = torch.cat(parts, dim=1).transpose(1, 2)
        return x.cpu().numpy() if host else x

    def close(self):
        if not self._closed:
            self._closed = True
            self._group.batch.close()
            self._group.layout.close()

    def __enter__(self):
        if self._closed:
            raise RuntimeError("batched factorization is closed")
        return self

    def __exit__(self, *exc):
        self.close()
