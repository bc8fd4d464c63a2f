"""Dense batched LU drop-in (``BatchedFactorization`` of ``execution/cuda.py:165-288``).

The argument, singularity and right-hand-side cases mirror the reference's own
``tests/test_gpu_batch.py:408-446`` (there against a simulated CuPy; here on
the device).  Solutions are compared with LAPACK (``numpy.linalg.solve``) to
1e-10 relative per matrix; matrices are seeded and well conditioned, so the
comparison measures the solver, not the conditioning.
"""

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from pymhm_amd import _lib, batched, dense  # noqa: E402
from pymhm_amd._lib import LinearSolveError  # noqa: E402

TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    num = np.linalg.norm((a - b).reshape(len(a), -1), axis=1)
    den = np.linalg.norm(b.reshape(len(b), -1), axis=1)
    return float(np.max(num / den))


def general(batch, n, seed):
    """Nonsymmetric, row-scaled matrices whose first rows need pivoting.

    The row scales stay within a factor 100: the acceptance test measures the
    residual in original units against ||b||, which no LU meets at 1e-10 once
    the rows differ by ~1e6 (a LAPACK run of the same algorithm fails there too)."""
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(batch, n, n)) + 2 * n ** 0.5 * np.eye(n)
    a[:, 0, 0] = 0.0                         # the first pivot must come from another row
    a *= 10.0 ** rng.uniform(-1, 1, size=(batch, n, 1))
    return a


@pytest.mark.parametrize("n", [1, 2, 7, 91, 141, 142, 325])
@pytest.mark.parametrize("nrhs", [1, 3, 13])
def test_dense_matches_lapack(n, nrhs):
    """LDS-resident (n <= 141) and global-memory eliminations against LAPACK."""
    print(f"dense n={n} nrhs={nrhs}", flush=True)
    batch = 24 if n < 300 else 8
    a = general(batch, n, seed=n * 100 + nrhs)
    if n == 1:
        a[:, 0, 0] = 3.0
    b = np.random.default_rng(7).normal(size=(batch, n, nrhs))
    with batched.BatchedFactorization(a) as fac:
        x = fac.solve(b, host=True)
        x2 = fac.solve(b[:, :, :1], host=True)
    ref = np.linalg.solve(a, b)
    assert rel(x, ref) < TOL
    assert np.array_equal(x2[:, :, 0], x[:, :, 0])  # a column's result does not depend on the others
    xo = dense.factor_solve(a, b, host=True)
    assert np.array_equal(xo, x)  # the one-shot entry runs the same kernels


def test_dense_augmented_local_systems():
    """The reference's condense_batched input: dense augmented saddle matrices
    of local problems (cuda.py:304-309) with their [f, B, A Z] right-hand sides."""
    from oracle import workloads as ow
    from pymhm_amd import workloads as pw

    case = pw.config1(nx=4)
    ref = ow.oracle_config(case)
    problems = [p for _, p, _ in ref["problems"]]
    mats = np.stack([p.condensation_matrix().toarray() for p in problems])
    rhs = np.stack([p.condensation_system()[1] for p in problems])
    with batched.BatchedFactorization(mats) as fac:
        x = fac.solve(rhs, host=True)
    assert rel(x, np.linalg.solve(mats, rhs)) < TOL


def test_local_solver_callable():
    """HipLocalSolver as the per-cell LocalLinearSolver (core/assembly.py:150-156)."""
    import pickle

    from oracle import workloads as ow
    from pymhm_amd import workloads as pw

    case = pw.config1(nx=2)
    ref = ow.oracle_config(case)
    solver = pickle.loads(pickle.dumps(dense.HipLocalSolver()))
    for _, p, _ in ref["problems"]:
        s, r = p.condensation_system()
        x = solver(s.copy(), r.copy())
        assert x.shape == r.shape
        res = np.linalg.norm(r - s @ x, axis=0) / np.linalg.norm(r, axis=0)
        assert res.max() < 1e-12
    solver.close()


@pytest.mark.parametrize(
    "matrix",
    [[], [[1.0]], np.empty((0, 1, 1)), np.empty((1, 0, 0)), np.ones((1, 2, 3)), [[[np.nan]]], [[[1j]]]],
)
def test_batch_matrix_validation(matrix):
    with pytest.raises(ValueError, match="matrices"):
        batched.BatchedFactorization(matrix)


@pytest.mark.parametrize(
    ("matrix", "match"),
    [
        ([[[0.0, 0.0], [1.0, 2.0]]], "zero row"),
        ([[[0.0, 1.0], [0.0, 2.0]]], "zero column"),
        ([[[1.0, 1.0], [1.0, 1.0]]], "singular pivot"),
        ([[[1.0, 1.0], [1.0, 1.0 + 2e-16]]], "rank deficient"),
    ],
)
def test_batch_singularity_checks(matrix, match):
    with pytest.raises(LinearSolveError, match=match):
        batched.BatchedFactorization(matrix)


@pytest.mark.parametrize("rhs", [[], [[1.0]], np.zeros((1, 2, 1)), np.zeros((1, 1, 0)), [[[np.inf]]], [[[1j]]]])
def test_batch_rhs_validation(rhs):
    with batched.BatchedFactorization([[[1.0]]]) as prepared, pytest.raises(ValueError, match="rhs"):
        prepared.solve(rhs)


def test_batch_residual_and_nonfinite_failures():
    # a residual of rounding size cannot meet rtol = 1e-300
    with batched.BatchedFactorization(general(1, 6, seed=5), rtol=1e-300) as prepared:
        with pytest.raises(LinearSolveError, match="residual"):
            prepared.solve(np.random.default_rng(1).normal(size=(1, 6, 2)))
    with batched.BatchedFactorization([[[1e-200]]]) as prepared:
        with pytest.raises(LinearSolveError, match="nonfinite"):
            prepared.solve([[[1e200]]])
    with batched.BatchedFactorization([[[2.0]]]) as prepared:
        np.testing.assert_allclose(prepared.solve([[[2.0]]], host=True), 1.0)
    with pytest.raises(RuntimeError, match="closed"):
        prepared.solve([[[2.0]]])
    with pytest.raises(ValueError, match="tolerance"):
        batched.BatchedFactorization([[[1.0]]], rtol=0.0, atol=0.0)


def test_failure_stays_with_its_matrix():
    """A singular matrix and an infinite right-hand-side column in one batch:
    only those matrices are flagged, and every other matrix's solution is
    bitwise the one it gets without them (the kernels keep no cross-matrix
    state).  Calls the C ABI directly to see per-matrix status."""
    lib = _lib.load()
    n, nrhs = 20, 3
    a = general(6, n, seed=11)
    b = np.random.default_rng(12).normal(size=(6, n, nrhs))
    bad_a, bad_b = a.copy(), b.copy()
    bad_a[2] = 1.0                         # rank one: singular
    bad_b[4, 5, 1] = np.inf                # one infinite entry in one column
    outs = []
    for aa, bb in ((a, b), (bad_a, bad_b)):
        ta = torch.as_tensor(aa, device="cuda")
        tb = torch.as_tensor(bb, device="cuda")
        tx = torch.zeros_like(tb)
        st = torch.zeros(6, dtype=torch.int32, device="cuda")
        rc = lib.mhm_dense_factor_solve(0, 6, n, nrhs, ctypes.c_void_p(ta.data_ptr()), ctypes.c_void_p(tb.data_ptr()),
                                        ctypes.c_void_p(tx.data_ptr()), 1e-10, 0.0, ctypes.c_void_p(st.data_ptr()),
                                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == 0, _lib.last_error()
        outs.append((tx.cpu().numpy(), st.cpu().numpy()))
    (x0, s0), (x1, s1) = outs
    assert not s0.any()
    assert s1[2] & (_lib.DENSE_SINGULAR | _lib.DENSE_RANK)
    assert s1[4] & _lib.DENSE_NONFINITE_RHS
    assert np.count_nonzero(s1) == 2
    for i in (0, 1, 3, 5):
        assert np.array_equal(x0[i], x1[i])
