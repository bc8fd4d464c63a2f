"""Batch-object contracts on the device: per-call status, invalidation after
new coupling data, and fault isolation between co-batched macros.

* A solve's status is its own: a call that misses its residual criterion does
  not make the next call on the same resident factors fail.
* New couplings, constraints or retained modes invalidate the factors (factor
  and condense again), and new retained modes after assembly update A Z.
* The condensation's first pass, built beside the factor, serves only the
  condensation that follows: a solve of other columns in between, or a second
  factorization, leaves every result what the plain order gives.
* A NaN coefficient in one macro, or an infinite value in one source field,
  flags only that macro, and every other macro's result is bitwise the result
  it gets in a batch without the fault (reference: ``tests/test_gpu_batch.py``
  isolation of failing local systems, forced from the host).
"""

import numpy as np
import pytest
from scipy import sparse

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from pymhm_amd import _lib, serving  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402
from pymhm_amd._lib import LinearSolveError  # noqa: E402


def _fields(prob, seeds):
    pts = torch.as_tensor(prob.quadrature_points(), dtype=torch.float64, device="cuda")
    return serving.bump_sources(pts, seeds)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_status_is_per_call(precision):
    case = pw.config1(nx=4)
    prob = case.problem
    fq = _fields(prob, [5, 6])
    with prob.batch(precision, max_columns=2) as b:
        b.assemble(case.kq, case.fq, case.f_const)
        b.factor()
        b.condense()
        first = b.sources(fq).cpu().numpy()
        with pytest.raises(LinearSolveError, match="residual"):
            b.sources(fq, rtol=1e-300)        # cannot be met: rounding-size residuals remain
        again = b.sources(fq).cpu().numpy()   # the failed call's bits do not carry over
        assert not b.status().any()
    assert np.array_equal(first, again)


def test_new_coupling_data_invalidates_factors_and_updates_az():
    case = pw.config1(nx=4)
    prob = case.problem
    lay = prob.layout
    with prob.batch("fp64") as b:
        b.assemble(case.kq, case.fq, case.f_const)
        b.factor()
        b.set_coupling(colmap=prob.colmap, colscale=prob.colscale)
        with pytest.raises(ValueError, match="not factored"):
            b.condense()
        b.factor()
        b.condense()
        z = np.random.default_rng(4).uniform(0.5, 1.5, size=(prob.M, prob.n))
        b.set_coupling(kernels=z)
        vals, _, az = (t.cpu().numpy() for t in b.operator())
        with pytest.raises(ValueError, match="not factored"):
            b.condense()
    for c in range(prob.M):
        a = sparse.csr_matrix((vals[c], lay.cols, lay.rowptr), shape=(lay.n, lay.n))
        expect = np.empty(lay.n)
        expect[lay.perm] = a @ z[c][lay.perm]
        assert np.allclose(az[c], expect, rtol=1e-13, atol=1e-13 * np.abs(vals[c]).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_faults_stay_with_their_macro(precision):
    case = pw.config1(nx=4)
    prob = case.problem
    fq = _fields(prob, [7, 8, 9])
    bad_kq = case.kq.copy()
    bad_kq[5, 3, 2] = np.nan                  # one coefficient value of macro 5
    bad_fq = fq.clone()
    bad_fq[1, 9, 0, 0] = np.inf               # one source value of macro 9, field 1
    out = {}
    for tag, kq, f in (("clean", case.kq, fq), ("fault", bad_kq, bad_fq)):
        with prob.batch(precision, max_columns=3) as b:
            b.assemble(kq, case.fq, case.f_const)
            b.factor(check=False)
            blocks, loads = b.condense(check=False)
            st_cond = b.status()
            src = b.sources(f, check=False)
            st_src = b.status()
            out[tag] = (blocks.cpu().numpy(), loads.cpu().numpy(), st_cond, src.cpu().numpy(), st_src)
    blocks0, loads0, s0, src0, t0 = out["clean"]
    blocks1, loads1, s1, src1, t1 = out["fault"]
    assert not s0.any() and not t0.any()
    assert s1[5] & _lib.STATUS_NONFINITE and np.count_nonzero(s1) == 1
    others = np.arange(prob.M) != 5
    assert np.array_equal(blocks1[others], blocks0[others])
    assert np.array_equal(loads1[others], loads0[others])
    # the source call: macro 9 (infinite field value) and macro 5 (bad factor) only
    assert t1[9] & _lib.STATUS_NONFINITE
    assert set(np.flatnonzero(t1)) <= {5, 9}
    keep = np.ones(prob.M, dtype=bool)
    keep[[5, 9]] = False
    assert np.array_equal(src1[:, keep], src0[:, keep])
    # fields 0 and 2 of macro 9 are untouched by field 1's infinity
    assert np.array_equal(src1[[0, 2], 9], src0[[0, 2], 9])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_calls_between_factor_and_condense(precision):
    case = pw.config2(nx=4)
    prob = case.problem
    print(f"factor/solve/condense order config2 4x4 {precision}", flush=True)
    with prob.batch(precision, max_columns=3) as b:
        b.assemble(case.kq, None, case.f_const)
        rhs = torch.randn((b.macros, 3, b.N), dtype=torch.float64,
                          generator=torch.Generator().manual_seed(7)).cuda()
        b.factor()
        blocks0, loads0 = (t.cpu().numpy().copy() for t in b.condense())
        x0 = b.solve(rhs).cpu().numpy()
        # a solve of other columns right after the factor, then the condensation
        b.factor()
        x1 = b.solve(rhs).cpu().numpy()
        blocks1, loads1 = (t.cpu().numpy().copy() for t in b.condense())
        # two factorizations in a row, then the condensation
        b.factor()
        b.factor()
        blocks2, loads2 = (t.cpu().numpy().copy() for t in b.condense())
        x2 = b.solve(rhs).cpu().numpy()
    # same calls in the same order: same bits
    assert np.array_equal(blocks2, blocks0) and np.array_equal(loads2, loads0)
    assert np.array_equal(x2, x0)
    # u and w computed by the first solve instead of the condensation: the
    # bordered correction may differ in its last bits
    assert _rel(x1, x0) < 1e-12
    assert _rel(blocks1, blocks0) < 1e-12 and _rel(loads1, loads0) < 1e-12
