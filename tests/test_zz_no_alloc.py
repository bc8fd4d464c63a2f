"""The no-allocation entry point of the dense batched LU.

``mhm_dense_factor_solve_noalloc`` must give, through the C ABI, exactly the
results and status of ``mhm_dense_factor_solve`` (the same kernels on the same
inputs, only the workspace differs) on two calls in a row, the second smaller
than the first, so the fixed per-device scratch is reused with the first call's
contents still in it.  A call whose workspace exceeds the scratch is rejected
before any work is enqueued.
"""

import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from pymhm_amd import _lib  # noqa: E402

P = ctypes.c_void_p


def _p(t):
    return P(t.data_ptr())


def _call(name, a, b, rtol=1e-10):
    lib = _lib.load()
    batch, n, _ = a.shape
    nrhs = b.shape[2]
    ta = torch.as_tensor(a, device="cuda")
    tb = torch.as_tensor(b, device="cuda")
    tx = torch.full_like(tb, -7.0)
    st = torch.full((batch,), 99, dtype=torch.int32, device="cuda")
    rc = getattr(lib, name)(0, batch, n, nrhs, _p(ta), _p(tb), _p(tx), rtol, 0.0, _p(st),
                            P(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return tx.cpu().numpy(), st.cpu().numpy()


def _case(batch, n, nrhs, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(batch, n, n)) + n ** 0.5 * np.eye(n)
    a[:, 0, 0] = 0.0
    b = rng.normal(size=(batch, n, nrhs))
    return a, b


@pytest.mark.parametrize("first,second", [((40, 120, 9), (7, 33, 2)), ((6, 300, 4), (3, 64, 5))])
def test_noalloc_matches_allocating_entry(first, second):
    print(f"noalloc {first} then {second}", flush=True)
    for shape, seed in ((first, 1), (second, 2)):
        a, b = _case(*shape, seed)
        if shape == second:
            a[1] = 1.0  # a singular matrix in the second call: its status, not its neighbours'
        x_ref, s_ref = _call("mhm_dense_factor_solve", a, b)
        x, s = _call("mhm_dense_factor_solve_noalloc", a, b)
        assert np.array_equal(s, s_ref)
        good = s_ref == 0
        assert np.array_equal(x[good], x_ref[good])
        # matrices whose factor failed are skipped by both entry points
        assert np.array_equal(x[~good], np.full_like(x[~good], -7.0))
        np.testing.assert_allclose(x[good], np.linalg.solve(a[good], b[good]), rtol=1e-9, atol=1e-12)


def test_noalloc_rejects_oversized_workspace():
    lib = _lib.load()
    need, limit = ctypes.c_int64(), ctypes.c_int64()
    assert lib.mhm_dense_scratch_bytes(64, 1100, 34, ctypes.byref(need), ctypes.byref(limit)) == 0
    assert need.value > limit.value
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = torch.zeros(64, dtype=torch.int32, device="cuda")
    rc = lib.mhm_dense_factor_solve_noalloc(0, 64, 1100, 34, _p(dummy), _p(dummy), _p(dummy), 1e-10, 0.0, _p(st),
                                            P(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.MHM_EINVAL
    assert "scratch" in _lib.last_error()
