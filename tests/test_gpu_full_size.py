"""The local stage at the full sizes of the 2D and 3D workloads (GPU).

64 x 64 square macros (P2 triangles, n = 1089, bandwidth 66, P1 multipliers)
and 16^3 cube macros (P2 tetrahedra, n = 729, bandwidth 189, P0 multipliers),
4096 macros each, both precisions:

* every macro's factor and solves succeed (no status bit);
* a seeded sample of macros matches the FP64 oracle within 1e-10;
* every condensed block is symmetric to 1e-10 of its largest entry (B^T S^-1 B
  with S symmetric: a size-independent check of all 4096 results);
* a macro's blocks and loads are bit for bit the same in the full batch as in
  a batch of its own.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from oracle import workloads as ow  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402

TOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    axes = tuple(range(1, a.ndim))
    num = np.sqrt(np.sum((a - b) ** 2, axis=axes))
    den = np.sqrt(np.sum(b ** 2, axis=axes))
    return float(np.max(num / np.where(den > 0, den, 1.0)))


def _condense(prob, kq, precision, macros=None):
    with prob.batch(precision, macros=macros) as b:
        b.assemble(kq if macros is None else kq[macros], None, 1.0)
        b.factor()
        blocks, loads = b.condense()
        return blocks.cpu().numpy().copy(), loads.cpu().numpy().copy()


@pytest.fixture(scope="module", params=["config2", "config3"])
def full_case(request):
    print(f"{request.param}: building the full workload", flush=True)
    case = pw.config2() if request.param == "config2" else pw.config3()
    rng = np.random.default_rng(20261204)
    sel = sorted(rng.choice(case.problem.M, size=4, replace=False).tolist())
    print(f"{request.param}: oracle on macros {sel}", flush=True)
    return request.param, case, sel, ow.oracle_config(case, macros=sel)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_stage(full_case, precision):
    which, case, sel, ref = full_case
    prob = case.problem
    assert case.fq is None and case.f_const == 1.0
    kq = torch.as_tensor(case.kq, device="cuda")
    print(f"{which} {precision}: {prob.M} macros", flush=True)
    blocks, loads = _condense(prob, kq, precision)
    assert blocks.shape[0] == prob.M == 4096
    assert np.all(np.isfinite(blocks)) and np.all(np.isfinite(loads))
    assert rel(blocks[sel], ref["blocks"]) < TOL
    assert rel(loads[sel], ref["loads"]) < TOL
    asym = np.abs(blocks - blocks.transpose(0, 2, 1)).max(axis=(1, 2))
    assert np.all(asym <= TOL * np.abs(blocks).max(axis=(1, 2)))
    m = sel[1]
    alone_b, alone_l = _condense(prob, kq, precision, macros=slice(m, m + 1))
    assert np.array_equal(alone_b[0], blocks[m]) and np.array_equal(alone_l[0], loads[m])
