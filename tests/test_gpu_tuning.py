"""Band solve kernel families and their autotuner on the GPU (marked ``gpu``).

Every family the tuner may choose, forced for the first pass and for the
refinement passes, must give the oracle's blocks and loads to 1e-10; a family
that cannot serve a batch fails with ``ValueError`` (the library's
MHM_EINVAL) and the tuner skips it; a tuned choice is cached by batch shape.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

from oracle import workloads as ow  # noqa: E402
from pymhm_amd import _lib, tuning  # noqa: E402
from pymhm_amd import workloads as pw  # noqa: E402

TOL = 1e-10


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def square():
    """16 macros of the config-2 shape (n = 1089, half-bandwidth 66)."""
    case = pw.config2(nx=4)
    print("config2 4x4: oracle", flush=True)
    return case, ow.oracle_config(case)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_solve_family_matches_oracle(square, precision):
    case, ref = square
    fams = tuning.CANDIDATES[precision]
    for first in fams:
        for refine in fams:
            print(f"{precision} solve families {tuning.describe(first, refine)}", flush=True)
            with case.problem.batch(precision) as b:
                b.set_solve_variants(first, refine)
                b.assemble(case.kq, None, case.f_const)
                b.factor()
                blocks, loads = (t.cpu().numpy() for t in b.condense())
            assert rel(blocks, ref["blocks"]) < TOL, (first, refine)
            assert rel(loads, ref["loads"]) < TOL, (first, refine)


def test_unservable_family_fails_and_tuner_skips_it():
    """A 3D macro of half-bandwidth 189 has no matrix-core solve: forcing one
    fails with ValueError, and the tuner settles on a family that serves."""
    case = pw.config3(nx=2, refinement=4)
    print("config3 2^3: forced matrix-core solve on b = 189", flush=True)
    with case.problem.batch("fp64") as b:
        assert b.layout.bandwidth > 111
        b.assemble(case.kq, None, case.f_const)
        b.factor()
        b.set_solve_variants(_lib.SOLVE_MFMA, _lib.SOLVE_AUTO)
        with pytest.raises(ValueError):
            b.condense()
        b.set_solve_variants()
        b.factor()
        want = [t.cpu().numpy().copy() for t in b.condense()]

        def run():
            b.factor(check=False)
            b.condense(check=False)

        first, refine, timings = tuning.tune_batch(b, run, repeats=1, use_cache=False)
        assert timings[(_lib.SOLVE_MFMA, _lib.SOLVE_AUTO)] is None
        assert timings[(first, refine)] is not None
        run()
        got = [t.cpu().numpy() for t in b.condense()]
        for g, w in zip(got, want):
            assert rel(g, w) < TOL


def test_tuned_choice_is_cached(square):
    case, _ = square
    with case.problem.batch("fp64") as b:
        b.assemble(case.kq, None, case.f_const)

        def run():
            b.factor(check=False)
            b.condense(check=False)

        tuning._CACHE.pop(tuning.shape_key(b), None)
        first, refine, timings = tuning.tune_batch(b, run, repeats=1)
        assert timings and b.solve_variants == (first, refine)
        again = tuning.tune_batch(b, run, repeats=1)
        assert again == (first, refine, {})
