"""The oracle's serving-cost helper (the CPU baseline of the serving workload)."""

import numpy as np

from oracle import workloads as ow
from pymhm_amd import workloads as pw


def test_load_only_equals_operator_load():
    """The load-only restatement gives the operators' load bit for bit (2D
    P1/P2 and 3D P2), so the CPU source request times the same arithmetic."""
    for case, spec_of in ((pw.config1(nx=4), ow.grid_spec), (pw.config2(nx=2), ow.grid_spec),
                          (pw.config3(nx=2, refinement=2), ow.cube_spec)):
        sel = [0, 3]
        src = np.random.default_rng(1).random(case.kq[sel].shape)
        spec = spec_of(case, sel)
        spec["fq"] = src
        for r, (_, p, mesh) in enumerate(ow._macro_problems(spec)):
            assert np.array_equal(ow._load_only(spec, mesh, src[r]), p.load), case.name


def test_serving_cost_counts_macros():
    case = pw.config1(nx=4)
    spec = ow.grid_spec(case, [1, 2])
    spec["src"] = np.ones(case.kq[[1, 2]].shape)
    n, t_coef, t_src = ow.serving_cost_spec(spec)
    assert n == 2 and t_coef > 0 and t_src > 0
