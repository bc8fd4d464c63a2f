"""Host logic of the serving loop: request stream and scheduler (CPU)."""

import numpy as np

from pymhm_amd.serving import COEFFICIENT, SOURCE, Request, Scheduler, poisson_requests


def test_poisson_requests_shape_and_mix():
    reqs = poisson_requests(4000, rate=50.0, nproblems=3, coefficient_fraction=0.3, seed=7)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(np.mean(np.diff(t)) - 1 / 50.0) < 0.1 / 50.0  # exponential gaps, mean 1/rate
    kinds = np.array([r.kind == COEFFICIENT for r in reqs])
    assert abs(kinds.mean() - 0.3) < 0.03
    assert {r.problem for r in reqs} == {0, 1, 2}
    assert poisson_requests(50, 10.0, 3, seed=7)[10] == poisson_requests(50, 10.0, 3, seed=7)[10]


def _req(rid, t, p, kind):
    return Request(rid, t, p, kind, rid)


def test_scheduler_fifo_batches_and_coefficient_barrier():
    s = Scheduler(2, max_columns=3)
    seq = [_req(0, 0.0, 0, SOURCE), _req(1, 0.1, 0, SOURCE), _req(2, 0.2, 1, SOURCE), _req(3, 0.3, 0, COEFFICIENT),
           _req(4, 0.4, 0, SOURCE), _req(5, 0.5, 0, SOURCE), _req(6, 0.6, 0, SOURCE), _req(7, 0.7, 0, SOURCE),
           _req(8, 0.8, 0, SOURCE)]
    for r in seq:
        s.push(r)
    out = []
    while (b := s.next_batch()) is not None:
        p, kind, reqs = b
        assert all(r.problem == p and r.kind == kind for r in reqs)
        out.append((p, kind, [r.rid for r in reqs]))
    # the run of sources stops at the coefficient request, which goes alone;
    # runs are cut at max_columns; problems are served oldest head first
    assert out == [(0, SOURCE, [0, 1]), (1, SOURCE, [2]), (0, COEFFICIENT, [3]), (0, SOURCE, [4, 5, 6]),
                   (0, SOURCE, [7, 8])]
    assert s.pending() == 0


def test_scheduler_preserves_per_problem_order():
    reqs = poisson_requests(300, rate=100.0, nproblems=3, coefficient_fraction=0.3, seed=3)
    s = Scheduler(3, max_columns=4)
    for r in reqs:
        s.push(r)
    served = {p: [] for p in range(3)}
    while (b := s.next_batch()) is not None:
        served[b[0]].extend(r.rid for r in b[2])
    for p in range(3):
        assert served[p] == [r.rid for r in reqs if r.problem == p]
