"""Online serving of local-stage requests against resident factors.

The online half of the reference's offline/online split
(``core/offline.py:121-127``, ``OfflineLocalProblem.response``, over the
resident factors of ``execution/cuda.py:165-288``): several problems (macro
grids) stay factored on the device, and requests arrive as a Poisson stream.
A request names one problem and either

* brings a new diffusion coefficient - the stage runs again for that problem
  (assembly, factorisation, condensation; the new blocks and loads are its
  result), or
* brings a new source field - the condensed loads of a re-solve against the
  resident factors (``LocalBatch.sources``) are its result.

The scheduler keeps each problem's requests in arrival order (a source request
sees the coefficient of the last coefficient request before it) and batches
consecutive source requests of one problem into a single multi-column re-solve
(continuous batching).  The device computes every column of a re-solve
independently, so a request's result is the same alone and in any batch.
"""

from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

COEFFICIENT, SOURCE = "coefficient", "source"


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start of the run
    problem: int    # index of the served problem
    kind: str       # COEFFICIENT or SOURCE
    seed: int
    done: float = float("nan")
    result: object = field(default=None, repr=False)


def poisson_requests(count: int, rate: float, nproblems: int, coefficient_fraction: float = 0.3,
                     seed: int = 20261205) -> list[Request]:
    """``count`` requests with exponential inter-arrival times (mean 1/rate),
    problems uniform over ``nproblems``, a ``coefficient_fraction`` share of
    coefficient requests and source requests otherwise."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    probs = rng.integers(0, nproblems, count)
    kinds = rng.random(count) < coefficient_fraction
    seeds = rng.integers(0, 2**31 - 1, count)
    return [Request(i, float(t[i]), int(probs[i]), COEFFICIENT if kinds[i] else SOURCE, int(seeds[i]))
            for i in range(count)]


class Scheduler:
    """Per-problem FIFO queues.  ``next_batch`` serves the problem whose head
    request arrived first: a coefficient request alone, or the run of source
    requests at the head of the queue (up to ``max_columns``, stopping at the
    next coefficient request)."""

    def __init__(self, nproblems: int, max_columns: int):
        if max_columns < 1:
            raise ValueError("max_columns must be >= 1")
        self.queues = [deque() for _ in range(nproblems)]
        self.max_columns = int(max_columns)

    def push(self, req: Request) -> None:
        self.queues[req.problem].append(req)

    def pending(self) -> int:
        return sum(len(q) for q in self.queues)

    def next_batch(self):
        heads = [(q[0].arrival, q[0].rid, p) for p, q in enumerate(self.queues) if q]
        if not heads:
            return None
        _, _, p = min(heads)
        q = self.queues[p]
        if q[0].kind == COEFFICIENT:
            return p, COEFFICIENT, [q.popleft()]
        out = []
        while q and q[0].kind == SOURCE and len(out) < self.max_columns:
            out.append(q.popleft())
        return p, SOURCE, out


def bump_sources(points, seeds, bumps: int = 8, width=(0.02, 0.08)):
    """Source fields at quadrature points: sums of ``bumps`` Gaussian bumps
    with seeded centres, widths and amplitudes in the unit box.  ``points`` is
    a device tensor ``(M, elements, q, d)``; returns ``(len(seeds), M,
    elements, q)``."""
    import torch

    d = points.shape[-1]
    out = torch.empty((len(seeds),) + tuple(points.shape[:-1]), dtype=torch.float64, device=points.device)
    for c, s in enumerate(seeds):
        rng = np.random.default_rng(int(s))
        centres = rng.random((bumps, d))
        widths = rng.uniform(width[0], width[1], bumps)
        amps = rng.uniform(0.5, 2.0, bumps)
        f = torch.zeros(points.shape[:-1], dtype=torch.float64, device=points.device)
        for k in range(bumps):
            ctr = torch.as_tensor(centres[k], dtype=torch.float64, device=points.device)
            r2 = ((points - ctr) ** 2).sum(-1)
            f += amps[k] * torch.exp(-0.5 * r2 / widths[k] ** 2)
        out[c] = f
    return out


def coefficient_field(base_kq, seed: int):
    """A new coefficient from the base field's family: ``exp(s log K + b)`` with
    a seeded sign s and shift b in [-1, 1].  A centred Gaussian log K keeps its
    covariance under the sign flip, and a two-valued K keeps its contrast, so
    the new field is as well conditioned as the base one."""
    import torch

    rng = np.random.default_rng(int(seed))
    sgn, b = (1.0 if rng.random() < 0.5 else -1.0), rng.uniform(-1.0, 1.0)
    return torch.exp(sgn * torch.log(base_kq) + b)


class Server:
    """Resident batches of several problems and the request loop.

    ``cases`` are workload cases (``pymhm_amd.workloads``); each gets one
    batch with ``max_columns`` re-solve columns, assembled, factored and
    condensed with its base coefficient before the first request.  With
    ``source_pool`` > 0 each problem keeps that many seeded source fields on
    the device, built up front, and a source request takes field
    ``seed mod source_pool`` (the field evaluation stays out of the served
    time); otherwise the field is evaluated per request.
    """

    def __init__(self, cases, precision: str = "fp64", max_columns: int = 8, source_pool: int = 0):
        import torch

        self.cases = list(cases)
        self.precision = precision
        self.max_columns = int(max_columns)
        self.source_pool = int(source_pool)
        self.pools = []
        self.batches, self.base_kq, self.points, self.current_kq, self.fq = [], [], [], [], []
        for case in self.cases:
            prob = case.problem
            kq = torch.as_tensor(case.kq, dtype=torch.float64, device="cuda")
            self.fq.append(None if case.fq is None else torch.as_tensor(case.fq, dtype=torch.float64, device="cuda"))
            pts = torch.as_tensor(prob.quadrature_points(), dtype=torch.float64, device="cuda")
            b = prob.batch(precision, max_columns=self.max_columns)
            self.batches.append(b)
            self.base_kq.append(kq)
            self.points.append(pts)
            self.current_kq.append(kq)
            if self.source_pool:
                self.pools.append(bump_sources(pts, [1000003 * (len(self.pools) + 1) + k
                                                     for k in range(self.source_pool)]))
            self._stage(len(self.batches) - 1, kq)
        torch.cuda.synchronize()

    def source_fields(self, p: int, seeds):
        """Source fields of the given request seeds for problem p."""
        if self.source_pool:
            idx = [int(s) % self.source_pool for s in seeds]
            return self.pools[p][idx]
        return bump_sources(self.points[p], seeds)

    def _stage(self, p: int, kq):
        case, b = self.cases[p], self.batches[p]
        b.assemble(kq, self.fq[p], case.f_const)
        b.factor()
        blocks, loads = b.condense()
        return blocks.clone(), loads.clone()

    def execute(self, p: int, kind: str, reqs: list[Request]):
        """Run one scheduled batch; returns one result per request."""
        if kind == COEFFICIENT:
            (req,) = reqs
            kq = coefficient_field(self.base_kq[p], req.seed)
            self.current_kq[p] = kq
            return [self._stage(p, kq)]
        fq = self.source_fields(p, [r.seed for r in reqs])
        loads = self.batches[p].sources(fq)
        return [loads[c] for c in range(len(reqs))]

    def run(self, requests: list[Request], realtime: bool = True) -> list[Request]:
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` arrivals
        follow the wall clock (a request is only scheduled once its arrival
        time has passed); otherwise all are queued at once.  Sets ``done``
        (seconds after the start) and ``result`` on every request."""
        import torch

        sched = Scheduler(len(self.cases), self.max_columns)
        order = sorted(requests, key=lambda r: (r.arrival, r.rid))
        nxt = 0
        t0 = time.perf_counter()
        finished = 0
        while finished < len(order):
            now = time.perf_counter() - t0
            while nxt < len(order) and (not realtime or order[nxt].arrival <= now):
                sched.push(order[nxt])
                nxt += 1
            batch = sched.next_batch()
            if batch is None:
                time.sleep(max(0.0, min(order[nxt].arrival - now, 0.01)))
                continue
            p, kind, reqs = batch
            results = self.execute(p, kind, reqs)
            torch.cuda.synchronize()
            done = time.perf_counter() - t0
            for r, res in zip(reqs, results):
                r.done, r.result = done, res
            finished += len(reqs)
        return order

    def close(self):
        for b in self.batches:
            b.close()
        self.batches, self.pools = [], []


def latency_stats(requests: list[Request]) -> dict:
    lat = np.array([r.done - max(r.arrival, 0.0) for r in requests])
    span = max(r.done for r in requests)
    return dict(requests=len(requests), span_s=span, requests_per_s=len(requests) / span,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3))
