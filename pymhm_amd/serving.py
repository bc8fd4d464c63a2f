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
consecutive source requests of one problem into a single multi-column re-solve.
The device computes every column of a re-solve independently, so a request's
result is the same alone and in any batch.

Continuous batching: every problem has its own HIP stream (default priority)
and at most one batch in flight; batches of different problems run
concurrently on the device (a refactorisation of one problem overlaps the
re-solves of another).  Completion is observed by polling HIP events, never
by synchronising the device, so new arrivals are admitted and launched while
earlier batches still run.  Admission is by HBM footprint: the transient
device memory of the batches in flight (source fields, coefficient fields,
results) stays within ``hbm_budget`` bytes - a source batch takes as many
columns as fit, and a batch that does not fit waits for running ones to
retire.
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
    error: str | None = None


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

    def ready(self, busy=()):
        """Problems with queued requests and no batch in flight, oldest head first."""
        heads = sorted((q[0].arrival, q[0].rid, p) for p, q in enumerate(self.queues) if q and p not in busy)
        return [p for _, _, p in heads]

    def head_run(self, p: int):
        """``(kind, requests available for one batch)`` at the head of problem p's queue."""
        q = self.queues[p]
        if q[0].kind == COEFFICIENT:
            return COEFFICIENT, 1
        run = 0
        for r in q:
            if r.kind != SOURCE or run == self.max_columns:
                break
            run += 1
        return SOURCE, run

    def take(self, p: int, count: int):
        q = self.queues[p]
        return [q.popleft() for _ in range(count)]

    def next_batch(self, busy=()):
        ready = self.ready(busy)
        if not ready:
            return None
        p = ready[0]
        kind, run = self.head_run(p)
        return p, kind, self.take(p, run)


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


@dataclass
class _Job:
    problem: int
    kind: str
    reqs: list
    results: list
    event: object
    nbytes: int


class Server:
    """Resident batches of several problems and the request loop.

    ``cases`` are workload cases (``pymhm_amd.workloads``); each gets one
    batch with ``max_columns`` re-solve columns, assembled, factored and
    condensed with its base coefficient before the first request.  With
    ``source_pool`` > 0 each problem keeps that many seeded source fields on
    the device, built up front, and a source request takes field
    ``seed mod source_pool`` (the field evaluation stays out of the served
    time); otherwise the field is evaluated per request.  ``hbm_budget``
    bounds the transient device bytes of the batches in flight (default: half
    of the device memory free after the set-up); ``resident_bytes`` is what
    the resident factors and pools took.
    """

    def __init__(self, cases, precision: str = "fp64", max_columns: int = 8, source_pool: int = 0,
                 hbm_budget: int | None = None):
        import torch

        self.cases = list(cases)
        self.precision = precision
        self.max_columns = int(max_columns)
        self.source_pool = int(source_pool)
        self.pools = []
        self.batches, self.base_kq, self.points, self.current_kq, self.fq = [], [], [], [], []
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
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
            b.raise_on_failure("set-up")
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        self.resident_bytes = max(0, free0 - free1)
        self.hbm_budget = int(0.5 * free1) if hbm_budget is None else int(hbm_budget)
        if self.hbm_budget <= 0:
            raise ValueError("hbm_budget must be positive")
        self.streams = [torch.cuda.Stream() for _ in self.cases]  # default priority
        self.peak_bytes = 0

    def request_bytes(self, p: int, kind: str, ncols: int = 1) -> int:
        """Transient device bytes of one batch of problem p."""
        b, lay = self.batches[p], self.batches[p].layout
        field = self.batches[p].macros * lay.ne * lay.nq * 8
        if kind == COEFFICIENT:
            return field + 2 * b.macros * (b.width * b.width + b.width) * 8
        return ncols * ((0 if self.source_pool else field) + b.macros * b.width * 8)

    def source_fields(self, p: int, seeds):
        """Source fields of the given request seeds for problem p."""
        if self.source_pool:
            idx = [int(s) % self.source_pool for s in seeds]
            return self.pools[p][idx]
        return bump_sources(self.points[p], seeds)

    def _stage(self, p: int, kq):
        case, b = self.cases[p], self.batches[p]
        b.assemble(kq, self.fq[p], case.f_const)
        b.factor(check=False)
        blocks, loads = b.condense(check=False)
        return blocks.clone(), loads.clone()

    def execute(self, p: int, kind: str, reqs: list[Request]):
        """Enqueue one scheduled batch on the current stream; returns one
        result per request (device tensors, complete once the stream is)."""
        if kind == COEFFICIENT:
            (req,) = reqs
            kq = coefficient_field(self.base_kq[p], req.seed)
            self.current_kq[p] = kq
            return [self._stage(p, kq)]
        fq = self.source_fields(p, [r.seed for r in reqs])
        loads = self.batches[p].sources(fq, check=False)
        return [loads[c] for c in range(len(reqs))]

    def _retire(self, job: _Job, done: float):
        import torch

        with torch.cuda.stream(self.streams[job.problem]):
            st = self.batches[job.problem].status()
        err = None
        if st.any():
            bad = np.flatnonzero(st)
            err = f"{job.kind} batch failed (status bits {int(np.bitwise_or.reduce(st))}) in macros {bad[:8].tolist()}"
        for r, res in zip(job.reqs, job.results):
            r.done, r.result, r.error = done, (None if err else res), err

    def run(self, requests: list[Request], realtime: bool = True, progress=None) -> list[Request]:
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` arrivals
        follow the wall clock (a request is only scheduled once its arrival
        time has passed); otherwise all are queued at once.  Sets ``done``
        (seconds after the start, when the host saw the batch complete),
        ``result`` and ``error`` on every request.  ``progress(finished,
        total, seconds)`` (optional) is called after every thousandth
        completed request."""
        import torch

        sched = Scheduler(len(self.cases), self.max_columns)
        order = sorted(requests, key=lambda r: (r.arrival, r.rid))
        nxt = finished = used = 0
        reported = 0
        inflight: dict[int, _Job] = {}
        t0 = time.perf_counter()
        while finished < len(order):
            now = time.perf_counter() - t0
            while nxt < len(order) and (not realtime or order[nxt].arrival <= now):
                sched.push(order[nxt])
                nxt += 1
            progressed = False
            for p in list(inflight):
                job = inflight[p]
                if job.event.query():
                    self._retire(job, time.perf_counter() - t0)
                    used -= job.nbytes
                    finished += len(job.reqs)
                    del inflight[p]
                    progressed = True
                    if progress is not None and finished // 1000 > reported:
                        reported = finished // 1000
                        progress(finished, len(order), time.perf_counter() - t0)
            for p in sched.ready(inflight):
                kind, run = sched.head_run(p)
                per = self.request_bytes(p, kind, 1)
                count = 1 if kind == COEFFICIENT else min(run, max(0, (self.hbm_budget - used) // per))
                need = self.request_bytes(p, kind, count) if count else per
                if count == 0 or used + need > self.hbm_budget:
                    if not inflight:
                        raise MemoryError(f"a {kind} batch of problem {p} needs {need} bytes, over the "
                                          f"hbm_budget of {self.hbm_budget}")
                    continue
                reqs = sched.take(p, count)
                with torch.cuda.stream(self.streams[p]):
                    results = self.execute(p, kind, reqs)
                    ev = torch.cuda.Event()
                    ev.record()
                inflight[p] = _Job(p, kind, reqs, results, ev, need)
                used += need
                self.peak_bytes = max(self.peak_bytes, used)
                progressed = True
            if progressed:
                continue
            if inflight:
                if realtime and nxt < len(order):
                    time.sleep(max(0.0, min(order[nxt].arrival - (time.perf_counter() - t0), 5e-5)))
                else:
                    time.sleep(2e-5)
            elif nxt < len(order):
                time.sleep(max(0.0, min(order[nxt].arrival - now, 0.01)))
        return order

    def close(self):
        import torch

        torch.cuda.synchronize()
        for b in self.batches:
            b.close()
        self.batches, self.pools = [], []


def latency_stats(requests: list[Request]) -> dict:
    lat = np.array([r.done - max(r.arrival, 0.0) for r in requests])
    span = max(r.done for r in requests)
    return dict(requests=len(requests), span_s=span, requests_per_s=len(requests) / span,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3))
