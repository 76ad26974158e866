"""Continuous-batching request loop over a Batch (no counterpart in the reference, whose
engine serves one synchronous call at a time: src/native_api.zig:9-19).

Requests (relax, measure or spacing of a seeded Haar sample) arrive over time.  Each loop
turn admits the requests that have arrived, in arrival order, while their device bytes fit
the HBM budget, then advances every admitted job by one step; jobs that end are read out
and removed, which frees their memory for the next admissions.  Job state (positions, rows,
BB history, near-field sites) stays on the device between steps.  Latency is the time from
a request's arrival to the end of the step that finished it.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .api import EXTENDED_GROUPS, haar
from .batch import Batch

ORDERS = {"C2": 2, "C3": 3, "C4": 4, "C6": 6, "D2": 4, "D3": 6, "D4": 8, "D6": 12, "T": 12, "O": 24, "I": 60,
          "C1": 1}


@dataclass
class Request:
    group: str
    count: int
    seed: int
    kind: str = "relax"      # relax, measure or spacing
    steps: int = 0           # BB steps of a relax request
    neighbourhood: float = 0.0
    arrival: float = 0.0     # seconds after the start of the run

    def estimate_bytes(self):
        """Device bytes the job will hold (a bound used before its first step)."""
        near = self.neighbourhood > 0 or (self.kind == "spacing" and self.count > 20000)
        per_point = 640 if near else 320
        return self.count * per_point + (16 << 20)


@dataclass
class Outcome:
    request: Request
    state: str
    error: str
    admitted: float
    finished: float
    stats: dict = field(default_factory=dict)
    quaternions: np.ndarray | None = None
    separation: np.ndarray | None = None
    indices: np.ndarray | None = None

    @property
    def latency(self):
        return self.finished - self.request.arrival


def poisson_requests(total, rate, seed=0, count_range=(1_000, 1_000_000), steps_range=(20, 100),
                     relax_share=0.7, measure_share=0.15, near_from=5_000, groups=EXTENDED_GROUPS):
    """A seeded request stream: exponential inter-arrival times at `rate` per second, sizes
    log-uniform in count_range, groups uniform; relax requests take steps_range steps, all
    pairs below near_from points and neighbourhood 20 above."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for index in range(int(total)):
        t += rng.exponential(1.0 / rate)
        count = int(np.exp(rng.uniform(np.log(count_range[0]), np.log(count_range[1]))))
        group = groups[int(rng.integers(len(groups)))]
        u = rng.uniform()
        kind = "relax" if u < relax_share else ("measure" if u < relax_share + measure_share else "spacing")
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1)) if kind == "relax" else 0
        neighbourhood = 20.0 if count >= near_from else 0.0
        out.append(Request(group, count, index, kind, steps, neighbourhood, t))
    return out


def mixed_relax_jobs(total=256, seed=193, count_range=(500, 200_000), steps=50, near_from=5_000,
                     groups=EXTENDED_GROUPS):
    """Independent relax jobs of uneven size: group uniform over `groups`, N log-uniform in
    count_range (drawn as floor(exp(U(ln lo, ln hi)))), Haar seed = job index, all pairs
    below near_from points and neighbourhood 20 from there, `steps` BB steps each."""
    rng = np.random.default_rng(seed)
    out = []
    for index in range(int(total)):
        group = groups[int(rng.integers(len(groups)))]
        count = int(np.floor(np.exp(rng.uniform(np.log(count_range[0]), np.log(count_range[1])))))
        out.append(Request(group, count, index, "relax", int(steps), 20.0 if count >= near_from else 0.0, 0.0))
    return out


def run_batched(jobs, *, graphs=True, keep_results=False, clock=time.perf_counter):
    """All jobs admitted at once into one Batch and stepped to the end (no budget)."""
    with Server(hbm_budget=1 << 62, graphs=graphs, keep_results=keep_results) as server:
        outcomes, stats = server.run(jobs, realtime=False, clock=clock)
        stats["graph_captures"], stats["graph_replays"] = server.batch.graphs
    return outcomes, stats


def _percentiles(values):
    if not values:
        return None, None
    v = np.asarray(values, dtype=np.float64)
    return float(np.percentile(v, 50)), float(np.percentile(v, 99))


def summary(outcomes, elapsed):
    """p50/p99 latency (ms) and throughput of a finished run."""
    lat = [o.latency * 1e3 for o in outcomes if o.state == "done"]
    p50, p99 = _percentiles(lat)
    points = sum(o.request.count * max(1, o.request.steps) for o in outcomes if o.state == "done")
    return {"jobs": len(outcomes), "done": len(lat), "failed": sum(o.state != "done" for o in outcomes),
            "elapsed_s": elapsed, "jobs_per_s": len(lat) / elapsed if elapsed > 0 else None,
            "point_steps_per_s": points / elapsed if elapsed > 0 else None, "p50_ms": p50, "p99_ms": p99}


class Server:
    """Admits requests by HBM budget into one Batch and steps it until all are served."""

    def __init__(self, hbm_budget=64 << 30, *, graphs=True, keep_results=False):
        self.hbm_budget = int(hbm_budget)
        self.keep_results = keep_results
        self.batch = Batch(graphs=graphs)
        self.peak_bytes = 0
        self.max_live = 0

    def close(self):
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, requests, *, realtime=True, clock=time.perf_counter):
        """Serves `requests` (sorted by arrival); with realtime=False every request counts
        as arrived at the start.  Returns (outcomes in request order, summary)."""
        pending = sorted(range(len(requests)), key=lambda k: requests[k].arrival)
        outcomes = [None] * len(requests)
        live = {}  # job id -> (request index, admitted time, bytes held)
        start = clock()
        head = 0
        while head < len(pending) or live:
            now = clock() - start
            held = sum(b for _, _, b in live.values())
            while head < len(pending):
                req = requests[pending[head]]
                if realtime and req.arrival > now:
                    break
                need = req.estimate_bytes()
                if live and held + need > self.hbm_budget:
                    break
                job = self.batch.add(req.group, count=req.count, seed=req.seed, kind=req.kind, steps=req.steps,
                                     neighbourhood=req.neighbourhood)
                live[job] = (pending[head], now, need)
                held += need
                head += 1
            if not live:
                # Idle until the next arrival.
                wait = requests[pending[head]].arrival - (clock() - start)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                continue
            self.max_live = max(self.max_live, len(live))
            self.batch.step()
            t = clock() - start
            for job in list(live):
                state = self.batch.status(job)
                index, admitted, need = live[job]
                if state == "running":
                    live[job] = (index, admitted, max(need, self.batch.device_bytes(job)))
                    continue
                req = requests[index]
                out = Outcome(req, state, self.batch.error(job) if state == "failed" else "", admitted, t)
                if state == "done":
                    out.stats = self.batch.stats(job)
                    if self.keep_results:
                        out.quaternions = self.batch.quaternions(job)
                        out.separation = self.batch.point_values(job, "separation")
                        if req.kind == "spacing":
                            out.indices = self.batch.point_indices(job)
                outcomes[index] = out
                self.batch.remove(job)
                del live[job]
            self.peak_bytes = max(self.peak_bytes, sum(b for _, _, b in live.values()))
        return outcomes, summary(outcomes, clock() - start)


def serve_plain(requests, *, realtime=True, keep_results=False, clock=time.perf_counter):
    """The per-sample path: each request on its own Grid, one at a time in arrival order."""
    outcomes = [None] * len(requests)
    start = clock()
    for index in sorted(range(len(requests)), key=lambda k: requests[k].arrival):
        req = requests[index]
        if realtime:
            wait = req.arrival - (clock() - start)
            if wait > 0:
                time.sleep(wait)
        admitted = clock() - start
        grid = haar(req.group, req.count, seed=req.seed)
        try:
            state, error = "done", ""
            try:
                if req.kind == "relax":
                    grid.relax(req.steps, neighbourhood=req.neighbourhood)
                elif req.kind == "measure":
                    grid.relax(0, neighbourhood=req.neighbourhood)
                    grid.measure()
                else:
                    grid.relax(0, neighbourhood=req.neighbourhood)
                    grid.spacing()
            except RuntimeError as e:
                state, error = "failed", str(e)
            out = Outcome(req, state, error, admitted, clock() - start)
            if state == "done":
                out.stats = grid.stats
                if keep_results:
                    out.quaternions = grid.quaternions
                    out.separation = grid.point_values("separation")
                    if req.kind == "spacing":
                        idx = grid.nearest_indices()
                        order = ORDERS[req.group.upper()]
                        out.indices = (idx[0] * 2 * order + idx[1] * 2 + idx[2]).astype(np.uint32)
            outcomes[index] = out
        finally:
            grid.close()
    return outcomes, summary(outcomes, clock() - start)


__all__ = ["Request", "Outcome", "Server", "mixed_relax_jobs", "poisson_requests", "run_batched", "serve_plain", "summary"]
