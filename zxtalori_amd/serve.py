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

import math
import threading
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


def run_batched(jobs, *, graphs=True, keep_results=False, clock=time.perf_counter, span=64):
    """All jobs admitted at once and stepped to the end in one batch (no memory budget, no
    lanes: with every job admitted at the start, nothing waits on a costlier job's steps).

    Each build stage and force variant of all near-field jobs is one launch, and the steps
    that end no job run as one launch of a captured graph of up to `span` steps
    (Batch.advance).  A job's latency runs from the start to the end of the step that
    finished it."""
    outcomes = [None] * len(jobs)
    with Batch(graphs=graphs, streams=2) as batch:
        start = clock()
        ids = [batch.add(req.group, count=req.count, seed=req.seed, kind=req.kind, steps=req.steps,
                         neighbourhood=req.neighbourhood) for req in jobs]
        open_ = set(range(len(jobs)))

        def collect():
            t = clock() - start
            for index in sorted(open_):
                state = batch.status(ids[index])
                if state == "running":
                    continue
                req = jobs[index]
                out = Outcome(req, state, batch.error(ids[index]) if state == "failed" else "", 0.0, t)
                if state == "done":
                    out.stats = batch.stats(ids[index])
                    if keep_results:
                        out.quaternions = batch.quaternions(ids[index])
                        out.separation = batch.point_values(ids[index], "separation")
                        if req.kind == "spacing":
                            out.indices = batch.point_indices(ids[index])
                outcomes[index] = out
                open_.discard(index)

        while batch.running:
            ended, _ = batch.advance(span)
            if ended:
                collect()
        batch.wait(0)
        collect()
        stats = summary(outcomes, clock() - start)
        stats["graph_captures"], stats["graph_replays"] = batch.graphs
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


class _Lane:
    """One Batch and the host thread that steps it."""

    def __init__(self, name, graphs, streams, limit=0):
        self.name = name
        self.limit = limit  # most jobs at once (0: no limit)
        self.batch = Batch(graphs=graphs, streams=streams)
        self.queue = []   # request indices, in arrival order
        self.live = {}    # job id -> (request index, admitted time, bytes held)
        self.max_live = 0
        self.steps = 0


class Server:
    """Admits requests by HBM budget into batches and steps them until all are served.

    Requests go to a lane by the pair evaluations one of their steps costs: up to
    `bounds[0]` the small lane (by default 6e8: a near-field N of about 80,000 at
    neighbourhood 20), up to each further bound a further lane, beyond the last the large
    lane.  Each lane is a Batch stepped by its own host thread and kept at most `ahead`
    steps ahead of the device, so the lanes' kernels overlap on the GPU and a small job does
    not wait for a larger job's steps.  Jobs within a lane advance together, one BB step per
    lane step, so a lane's step lasts as long as its jobs' steps together: lanes of similar
    costs keep a job from being slowed by much costlier ones.  The large lane runs one relax
    job at a time, in arrival order, as the per-sample path does: a large job nearly fills
    the device by itself, and sharing it among several large jobs (a medium lane batching
    jobs up to 2e9) made the slowest request slower than on the per-sample path.  Its
    single-pass requests (measure, spacing) join beside the running relax job instead of
    waiting for all of its steps.
    """

    def __init__(self, hbm_budget=64 << 30, *, graphs=True, keep_results=False, bounds=(6e8,), ahead=1,
                 near_streams=1):
        self.hbm_budget = int(hbm_budget)
        self.keep_results = keep_results
        self.bounds = tuple(float(b) for b in bounds)
        self.ahead = int(ahead)
        # The large lane runs one job; the others run their near-field launches on
        # `near_streams` extra streams beside their all-pairs launch (two: half the jobs
        # each, one set's build overlapping the other's force pass).  A batch of all jobs at
        # once gains from two (the 256-job mix: 151 against 159 ms per step); the serving
        # lanes' small jobs do not (p50 68/53/432 ms with one against 87/57/443 with two,
        # p99 alike, on three seeds of the 40-request stream).
        self.lanes = [_Lane(f"lane{k}", graphs, near_streams) for k in range(len(self.bounds))]
        self.lanes.append(_Lane("large", graphs, 0, limit=1))
        self.peak_bytes = 0
        self.max_live = 0
        self._lock = threading.Lock()
        self._held = 0

    @property
    def batch(self):
        """The small lane's batch (most jobs run there)."""
        return self.lanes[0].batch

    def graphs(self):
        captured = replayed = 0
        for lane in self.lanes:
            c, r = lane.batch.graphs
            captured, replayed = captured + c, replayed + r
        return captured, replayed

    def close(self):
        for lane in self.lanes:
            lane.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def lane_of(self, req):
        if req.kind == "relax":
            per_step = req.count * 7500.0 if req.neighbourhood > 0 else req.count ** 2 * ORDERS[req.group.upper()]
        else:
            per_step = req.count * 7500.0
        for lane, bound in zip(self.lanes, self.bounds):
            if per_step <= bound:
                return lane
        return self.lanes[-1]

    def run(self, requests, *, realtime=True, clock=time.perf_counter):
        """Serves `requests`; with realtime=False every request counts as arrived at the
        start.  Returns (outcomes in request order, summary)."""
        for lane in self.lanes:
            lane.queue, lane.live = [], {}
        for index in sorted(range(len(requests)), key=lambda k: requests[k].arrival):
            self.lane_of(requests[index]).queue.append(index)
        outcomes = [None] * len(requests)
        start = clock()
        errors = []

        def serve(lane):
            try:
                self._serve_lane(lane, requests, outcomes, realtime, clock, start)
            except BaseException as e:  # surfaced after the join
                errors.append(e)

        threads = [threading.Thread(target=serve, args=(lane,), name=f"lane-{lane.name}") for lane in self.lanes]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        self.max_live = max(lane.max_live for lane in self.lanes)
        return outcomes, summary(outcomes, clock() - start)

    def _admit(self, lane, requests, index, now):
        """Adds request `index` to the lane's batch if its bytes fit; False if they do not."""
        req = requests[index]
        need = req.estimate_bytes()
        with self._lock:
            if lane.live and self._held + need > self.hbm_budget:
                return False
            self._held += need
            self.peak_bytes = max(self.peak_bytes, self._held)
        job = lane.batch.add(req.group, count=req.count, seed=req.seed, kind=req.kind, steps=req.steps,
                             neighbourhood=req.neighbourhood)
        lane.live[job] = (index, now, need)
        return True

    def _serve_lane(self, lane, requests, outcomes, realtime, clock, start):
        pending = list(lane.queue)  # request indices in arrival order, not yet admitted
        batch = lane.batch
        while pending or lane.live:
            now = clock() - start
            while pending:
                req = requests[pending[0]]
                if realtime and req.arrival > now:
                    break
                if lane.limit and len(lane.live) >= lane.limit:
                    break
                if not self._admit(lane, requests, pending[0], now):
                    break
                pending.pop(0)
            if lane.limit and len(lane.live) >= lane.limit:
                # A lane limited to one job at a time still takes the single-pass requests
                # (measure, spacing) that have arrived, in arrival order, beside the running
                # relax job: they would otherwise wait for all of its steps.
                for index in [i for i in pending if requests[i].kind != "relax"
                              and not (realtime and requests[i].arrival > now)]:
                    if not self._admit(lane, requests, index, now):
                        break
                    pending.remove(index)
            if not lane.live:
                # Idle until the next arrival (or until memory frees up).
                wait = requests[pending[0]].arrival - (clock() - start) if realtime else 0.0
                time.sleep(min(max(wait, 0.0005), 0.02))
                continue
            lane.max_live = max(lane.max_live, len(lane.live))
            batch.wait(self.ahead)
            batch.step()
            lane.steps += 1
            t = clock() - start
            for job in list(lane.live):
                state = batch.status(job)
                index, admitted, need = lane.live[job]
                if state == "running":
                    held = max(need, batch.device_bytes(job))
                    if held != need:
                        with self._lock:
                            self._held += held - need
                            self.peak_bytes = max(self.peak_bytes, self._held)
                        lane.live[job] = (index, admitted, held)
                    continue
                req = requests[index]
                out = Outcome(req, state, batch.error(job) if state == "failed" else "", admitted, t)
                if state == "done":
                    out.stats = batch.stats(job)
                    if self.keep_results:
                        out.quaternions = batch.quaternions(job)
                        out.separation = batch.point_values(job, "separation")
                        if req.kind == "spacing":
                            out.indices = batch.point_indices(job)
                outcomes[index] = out
                batch.remove(job)
                del lane.live[job]
                with self._lock:
                    self._held -= need


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
