"""GPU: batches and the continuous-batching loop give each job the result it has alone.

A job in a batch runs the same kernels on the same data as a Grid running it alone (its
all-pairs pass, BB step and fold are one row of a multi-job launch; near-field passes are
its own), so the bar is bit-identity: positions, stats and per-point values equal to the
last bit, with graph replay on and off, with jobs joining mid-run and finishing at
different steps.
"""
import sys

import numpy as np
import pytest

import zxtalori_amd as zo
from zxtalori_amd.batch import Batch
from zxtalori_amd.serve import Request, Server, poisson_requests, serve_plain

pytestmark = pytest.mark.gpu


def note(*args):
    print(*args, file=sys.stderr, flush=True)


# (group, count, seed, kind, steps, neighbourhood): all pairs and near field, several
# groups, uneven sizes, jobs that end at different steps.
MIX = [
    ("O", 1000, 193, "relax", 12, 0.0),
    ("C2", 3000, 5, "relax", 7, 0.0),
    ("I", 9000, 11, "relax", 9, 20.0),
    ("D4", 6000, 2, "relax", 3, 20.0),
    ("C1", 700, 9, "relax", 15, 0.0),
    ("T", 500, 4, "measure", 0, 0.0),
    ("D6", 30000, 8, "spacing", 0, 0.0),
    ("C3", 2500, 6, "relax", 0, 0.0),
]


def alone(group, count, seed, kind, steps, neighbourhood):
    with zo.haar(group, count, seed) as grid:
        if kind == "relax":
            grid.relax(steps, neighbourhood=neighbourhood)
            if steps == 0:  # zo_relax(0): no step, only the measure that ends a relax call
                grid._check(grid._lib.zo_relax(grid._handle, 0))
        elif kind == "measure":
            grid.relax(0, neighbourhood=neighbourhood)
            grid.measure()
        else:
            grid.spacing()
        idx = grid.nearest_indices() if kind == "spacing" else None
        if idx is not None:
            order = grid.stats["order"]
            idx = (idx[0] * 2 * order + idx[1] * 2 + idx[2]).astype(np.uint32)
        return grid.quaternions, grid.stats, grid.point_values("separation"), idx


def check_same(batch, job, expected, label):
    q, stats, sep, idx = expected
    assert batch.status(job) == "done", (label, batch.error(job))
    np.testing.assert_array_equal(batch.quaternions(job), q, err_msg=label)
    assert batch.stats(job) == stats, label
    got_sep = batch.point_values(job, "separation")
    if sep is None:
        assert got_sep is None, label
    else:
        np.testing.assert_array_equal(got_sep, sep, err_msg=label)
    if idx is not None:
        np.testing.assert_array_equal(batch.point_indices(job), idx, err_msg=label)


@pytest.fixture(scope="module")
def expected():
    out = []
    for spec in MIX:
        note("alone", spec)
        out.append(alone(*spec))
    return out


# Jobs with long steady runs between their ends: steps that end no job run as one launch
# of a graph of several steps (Batch.advance).
SPAN_MIX = [
    ("O", 1500, 3, "relax", 20, 0.0),
    ("I", 12000, 7, "relax", 20, 20.0),
    ("D3", 8000, 1, "relax", 10, 20.0),
    ("C4", 900, 2, "relax", 14, 0.0),
]


@pytest.mark.parametrize("streams", [0, 2])
def test_spans_of_steps_equal_single_steps(streams):
    """Spans of steps replayed as one graph launch give the positions, stats and per-point
    values of each job relaxed alone (bit for bit), as single steps do."""
    note("batch spans, streams", streams)
    with Batch(graphs=True, streams=streams) as batch:
        jobs = [batch.add(g, count=n, seed=s, kind=k, steps=t, neighbourhood=d) for g, n, s, k, t, d in SPAN_MIX]
        taken, calls = [], 0
        while batch.running:
            _, steps = batch.advance(64)
            taken.append(steps)
            calls += 1
            assert calls <= 40
        assert sum(taken) == 20 and max(taken) > 1, taken
        for spec, job in zip(SPAN_MIX, jobs):
            check_same(batch, job, alone(*spec), spec)
        captured, replayed = batch.graphs
        assert replayed >= max(taken), (captured, replayed, taken)


@pytest.mark.parametrize("graphs,streams", [(False, 0), (True, 0), (False, 3), (True, 3)])
def test_alone_equals_in_mix(expected, graphs, streams):
    note("batch mix, graphs", graphs, "streams", streams)
    with Batch(graphs=graphs, streams=streams) as batch:
        jobs = [batch.add(g, count=n, seed=s, kind=k, steps=t, neighbourhood=d) for g, n, s, k, t, d in MIX]
        steps = 0
        while batch.running:
            batch.step()
            steps += 1
            assert steps <= 100
        assert steps == max(t for *_, t, _ in MIX)
        for spec, job, exp in zip(MIX, jobs, expected):
            check_same(batch, job, exp, spec)
        captured, replayed = batch.graphs
        if graphs:
            # Job sets change at steps 3, 7, 9 and 12: each new set is captured on its second step.
            assert captured >= 3 and replayed >= 5, (captured, replayed)
        else:
            assert (captured, replayed) == (0, 0)


def test_jobs_join_mid_run(expected):
    note("batch join mid run")
    with Batch(graphs=True) as batch:
        first = batch.add("O", count=1000, seed=193, steps=12)
        for _ in range(4):
            batch.step()
        second = batch.add("I", count=9000, seed=11, steps=9, neighbourhood=20.0)
        for _ in range(3):
            batch.step()
        third = batch.add("C2", count=3000, seed=5, steps=7)
        batch.run()
        check_same(batch, first, expected[0], "first")
        check_same(batch, second, expected[2], "second")
        check_same(batch, third, expected[1], "third")
        for job in (first, second, third):
            batch.remove(job)
        assert batch.running == 0
        with pytest.raises(KeyError):
            batch.status(first)


def test_batch_errors():
    note("batch errors")
    with Batch() as batch:
        with pytest.raises(ValueError):
            batch.add("X", count=10)
        with pytest.raises(RuntimeError, match="NeighbourhoodOutside2To1000"):
            batch.add("O", count=100, neighbourhood=1.0)
        with pytest.raises(RuntimeError, match="InvalidCount"):
            batch.add("O", quaternions=np.zeros((0, 4), np.float32))
        # A failing job fails alone: duplicate orientations stop its BB step.
        dup = np.tile(np.array([[1.0, 0.0, 0.0, 0.0]], np.float32), (50, 1))
        bad = batch.add("O", quaternions=dup, steps=3)
        good = batch.add("O", count=1000, seed=193, steps=3)
        batch.run()
        assert batch.status(bad) == "failed"
        assert "Duplicate" in batch.error(bad) or "NonFinite" in batch.error(bad), batch.error(bad)
        assert batch.status(good) == "done"
        exp = alone("O", 1000, 193, "relax", 3, 0.0)
        check_same(batch, good, exp, "good next to failed")
        with pytest.raises(RuntimeError, match="UnknownJob"):
            batch.remove(10_000)


@pytest.mark.parametrize("bounds", [(2e9,), (2e6, 2e7)])
def test_server_matches_plain_path(bounds):
    # With bounds of 2e6 and 2e7 pair evaluations the requests spread over three lanes:
    # their host threads then step, capture graphs and create and free jobs at once.
    note(f"server vs plain, lane bounds {bounds}")
    requests = poisson_requests(10, rate=1e6, seed=3, count_range=(500, 40_000), steps_range=(3, 12))
    several = len(bounds) > 1
    with Server(hbm_budget=(256 if several else 48) << 20, keep_results=True, bounds=bounds) as server:
        served, summary = server.run(requests, realtime=False)
        if several:
            assert sum(lane.steps > 0 for lane in server.lanes) >= 2
    plain, plain_summary = serve_plain(requests, realtime=False, keep_results=True)
    assert summary["done"] == len(requests) and plain_summary["done"] == len(requests)
    assert server.max_live > 1  # requests overlapped on the device (the small lane's)
    for a, b in zip(served, plain):
        label = (a.request.group, a.request.count, a.request.kind)
        assert a.state == b.state == "done", label
        np.testing.assert_array_equal(a.quaternions, b.quaternions, err_msg=str(label))
        assert a.stats == b.stats, label
        if a.request.kind == "spacing":
            np.testing.assert_array_equal(a.indices, b.indices, err_msg=str(label))
    assert summary["p50_ms"] is not None and summary["p99_ms"] >= summary["p50_ms"]
