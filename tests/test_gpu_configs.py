"""GPU: the batched and serving workloads at their full sizes, and the reference's
spacing-against-all-pairs check.

* 256 relax jobs (12 groups with C1, N log-uniform in [500, 2e5], Haar seed = job index,
  all pairs below 5,000 points and neighbourhood 20 above, 50 BB steps) batched with graph
  replay give every job the positions and stats the plain per-sample loop gives it, to
  the last bit.
* A Poisson request stream (40 requests: relax 20-100 steps, measure, spacing; N
  log-uniform in [1e3, 1e6]) served by the continuous-batching loop gives every request
  the plain path's result, bitwise; a served near-field relax's energy is then checked
  against FP64 pair terms over all of its sources (relative error <= 1e-6).
* tests/test_measures.py:121-130 of the reference: at O, N = 60,000 (seed 3) the
  near-field spacing (4 mean radii) equals the all-pairs separation per point within
  1e-5 degrees, so the cutoff never misses the true nearest orbit point.
"""
import sys

import numpy as np
import pytest

import oracle
import zxtalori_amd as zo
from oracle import dense_f64 as dense
from tests.conftest import ops32
from zxtalori_amd import serve

pytestmark = pytest.mark.gpu


def note(*args):
    print(*args, file=sys.stderr, flush=True)


def same_outcome(a, b, label):
    assert a.state == b.state == "done", (label, a.error, b.error)
    np.testing.assert_array_equal(a.quaternions, b.quaternions, err_msg=str(label))
    assert a.stats == b.stats, label
    if a.separation is None:
        assert b.separation is None, label
    else:
        np.testing.assert_array_equal(a.separation, b.separation, err_msg=str(label))
    if a.request.kind == "spacing":
        np.testing.assert_array_equal(a.indices, b.indices, err_msg=str(label))


def operators(fixture, group):
    if group == "C1":
        return np.array([[1.0, 0.0, 0.0, 0.0]], np.float32), np.array([[1.0, 0.0, 0.0, 0.0]])
    return ops32(fixture, group), np.asarray(fixture[group], np.float64)


def test_batched_256_jobs_equal_plain_loop():
    jobs = serve.mixed_relax_jobs(256, seed=193, steps=50)
    assert sum(j.neighbourhood > 0 for j in jobs) > 100 and min(j.count for j in jobs) < 1000
    note("batched: 256 jobs x 50 steps")
    batched, b = serve.run_batched(jobs, keep_results=True)
    note(f"batched {b['elapsed_s']:.2f} s, graphs {b['graph_captures']}/{b['graph_replays']}; plain loop")
    plain, p = serve.serve_plain(jobs, realtime=False, keep_results=True)
    note(f"plain {p['elapsed_s']:.2f} s")
    assert b["done"] == p["done"] == len(jobs)
    assert b["graph_replays"] > 0
    for index, (x, y) in enumerate(zip(batched, plain)):
        same_outcome(x, y, (index, x.request.group, x.request.count))
        assert x.stats["iteration"] == 50


def test_serving_stream_equals_plain_path(operator_fixture):
    requests = serve.poisson_requests(40, 4.0, seed=5)
    counts = [r.count for r in requests]
    assert max(counts) > 500_000 and {r.kind for r in requests} == {"relax", "measure", "spacing"}
    note("serving: 40 requests, all queued at the start")
    with serve.Server(hbm_budget=64 << 30, keep_results=True) as server:
        served, s = server.run(requests, realtime=False)
        assert server.max_live > 1
    note(f"served in {s['elapsed_s']:.2f} s, p50 {s['p50_ms']:.0f} ms p99 {s['p99_ms']:.0f} ms; plain")
    plain, p = serve.serve_plain(requests, realtime=False, keep_results=True)
    assert s["done"] == p["done"] == len(requests)
    for x, y in zip(served, plain):
        same_outcome(x, y, (x.request.group, x.request.count, x.request.kind))
    # A served near-field relax against FP64 pair terms over all of its sources.
    near = [o for o in served if o.request.kind == "relax" and o.request.neighbourhood > 0
            and o.request.count <= 60_000]
    out = max(near, key=lambda o: o.request.count)
    req = out.request
    ops, ops64 = operators(operator_fixture, req.group)
    cut, r = dense.near_cut(req.count, len(ops), req.neighbourhood)
    note(f"fp64 energy of served {req.group} N={req.count} after {req.steps} steps")
    ref64 = oracle.near_grid64(out.quaternions, ops, ops64, cut, r)
    n = req.count
    energy64 = ref64[:, 4].sum() / (n * len(ops) * (n - 1))
    rel = abs(out.stats["energy"] - energy64) / energy64
    note("served energy rel err vs fp64", rel)
    assert rel <= 1e-6


def test_spacing_near_field_equals_all_pairs():
    """tests/test_measures.py:121-130: O, 60,000 Haar orientations (seed 3)."""
    note("spacing O 60000 vs all pairs")
    with zo.haar("O", 60_000, 3) as grid:
        spacing = grid.spacing()
        near = grid.point_values("separation").copy()
        src, op, sign = grid.nearest_indices()
        assert spacing["quey_energy"] is None
        assert (src >= 0).all() and (src < 60_000).all() and (op < 24).all()
        grid.measure()
        full = grid.point_values("separation")
        assert np.abs(near - full).max() < 1e-5
        assert abs(spacing["separation_deg"] - grid.stats["separation_deg"]) < 1e-5
