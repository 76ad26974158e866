"""Host logic of the request loop (no GPU): the seeded request stream and admission sizes."""
import numpy as np

from zxtalori_amd.serve import Request, poisson_requests, summary, Outcome


def test_poisson_stream_is_seeded_and_in_range():
    a = poisson_requests(500, rate=20.0, seed=7)
    b = poisson_requests(500, rate=20.0, seed=7)
    assert a == b
    arrivals = np.array([r.arrival for r in a])
    assert np.all(np.diff(arrivals) > 0)
    # mean inter-arrival 1/rate within sampling error (500 draws: sd 1/rate/sqrt(500))
    assert abs(np.diff(arrivals, prepend=0.0).mean() - 0.05) < 0.05 * 5 / np.sqrt(500)
    counts = np.array([r.count for r in a])
    assert counts.min() >= 1_000 and counts.max() < 1_000_000
    kinds = {r.kind for r in a}
    assert kinds == {"relax", "measure", "spacing"}
    for r in a:
        assert (r.neighbourhood == 20.0) == (r.count >= 5_000)
        assert (20 <= r.steps <= 100) if r.kind == "relax" else r.steps == 0


def test_estimates_bound_near_field_jobs_above_all_pairs():
    small = Request("O", 4000, 0, "relax", 10, 0.0)
    near = Request("O", 4000, 0, "relax", 10, 20.0)
    assert near.estimate_bytes() > small.estimate_bytes() > 4000 * 100


def test_summary_percentiles():
    reqs = [Request("O", 1000, k, arrival=0.0) for k in range(100)]
    outs = [Outcome(r, "done", "", 0.0, (k + 1) / 1000.0) for k, r in enumerate(reqs)]
    s = summary(outs, 0.1)
    assert s["done"] == 100 and s["failed"] == 0
    assert abs(s["p50_ms"] - 50.5) < 1e-9
    assert abs(s["p99_ms"] - 99.01) < 1e-9
    assert abs(s["jobs_per_s"] - 1000.0) < 1e-9
