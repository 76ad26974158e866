"""measure() at the reference's own sizes: chunk invariance and the log ratio.

Restates tests/test_pair_chunks.py:11-24 (splitting the all-pairs sums over operator and
site ranges changes only rounding, for measure() and one relaxation step) and
tests/test_log_scale.py:19-49 (the log-energy ratio of about a thousand points against a
dense FP64 sum, within 1e-5, for three transports and two random samples).
"""
import numpy as np
import pytest

import zxtalori_amd as zo
from oracle import dense_f64 as dense

pytestmark = pytest.mark.gpu
KEYS = ("quey_energy", "riesz_1", "riesz_2", "riesz_3", "green_ratio", "log_ratio", "separation_deg")


@pytest.mark.parametrize("group,count", [("T", 700), ("O", 1000), ("D4", 400)])
def test_chunked_pair_sums_match_single_dispatch(group, count):
    start = zo.haar(group, count, 5)
    q0 = start.quaternions
    start.close()
    with zo.from_quaternions(q0, group) as grid:
        grid.pair_budget(1)
        whole = dict(grid.measure())
        grid.relax(1)
        moved_whole = grid.quaternions
        grid.set(q0, group)
        grid.pair_budget(1e-5)  # chunks of 64 sites, one operator per dispatch
        chunked = dict(grid.measure())
        grid.relax(1)
        moved_chunked = grid.quaternions
    for key in KEYS:
        assert chunked[key] == pytest.approx(whole[key], rel=2e-5, abs=2e-6), key
    assert np.max(np.abs(moved_chunked - moved_whole)) < 2e-5


RANDOM = np.random.default_rng(936).normal(size=(1000, 4))
RANDOM /= np.linalg.norm(RANDOM, axis=1, keepdims=True)


@pytest.mark.parametrize("group,h", [("T", 5), ("O", 6), ("I", 6), ("O", None), ("I", None)])
def test_log_ratio_against_dense(group, h):
    if h is None:
        grid = zo.from_quaternions(RANDOM, group)
    else:
        grid = zo.generate(group, h=h, kind="transport")
    with grid:
        actual = grid.measure()["log_ratio"]
        expected = dense.log_ratio(grid.quaternions, grid.operations)
    assert abs(actual - expected) < 1e-5, (group, h, actual, expected)
