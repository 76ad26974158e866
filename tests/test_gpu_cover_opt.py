"""Covering optimisation on the GPU (restates the reference's tests/test_covering_optimizer.py).

The optimiser here is greedy cap moves measured by the full covering pass, not the
reference's local LP re-tessellation, so the reference's checks on that machinery's
counters (local error, rounds, skipped jobs, full-pass steps) have no counterpart; every
outcome it pins is kept: the radius drops, the zone, the norms and a held identity stay,
the best sample is the one kept and matches the full-orbit Qhull oracle, separation floors
hold through kicks, and the kick settings and their validation.
"""
import numpy as np
import pytest

import zxtalori_amd as zo
from tests.test_gpu_cover import full_orbit_covering

pytestmark = pytest.mark.gpu


def in_zone(grid):
    q = grid.quaternions.astype(np.float64)
    ops = grid.operations.astype(np.float64)
    return np.max(np.abs(q @ ops.T) - q[:, :1]) < 1e-6


@pytest.mark.parametrize("group,mode", [("O", "tet"), ("D4", "face"), ("I", "face")])
def test_optimizer_lowers_covering_and_keeps_zone(group, mode):
    with zo.haar(group, 300, 11) as grid:
        grid.relax(20)
        before = grid.covering()["covering_radius_deg"]
        grid.optimize_covering(250, mode=mode)
        stats = grid.stats
        assert len(grid) == 300
        assert stats["covering_current"]
        assert stats["covering_radius_deg"] < before
        assert stats["measures_current"] is False
        assert grid.measure()["measures_current"]
        assert stats["covering_radius_deg"] == pytest.approx(stats["covering_optimizer_best_deg"], abs=1e-5)
        assert np.max(np.abs(np.linalg.norm(grid.quaternions, axis=1) - 1)) < 1e-5
        assert in_zone(grid)
        assert full_orbit_covering(grid.quaternions, grid.operations) == pytest.approx(stats["covering_radius_deg"], abs=2e-3)


def test_optimizer_holds_fixed_identity_and_stops_on_best():
    with zo.generate("O", count=400, exact=True, fixed_identity=True) as grid:
        index = grid.stats["identity_index"]
        calls = []
        grid.optimize_covering(600, batch=50, callback=lambda g, s: calls.append(s["covering_optimizer_steps"]) or len(calls) < 3)
        stats = grid.stats
        assert len(calls) == 3 and stats["covering_optimizer_steps"] == 150
        np.testing.assert_array_equal(grid.quaternions[index], [1, 0, 0, 0])
        assert stats["covering_radius_deg"] == pytest.approx(stats["covering_optimizer_best_deg"], abs=1e-5)
        # The reference's local re-tessellation counters have no counterpart: present, null.
        for key in ("covering_optimizer_jobs", "covering_optimizer_mismatches", "covering_optimizer_local_rounds",
                    "covering_optimizer_full_steps", "covering_optimizer_phase_ms"):
            assert key in stats and stats[key] is None, key
        with pytest.raises(ValueError):
            grid.optimize_covering(1, mode="everything")


def test_full_pass_variant_still_optimizes():
    with zo.haar("T", 300, 3) as grid:
        grid.relax(20)
        before = grid.covering()["covering_radius_deg"]
        grid.optimize_covering(150, local=False)
        assert grid.stats["covering_radius_deg"] < before and in_zone(grid)


def test_separation_floor_holds():
    with zo.haar("D4", 400, 9) as grid:
        grid.relax(30)
        before = grid.covering()["covering_radius_deg"]
        floor = 0.95 * grid.measure()["separation_deg"]
        grid.optimize_covering(200, min_separation_deg=floor)
        assert grid.stats["covering_radius_deg"] < before
        assert grid.measure()["separation_deg"] >= floor * (1 - 1e-5)
        with pytest.raises(ValueError):
            grid.optimize_covering(1, min_separation_deg=200)


def test_separation_floor_holds_through_kicks():
    with zo.generate("O", count=300) as grid:
        grid.relax(100)
        start = grid.covering()["covering_radius_deg"]
        floor = 0.99 * grid.measure()["separation_deg"]
        grid.optimize_covering(300, min_separation_deg=floor, patience=50, plateau_gain=0.5)
        assert grid.stats["covering_optimizer_kicks"] >= 4
        assert grid.stats["covering_radius_deg"] < start
        assert grid.measure()["separation_deg"] >= floor * (1 - 1e-5)


def test_kick_settings():
    with zo.haar("O", 300, 4) as grid:
        grid.relax(20)
        grid.optimize_covering(60, patience=5, plateau_gain=0.5, kick=0.1, kick_trust=0.05)
        stats = grid.stats
        assert stats["covering_kick_settings"] == pytest.approx({"patience": 5, "progress": 0.5, "size": 0.1, "trust": 0.05})
        assert stats["covering_optimizer_kicks"] >= 10
        for bad in ({"patience": 0}, {"plateau_gain": 1}, {"kick": 2}, {"kick_trust": 0}):
            with pytest.raises(ValueError):
                grid.optimize_covering(1, **bad)
