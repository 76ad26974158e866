"""zo_cover against the reference's own covering numbers.

tests/fixtures/covering-reference.json holds full-orbit Qhull radii for the eleven groups
(inverse Hopf f = 2 for the cyclic and dihedral groups, transports h = 2 for T, O, I), and
the reference's covering QA pages (tests/browser-covering.html:11-35,
tests/browser-parity.html:73-81) pin two larger transports: O at h = 12 (7637 points,
4.0472863997 deg) and O at h = 29 (99,137 points, 1.6397905669 deg), all within 2e-4 deg.
Both large ones run in the near-field site-grid mode.
"""
import json
import sys

import pytest

import zxtalori_amd as zo
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu
REFERENCE = json.loads((GOLDEN / "covering-reference.json").read_text())


def note(*args):
    print(*args, file=sys.stderr, flush=True)


@pytest.mark.parametrize("group", list(REFERENCE))
def test_qhull_radii(group):
    expected = REFERENCE[group]
    kind = "transport" if group in ("T", "O", "I") else None
    with zo.generate(group, h=2, kind=kind) as grid:
        assert len(grid) == expected["count"]
        measured = grid.covering()
        assert abs(measured["covering_radius_deg"] - expected["radius_deg"]) < 2e-4, (
            group, measured["covering_radius_deg"], expected["radius_deg"])
        assert measured["covering_current"]
        grid.relax(1)
        assert not grid.stats["covering_current"]
        refreshed = grid.covering()
        assert refreshed["covering_current"] and refreshed["covering_iteration"] == 1


@pytest.mark.parametrize("h,count,radius", [(12, 7637, 4.0472863997), (29, 99137, 1.6397905669)])
def test_large_transports(h, count, radius):
    note(f"cover: O transport h={h}")
    with zo.generate("O", h=h, kind="transport") as grid:
        assert len(grid) == count
        for _ in range(2):
            grid.set(grid.quaternions, "O")  # a fresh measurement each time, not the cached one
            stats = grid.covering()
            assert not stats["covering_cached"] and stats["covering_current"]
            assert abs(stats["covering_radius_deg"] - radius) < 2e-4, (h, stats["covering_radius_deg"], radius)
            # Site grid, not every signed image.
            assert stats["covering_shell_points"] < 2 * 24 * count
