"""zo_generate on the GPU: inverse Hopf product grids folded into the zone.

The host builds the grid (checked against the reference's fixture on the CPU, in
test_hopf_host.py); here the folded result must equal the oracle's fold of those rows bit
for bit, match the fixture row by row up to the orbit, and satisfy the reference's
geometry checks.  Restates tests/test_hopf.py:38-51, :78-103, :106-115, :118-126.
"""
import math
import sys

import numpy as np
import pytest

import oracle
import zxtalori_amd as zo
from tests.conftest import GROUPS, ops32
from tests.test_hopf_host import CASES, REFERENCE, count, orbit_chord

pytestmark = pytest.mark.gpu
GROUP_ID = {g: i for i, g in enumerate(GROUPS)}


def note(*args):
    print(*args, file=sys.stderr, flush=True)


def check_geometry(q, operations):
    """tests/test_generation.py:30-35: unit, scalar part >= 0, and the representative
    nearest the identity among its orbit's images."""
    assert np.isfinite(q).all()
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1, atol=4e-7, rtol=0)
    assert (q[:, 0] >= -3e-7).all()
    assert np.max(np.abs(q.astype(np.float64) @ operations.astype(np.float64).T) - q[:, :1]) < 3e-6


def regenerate(grid, group, h=0, count=2048):
    grid._check(grid._lib.zo_generate(grid._handle, GROUP_ID[group], h, count, 0, 0))


def test_fixture_grids_fold_like_the_oracle(operator_fixture):
    note("generate: fixture cases")
    with zo.haar("O", 10, 1) as grid:
        for group, f in CASES:
            regenerate(grid, group, h=f)
            q = grid.quaternions
            stats = grid.stats
            assert stats["grid"] == "hopf" and stats["h"] == f and stats["group"] == group
            assert stats["base_count"] == len(q) == count(group, f)
            assert stats["generator"].startswith("Inverse Hopf product grid")
            ops = ops32(operator_fixture, group)
            np.testing.assert_array_equal(q, oracle.fold(zo.hopf_points(group, f), ops, GROUP_ID[group]),
                                          err_msg=f"{group} f={f}")
            check_geometry(q, ops)
            chord = orbit_chord(q.astype(np.float64), REFERENCE[f"{group}_{f}"], operator_fixture[group])
            assert chord.max() < 1e-6, (group, f, chord.max())


@pytest.mark.parametrize("group", GROUPS)
def test_counts_and_selection(group):
    note(f"generate: counts {group}")
    sizes = {f: count(group, f) for f in range(1, 513) if count(group, f) <= 10_000_000}
    with zo.generate(group, count=1000, exact=True, identity=True) as grid:
        assert len(grid) == 1000 and grid.stats["grid"] == "hopf"
        assert grid.stats["base_count"] == max(n for n in sizes.values() if n <= 999)
        for f in [f for f, n in sizes.items() if n <= 1100]:
            regenerate(grid, group, h=f)
            assert len(grid) == count(group, f) and grid.stats["h"] == f
        for target in (300, 900):
            regenerate(grid, group, count=target)
            assert abs(math.log(len(grid) / target)) == min(abs(math.log(n / target)) for n in sizes.values())


def test_identity_augmentation_keeps_the_grid_record():
    with zo.generate("T", h=3, identity=True) as grid:
        q = grid.quaternions
        stats = grid.stats
        assert stats["grid"] == "hopf" and stats["h"] == 3 and stats["base_count"] == count("T", 3)
        assert np.abs(q[stats["identity_index"], 1:]).sum() < 1e-6


def test_frequency_beyond_the_count_limit():
    with pytest.raises(RuntimeError, match="GridTooLarge"):
        zo.generate("C2", h=100)


def test_transport_kind():
    with pytest.raises(ValueError):
        zo.generate("D4", kind="transport")
    with pytest.raises(ValueError):
        zo.generate("O", kind="lattice")
    with zo.generate("O", h=6, kind="transport") as grid:
        assert len(grid) == 1043 and grid.stats["grid"] == "transport"


def test_generated_grid_relaxes_and_measures():
    with zo.generate("O", h=5) as grid:
        before = grid.measure()
        grid.relax(5)
        after = grid.stats
        assert after["iteration"] == 5 and after["grid"] == "hopf"
        assert before["separation_deg"] > 0 and after["separation_deg"] > 0


def test_ten_million(operator_fixture):
    # tests/test_hopf.py:118-124: near the count limit, in the zone and nearest to the request.
    note("generate: O at 1e7")
    with zo.generate("O", count=10_000_000) as grid:
        q = grid.quaternions
        assert grid.stats["h"] == 281 and len(grid) == count("O", 281)
        assert np.all(q[:, 0] >= -1e-6) and np.abs(np.linalg.norm(q, axis=1) - 1).max() < 1e-5
        rows = zo.hopf_points("O", 281)[:200_000]
        np.testing.assert_array_equal(q[:200_000], oracle.fold(rows, ops32(operator_fixture, "O"), GROUP_ID["O"]))
        assert grid.spacing()["separation_deg"] > 0.3
