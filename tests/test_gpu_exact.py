"""zo_generate_exact on the GPU: exact counts by empty-cap insertion.

Restates the reference's tests/test_hopf.py:84-94 and :106-110 (base grid and count) and
tests/test_workbench.py:7-50 (every group at 1000 with a fixed identity: no clustered
insertions, the zone and the identity kept through relaxation; thin-shell counts).
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import zxtalori_amd as zo
from tests.conftest import GROUPS
from tests.test_hopf_host import count as hopf_count

pytestmark = pytest.mark.gpu


def orbit_cloud(q, ops):
    w = q[:, None, :1] * ops[None, :, :1] - np.sum(q[:, None, 1:] * ops[None, :, 1:], axis=-1, keepdims=True)
    v = q[:, None, :1] * ops[None, :, 1:] + q[:, None, 1:] * ops[None, :, :1] + np.cross(q[:, None, 1:], ops[None, :, 1:])
    cloud = np.concatenate((w, v), axis=-1).reshape(-1, 4)
    return np.concatenate((cloud, -cloud))


@pytest.mark.parametrize("group", GROUPS)
def test_exact_count_identity_and_zone_after_relaxation(group):
    with zo.generate(group, count=1000, exact=True, fixed_identity=True) as grid:
        assert len(grid) == 1000
        stats = grid.stats
        assert stats["base_count"] + stats["inserted_count"] == 1000 and stats["inserted_count"] > 0
        assert stats["grid"] == "hopf" and stats["base_count"] == max(
            n for n in (hopf_count(group, f) for f in range(1, 513)) if n <= 999)
        before = grid.quaternions.copy()
        q = before.astype(np.float64)
        ops = grid.operations.astype(np.float64)
        # Insertions do not cluster nearly co-spherical cell centres.
        assert cKDTree(orbit_cloud(q, ops)).query(q, k=2)[0][:, 1].min() > 0.01
        identity = stats["identity_index"]
        assert identity is not None
        grid.relax(12)
        q = grid.quaternions.astype(np.float64)
        assert np.max(np.abs(q @ ops.T) - q[:, :1]) < 1e-6
        np.testing.assert_array_equal(q[identity], [1, 0, 0, 0])
        assert np.max(np.abs(q - before)) > 1e-5
        assert grid.stats["separation_deg"] > 0
        grid.identity(fixed=False)
        assert not grid.stats["fixed_identity"]


def test_exact_transport():
    with zo.generate("O", count=1000, exact=True, kind="transport") as grid:
        assert len(grid) == 1000 and grid.stats["grid"] == "transport"
        # The largest transport within the count: h = 5, 641 points (h = 6 has 1043).
        assert grid.stats["h"] == 5 and grid.stats["base_count"] == 641 and grid.stats["inserted_count"] == 359


@pytest.mark.parametrize("group,count", [("C6", 84), ("D3", 62), ("D3", 84), ("D4", 84)])
def test_exact_count_thin_shell_hull(group, count):
    with zo.generate(group, count=count, exact=True, identity=True) as grid:
        assert len(grid) == count


@pytest.mark.parametrize("group", ["D2", "T", "O", "I"])
def test_exact_ten_thousand_relaxes(group):
    """tests/browser-parity.html:61-66: 10,000 with the identity, fixed, two steps."""
    with zo.generate(group, count=10_000, exact=True, identity=True) as grid:
        grid.identity(fixed=True)
        grid.relax(2)
        assert len(grid) == 10_000
        assert grid.stats["separation_deg"] > 1


def test_exact_range():
    with pytest.raises(RuntimeError, match="ExactCountRange32To1000000"):
        zo.generate("O", count=31, exact=True)
