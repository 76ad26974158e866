"""zo_generate(kind=2) on the GPU: the T/O/I FCC transports.

Against the reference's CPU-f64 fixtures at h = 2, 6, 8, 12 (the reference's own bar is
3e-6 per quaternion, tests/test_grid_shaders.py:39-64; the device path here computes in
FP64 and is held to 3e-7, FP32 rounding of the output included), against the FP64 oracle
at h = 3 (no fixture), and by the geometry checks of tests/test_generation.py:30-58
at h = 12 and 64.
"""
import sys

import numpy as np
import pytest
from scipy.spatial import cKDTree

import zxtalori_amd as zo
from oracle import transport as tr
from tests.conftest import GOLDEN, GROUPS

pytestmark = pytest.mark.gpu
GROUP_ID = {g: i for i, g in enumerate(GROUPS)}


def note(*args):
    print(*args, file=sys.stderr, flush=True)


def fixtures():
    out = {}
    with np.load(GOLDEN / "reference-small.npz", allow_pickle=False) as d:
        out.update({(g, 2): d[g].copy() for g in "TOI"})
    for name in ("transports-h6.npz", "transports-h8-h12.npz"):
        with np.load(GOLDEN / name, allow_pickle=False) as d:
            out.update({(k.split("_")[0], int(k.split("_")[1])): d[k].copy() for k in d.files})
    return out


def multiply(q, g):
    w = q[..., :1] * g[..., :1] - np.sum(q[..., 1:] * g[..., 1:], axis=-1, keepdims=True)
    v = q[..., :1] * g[..., 1:] + g[..., :1] * q[..., 1:] + np.cross(q[..., 1:], g[..., 1:])
    return np.concatenate((w, v), axis=-1)


def check_geometry(q, operations):
    assert np.isfinite(q).all()
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 1, atol=4e-7, rtol=0)
    assert (q[:, 0] >= -3e-7).all()
    assert np.max(np.abs(q.astype(np.float64) @ operations.astype(np.float64).T) - q[:, :1]) < 3e-6


def minimum_orbit_chord(q, operations):
    q = q.astype(np.float64)
    cloud = multiply(q[:, None, :], operations.astype(np.float64)[None, :, :]).reshape(-1, 4)
    cloud /= np.linalg.norm(cloud, axis=1, keepdims=True)
    cloud = np.concatenate((cloud, -cloud))
    return cKDTree(cloud).query(q, k=2)[0][:, 1].min()


def regenerate(grid, group, h):
    grid._check(grid._lib.zo_generate(grid._handle, GROUP_ID[group], h, 2048, 0, 2))


def test_fixtures():
    note("transport fixtures h = 2, 6, 8, 12")
    with zo.haar("O", 10, 1) as grid:
        for (group, h), expected in sorted(fixtures().items()):
            regenerate(grid, group, h)
            q = grid.quaternions
            stats = grid.stats
            assert stats["grid"] == "transport" and stats["h"] == h and stats["group"] == group
            assert len(q) == len(expected) == stats["base_count"]
            err = np.linalg.norm(q.astype(np.float64) - expected, axis=1)
            note(group, h, len(q), f"max error {err.max():.2e}")
            assert err.max() < 3e-7, (group, h, err.max())


@pytest.mark.parametrize("group", ["T", "O", "I"])
def test_matches_oracle_off_fixture(group):
    with zo.haar("O", 10, 1) as grid:
        for h in (3,):
            regenerate(grid, group, h)
            expected = tr.transport(group, h)
            err = np.linalg.norm(grid.quaternions.astype(np.float64) - expected, axis=1)
            assert err.max() < 3e-7, (group, h, err.max())


@pytest.mark.parametrize("group", ["T", "O", "I"])
def test_geometry_and_no_duplicate_orbits(group):
    """tests/test_generation.py:44-58 at h = 12."""
    with zo.generate(group, h=12, kind="transport") as grid:
        q = grid.quaternions
        ops = grid.operations
        check_geometry(q, ops)
        assert minimum_orbit_chord(q, ops) > 1e-4
        assert grid.stats["generator"].startswith("FCC")


def test_large_lattice():
    note("transport T h = 64")
    with zo.generate("T", h=64, kind="transport") as grid:
        q = grid.quaternions
        assert len(q) == (16 * 64 ** 3 + 24 * 64 ** 2 + 14 * 64 + 3) // 3
        check_geometry(q, grid.operations)
        sep = grid.spacing()["separation_deg"]
        assert 0.5 < sep < 3.0, sep
        # A sample of points against the oracle.
        pick = np.random.default_rng(64).choice(len(q), 500, replace=False)
        err = np.linalg.norm(q[pick].astype(np.float64) - tr.transport("T", 64, pick), axis=1)
        assert err.max() < 3e-7, err.max()


def test_polyhedral_only():
    """tests/test_hopf.py:106-115."""
    with zo.generate("O", h=6, kind="transport") as grid:
        assert len(grid) == 1043 and grid.stats["grid"] == "transport"
    with pytest.raises(ValueError):
        zo.generate("D4", kind="transport")
    with pytest.raises(ValueError):
        zo.generate("O", kind="lattice")
    with zo.generate("O", count=1000, exact=True, kind="transport") as grid:
        assert len(grid) == 1000 and grid.stats["grid"] == "transport"
    with zo.haar("O", 10, 1) as grid:
        with pytest.raises(RuntimeError, match="TransportNeedsTOrOOrI"):
            regenerate(grid, "D4", 3)
        with pytest.raises(RuntimeError, match="InvalidGridResolution"):
            regenerate(grid, "O", 129)
