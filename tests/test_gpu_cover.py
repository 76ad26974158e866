"""zo_cover on the GPU: the covering radius against a full-orbit 4D Qhull oracle.

Restates the reference's tests/test_covering.py (full_orbit_covering, the stats record,
stale and cached results) and the point values of api.py:155-170.  The oracle is the
reference's own test oracle: scipy's ConvexHull of every signed orbit image on S3, whose
facet farthest from the origin bounds the largest empty cap.  Both cutoff modes are
checked: every signed image (clouds up to 131,072 images) and the near-field site grid
(O, N = 6000: 144,000 images).
"""
import sys

import numpy as np
import pytest
from scipy.spatial import ConvexHull

import zxtalori_amd as zo
from tests.conftest import GROUPS

pytestmark = pytest.mark.gpu
COEFFICIENTS = [1.889832101454757, .573930763512449, .581069217539205, .859789414116394, 1.490131881944242,
                2.791922498626376, 5.481645765756282, 11.12163594341394]


def note(*args):
    print(*args, file=sys.stderr, flush=True)


def full_orbit_covering(q, operations):
    q = q.astype(np.float64)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    ops = operations.astype(np.float64)
    ops /= np.linalg.norm(ops, axis=1, keepdims=True)
    w = q[:, None, :1] * ops[None, :, :1] - np.sum(q[:, None, 1:] * ops[None, :, 1:], axis=-1, keepdims=True)
    v = q[:, None, :1] * ops[None, :, 1:] + q[:, None, 1:] * ops[None, :, :1] + np.cross(q[:, None, 1:], ops[None, :, 1:])
    cloud = np.concatenate((w, v), axis=-1).reshape(-1, 4)
    cloud /= np.linalg.norm(cloud, axis=1, keepdims=True)
    hull = ConvexHull(np.concatenate((cloud, -cloud)))
    return 2 * np.degrees(np.arccos(np.clip(np.min(-hull.equations[:, -1]), -1, 1)))


@pytest.mark.parametrize("group", GROUPS)
def test_full_orbit_and_stale_state(group):
    kind = "transport" if group in ("T", "O", "I") else None
    with zo.generate(group, h=2, count=100, kind=kind) as grid:
        assert grid.stats["covering_radius_deg"] is None
        for stage in range(2):
            if stage:
                previous = grid.stats["covering_radius_deg"]
                grid.relax(5)
                assert grid.stats["covering_radius_deg"] == previous
                assert not grid.stats["covering_current"]
            measured = grid.covering()
            expected = full_orbit_covering(grid.quaternions, grid.operations)
            assert abs(measured["covering_radius_deg"] - expected) < 2e-4, (group, stage, measured["covering_radius_deg"], expected)
            count = 2 * len(grid.operations) * len(grid)
            bound = sum(c * count ** (-(2 * i + 1) / 3) for i, c in enumerate(COEFFICIENTS)) * 360 / np.pi
            assert abs(measured["simplex_bound_deg"] - bound) < 1e-12
            assert abs(measured["covering_excess"] - (measured["covering_radius_deg"] / measured["simplex_bound_deg"] - 1)) < 1e-12
            if stage:
                assert abs(measured["mesh_ratio"] - measured["covering_radius_deg"] / measured["separation_deg"]) < 1e-12
            assert measured["covering_current"]
            assert measured["covering_iteration"] == grid.stats["iteration"]
            assert measured["triangulation_backend"] == "hip"
            assert measured["covering_certificate_probes"] > 0
            assert measured["covering_domain_bound_deg"] >= measured["covering_radius_deg"]
            assert np.isfinite(measured["covering_witness"]).all()


def test_imported_grid_covering_and_regeneration_invalidates():
    q = np.random.default_rng(197).normal(size=(65, 4))
    with zo.from_quaternions(q, "I") as grid:
        result = grid.covering()
        assert abs(result["covering_radius_deg"] - full_orbit_covering(grid.quaternions, grid.operations)) < 2e-4
        grid.set(q, "O")
        assert grid.stats["covering_radius_deg"] is None


def test_grid_mode_against_hull():
    note("cover: O N=6000 (grid mode) vs Qhull")
    with zo.haar("O", 6000, 41) as grid:
        result = grid.covering()
        expected = full_orbit_covering(grid.quaternions, grid.operations)
        assert result["covering_shell_points"] < 2 * 24 * 6000  # the site grid, not every image
        assert abs(result["covering_radius_deg"] - expected) < 2e-4, (result["covering_radius_deg"], expected)


def test_point_values_and_witness():
    with zo.haar("D6", 800, 5) as grid:
        assert grid.point_values("covering") is None
        result = grid.covering()
        radius = grid.point_values("covering")
        faces = grid.point_values("faces")
        assert grid.stats["point_covering_iteration"] == result["iteration"]
        assert abs(radius.max() - result["covering_radius_deg"]) < 1e-4
        assert (radius > 0).all() and (faces >= 4).all() and faces.mean() > 10
        # The witness is a Voronoi vertex: its nearest orbit points lie at the covering radius.
        w = np.asarray(result["covering_witness"], np.float64)
        q = grid.quaternions.astype(np.float64)
        ops = grid.operations.astype(np.float64)
        cloud_w = np.concatenate([q[:, None, :1] * ops[None, :, :1] - np.sum(q[:, None, 1:] * ops[None, :, 1:], -1, keepdims=True),
                                  q[:, None, :1] * ops[None, :, 1:] + q[:, None, 1:] * ops[None, :, :1]
                                  + np.cross(q[:, None, 1:], ops[None, :, 1:])], -1).reshape(-1, 4)
        cloud_w /= np.linalg.norm(cloud_w, axis=1, keepdims=True)
        angles = 2 * np.degrees(np.arccos(np.clip(np.abs(cloud_w @ w), -1, 1)))
        assert abs(angles.min() - result["covering_radius_deg"]) < 1e-3


def test_covering_result_is_kept_for_an_unchanged_sample():
    with zo.haar("O", 1000, 11) as grid:
        first = grid.covering()
        assert not first["covering_cached"] and first["covering_unresolved_caps"] == 0
        again = grid.covering()
        assert again["covering_cached"] and again["covering_radius_deg"] == first["covering_radius_deg"]
        grid.relax(1)
        moved = grid.covering()
        assert not moved["covering_cached"] and moved["covering_radius_deg"] != first["covering_radius_deg"]
        grid.set(grid.quaternions, "T")
        assert not grid.covering()["covering_cached"]


def test_covering_at_half_a_million():
    """tests/test_covering.py:48-53: O at 500,000 orientations (12 million images)."""
    note("cover: O count=500000")
    with zo.generate("O", count=500_000) as grid:
        result = grid.covering()
        assert 0.8 < result["covering_radius_deg"] < 0.95, result["covering_radius_deg"]


def test_too_few_orientations():
    with zo.from_quaternions(np.eye(4, dtype=np.float32)[:3], "O") as grid:
        with pytest.raises(RuntimeError, match="TooFewOrientations"):
            grid.covering()


def test_mesh_export():
    """tests/test_explorer.py:17-22, and every tetrahedron's circumcap is empty with the
    largest at the covering radius."""
    with zo.generate("O", h=2, kind="transport") as grid:
        mesh = grid.triangulation()
        p = mesh["stereographic"].astype(np.float64)
        tets = mesh["tetrahedra"]
        assert len(p) >= len(grid) and len(tets) > 0 and tets.max() < len(p)
        assert grid.stats["covering_current"]
        radius = grid.stats["covering_radius_deg"]
        # Back to S3: q = ((1 - r^2), 2p) / (1 + r^2).
        r2 = np.sum(p * p, axis=1, keepdims=True)
        q = np.concatenate([(1 - r2), 2 * p], axis=1) / (1 + r2)
        np.testing.assert_allclose(q[:len(grid)], grid.quaternions, atol=2e-6)
        a = q[tets]                                   # (t, 4, 4)
        # The circumcap centre c: c . a_i equal for the four points, |c| = 1.
        e = a[:, 1:] - a[:, :1]                       # (t, 3, 4)
        c = np.stack([(-1) ** k * np.linalg.det(np.delete(e, k, axis=2)) for k in range(4)], axis=1)
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        h = np.einsum("tk,tk->t", c, a[:, 0])
        c *= np.sign(h)[:, None]
        caps = 2 * np.degrees(np.arccos(np.clip(np.abs(h), -1, 1)))
        assert abs(caps.max() - radius) < 1e-3, (caps.max(), radius)
        cloud = np.concatenate([q, -q])
        # Empty: no vertex of the mesh strictly inside any cap.
        inside = (c @ cloud.T) > np.abs(h)[:, None] + 1e-6
        assert not inside.any()


def test_mesh_stars_are_closed():
    """tests/test_gpu_delaunay.py:45-60 on S3: no repeated tetrahedron, and around every
    sample point each face is shared by exactly two tetrahedra (a closed star), with every
    circumcap empty."""
    with zo.haar("D4", 400, 23) as grid:
        mesh = grid.triangulation()
        n = len(grid)
        tets = np.sort(mesh["tetrahedra"].astype(np.int64), axis=1)
        assert len(np.unique(tets, axis=0)) == len(tets)
        faces = np.concatenate([np.delete(tets, k, axis=1) for k in range(4)])
        faces = faces[(faces < n).any(axis=1)]
        _, counts = np.unique(faces, axis=0, return_counts=True)
        assert (counts == 2).all(), np.bincount(counts)
        # Every sample point has a star.
        assert set(np.unique(tets[(tets < n).any(axis=1)])) >= set(range(n))


@pytest.mark.parametrize("group", GROUPS)
def test_scheduler_mode_is_accepted(group):
    """tests/test_gao_covering.py:10-21: with the reference's scheduler mode 32 set, the
    covering still matches the full-orbit hull before and after three steps."""
    import ctypes
    kind = "transport" if group in ("T", "O", "I") else None
    with zo.generate(group, h=2, count=100, kind=kind) as grid:
        grid._lib.zo_delaunay_mode.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        grid._lib.zo_delaunay_mode(grid._handle, 32)
        for step in (0, 3):
            if step:
                grid.relax(step)
            measured = grid.covering()
            assert abs(measured["covering_radius_deg"] - full_orbit_covering(grid.quaternions, grid.operations)) < 2e-4
            assert measured["covering_current"] and np.isfinite(measured["covering_witness"]).all()
