"""FCC transports: the FP64 oracle (oracle/transport.py) against the reference's fixtures.

The reference checks its transports against CPU-f64 grids from SO3K_Transports with a
bar of 3e-6 per quaternion (tests/test_grid_shaders.py:39-64, tests/test_generation.py:
44-58).  The oracle, derived from the maps' definitions with no fitted tables, meets
them at the 1e-8 level; the GPU tests compare the device path with both.
"""
import math

import numpy as np
import pytest

from oracle import transport as tr
from tests.conftest import GOLDEN

with np.load(GOLDEN / "reference-small.npz", allow_pickle=False) as _small:
    H2 = {g: _small[g].copy() for g in "TOI"}


def t_count(h):
    return (16 * h ** 3 + 24 * h ** 2 + 14 * h + 3) // 3  # grids.zig:20


@pytest.mark.parametrize("group", ["T", "O", "I"])
def test_oracle_matches_h2_fixture(group):
    q = tr.transport(group, 2)
    assert q.shape == H2[group].shape
    assert np.linalg.norm(q - H2[group], axis=1).max() < 1e-9


def test_planted_mistake_fails_the_fixture():
    """The O source truncation plane taken 1% too far out moves points well past the bar."""
    cu, beta = tr.lattice("O", 2)
    zero = (cu == 0).all(1)
    ho = tr._o_map(np.where(zero[:, None], 1.0, cu), beta * 1.01)
    q = tr.homochoric_to_quaternion(ho)
    q[zero] = [1, 0, 0, 0]
    assert np.linalg.norm(q - H2["O"], axis=1).max() > 3e-6


def test_lattice_counts():
    for h in range(1, 20):
        assert len(tr.lattice("T", h)[0]) == t_count(h)
        assert len(tr.lattice("I", h)[0]) == sum(1 for i in range(-h, h + 1) for j in range(-h, h + 1)
                                                 for k in range(-h, h + 1) if (i + j + k) % 2 == 0)
    assert len(tr.lattice("O", 6)[0]) == 1043  # tests/test_hopf.py:107


def test_cut_index():
    for h in range(1, 129):
        cut = tr.cut_index(h)
        wanted = (1 + math.sqrt(2)) * (h + 0.5)
        assert cut % 2 == 1 and abs(cut - wanted) <= 1 + 1e-12


def test_homochoric_round_trip():
    rng = np.random.default_rng(4)
    q = rng.normal(size=(500, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1
    w = 2 * np.arccos(np.clip(q[:, 0], -1, 1))
    axis = q[:, 1:] / np.linalg.norm(q[:, 1:], axis=1, keepdims=True)
    ho = axis * (0.75 * (w - np.sin(w)))[:, None] ** (1 / 3)
    np.testing.assert_allclose(tr.homochoric_to_quaternion(ho), q, atol=1e-12)


def test_cube_to_ball_preserves_volume_and_bounds():
    """The cubochoric map sends the cube of edge pi^(2/3) onto the ball of radius
    (3 pi / 4)^(1/3): uniform cube samples give uniform ball radii (r^3 uniform)."""
    rng = np.random.default_rng(9)
    x = (rng.random((200_000, 3)) - 0.5) * math.pi ** (2 / 3)
    b = tr.cube_to_ball(x)
    r = np.linalg.norm(b, axis=1) / tr.R_BALL
    assert r.max() <= 1 + 1e-12
    hist, _ = np.histogram(r ** 3, bins=10, range=(0, 1))
    assert np.abs(hist / len(r) - 0.1).max() < 0.005
