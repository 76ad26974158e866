"""Inverse Hopf product grids on the host (zo_hopf_points, zo_hopf_frequency), no device.

Restates the reference's tests/test_hopf.py:
* rows against tests/golden/hopf-reference.npz (the reference's fixture: SO3K_Grids
  product_grid(group, f) in float64 for all eleven groups at small frequencies), row by
  row up to the orbit, as the reference compares them (test_hopf.py:38-51);
* counts against the closed form n |V| / |K| (test_hopf.py:23-35);
* the count-to-frequency choice (test_hopf.py:78-98), the count limit (:101-103).
The rows here are the grid before the zone fold; the GPU tests check the folded grid.
"""
import math

import numpy as np
import pytest

import zxtalori_amd as zo
from tests.conftest import GOLDEN, GROUPS

with np.load(GOLDEN / "hopf-reference.npz", allow_pickle=False) as _data:
    REFERENCE = {key: _data[key].copy() for key in _data.files}
CASES = [(key.split("_")[0], int(key.split("_")[1])) for key in REFERENCE if not key.endswith("bound")]
ROOT = {"C2": "I", "C3": "I", "D2": "I", "D3": "I", "T": "I", "I": "I", "C4": "O", "D4": "O", "O": "O",
        "C6": "D6", "D6": "D6", "C1": "I"}
ORDER = {"C2": 2, "C3": 3, "C4": 4, "C6": 6, "D2": 4, "D3": 6, "D4": 8, "D6": 12, "T": 12, "O": 24, "I": 60, "C1": 1}


def ring(f, j):
    mult = 12 if j == f else 6
    return max(mult, mult * int(np.floor(4 * f * np.sin(np.pi * j / (2 * f)) / mult + 0.5)))


def count(group, f):
    """Fibre rotations times sphere nodes over |K| for the group's root grid."""
    root = ROOT[group]
    if root == "O":
        n, nodes = (4 * f if f < 6 else 4 * f - 12 * ((2 * f + 9) // 18)), 4 * f * f + 2
    elif root == "I":
        n, nodes = 5 * f, 10 * f * f + 2
    else:
        n, nodes = 6 * int(np.ceil(4 * f / 6)), 2 + ring(f, f) + 2 * sum(ring(f, j) for j in range(1, f))
    return n * nodes // ORDER[group]


def multiply(q, g):
    w = q[..., :1] * g[..., :1] - np.sum(q[..., 1:] * g[..., 1:], axis=-1, keepdims=True)
    v = q[..., :1] * g[..., 1:] + g[..., :1] * q[..., 1:] + np.cross(q[..., 1:], g[..., 1:])
    return np.concatenate((w, v), axis=-1)


def orbit_chord(actual, expected, operations):
    """Per row: the chord from actual[i] to the nearest point of expected[i]'s orbit."""
    images = multiply(expected[None], operations[:, None])
    return np.minimum(np.linalg.norm(actual[None] - images, axis=-1),
                      np.linalg.norm(actual[None] + images, axis=-1)).min(0)


@pytest.fixture(scope="module")
def operations(operator_fixture):
    return {g: operator_fixture[g].astype(np.float64) for g in GROUPS}


@pytest.mark.parametrize("group,f", CASES)
def test_rows_match_reference_fixture(group, f, operations):
    expected = REFERENCE[f"{group}_{f}"]
    actual = zo.hopf_points(group, f).astype(np.float64)
    assert len(actual) == len(expected) == count(group, f)
    np.testing.assert_allclose(np.linalg.norm(actual, axis=1), 1, atol=3e-7, rtol=0)
    chord = orbit_chord(actual, expected, operations[group])
    assert chord.max() < 1e-6, chord.max()


def test_fixture_check_catches_a_wrong_row(operations):
    """The row comparison fails when one row is replaced by a point of another orbit."""
    expected = REFERENCE["O_5"]
    actual = zo.hopf_points("O", 5).astype(np.float64)
    actual[7] = actual[8]
    assert orbit_chord(actual, expected, operations["O"]).max() > 1e-3


def test_no_orbit_twice(operations):
    """Distinct rows are distinct orbits: the nearest other orbit point is far (the
    reference's minimum_orbit_chord > 1e-4, tests/test_generation.py:38-41)."""
    for group, f in [("T", 3), ("O", 7), ("I", 3), ("D6", 4), ("C6", 3), ("C3", 2), ("D4", 3), ("C1", 2)]:
        ops = operations[group] if group != "C1" else np.array([[1.0, 0, 0, 0]])
        q = zo.hopf_points(group, f).astype(np.float64)
        cloud = multiply(q[:, None], ops[None]).reshape(-1, 4)
        owner = np.repeat(np.arange(len(q)), len(ops))
        chord = np.minimum(np.linalg.norm(q[:, None] - cloud[None], axis=-1),
                           np.linalg.norm(q[:, None] + cloud[None], axis=-1))
        chord[owner[None] == np.arange(len(q))[:, None]] = np.inf
        assert chord.min() > 1e-4, (group, f, chord.min())


@pytest.mark.parametrize("group", GROUPS + ("C1",))
def test_counts_follow_the_closed_form(group):
    for f in range(1, 513):
        n = count(group, f)
        got = zo.api._library().zo_hopf_points(zo.api._group_id(group), f, None, 0)
        assert got == (n if n <= 10_000_000 else 0), (group, f)


@pytest.mark.parametrize("group", GROUPS)
def test_frequency_nearest_the_requested_count(group):
    sizes = {f: count(group, f) for f in range(1, 513) if count(group, f) <= 10_000_000}
    for target in (1, 300, 900, 2048, 123_457, 10_000_000):
        f = zo.hopf_frequency(group, target)
        best = min(abs(math.log(n / target)) for n in sizes.values())
        assert abs(math.log(sizes[f] / target)) == best
        # The lowest such frequency.
        assert all(abs(math.log(sizes[g] / target)) > best for g in sizes if g < f)
    assert zo.hopf_frequency(group, 0) == zo.hopf_frequency(group, 2048)


def test_o_ten_million_takes_frequency_281():
    assert zo.hopf_frequency("O", 10_000_000) == 281 and count("O", 281) <= 10_000_000


def test_frequency_beyond_the_count_limit():
    assert count("C2", 100) > 10_000_000
    with pytest.raises(ValueError):
        zo.hopf_points("C2", 100)
    with pytest.raises(ValueError):
        zo.hopf_points("O", 513)
