"""Pins the CPU oracle (oracle/qvm_oracle.cpp) before it is trusted as the GPU's checker.

Each check is chosen so that a plausible slip (a dropped regulariser, a sign, a swapped
operand, a wrong index) fails it: the reference's operator fixture, its dense FP64 test
models (restated in oracle/dense_f64.py), finite differences of the energy, antipodal
invariance, brute-force pair sets, an exact median by sort, fold invariants, and Haar
uniformity.  Reference tests restated: tests/test_pair_kernels.py:118-159,
tests/test_near_field.py:43-54, tests/test_measures.py:99-115, tests/test_explorer.py:5-15.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import dense_f64 as dense
from tests.conftest import GROUPS, ops32

GROUP_ID = {g: i for i, g in enumerate(GROUPS)}


def random_unit(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=-1, keepdims=True)).astype(np.float32)


@pytest.mark.parametrize("group", GROUPS)
def test_allpairs_rows_match_dense_f64(operator_fixture, group):
    q = random_unit(37, 752)  # test_pair_kernels.py:119-120
    ops = ops32(operator_fixture, group)
    rows = oracle.allpairs(q, ops)
    ref = dense.rows(q, ops)
    np.testing.assert_allclose(rows[:, 4], ref[:, 5], rtol=1e-6)
    np.testing.assert_allclose(rows[:, 3], ref[:, 11], rtol=1e-5)
    assert np.linalg.norm(rows[:, :3] - ref[:, 8:11]) <= 1e-5 * np.linalg.norm(ref[:, 8:11])
    np.testing.assert_allclose(rows[:, 5], ref[:, 6], rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("group", ["C2", "D6", "O", "I"])
def test_antipodal_invariance(operator_fixture, group):
    q = random_unit(37, 752)
    ops = ops32(operator_fixture, group)
    signed = q.copy()
    signed[::2] *= -1
    a, b = oracle.allpairs(q, ops), oracle.allpairs(signed, ops)
    np.testing.assert_allclose(b[:, 3:], a[:, 3:], rtol=3e-6)
    assert np.linalg.norm(b[:, :3] - a[:, :3]) <= 3e-6 * np.linalg.norm(a[:, :3])


def test_force_is_minus_half_energy_gradient(operator_fixture):
    """Sign, left-tangent frame and the factor 2 (test_pair_kernels.py:148-159)."""
    base = random_unit(9, 752).astype(np.float64)
    ops = ops32(operator_fixture, "C2")
    force = oracle.allpairs(base.astype(np.float32), ops)[:, :3]
    h = 1e-6
    for axis in range(3):
        plus, minus = base.copy(), base.copy()
        d = np.zeros(4)
        d[0] = math.cos(h)
        d[axis + 1] = math.sin(h)
        plus[0] = dense.multiply(d, base[0])
        d[axis + 1] *= -1
        minus[0] = dense.multiply(d, base[0])
        derivative = (dense.rows(plus, ops)[:, 5].sum() - dense.rows(minus, ops)[:, 5].sum()) / (2 * h)
        np.testing.assert_allclose(derivative, -2 * force[0, axis], rtol=2e-5, atol=2e-5)


def test_duplicate_and_singleton(operator_fixture):
    q = random_unit(2, 5)
    ops = ops32(operator_fixture, "C2")
    dup = np.stack([q[0], q[0]])
    rows = oracle.allpairs(dup, ops)
    assert (rows[:, 5] == 0).all() and np.isfinite(rows[:, :5]).all()
    single = oracle.allpairs(q[:1], ops)
    assert (single[:, :3] == 0).all()


def haar_folded(group, n, seed):
    return oracle.fold(oracle.haar(n, seed), OPS[group], GROUP_ID[group])


OPS = {}


@pytest.fixture(autouse=True)
def _ops(operator_fixture):
    for g in GROUPS:
        OPS[g] = ops32(operator_fixture, g)


@pytest.mark.parametrize("group,count", [("O", 1000), ("D6", 600), ("C2", 500)])
def test_near_brute_matches_dense(group, count):
    """Windowed near-field sums against the dense FP64 model (test_near_field.py:23-51)."""
    q = haar_folded(group, count, 11)
    ops = OPS[group]
    cut, _ = dense.near_cut(count, len(ops), 20)
    rows = oracle.near_brute(q, ops, cut)
    ref = dense.near_rows(q, ops, 20)
    assert abs(rows[:, 4].sum() - ref[:, 0].sum()) <= 2e-7 * ref[:, 0].sum()
    np.testing.assert_allclose(rows[:, 5], ref[:, 1], rtol=1e-6, atol=5e-9)
    assert np.linalg.norm(rows[:, :3] - ref[:, 2:5]) <= 2e-5 * np.linalg.norm(ref[:, 2:5])
    np.testing.assert_allclose(rows[:, 3], ref[:, 5], rtol=1e-4)


@pytest.mark.parametrize("group,count,delta", [("O", 1500, 20), ("I", 800, 20), ("C2", 2000, 20),
                                               ("D4", 3000, 4), ("T", 1200, 8)])
def test_near_grid_pair_set_is_brute_force(group, count, delta):
    """The cell grid with the shell finds exactly the orbit points within the cutoff:
    the same FP32 pair terms, so only the FP64 summation order differs."""
    q = haar_folded(group, count, 3)
    ops = OPS[group]
    cut, r = dense.near_cut(count, len(ops), delta)
    brute = oracle.near_brute(q, ops, cut)
    grid, stats = oracle.near_grid(q, ops, cut, r)
    np.testing.assert_array_equal(grid[:, 5], brute[:, 5])
    np.testing.assert_allclose(grid[:, :5], brute[:, :5], rtol=1e-11, atol=1e-9)
    assert stats["candidates"] >= stats["accepted"] > 0


@pytest.mark.parametrize("group,count,delta", [("O", 3000, 4), ("I", 2000, 4), ("C2", 4000, 4), ("D6", 2500, 8),
                                               ("T", 1500, 20)])
def test_nearest_grid_is_brute_force_within_the_cutoff(group, count, delta):
    """The spacing pass's grid search returns exactly what the brute-force scan of every
    orbit point within the cutoff returns: ranks and FP32 chords, ties to the lowest rank."""
    q = haar_folded(group, count, 21)
    ops = OPS[group]
    cut, r = dense.near_cut(count, len(ops), delta)
    rank, chord = oracle.nearest_grid(q, ops, cut[0], r)
    ref_rank, ref_chord = oracle.nearest(q, ops, cut[0])
    np.testing.assert_array_equal(chord, ref_chord)
    np.testing.assert_array_equal(rank, ref_rank)
    # and a sub-range of sources gives the same answers
    r2, c2 = oracle.nearest_grid(q, ops, cut[0], r, first=count // 3, count=count // 4)
    np.testing.assert_array_equal(r2, ref_rank[count // 3:count // 3 + count // 4])
    np.testing.assert_array_equal(c2, ref_chord[count // 3:count // 3 + count // 4])


@pytest.mark.parametrize("group,count", [("O", 700), ("C2", 900), ("I", 400)])
def test_near_grid64_is_the_dense_fp64_model(operator_fixture, group, count):
    """FP64 pair terms over the grid's pair set agree with the dense FP64 near-field model
    (tests/test_near_field.py's reference, restated in dense_f64.near_rows) to FP64
    rounding, and the FP32 rows to FP32 rounding."""
    q = haar_folded(group, count, 5)
    ops = OPS[group]
    ops64 = np.asarray(operator_fixture[group], np.float64)
    cut, r = dense.near_cut(count, len(ops), 20)
    rows64 = oracle.near_grid64(q, ops, ops64, cut, r)
    ref = dense.near_rows(q, ops64, 20)  # energy, nearest, fx, fy, fz, curvature
    # The dense model also normalises the identity images; a 1e-7 change of a point moves a
    # close pair's chord by that much, so agreement is at that level, and FP64-tight in sums.
    assert abs(rows64[:, 4].sum() - ref[:, 0].sum()) <= 1e-9 * ref[:, 0].sum()
    np.testing.assert_allclose(rows64[:, 4], ref[:, 0], rtol=1e-6)
    np.testing.assert_allclose(rows64[:, 3], ref[:, 5], rtol=1e-5)
    assert np.linalg.norm(rows64[:, :3] - ref[:, 2:5]) <= 1e-6 * np.linalg.norm(ref[:, 2:5])
    rows32, _ = oracle.near_grid(q, ops, cut, r)
    np.testing.assert_allclose(rows32[:, 4], rows64[:, 4], rtol=2e-6)
    assert np.linalg.norm(rows32[:, :3] - rows64[:, :3]) <= 2e-5 * np.linalg.norm(rows64[:, :3])


@pytest.mark.parametrize("group,count", [("O", 300), ("I", 120), ("C2", 500), ("D3", 400)])
def test_nearest_ranks_match_fp64(group, count):
    q = haar_folded(group, count, 9)
    ops = OPS[group]
    rank, chord = oracle.nearest(q, ops)
    ref_rank, ref_chord = dense.nearest_orbit(q, ops)
    np.testing.assert_allclose(chord, ref_chord, rtol=1e-5, atol=1e-9)
    differ = rank != ref_rank
    # Disagreement only where FP32 cannot order two candidates: equal FP32 chords or an
    # FP64 gap within FP32 resolution.
    if differ.any():
        q64 = q.astype(np.float64)
        for i in np.flatnonzero(differ):
            j, k, s = rank[i] // (2 * len(ops)), (rank[i] % (2 * len(ops))) // 2, rank[i] % 2
            y = dense.multiply(q64[j], ops[k].astype(np.float64))
            y /= np.linalg.norm(y)
            y = -y if s else y
            alt = ((q64[i] - y) ** 2).sum()
            assert abs(alt - ref_chord[i]) <= 4e-7 * ref_chord[i] + 1e-12


def test_median_is_exact():
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 10, 1001, 4096):
        c = rng.normal(size=n).astype(np.float32) * np.float32(1e3)
        c[::7] = 0
        v = np.sort(np.maximum(c, 0))
        expected = max(np.float32(1e-12), np.float32(0.5) * (v[(n - 1) // 2] + v[n // 2]))
        assert oracle.median(c) == expected


def test_fold_scale_invariance():
    """test_measures.py:99-115: rescaled rows fold to the same representatives."""
    q = np.random.default_rng(198).normal(size=(29, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    expected = oracle.fold(q.astype(np.float32), OPS["I"], GROUP_ID["I"])
    scaled = (q * np.geomspace(1e-8, 3e38, len(q))[:, None]).astype(np.float32)
    actual = oracle.fold(scaled, OPS["I"], GROUP_ID["I"])
    assert np.isfinite(actual).all()
    np.testing.assert_allclose(actual, expected, rtol=0, atol=4e-7)
    np.testing.assert_allclose(np.linalg.norm(actual, axis=1), 1, rtol=0, atol=3e-7)


@pytest.mark.parametrize("group", GROUPS)
def test_fold_lands_in_zone_and_is_orbit_invariant(group):
    ops = OPS[group]
    q = random_unit(500, 21)
    folded = oracle.fold(q, ops, GROUP_ID[group])
    f64 = folded.astype(np.float64)
    # The representative is the orbit point nearest the identity (test_explorer.py:12).
    assert np.max(np.abs(f64 @ ops.astype(np.float64).T) - f64[:, :1]) < 1e-6
    assert (folded[:, 0] >= 0).all()
    # Any other orbit point folds to the same representative.
    k = np.random.default_rng(2).integers(0, len(ops), size=len(q))
    moved = dense.multiply(q.astype(np.float64), ops[k].astype(np.float64))
    moved[::3] *= -1
    again = oracle.fold(moved.astype(np.float32), ops, GROUP_ID[group])
    close = np.minimum(np.abs(again - folded).max(1), np.abs(again + folded).max(1))
    # Points within rounding of a zone face may pick the face's other copy.
    assert np.mean(close < 1e-5) > 0.99


def test_haar_repeatable_and_uniform():
    """test_explorer.py:5-15 on the oracle's draw."""
    q = oracle.haar(1000, 48)
    np.testing.assert_array_equal(q, oracle.haar(1000, 48))
    n = len(q)
    assert np.max(np.abs(np.linalg.norm(q, axis=1) - 1)) < 3e-6
    invariant = (q[:, 0].astype(np.float64) ** 2 + q[:, 3].astype(np.float64) ** 2)
    assert abs(invariant.mean() - .5) < 4 / np.sqrt(12 * n)
    assert np.max(np.abs(np.sort(invariant) - np.arange(n) / n)) < 1.95 / np.sqrt(n)


def test_haar_matches_fp64_formula():
    """The integer hash is exact; the float draw within 2 ulp of an FP64 evaluation."""
    def hash32(v):
        x = np.uint64(v)
        m = np.uint64(0xffffffff)
        x ^= x >> np.uint64(16)
        x = (x * np.uint64(0x7feb352d)) & m
        x ^= x >> np.uint64(15)
        x = (x * np.uint64(0x846ca68b)) & m
        return int(x ^ (x >> np.uint64(16)))

    seed = 193
    hs = hash32(seed)
    q = oracle.haar(64, seed)
    for i in range(64):
        base = hash32(i ^ hs)
        ru = [(float(np.float32(hash32((base + k) & 0xffffffff) >> 8)) + .5) / 16777216 for k in range(3)]
        u, a, b = ru[0], 2 * math.pi * ru[1], 2 * math.pi * ru[2]
        expected = [math.sqrt(u) * math.cos(b), math.sqrt(1 - u) * math.sin(a), math.sqrt(1 - u) * math.cos(a),
                    math.sqrt(u) * math.sin(b)]
        np.testing.assert_allclose(q[i], expected, atol=4e-7)


def test_mean_radius_closed_form():
    for count in (2, 24000, 6_000_000, 600_000_000):
        r = oracle.mean_radius(count)
        assert r == pytest.approx(dense.mean_radius(count), rel=1e-14)
        m = 2 * r
        assert math.pi * (m - math.sin(m)) == pytest.approx(math.pi ** 2 / count, rel=1e-9)


def test_bb_step_rules():
    """relax.wgsl:115-165: alpha0 on the first step, BB1 on odd iterations, moves clipped
    to half the mean radius, the fixed identity exact, a flagged step moves nothing."""
    group = "O"
    ops = OPS[group]
    q = haar_folded(group, 200, 1)
    radius = oracle.mean_radius(len(q) * len(ops))
    bb = oracle.BB(len(q))
    q[7] = [1, 0, 0, 0]
    before = q.copy()
    rows = oracle.allpairs(q, ops)
    s0 = bb.step(q, rows, radius, fixed=8)
    assert s0["alpha"] == pytest.approx((0.8 / 3.4) * radius ** 3, rel=1e-12)
    assert s0["step_max"] <= 0.5 * radius * (1 + 1e-6)
    np.testing.assert_array_equal(q[7], [1, 0, 0, 0])
    angle = 2 * np.arccos(np.clip(np.abs((q.astype(np.float64) * before).sum(1)), 0, 1))
    assert angle.max() <= 2 * 0.5 * radius * (1 + 1e-4) + 1e-6
    rows = oracle.allpairs(q, ops)
    s1 = bb.step(q, rows, radius, fixed=8)
    expect = min(max(abs(s1["sy"] / s1["yy"]), 1e-12), max(1.0, 10 * radius))
    assert s1["alpha"] == pytest.approx(expect, rel=1e-12)
    # A duplicate orientation flags the step and nothing moves.
    dup = q.copy()
    dup[1] = dup[0]
    rows = oracle.allpairs(dup, ops)
    frozen = dup.copy()
    oracle.BB(len(q)).step(frozen, rows, radius)
    np.testing.assert_array_equal(frozen, dup)


def test_relax_descends():
    """Three steps do not raise the Quey energy (test_measures.py:73-74)."""
    group = "O"
    ops = OPS[group]
    q = haar_folded(group, 256, 2)
    e0 = oracle.allpairs(q, ops)[:, 4].sum()
    q3, rows, _ = oracle.relax(q, ops, GROUP_ID[group], 3)
    assert rows[:, 4].sum() <= e0
    np.testing.assert_allclose(np.linalg.norm(q3, axis=1), 1, atol=3e-7)


@pytest.mark.parametrize("group", ["C2", "D4", "T", "I"])
def test_gram_log_ratio_matches_explicit_differences(operator_fixture, group):
    """dense.log_ratio (4(1 - c^2) per pair) against measures() (|x-y|^2 |x+y|^2 explicitly)."""
    q = random_unit(120, 936).astype(np.float64)
    ops = operator_fixture[group]
    assert abs(dense.log_ratio(q, ops) - dense.measures(q, ops)["log_ratio"]) < 1e-9
