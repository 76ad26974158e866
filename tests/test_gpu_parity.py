"""GPU parity: the gfx950 kernels through the C ABI against the CPU oracle.

Bars: bit-exact for the fold, every nearest chord^2 and every spacing index (the image
and chord arithmetic is exact-order on both sides); FP32 pair terms with FP64 sums within
the stated relative tolerances for energies, curvatures and forces.  Reference tests
restated: tests/test_pair_kernels.py, tests/test_near_field.py, tests/test_measures.py,
tests/test_point_values.py, tests/test_workbench.py.
"""
import math
import sys

import numpy as np
import pytest

import oracle
import zxtalori_amd as zo
from oracle import dense_f64 as dense
from tests.conftest import GROUPS, ops32

pytestmark = pytest.mark.gpu
GROUP_ID = {g: i for i, g in enumerate(GROUPS)}


def note(*args):
    print(*args, file=sys.stderr, flush=True)


def random_unit(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=-1, keepdims=True)).astype(np.float32)


def check_rows(force, energy, nearest, ref, e_rtol=2e-6, f_rtol=2e-5):
    np.testing.assert_array_equal(nearest, ref[:, 5].astype(np.float32))
    np.testing.assert_allclose(energy, ref[:, 4], rtol=e_rtol)
    np.testing.assert_allclose(force[:, 3], ref[:, 3], rtol=5 * e_rtol)
    err = np.linalg.norm(force[:, :3] - ref[:, :3], axis=1)
    assert err.max() <= f_rtol * np.abs(ref[:, :3]).max() * 10
    assert np.linalg.norm(force[:, :3] - ref[:, :3]) <= f_rtol * np.linalg.norm(ref[:, :3])


def test_fold_is_bit_exact(operator_fixture):
    note("fold 11 groups x 4000")
    q = np.random.default_rng(5).normal(size=(4000, 4)).astype(np.float32)
    q[:100] *= np.geomspace(1e-8, 3e38, 100)[:, None].astype(np.float32)
    with zo.from_quaternions(q, "C2") as grid:
        for group in GROUPS:
            grid.set(q, group)
            expected = oracle.fold(q, ops32(operator_fixture, group), GROUP_ID[group])
            np.testing.assert_array_equal(grid.quaternions, expected)


def test_haar_matches_oracle(operator_fixture):
    note("haar O/I/C2 x 100000")
    for group, seed in (("O", 193), ("I", 7), ("C2", 48)):
        with zo.haar(group, 100_000, seed) as grid:
            q = grid.quaternions
            with zo.haar(group, 100_000, seed) as again:
                np.testing.assert_array_equal(again.quaternions, q)  # repeatable
            raw = oracle.haar(100_000, seed)
            folded = oracle.fold(raw, ops32(operator_fixture, group), GROUP_ID[group])
            # sin/cos differ by ulps between the GPU and libm; the fold may then pick
            # the other copy of a point within rounding of a zone face.
            close = np.abs(q - folded).max(1) < 1e-6
            assert close.mean() > 0.9999, group


@pytest.mark.parametrize("group", GROUPS)
def test_allpairs_rows_small(operator_fixture, group):
    q = random_unit(37, 752)
    with zo.from_quaternions(q, group) as grid:
        force, energy, nearest = grid.force_rows()
        ref = oracle.allpairs(grid.quaternions, ops32(operator_fixture, group))
    check_rows(force, energy, nearest, ref)


@pytest.mark.parametrize("group,count", [("O", 1000), ("I", 600), ("D4", 3000)])
def test_allpairs_rows(operator_fixture, group, count):
    note("allpairs", group, count)
    with zo.haar(group, count, 3) as grid:
        force, energy, nearest = grid.force_rows()
        ref = oracle.allpairs(grid.quaternions, ops32(operator_fixture, group))
        grid.pair_budget(1e-5)  # finer site splits: only rounding may change
        f2, e2, n2 = grid.force_rows()
    check_rows(force, energy, nearest, ref)
    np.testing.assert_array_equal(n2, nearest)
    np.testing.assert_allclose(e2, energy, rtol=1e-6)


@pytest.mark.parametrize("group,count,delta", [("O", 1000, 20), ("D6", 1000, 20), ("C2", 800, 20),
                                               ("I", 3000, 20), ("T", 2000, 4), ("D3", 5000, 8)])
def test_near_rows(operator_fixture, group, count, delta):
    note("near rows", group, count, delta)
    with zo.haar(group, count, 11) as grid:
        grid.relax(0, neighbourhood=delta)
        force, energy, nearest = grid.force_rows()
        q = grid.quaternions
        info = grid.near_info()
    ops = ops32(operator_fixture, group)
    cut, _ = dense.near_cut(count, len(ops), delta)
    ref = oracle.near_brute(q, ops, cut)
    assert info["sites"] >= count
    check_rows(force, energy, nearest, ref)


@pytest.mark.parametrize("copies", [1, 2])
def test_near_rows_with_exact_duplicates(operator_fixture, copies):
    """Duplicates in the near field (I, N=3000, neighbourhood 20: no own image in reach):
    the source's identity image is skipped, every other site at chord 0 is a pair
    (pairs.wgsl:404-408), so the duplicates' nearest chord is 0 and their rows hold the
    clamped pair terms; a relaxation step then fails like the reference's."""
    note("near duplicates", copies)
    q = random_unit(3000, 29)
    q[1:1 + copies] = q[0]
    with zo.from_quaternions(q, "I") as grid:
        grid.relax(0, neighbourhood=20)
        force, energy, nearest = grid.force_rows()
        q = grid.quaternions
        info = grid.near_info()
        assert info["radius"] < 0.62  # below I's own-orbit angle, 0.628 rad
        with pytest.raises(RuntimeError, match="DuplicateOrientations"):
            grid.relax(1)
    assert (nearest[:1 + copies] == 0).all() and (nearest[1 + copies:] > 0).all()
    ops = ops32(operator_fixture, "I")
    cut, _ = dense.near_cut(3000, len(ops), 20)
    check_rows(force, energy, nearest, oracle.near_brute(q, ops, cut))


def test_near_rows_repeat_bitwise_in_crowded_cells(operator_fixture):
    """A wide cutoff over a small sample (C2, N=5073, neighbourhood 20: r = 1.23 rad) crowds
    thousands of sites into the chart's central cells; their in-cell order must still be the
    same every pass, so fresh engines give bit-identical rows and trajectories."""
    note("near crowded repeat C2 5073")
    runs = []
    for _ in range(2):
        with zo.haar("C2", 5073, 67) as grid:
            grid.relax(0, neighbourhood=20)
            rows = grid.force_rows()
            grid.relax(3)
            runs.append((rows, grid.quaternions, grid.stats["energy"]))
    for a, b in zip(runs[0][0], runs[1][0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]
    # and the rows are the oracle's
    with zo.haar("C2", 5073, 67) as grid:
        q = grid.quaternions
    ops = ops32(operator_fixture, "C2")
    cut, _ = dense.near_cut(5073, len(ops), 20)
    force, energy, nearest = runs[0][0]
    check_rows(force, energy, nearest, oracle.near_brute(q, ops, cut))


def test_near_sums_match_dense_reference(operator_fixture):
    """test_near_field.py:43-54."""
    for group, count in (("O", 1000), ("D6", 1000), ("C2", 800)):
        note("near dense", group, count)
        with zo.haar(group, count, 11) as grid:
            grid.relax(20)
            grid.relax(1, neighbourhood=20)
            stats = grid.stats
            assert stats["energy_neighbourhood"] == 20 and stats["relax_neighbourhood"] == 20
            energy, separation = dense.near_energy(grid.quaternions, grid.operations, 20)
            assert stats["energy"] == pytest.approx(energy, rel=2e-6)
            assert stats["separation_deg"] == pytest.approx(separation, rel=1e-4)
            grid.relax(1, neighbourhood=0)
            assert grid.stats["energy_neighbourhood"] is None


def test_relax_matches_oracle(operator_fixture):
    """30 BB steps of the small all-pairs workload: O, N=1000 normals (seed 193)."""
    rng = np.random.default_rng(193)
    q0 = rng.standard_normal((1000, 4))
    q0 /= np.linalg.norm(q0, axis=1, keepdims=True)
    ops = ops32(operator_fixture, "O")
    with zo.from_quaternions(q0.astype(np.float32), "O") as grid:
        start = grid.quaternions
        grid.relax(30)
        stats = grid.stats
        q_gpu = grid.quaternions
    q_cpu, rows, last = oracle.relax(start, ops, GROUP_ID["O"], 30)
    n, order = len(q_cpu), len(ops)
    e_cpu = rows[:, 4].sum() / (n * order * (n - 1))
    assert stats["iteration"] == 30
    assert stats["energy"] == pytest.approx(e_cpu, rel=1e-5)
    sep_cpu = 4 * math.degrees(math.asin(math.sqrt(rows[:, 5].min()) / 2))
    assert stats["separation_deg"] == pytest.approx(sep_cpu, rel=1e-3)
    # Same representatives; positions agree to the BB amplification of FP32 rounding.
    diff = np.minimum(np.abs(q_gpu - q_cpu).max(1), np.abs(q_gpu + q_cpu).max(1))
    assert np.median(diff) < 1e-4
    np.testing.assert_allclose(np.linalg.norm(q_gpu, axis=1), 1, atol=3e-7)


def test_near_relaxation_approaches_all_pairs():
    """test_near_field.py:57-65."""
    outcomes = {}
    for delta in (0, 20):
        note("near vs all pairs", delta)
        with zo.haar("O", 1000, 7) as grid:
            grid.relax(200, neighbourhood=delta)
            grid.relax(0, neighbourhood=0)
            grid.measure()
            outcomes[delta] = grid.stats
    assert outcomes[20]["energy"] == pytest.approx(outcomes[0]["energy"], rel=1e-5)
    assert outcomes[20]["separation_deg"] > 0.95 * outcomes[0]["separation_deg"]


@pytest.mark.parametrize("group,count", [("O", 2000), ("I", 600), ("C2", 5000), ("D6", 3000)])
def test_spacing_indices_all_pairs(operator_fixture, group, count):
    with zo.haar(group, count, 9) as grid:
        stats = grid.spacing()
        src, op, sign = grid.nearest_indices()
        values = grid.point_values("separation")
        q = grid.quaternions
    rank_ref, chord_ref = oracle.nearest(q, ops32(operator_fixture, group))
    order = len(ops32(operator_fixture, group))
    np.testing.assert_array_equal(src * 2 * order + op * 2 + sign, rank_ref)
    np.testing.assert_allclose(values, oracle.degrees(chord_ref), rtol=2e-6, atol=1e-6)
    assert stats["separation_deg"] == pytest.approx(float(values.min()))


@pytest.mark.parametrize("group,count", [("C2", 25000), ("D2", 22000)])
def test_spacing_indices_near_field(operator_fixture, group, count):
    """Above 20,000 orientations spacing() scans a near field of 4 mean radii."""
    note("spacing near", group, count)
    ops = ops32(operator_fixture, group)
    with zo.haar(group, count, 5) as grid:
        grid.spacing()
        src, op, sign = grid.nearest_indices()
        q = grid.quaternions
    cut, _ = dense.near_cut(count, len(ops), 4)
    rank_ref, chord_ref = oracle.nearest(q, ops, float(cut[0]))
    order = len(ops)
    np.testing.assert_array_equal(src * 2 * order + op * 2 + sign, rank_ref)


def test_spacing_ten_million_indices_match_oracle(operator_fixture):
    """Spacing of 1e7 Haar orientations of I (seed 193): every one of the 1e7 indices
    bit-exact against the oracle's grid search of the same near field (4 mean radii), with
    its angle, and a sample of the reported orbit points rebuilt from their indices."""
    n, group = 10_000_000, "I"
    note("spacing I 1e7")
    ops = ops32(operator_fixture, group)
    order = len(ops)
    with zo.haar(group, n, 193) as grid:
        stats = grid.spacing()
        src, op, sign = grid.nearest_indices()
        values = grid.point_values("separation")
        q = grid.quaternions
    rank = (src * 2 * order + op * 2 + sign).astype(np.int64)
    assert (src >= 0).all() and (src < n).all() and (src != np.arange(n)).any()
    assert np.isfinite(values).all() and (values > 0).all()
    assert stats["separation_deg"] == pytest.approx(float(values.min()))
    cut, r = dense.near_cut(n, order, 4)
    note("oracle nearest grid, all 1e7 sources")
    ref_rank, ref_chord = oracle.nearest_grid(q, ops, float(cut[0]), r)
    assert (ref_rank != 0xFFFFFFFF).all()  # every source has an orbit point within the cutoff
    np.testing.assert_array_equal(rank, ref_rank.astype(np.int64))
    np.testing.assert_allclose(values, oracle.degrees(ref_chord), rtol=2e-6, atol=1e-6)
    # size-independent: recompute a sample of the reported chords from the indices
    pick = np.random.default_rng(1).choice(n, 2000, replace=False)
    images = oracle.images(q[src[pick]], ops)  # (order, len(pick), 4)
    y = images[op[pick], np.arange(len(pick))].astype(np.float64)
    y = np.where(sign[pick, None] == 1, -y, y)
    chord = ((q[pick].astype(np.float64) - y) ** 2).sum(1)
    np.testing.assert_allclose(oracle.degrees(chord.astype(np.float32)), values[pick], rtol=1e-4, atol=1e-4)


def test_relaxed_rows_against_fp64_subsample(operator_fixture):
    """Group O, N=1e5 Haar (seed 193), near-field relax (neighbourhood 20): after 100 BB
    steps the GPU's FP32 pair terms with FP64 sums are measured against FP64 pair terms
    over the same pair set for 10,000 sources.  Bars: energy sum within 1e-6 relative,
    every row's energy and curvature within 1e-5, the separation within 1e-3 degrees, and
    the force error within 2e-6 of the pair-force scale (sum of |pair force|: a relaxed
    row's net force is a small difference of large terms) and no more than twice the
    error of the reference's own FP32 formulas (the oracle's FP32 rows)."""
    n, group = 100_000, "O"
    note("fp64 subsample O 1e5")
    ops = ops32(operator_fixture, group)
    ops64 = np.asarray(operator_fixture[group], np.float64)
    with zo.haar(group, n, 193) as grid:
        grid.relax(100, neighbourhood=20)
        force, energy, nearest = grid.force_rows()
        q = grid.quaternions
    cut, r = dense.near_cut(n, len(ops), 20)
    ref = oracle.near_grid64(q, ops, ops64, cut, r, first=0, count=10_000)
    e_err = abs(energy[:10_000].sum() - ref[:, 4].sum()) / ref[:, 4].sum()
    note("energy sum rel err", e_err)
    assert e_err <= 1e-6
    np.testing.assert_allclose(energy[:10_000], ref[:, 4], rtol=1e-5)
    np.testing.assert_allclose(force[:10_000, 3], ref[:, 3], rtol=1e-5)
    diff = np.linalg.norm(force[:10_000, :3] - ref[:, :3], axis=1)
    f_err = np.linalg.norm(diff) / np.linalg.norm(ref[:, :3])
    scaled = np.linalg.norm(diff) / np.linalg.norm(ref[:, 6])
    rows32, _ = oracle.near_grid(q, ops, cut, r, first=0, count=10_000)
    ref_err = np.linalg.norm(rows32[:, :3] - ref[:, :3]) / np.linalg.norm(ref[:, 6])
    note("force rel err (net)", f_err, "vs pair-force scale", scaled, "oracle fp32 vs scale", ref_err)
    assert scaled <= 2e-6
    assert scaled <= 2 * ref_err
    sep = oracle.degrees(nearest[:10_000])
    sep64 = 4 * np.degrees(np.arcsin(np.minimum(1, np.sqrt(ref[:, 5]) / 2)))
    note("row energy max rel err", np.abs(energy[:10_000] / ref[:, 4] - 1).max(),
         "separation max abs err deg", np.abs(sep - sep64).max(), "min sep rel err", sep.min() / sep64.min() - 1)
    assert np.abs(sep - sep64).max() <= 1e-3


@pytest.mark.parametrize("group", GROUPS)
def test_measure_matches_dense(group):
    """test_measures.py:46-92: Quey energy, separation and force RMS of 29 points."""
    q = random_unit(29, 312).astype(np.float64)
    with zo.from_quaternions(q, group) as grid:
        actual = grid.measure()
        expected = dense.measures(grid.quaternions, grid.operations)
        for key in ("quey_energy", "separation_deg", "force_rms"):
            assert actual[key] == pytest.approx(expected[key], rel=1e-5), key
        # The ratios are FP64 on the GPU: far tighter than the reference's 4e-4.
        for key in ("riesz_1", "riesz_2", "riesz_3", "green_ratio", "log_ratio"):
            assert actual[key] == pytest.approx(expected[key], rel=1e-9, abs=1e-12), key
        assert actual["energy_ratios_current"] is True
        signed = grid.quaternions
        signed[::2] *= -1
        grid.set(signed, group)
        again = grid.measure()
        for key in ("quey_energy", "separation_deg", "force_rms"):
            assert again[key] == pytest.approx(actual[key], rel=1e-5), key


def test_close_pair_resolved_by_chords():
    """test_measures.py:79-85: a pair 5e-5 rad apart."""
    near = np.array([[1, 0, 0, 0], [math.cos(5e-5), math.sin(5e-5), 0, 0]], np.float64)
    with zo.from_quaternions(near, "C2") as grid:
        actual = grid.measure()
        expected = dense.measures(grid.quaternions, grid.operations)
        assert actual["separation_deg"] == pytest.approx(expected["separation_deg"], rel=1e-5)


def test_error_paths():
    with zo.from_quaternions(np.array([[1, 0, 0, 0], [1, 0, 0, 0]], np.float32), "C2") as grid:
        with pytest.raises(RuntimeError, match="DuplicateOrientations"):
            grid.measure()
        with pytest.raises(RuntimeError, match="DuplicateOrientations"):
            grid.relax(2)
    with zo.haar("O", 200, 3) as grid:
        with pytest.raises(RuntimeError):
            grid.relax(1, neighbourhood=1)
        with pytest.raises(RuntimeError, match="NonFiniteQuaternion"):
            grid.set(np.array([[np.nan, 0, 0, 0]], np.float32))
        with pytest.raises(RuntimeError, match="ZeroQuaternion"):
            grid.set(np.zeros((1, 4), np.float32))
        assert len(grid) == 200  # a failed import keeps the sample
    with zo.from_quaternions(np.array([[1, 0, 0, 0]], np.float32), "O") as grid:
        with pytest.raises(RuntimeError, match="TooFewOrientations"):
            grid.relax(1)
    with zo.haar("O", 200, 3) as grid:
        with pytest.raises(ValueError):
            grid.optimize_covering(1, mode="everything")


def test_fixed_identity_stays_exact(operator_fixture):
    """test_workbench.py:8-28: the held identity stays [1,0,0,0] and the zone holds."""
    with zo.haar("O", 500, 4) as grid:
        grid.identity(add=True, fixed=True)
        index = grid.stats["identity_index"]
        assert index is not None
        before = grid.quaternions
        grid.relax(12)
        q = grid.quaternions.astype(np.float64)
        ops = grid.operations.astype(np.float64)
        np.testing.assert_array_equal(q[index], [1, 0, 0, 0])
        assert np.max(np.abs(q @ ops.T) - q[:, :1]) < 1e-6
        assert np.max(np.abs(q - before)) > 1e-5
        grid.identity(fixed=False)
        assert not grid.stats["fixed_identity"]


def test_near_field_large_sample(operator_fixture):
    """Group I, N=1,000,000, neighbourhood 20 (the throughput workload): deterministic
    passes, the rows of a 2,000-source sample equal the oracle's grid restatement, and
    their energy sum is within 1e-6 relative of FP64 pair terms over the same pairs."""
    note("near I 1e6")
    ops = ops32(operator_fixture, "I")
    with zo.haar("I", 10_000_000, 193) as big:
        q = big.quaternions[:1_000_000].copy()
    with zo.from_quaternions(q, "I") as grid:
        grid.relax(0, neighbourhood=20)
        f1, e1, n1 = grid.force_rows()
        f2, e2, n2 = grid.force_rows()
        np.testing.assert_array_equal(f1, f2)
        np.testing.assert_array_equal(e1, e2)
        q = grid.quaternions
    cut, r = dense.near_cut(len(q), len(ops), 20)
    first = 123_456  # Haar indices carry no spatial order: a contiguous block is a random sample
    ref, stats = oracle.near_grid(q, ops, cut, r, first, 2000)
    f, e, nr = f1[first:first + 2000], e1[first:first + 2000], n1[first:first + 2000]
    np.testing.assert_array_equal(nr, ref[:, 5].astype(np.float32))
    np.testing.assert_allclose(e, ref[:, 4], rtol=2e-6)
    np.testing.assert_allclose(f[:, 3], ref[:, 3], rtol=1e-5)
    # Forces against FP64 pair terms: the kernel sums coef (x - y), exact for close pairs,
    # where the reference's direction x0 y - y0 x + x x y cancels in FP32; so the FP32
    # oracle is no closer to FP64 than the kernel is.
    ref64 = oracle.near_grid64(q, ops, np.asarray(operator_fixture["I"], np.float64), cut, r, first, 2000)
    e_err = abs(e.sum() - ref64[:, 4].sum()) / ref64[:, 4].sum()
    note("energy sum rel err vs fp64", e_err, "worst row", np.abs(e / ref64[:, 4] - 1).max())
    assert e_err <= 1e-6
    np.testing.assert_allclose(e, ref64[:, 4], rtol=1e-5)
    scale = np.linalg.norm(ref64[:, 6])
    gpu_err = np.linalg.norm(f[:, :3] - ref64[:, :3]) / scale
    cpu_err = np.linalg.norm(ref[:, :3] - ref64[:, :3]) / scale
    note("force error vs fp64 (pair-force scale): gpu", gpu_err, "fp32 oracle", cpu_err)
    assert gpu_err <= 2e-6 and gpu_err <= 2 * cpu_err
    assert (n1 > 0).all()
    assert 7000 < stats["accepted"] / 2000 < 9000


def test_cancel_stops_a_long_relaxation():
    """zo_cancel (exports.zig:79-81) from another thread ends relax between steps: the engine
    waits every ~2e10 pair evaluations, so the flag is seen mid-call (engine.zig:896-915)."""
    import threading
    import time
    note("cancel: I N=1e6, 1000 steps requested")
    with zo.haar("I", 1_000_000, 193) as grid:
        grid.relax(1, neighbourhood=20)
        timer = threading.Timer(0.5, lambda: grid._lib.zo_cancel(grid._handle))
        t0 = time.perf_counter()
        timer.start()
        grid.relax(1000)
        elapsed = time.perf_counter() - t0
        timer.join()
        done = grid.stats["iteration"] - 1
        assert 0 < done < 1000, done
        assert elapsed < 5.0, elapsed
        # The stopped sample is a normal one: measured, and the next call runs in full.
        assert grid.stats["separation_deg"] > 0
        grid.relax(2)
        assert grid.stats["iteration"] == done + 3


def test_near_overflow_is_sized_again_and_redone(operator_fixture):
    """The reference grows the near-field buffers when the shell outgrows them and redoes
    the pass (src/near.zig:102-116, :191-195; src/pairs.zig:391-394).  Forced here with the
    sizing hook: room for one site fewer than the sample has makes the first pass overflow,
    and the rows that come back are those of a pass sized again (equal to an engine that
    never overflowed); room for half the shell makes every pass of a relaxation overflow,
    and the rows it ends with are redone at the final positions (equal to the oracle's)."""
    group, count = "I", 5000
    ops = ops32(operator_fixture, group)
    cut, _ = dense.near_cut(count, len(ops), 20)
    note("near overflow: one site short")
    with zo.haar(group, count, 5) as plain:
        plain.relax(0, neighbourhood=20)
        expected = plain.force_rows()
        shell = plain.near_info()["shell"]
        assert plain._lib.zo_near_resizes(plain._handle) == 0
    assert shell > 100
    with zo.haar(group, count, 5) as grid:
        grid.relax(0, neighbourhood=20)
        grid._lib.zo_near_squeeze(grid._handle, -1)
        rows = grid.force_rows()
        assert grid._lib.zo_near_resizes(grid._handle) == 1
        assert grid.near_info()["shell"] == shell
    for a, b in zip(rows, expected):
        np.testing.assert_array_equal(a, b)
    note("near overflow: mid-relax, half the shell")
    with zo.haar(group, count, 5) as grid:
        grid._lib.zo_near_squeeze(grid._handle, -(shell // 2))
        grid.relax(4, neighbourhood=20)
        assert grid._lib.zo_near_resizes(grid._handle) == 1
        assert grid.stats["iteration"] == 4
        energy = grid.stats["energy"]
        rows = grid.force_rows()
        q = grid.quaternions
        assert grid._lib.zo_near_resizes(grid._handle) == 1
    ref = oracle.near_brute(q, ops, cut)
    force, energies, nearest = rows
    np.testing.assert_array_equal(nearest, ref[:, 5].astype(np.float32))
    np.testing.assert_allclose(energies, ref[:, 4], rtol=2e-6)
    np.testing.assert_allclose(force[:, 3], ref[:, 3], rtol=1e-5)
    assert energy == pytest.approx(ref[:, 4].sum() / (count * len(ops) * (count - 1)), rel=2e-6)
    # Relaxed rows: the net force is a small difference of large pair forces, so it is
    # measured against FP64 pair terms on the pair-force scale (as for the O, N=1e5 rows).
    _, r = dense.near_cut(count, len(ops), 20)
    ref64 = oracle.near_grid64(q, ops, np.asarray(operator_fixture[group], np.float64), cut, r)
    scaled = np.linalg.norm(force[:, :3] - ref64[:, :3]) / np.linalg.norm(ref64[:, 6])
    ref_err = np.linalg.norm(ref[:, :3] - ref64[:, :3]) / np.linalg.norm(ref64[:, 6])
    note("relaxed force err vs pair-force scale", scaled, "oracle fp32", ref_err)
    assert scaled <= 2e-6 and scaled <= 2 * ref_err


def test_batch_end_of_relax_redoes_an_overflowed_pass(operator_fixture):
    """The batch ends a job's relaxation in a pass shared with the other ending jobs; when the
    job's near field overflowed there (room for half the shell, the test hook), the job's
    pass is sized again and redone alone, as finish_relax would: one resize, and the energy
    and separation it reports are the oracle's at its final positions."""
    from zxtalori_amd.batch import Batch
    group, count = "I", 5000
    ops = ops32(operator_fixture, group)
    cut, _ = dense.near_cut(count, len(ops), 20)
    with zo.haar(group, count, 5) as plain:
        plain.relax(0, neighbourhood=20)
        plain.force_rows()
        shell = plain.near_info()["shell"]
    assert shell > 100
    note("batch end of relax after a near overflow")
    with Batch(graphs=False, streams=2) as batch:
        other = batch.add("O", count=3000, seed=1, kind="relax", steps=4, neighbourhood=20)
        job = batch.add(group, count=count, seed=5, kind="relax", steps=4, neighbourhood=20)
        assert batch._lib.zo_batch_near_squeeze(batch._handle, job, -(shell // 2)) == 0
        while batch.running:
            batch.step()
            batch.wait(0)
        assert batch.status(job) == "done" and batch.status(other) == "done"
        assert batch._lib.zo_batch_near_resizes(batch._handle, job) == 1
        assert batch._lib.zo_batch_near_resizes(batch._handle, other) == 0
        stats = batch.stats(job)
        q = batch.quaternions(job).copy()
    ref = oracle.near_brute(q, ops, cut)
    assert stats["energy"] == pytest.approx(ref[:, 4].sum() / (count * len(ops) * (count - 1)), rel=2e-6)
    least = 4 * math.degrees(math.asin(min(1.0, math.sqrt(ref[:, 5].min()) / 2)))
    assert stats["separation_deg"] == pytest.approx(least, abs=1e-5)
