"""Near-field pair-evaluation throughput of the QVM force pass on one MI355X.

Workload (the configuration the metric is quoted on): group I, the first 1,000,000 of
zo.haar("I", 10_000_000, seed 193), neighbourhood 20 (about 8,000 orbit points per
source).  A step is one near-field energy/force pass over the whole sample: shell of
images near the zone, cell sort of the sites, source order from the sorted sites, the pair
kernel.  Positions are resident in HBM before the timed region.

value = accepted pair evaluations per second over the timed passes (wall clock around
zo_force_passes, which ends with a device wait).  roofline = the pair kernel against the
VALU + transcendental roofline: the per-pair algorithmic work is counted by running the
reference's per-pair near-field function (shaders/pairs.wgsl:246-262, :349-361, plus the
candidate chord and the row sums) on a counting scalar type, and priced at the MI355X
peaks (7.86e13 FP32 lane-ops/s, 9.83e12 transcendentals/s).  Expressed in TFLOP/s-
equivalent against 157.3 TFLOP/s so that frac = pairs/s x (VALU/7.86e13 + TRANS/9.83e12).
cpu_baseline = the CPU restatement in oracle/ (the reference engine cannot run here) on a
bounded sample of the same workload, on all host cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "near-field pair-evals/s at N=1e6 (I) vs roofline; batched p50/p99 vs plain"
PEAK_VALU = 7.86e13       # FP32 lane-ops/s (256 CU x 128 lanes x 2.4 GHz, packed)
PEAK_TRANS = 9.83e12      # transcendental lane-ops/s (quarter rate)
PEAK_FP32_TFLOPS = 157.3  # published dense FP32 vector peak


# ---- the counting scalar type ----------------------------------------------------
class Count:
    valu = 0
    trans = 0

    @classmethod
    def reset(cls):
        cls.valu = cls.trans = 0


class F:
    """A float that counts the VALU and transcendental operations applied to it."""
    __slots__ = ("v",)

    def __init__(self, v):
        self.v = float(v.v if isinstance(v, F) else v)

    @staticmethod
    def _val(o):
        return o.v if isinstance(o, F) else float(o)

    def _op(self, o, fn, n=1):
        Count.valu += n
        return F(fn(self.v, F._val(o)))

    def __add__(self, o): return self._op(o, lambda a, b: a + b)
    def __radd__(self, o): return self._op(o, lambda a, b: b + a)
    def __sub__(self, o): return self._op(o, lambda a, b: a - b)
    def __rsub__(self, o): return self._op(o, lambda a, b: b - a)
    def __mul__(self, o): return self._op(o, lambda a, b: a * b)
    def __rmul__(self, o): return self._op(o, lambda a, b: b * a)

    def __truediv__(self, o):  # v_rcp + v_mul
        Count.trans += 1
        return self._op(o, lambda a, b: a / b)

    def __rtruediv__(self, o):
        Count.trans += 1
        return self._op(o, lambda a, b: b / a)

    def __neg__(self): return F(-self.v)  # a source modifier, free

    def __le__(self, o):
        Count.valu += 1
        return self.v <= F._val(o)

    def __gt__(self, o):
        Count.valu += 1
        return self.v > F._val(o)


def fma(a, b, c):
    Count.valu += 1
    return F(F._val(a) * F._val(b) + F._val(c))


def fmin(a, b):
    Count.valu += 1
    return F(min(F._val(a), F._val(b)))


def fmax(a, b):
    Count.valu += 1
    return F(max(F._val(a), F._val(b)))


def clamp01(a):  # v_med3
    Count.valu += 1
    return F(min(max(F._val(a), 0.0), 1.0))


def select(cond, a, b):
    Count.valu += 1
    return F(F._val(a) if cond else F._val(b))


def fsqrt(a):
    Count.trans += 1
    return F(math.sqrt(F._val(a)))


def fsin(a):  # v_sin takes revolutions: one scale, one transcendental
    Count.valu += 1
    Count.trans += 1
    return F(math.sin(F._val(a)))


def fatan(t, terms):
    """The build's atan as the pair kernel executes it for this cutoff: t times a Taylor
    series of `terms` terms in t^2, or (terms 0) the degree-8 minimax polynomial."""
    z = t * t
    if terms == 0:
        p = fma(z, 0.0028498897, -0.01606863)
        for c in (0.04269152, -0.07504295, 0.10640934, -0.14203644, 0.1999262, -0.33333072, 1.0):
            p = fma(z, p, c)
    else:
        coef = [1.0, -1 / 3, 1 / 5, -1 / 7, 1 / 9, -1 / 11]
        p = F(coef[terms])
        for c in reversed(coef[:terms]):
            p = fma(z, p, c)
    return t * p


def dot4(a, b):
    return fma(a[3], b[3], fma(a[2], b[2], fma(a[1], b[1], a[0] * b[0])))


def canonical_near_pair(x, y, taper_start, inv_width, terms=0):
    """The reference's near-field pair (candidate chord, quey_pair, near_pair, row sums),
    operation for operation, for counting; atan as the build executes it."""
    m = [x[i] - y[i] for i in range(4)]
    dm2 = dot4(m, m)
    p = [x[i] + y[i] for i in range(4)]
    dp2 = dot4(p, p)
    near = fmax(fmin(dm2, dp2), 1e-30)
    far = fmax(fmax(dm2, dp2), 1e-30)
    t = fsqrt(near / far)
    a = fmax(2.0 * fatan(t, terms), 1e-12)
    b = math.pi - a
    a2, b2 = fma(a, a, 1e-9), fma(b, b, 1e-9)
    a3, b3 = fma(a * a, a, 1e-9), fma(b * b, b, 1e-9)
    sine = (2.0 * t) / fma(t, t, 1.0) + 1e-9
    sign = select(dm2 <= dp2, 1.0, -1.0)
    d = [fma(x[0], y[1], -(y[0] * x[1])) + fma(x[2], y[3], -(x[3] * y[2])),
         fma(x[0], y[2], -(y[0] * x[2])) + fma(x[3], y[1], -(x[1] * y[3])),
         fma(x[0], y[3], -(y[0] * x[3])) + fma(x[1], y[2], -(x[2] * y[1]))]
    along = sign * (b2 - a2) / (a2 * b2 * sine)
    curvature = (a3 + b3) / (a3 * b3)
    energy = math.pi / (a * b)
    force = [-along * c for c in d]
    s = clamp01((a - taper_start) * inv_width)
    if s > 0:
        window = 1.0 - s * s * fma(-2.0, s, 3.0)
        slope = (-6.0 * inv_width) * (s * (1.0 - s))
        extra = sign * energy * slope / fmax(fsin(a), 1e-9)
        force = [fma(window, f, extra * c) for f, c in zip(force, d)]
        curvature = curvature * window
        energy = energy * window
    acc = [F(0)] * 5
    acc = [acc[0] + force[0], acc[1] + force[1], acc[2] + force[2], acc[3] + curvature, acc[4] + energy]
    return acc


def pair_op_counts(terms=0):
    """(VALU, transcendental) per accepted near-field pair, averaged over the cutoff ball:
    the window branch runs for the pairs in its outer half radius, 1 - (1/2)^3 = 7/8 of
    them at uniform density, and is skipped for the inner 1/8 (pairs.wgsl:349-361)."""
    x = [F(v) for v in (0.9, 0.1, -0.3, 0.3)]
    y = [F(v) for v in (0.89, 0.12, -0.31, 0.31)]
    Count.reset()
    canonical_near_pair(x, y, F(0.0), F(1.0), terms)  # taper from 0: the window runs
    windowed = (Count.valu, Count.trans)
    Count.reset()
    canonical_near_pair(x, y, F(10.0), F(1.0), terms)  # taper past the pair: no window
    plain = (Count.valu, Count.trans)
    return tuple(plain[k] + 7 / 8 * (windowed[k] - plain[k]) for k in range(2))


# ---- measurement ------------------------------------------------------------------
def cpu_baseline(q, ops, cut, radius, sample, repeats=5):
    """The CPU restatement's force pass over a bounded sample, on all host threads: the
    median pair-evals/s of `repeats` timed runs after one warm-up."""
    import oracle
    cores = oracle.threads(0)
    oracle.near_grid(q, ops, cut, radius, 0, 64)  # warm-up
    runs = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        _, stats = oracle.near_grid(q, ops, cut, radius, 0, sample)
        runs.append((stats["accepted"] / (time.perf_counter() - t0), stats["accepted"]))
    runs.sort()
    value, pairs = runs[len(runs) // 2]
    return {"value": value, "unit": "pair-evals/s", "cores": cores, "kind": "port",
            "sample": f"CPU restatement of the reference algorithm (oracle/qvm_oracle.cpp orc_near_grid): "
                      f"each run builds the sites, shell and cell grid of all {len(q)} sources, then the rows "
                      f"of the first {sample} sources ({pairs:.3e} pairs); median of {repeats} runs after a "
                      f"warm-up, runs {runs[0][0]:.3e} .. {runs[-1][0]:.3e} pair-evals/s"}


def traffic_from_profiles():
    """HBM bytes per near-field force launch from the committed counter summary
    (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, factors from a copy kernel of known size)."""
    path = ROOT / "profiles" / "near_force_i1e6_traffic.md"
    if path.is_file():
        for line in path.read_text().splitlines():
            if line.startswith("hbm_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    return None


def counters_from_profile(name):
    """Wave64 instruction counts per launch (SQ_INSTS_VALU including transcendentals,
    SQ_INSTS_VALU_TRANS_F32, SQ_INSTS_SALU) of a kernel at its workload, from the committed
    counter summary profiles/<name>_counters.md (the counts are fixed by the kernel and
    its input, so they hold for every run of the same workload)."""
    path = ROOT / "profiles" / f"{name}_counters.md"
    out = {}
    if path.is_file():
        for line in path.read_text().splitlines():
            for key in ("valu_per_launch", "trans_per_launch", "salu_per_launch"):
                if line.startswith(key + ":"):
                    out[key] = float(line.split(":", 1)[1])
    return out if len(out) == 3 else None


def mix_pairs(total=256, seed=193):
    """Accepted pairs and sites tested of one force pass of each near-field job of the
    batched mix (serve.mixed_relax_jobs), summed by arctangent variant."""
    import zxtalori_amd as zo
    from zxtalori_amd import serve
    out = {}
    for job in serve.mixed_relax_jobs(total, seed=seed):
        if job.neighbourhood <= 0:
            continue
        with zo.haar(job.group, job.count, job.seed) as grid:
            grid.relax(1, neighbourhood=job.neighbourhood)
            info = grid.near_info()
        row = out.setdefault(info["atan_terms"], {"jobs": 0, "points": 0, "pairs": 0, "candidates": 0, "sites": 0})
        row["jobs"] += 1
        row["points"] += job.count
        row["pairs"] += info["pairs"]
        row["candidates"] += info["candidates"]
        row["sites"] += info["sites"]
    for row in out.values():
        row["candidates_per_pair"] = row["candidates"] / max(1, row["pairs"])
    return out


def issue_rates():
    """The chip's measured issue rates (zo_issue_probe): non-packed FP32 FMA,
    transcendental and v_mfma_f32_16x16x4f32 wave64 instructions per second, each the best
    of three runs."""
    import ctypes
    from zxtalori_amd.api import _library
    lib = _library()
    rates = {}
    for kind, name in ((0, "fma"), (1, "trans"), (2, "mfma_f32_16x16x4")):
        best = 0.0
        for _ in range(3):
            value = ctypes.c_double(0.0)
            if lib.zo_issue_probe(kind, 20000, ctypes.addressof(value)) == 0:
                best = max(best, value.value)
        rates[name] = best
    return rates


# gfx950 issue model: one non-packed FP32 VALU wave64 instruction per 2 cycles per SIMD (the
# published 157.3 TFLOP/s), a transcendental per 8, 1024 SIMDs at 2.4 GHz.
VALU_CYCLES, TRANS_CYCLES, SIMDS, CLOCK_HZ = 2.0, 8.0, 1024, 2.4e9


def valu_busy(counts, kernel_s):
    """Fraction of the published VALU issue peak the kernel's executed instructions fill:
    the cycles its non-transcendental VALU and its transcendental wave64 instructions take at
    the peak issue rates, over the kernel's time on all SIMDs at the peak clock."""
    if not counts or not kernel_s:
        return None
    trans = counts["trans_per_launch"]
    cycles = (counts["valu_per_launch"] - trans) * VALU_CYCLES + trans * TRANS_CYCLES
    return cycles / (SIMDS * CLOCK_HZ * kernel_s)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--group", default="I")
    ap.add_argument("--delta", type=float, default=20.0)
    ap.add_argument("--cpu-sample", type=int, default=int(os.environ.get("ZXT_CPU_SAMPLE", 300000)))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spacing-n", type=int, default=10_000_000, help="0 skips the spacing pass")
    ap.add_argument("--batched-jobs", type=int, default=256, help="0 skips the batched comparison")
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--serving-requests", type=int, default=40, help="0 skips the serving run")
    ap.add_argument("--serving-rate", type=float, default=4.0)
    ap.add_argument("--no-graphs", action="store_true",
                    help="batched and serving runs launch every step eagerly (for kernel-trace profiles)")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the small all-pairs relax (O, 1e3, 30 steps) and near relax (O, 1e5, 100 steps)")
    ap.add_argument("--no-widened", action="store_true",
                    help="skip the generation, covering and exact-count measurements")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (the HBM counter calibration kernel)")
    ap.add_argument("--mix-pairs", action="store_true",
                    help="run only one force pass of each near-field job of the batched mix and print its "
                         "accepted pairs and sites tested by arctangent variant")
    ap.add_argument("--calibrate-issue", action="store_true",
                    help="run only the VALU issue-rate probe (FMA and transcendental chains)")
    args = ap.parse_args()
    if args.mix_pairs:
        print(json.dumps({"mix_pairs_by_atan_terms": mix_pairs()}))
        return
    if args.calibrate_issue:
        print(json.dumps({"issue_rates_wave64_per_s": issue_rates()}))
        return
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: the near-field pass is not sharded "
                                   "across GPUs here (see DESIGN.md for how it would shard)"}))
        sys.exit(2)

    import zxtalori_amd as zo

    print(f"[bench] sample {args.group} N={args.n} delta={args.delta}", file=sys.stderr, flush=True)
    with zo.haar(args.group, 10_000_000, 193) as big:
        q = big.quaternions[:args.n].copy()
    with zo.from_quaternions(q, args.group) as grid:
        q = grid.quaternions
        ops = grid.operations
        grid.relax(0, neighbourhood=args.delta)
        grid.force_passes(max(1, args.warmup))
        print(f"[bench] warm-up done {grid.near_info()}", file=sys.stderr, flush=True)
        grid.profile(True)
        grid.profile_report()  # clears nothing; totals start with this run
        t0 = time.perf_counter()
        grid.force_passes(args.steps)
        dt = time.perf_counter() - t0
        report = grid.profile_report()
        grid.profile(False)
        info = grid.near_info()
    pairs = info["pairs"]
    ms_per_step = dt * 1e3 / args.steps
    value = pairs / (dt / args.steps)
    valu, trans = pair_op_counts(info["atan_terms"])
    kernel_count, kernel_ms = report.get("near_force", (0, float("nan")))
    kernel_s = kernel_ms / max(kernel_count, 1) / 1e3
    flop_equiv = 2.0 * (valu + trans * PEAK_VALU / PEAK_TRANS)
    achieved = pairs * flop_equiv / kernel_s / 1e12
    breakdown = {k: round(ms / c, 4) for k, (c, ms) in sorted(report.items(), key=lambda kv: -kv[1][1])}
    # Instruction counts per launch hold for this workload only (I, N = 1e6, neighbourhood 20).
    default_workload = (args.group, args.n, args.delta) == ("I", 1_000_000, 20.0)
    near_counts = counters_from_profile("near_force_i1e6") if default_workload else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Haar orientations)",
        "config": {"workload": f"near-field force pass, group {args.group}, N={args.n} "
                               f"(first of zo.haar('{args.group}', 1e7, 193)), neighbourhood {args.delta:g}",
                   "group": args.group, "n": args.n, "neighbourhood": args.delta, "seed": 193},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_from_profiles(),
                     "traffic_source": "HBM bytes per launch of this kernel at this workload, read from "
                                       "profiles/near_force_i1e6_traffic.md (rocprofv3 FETCH_SIZE x 2.0 + "
                                       "WRITE_SIZE x 1.0, factors from a copy of known size), not measured "
                                       "in this run",
                     "kernel": "near_force", "kernel_ms": kernel_s * 1e3, "pairs_per_launch": pairs,
                     "alg_valu_per_pair": valu, "alg_trans_per_pair": trans,
                     "basis": "frac: speed on the reference's per-pair arithmetic (its VALU and "
                              "transcendental counts priced at the published peaks), not issue "
                              "utilisation; valu_busy: the kernel's executed VALU and transcendental "
                              "instructions (rocprof SQ counters, profiles/near_force_i1e6_counters.md) at "
                              "2 and 8 cycles per wave64 instruction over its measured time on 1024 SIMDs "
                              "at 2.4 GHz",
                     "valu_busy": valu_busy(near_counts, kernel_s),
                     "executed_per_64_pairs": None if not near_counts else {
                         "valu": (near_counts["valu_per_launch"] - near_counts["trans_per_launch"]) * 64 / pairs,
                         "trans": near_counts["trans_per_launch"] * 64 / pairs,
                         "salu": near_counts["salu_per_launch"] * 64 / pairs},
                     "candidates_per_pair": info["candidates"] / max(pairs, 1)},
        "breakdown_ms": breakdown,
        "sites": info["sites"],
        "shell": info["shell"],
        "sdim": info["sdim"],
    }
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        from oracle import dense_f64 as dense  # the checker, for the CPU baseline leg only
        cut, radius = dense.near_cut(len(q), len(ops), args.delta)
        out["cpu_baseline"] = cpu_baseline(q, ops, cut, radius, args.cpu_sample)
    if args.spacing_n > 0:
        out["spacing"] = spacing_section(args.spacing_n)
    if not args.no_small:
        out["small_relax"] = small_relax_section(with_cpu=not args.no_cpu)
        out["near_relax"] = near_relax_section(with_cpu=not args.no_cpu)
    if args.batched_jobs > 0:
        out["batched"] = batched_section(args.batched_jobs, args.batched_steps, graphs=not args.no_graphs)
    if args.serving_requests > 0:
        out["serving"] = serving_section(args.serving_requests, args.serving_rate, graphs=not args.no_graphs)
    if not args.no_widened:
        out["widened"] = widened_section()
    print(json.dumps(out), flush=True)


def _timed(grid, call):
    """Wall time of call() and the per-kernel device times (HIP events) it recorded."""
    grid.profile(True)
    before = grid.profile_report()
    t0 = time.perf_counter()
    call()
    dt = time.perf_counter() - t0
    after = grid.profile_report()
    grid.profile(False)
    kernels = {}
    for name, (count, ms) in after.items():
        c0, m0 = before.get(name, (0, 0.0))
        if count > c0:
            kernels[name] = round(ms - m0, 4)
    return dt * 1e3, kernels


def widened_section():
    """The rows next to the hot path, one workload each, timed after a warm-up call:
    inverse Hopf generation (O, 1e7 points: host FP64 construction, upload, GPU fold), the
    FCC transport (T, h = 64: 1.43e6 points, FP64 map per point on the GPU), the covering
    radius (O, 5e5 points = 1.2e7 images: per-point Voronoi cells in FP64 over the site
    grid) and an exact count (O, 1e5 points: base grid plus empty-cap insertion).  The fold
    is HBM-bound (32 B per point); k_transport and k_cover are FP64-VALU-bound."""
    import zxtalori_amd as zo
    out = {}
    print("[bench] widened: generation, transport, covering, exact", file=sys.stderr, flush=True)
    with zo.haar("O", 8, 1) as grid:
        lib, handle = grid._lib, grid._handle

        def generate(group_id, h, target, kind):
            grid._check(lib.zo_generate(handle, group_id, h, target, 0, kind))

        generate(9, 0, 10_000_000, 0)
        ms, kernels = _timed(grid, lambda: generate(9, 0, 10_000_000, 0))
        n = len(grid)
        fold_ms = kernels.get("fold", float("nan"))
        out["hopf"] = {"workload": "zo.generate('O', count=1e7): h = 281, inverse Hopf grid", "n": n, "ms": ms,
                       "points_per_s": n / (ms * 1e-3), "fold_ms": fold_ms,
                       "fold_roofline": {"bound": "hbm", "achieved": 32 * n / (fold_ms * 1e-3) / 1e9, "peak": 8000.0,
                                         "unit": "GB/s", "frac": 32 * n / (fold_ms * 1e-3) / 8e12}}
        generate(8, 64, 2048, 2)
        ms, kernels = _timed(grid, lambda: generate(8, 64, 2048, 2))
        n = len(grid)
        out["transport"] = {"workload": "zo.generate('T', h=64, kind='transport')", "n": n, "ms": ms,
                            "kernel_ms": kernels.get("transport", float("nan")),
                            "points_per_s_device": n / (kernels.get("transport", float("nan")) * 1e-3),
                            "bound": "fp64 valu"}
    with zo.generate("O", count=500_000) as grid:
        grid.covering()
        grid.set(grid.quaternions, "O")  # a new sample: the next call measures again
        ms, kernels = _timed(grid, grid.covering)
        stats = grid.stats
        out["covering"] = {"workload": "zo.generate('O', count=5e5).covering(), after a warm-up on the same points",
                           "n": len(grid),
                           "images": 2 * 24 * len(grid), "ms": ms, "kernels_ms": kernels,
                           "cells_per_s_device": len(grid) / (kernels.get("cover", float("nan")) * 1e-3),
                           "covering_radius_deg": stats["covering_radius_deg"], "bound": "fp64 valu"}
    t0 = time.perf_counter()
    with zo.generate("O", count=100_000, exact=True) as grid:
        dt = time.perf_counter() - t0
        out["exact"] = {"workload": "zo.generate('O', count=1e5, exact=True)", "n": len(grid), "ms": dt * 1e3,
                        "base_count": grid.stats["base_count"], "inserted_count": grid.stats["inserted_count"]}
    return out


def spacing_section(n):
    """Spacing of n Haar orientations of I (seed 193): nearest orbit point of every point
    (index and angle) over the 4-mean-radius near field; the second call is timed (the
    first sizes the near field).  Bytes: the algorithmic 76 B per point of a linear-time
    pass (site sort 52, positions 16, index and angle 8)."""
    import zxtalori_amd as zo
    print(f"[bench] spacing I N={n}", file=sys.stderr, flush=True)
    with zo.haar("I", n, 193) as grid:
        grid.spacing()
        grid.profile(True)
        t0 = time.perf_counter()
        stats = grid.spacing()
        dt = time.perf_counter() - t0
        report = grid.profile_report()
        grid.profile(False)
        info = grid.near_info()
    kernels = {k: round(ms / c, 4) for k, (c, ms) in sorted(report.items(), key=lambda kv: -kv[1][1])}
    device_ms = sum(ms for _, ms in report.values())
    out = {"workload": f"zo.haar('I', {n}, 193).spacing(): nearest orbit point of every point, "
                       "neighbourhood 4 mean radii", "n": n, "ms": dt * 1e3, "device_ms": device_ms,
           "points_per_s": n / dt, "alg_bytes_per_point": 76,
           "alg_gbps_device": 76 * n / (device_ms * 1e-3) / 1e9, "kernels_ms": kernels,
           "sites": info["sites"], "separation_deg": stats["separation_deg"]}
    if n == 10_000_000 and "near_nearest" in kernels:
        # The search kernel's executed VALU and transcendental instructions at this workload
        # (profiles/spacing_i1e7_counters.md) over its measured time, as roofline.valu_busy.
        out["search_valu_busy"] = valu_busy(counters_from_profile("spacing_i1e7"), kernels["near_nearest"] * 1e-3)
    return out


def _median_of(run, repeats):
    """Median wall time (s) of `repeats` calls of run() after one warm-up call, and the last
    call's result."""
    run()
    times, last = [], None
    for _ in range(repeats):
        t0 = time.perf_counter()
        last = run()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), last


def small_relax_section(repeats=5, with_cpu=True):
    """The small all-pairs workload: group O, N=1,000 normals (default_rng(193)), 30 BB steps
    (neighbourhood 0), FP32 pair terms with FP64 sums; wall time of relax(30) (median of 5
    after a warm-up), energy, separation, spacing, launches per step; the CPU restatement of
    the same 30 steps on all host cores beside it."""
    import oracle
    import zxtalori_amd as zo
    print("[bench] small relax O N=1000 x 30", file=sys.stderr, flush=True)
    rng = np.random.default_rng(193)
    q0 = rng.standard_normal((1000, 4))
    q0 = (q0 / np.linalg.norm(q0, axis=1, keepdims=True)).astype(np.float32)

    def gpu():
        with zo.from_quaternions(q0, "O") as grid:
            t0 = time.perf_counter()
            grid.relax(30)
            dt = time.perf_counter() - t0
            return dt, grid.stats
    gpu()
    runs = [gpu() for _ in range(repeats)]
    wall = float(np.median([r[0] for r in runs]))
    stats = runs[-1][1]
    with zo.from_quaternions(q0, "O") as grid:
        start, ops = grid.quaternions, grid.operations
        grid.profile(True)
        grid.relax(30)
        launches = sum(c for c, _ in grid.profile_report().values())
        grid.profile(False)
        spacing = grid.spacing()["separation_deg"]
    out = {"workload": "O, N=1000 (default_rng(193) normals), 30 BB steps, all pairs", "wall_ms": wall * 1e3,
           "steps_per_s": 30 / wall, "energy": stats["energy"], "separation_deg": stats["separation_deg"],
           "spacing_separation_deg": spacing, "launches_per_step": launches / 30,
           "pair_evals_per_s": 31 * 1000 * 999 * 24 / wall}
    if with_cpu:
        cores = oracle.threads(0)
        cpu_s, (_, rows, _) = _median_of(lambda: oracle.relax(start, ops, 9, 30), repeats)
        out["cpu_baseline"] = {"wall_ms": cpu_s * 1e3, "cores": cores, "kind": "port",
                               "energy": rows[:, 4].sum() / (1000 * 24 * 999),
                               "sample": "the CPU restatement (oracle.relax: orc_allpairs, orc_bb_step, orc_fold) "
                                         "of the same 30 steps, median of 5 after a warm-up"}
    return out


def near_relax_section(steps=100, cpu_repeats=5, with_cpu=True):
    """Group O, N=1e5 Haar (seed 193), neighbourhood 20, 100 BB steps: per-step time, then the
    rows of the first 10,000 sources against FP64 pair terms over the same pair set (energy,
    force, separation errors); beside it the CPU restatement's per-step time (one near-field
    pass over all sources, the BB step and the fold), median of 5 single steps after a
    warm-up step, on all host cores."""
    import oracle
    import zxtalori_amd as zo
    from oracle import dense_f64 as dense
    print(f"[bench] near relax O N=1e5 x {steps}", file=sys.stderr, flush=True)
    with zo.haar("O", 100_000, 193) as grid:
        grid.relax(2, neighbourhood=20)  # warm-up (sizing, code paths)
    with zo.haar("O", 100_000, 193) as grid:
        start, ops = grid.quaternions, grid.operations
        grid.relax(0, neighbourhood=20)
        t0 = time.perf_counter()
        grid.relax(steps)
        dt = time.perf_counter() - t0
        stats = grid.stats
        force, energy, nearest = grid.force_rows()
        q = grid.quaternions
    print("[bench] near relax: fp64 subsample", file=sys.stderr, flush=True)
    with np.load(ROOT / "tests" / "golden" / "pair-operators.npz", allow_pickle=False) as data:
        ops64 = np.asarray(data["O"], np.float64)
    cut, r = dense.near_cut(len(q), len(ops), 20)
    ref = oracle.near_grid64(q, ops, ops64, cut, r, 0, 10_000)
    e_err = abs(energy[:10_000].sum() - ref[:, 4].sum()) / ref[:, 4].sum()
    f_err = np.linalg.norm(force[:10_000, :3] - ref[:, :3]) / np.linalg.norm(ref[:, 6])
    sep = oracle.degrees(nearest[:10_000])
    sep64 = 4 * np.degrees(np.arcsin(np.minimum(1, np.sqrt(ref[:, 5]) / 2)))
    out = {"workload": f"O, N=1e5 (zo.haar seed 193), neighbourhood 20, {steps} BB steps",
           "ms_per_step": dt * 1e3 / steps, "energy": stats["energy"], "separation_deg": stats["separation_deg"],
           "fp64_subsample": {"sources": 10_000, "energy_sum_rel_err": e_err,
                              "force_err_vs_pair_force_scale": f_err,
                              "separation_max_abs_err_deg": float(np.abs(sep - sep64).max())}}
    if with_cpu:
        print("[bench] near relax: cpu steps", file=sys.stderr, flush=True)
        cores = oracle.threads(0)
        n, order = len(start), len(ops)
        radius = oracle.mean_radius(n * order)
        state = {"q": start.copy(), "bb": oracle.BB(n)}

        def cpu_step():
            rows, _ = oracle.near_grid(state["q"], ops, cut, r)
            state["bb"].step(state["q"], rows, radius)
            state["q"] = oracle.fold(state["q"], ops, 9)
        step_s, _ = _median_of(cpu_step, cpu_repeats)
        out["cpu_baseline"] = {"ms_per_step": step_s * 1e3, "cores": cores, "kind": "port",
                               "sample": "the CPU restatement (orc_near_grid over all sources, orc_bb_step, "
                                         "orc_fold) of single steps from the same start, median of 5 after a "
                                         "warm-up step"}
    return out


def _same(a, b):
    """Whether two runs of the same jobs gave bit-identical results (stats and positions)."""
    return all(x.state == y.state == "done" and x.stats == y.stats for x, y in zip(a, b))


def batched_section(total, steps, graphs=True):
    """Mixed-size relax jobs (12 groups with C1, N log-uniform in [500, 2e5], all pairs below
    5,000 points, neighbourhood 20 above, Haar seed = job index) run in one batch with graph
    replay, then one after another on their own Grids (the plain per-sample path)."""
    from zxtalori_amd import serve
    jobs = serve.mixed_relax_jobs(total, seed=193, steps=steps)
    print(f"[bench] batched {total} jobs x {steps} steps", file=sys.stderr, flush=True)
    batched, b = serve.run_batched(jobs, graphs=graphs)
    print(f"[bench] batched {b['elapsed_s']:.2f} s; plain loop", file=sys.stderr, flush=True)
    plain, p = serve.serve_plain(jobs, realtime=False)
    points = sum(j.count for j in jobs)
    return {"workload": f"{total} relax jobs, 12 groups (C1 included), N log-uniform [500, 2e5] "
                        f"(rng 193), all pairs below 5000 points, neighbourhood 20 above, {steps} BB steps",
            "points": points, "near_jobs": sum(j.neighbourhood > 0 for j in jobs),
            "batched_s": b["elapsed_s"], "plain_s": p["elapsed_s"],
            "batched_point_steps_per_s": b["point_steps_per_s"], "plain_point_steps_per_s": p["point_steps_per_s"],
            "speedup": p["elapsed_s"] / b["elapsed_s"], "graph_captures": b["graph_captures"],
            "graph_replays": b["graph_replays"], "identical_results": _same(batched, plain),
            "failed": b["failed"] + p["failed"]}


def serving_section(total, rate, graphs=True):
    """Poisson request stream (relax 20-100 steps, measure, spacing; N log-uniform in
    [1e3, 1e6]; 12 groups) served by the continuous-batching loop and by the plain path, at
    the arrival rate (latency) and with every request queued at the start (throughput)."""
    from zxtalori_amd import serve
    reqs = serve.poisson_requests(total, rate, seed=5)
    print(f"[bench] serving {total} requests at {rate}/s", file=sys.stderr, flush=True)
    with serve.Server(hbm_budget=64 << 30, graphs=graphs) as server:
        served, s = server.run(reqs, realtime=True)
        max_live = server.max_live
    print(f"[bench] served p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms; plain", file=sys.stderr, flush=True)
    plain, p = serve.serve_plain(reqs, realtime=True)
    print("[bench] saturated throughput", file=sys.stderr, flush=True)
    with serve.Server(hbm_budget=64 << 30, graphs=graphs) as server:
        _, s_sat = server.run(reqs, realtime=False)
    _, p_sat = serve.serve_plain(reqs, realtime=False)
    return {"workload": f"{total} requests, Poisson rate {rate}/s (seed 5): 70% relax (20-100 steps), 15% "
                        f"measure, 15% spacing; N log-uniform [1e3, 1e6], 12 groups; HBM budget 64 GiB",
            "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"], "plain_p50_ms": p["p50_ms"], "plain_p99_ms": p["p99_ms"],
            "jobs_per_s": s["jobs_per_s"], "plain_jobs_per_s": p["jobs_per_s"], "max_live": max_live,
            "saturated_jobs_per_s": s_sat["jobs_per_s"], "plain_saturated_jobs_per_s": p_sat["jobs_per_s"],
            "saturated_speedup": p_sat["elapsed_s"] / s_sat["elapsed_s"],
            "identical_results": _same(served, plain), "failed": s["failed"] + p["failed"]}


if __name__ == "__main__":
    main()
