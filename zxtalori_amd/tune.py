"""Autotuner for the all-pairs force pass's tile plan; writes csrc/allpairs_tuned.h.

    python -m zxtalori_amd.tune [--out PATH] [--report PATH] [--reps 3]

For every operator count |K| of the twelve groups and every point count 2^8 .. 2^15 (cells
above 2e9 pair evaluations per pass take the nearest smaller tuned size), it times the
all-pairs kernels (k_allpairs and its reduction, HIP events via zo_profile) over a grid of
plans: about `target` workgroups (512 .. 16384) and operators batched 1 .. 8 to a
workgroup (0: the automatic rule).  A plan replaces the default (2048, automatic) only when
it is at least 3% faster.  The table is compiled into libzxtalori.so, so a plan never
depends on the timing of the run that uses it (results of a job are reproducible); plans
change the summation order only.  Rebuild after writing the table.
"""
import argparse
import datetime
import json
import sys
from pathlib import Path

from .api import _library, haar
from .build import ARCH

HERE = Path(__file__).resolve().parent
ORDERS = [(1, "C1"), (2, "C2"), (3, "C3"), (4, "C4"), (6, "C6"), (8, "D4"), (12, "D6"), (24, "O"), (60, "I")]
LOG_MIN, LOG_MAX = 8, 15
TARGETS = (512, 1024, 2048, 4096, 8192, 16384)
OPS = (0, 1, 2, 4, 8)
DEFAULT = (2048, 0)
MAX_PAIRS = 2e9


def kernel_ms(grid, passes):
    before = grid.profile_report()
    grid.force_passes(passes)
    after = grid.profile_report()
    total = 0.0
    for name in ("allpairs", "allpairs_reduce"):
        total += after.get(name, (0, 0.0))[1] - before.get(name, (0, 0.0))[1]
    return total / passes


def tune_cell(lib, group, n, reps):
    times = {}
    with haar(group, n, 7) as grid:
        grid.profile(True)
        for target in TARGETS:
            for ops in OPS:
                lib.zo_tune_allpairs(target, ops)
                grid.force_passes(1)  # buffers sized, caches warm
                times[(target, ops)] = min(kernel_ms(grid, 1) for _ in range(reps))
        lib.zo_tune_allpairs(0, 0)
    best = min(times, key=times.get)
    if times[best] > 0.97 * times[DEFAULT]:
        best = DEFAULT
    return best, times


def write_header(path, table, device):
    rows_t = ",\n    ".join("{" + ", ".join(str(table[o][lg][0]) for lg in range(LOG_MIN, LOG_MAX + 1)) + "}"
                           for o, _ in ORDERS)
    rows_o = ",\n    ".join("{" + ", ".join(str(table[o][lg][1]) for lg in range(LOG_MIN, LOG_MAX + 1)) + "}"
                           for o, _ in ORDERS)
    text = f"""// All-pairs tile plans per operator count and point count (2^{LOG_MIN} .. 2^{LOG_MAX}), written by
// zxtalori_amd/tune.py on {device} ({datetime.date.today().isoformat()}).  target: workgroups per
// pass; ops: operators per workgroup (0: automatic).  Regenerate with
// `python -m zxtalori_amd.tune` on the GPU and rebuild.
#pragma once
#include <cstdint>

namespace zxt {{

constexpr int kTunedOrderCount = {len(ORDERS)};
constexpr int kTunedLogMin = {LOG_MIN};
constexpr int kTunedSizeCount = {LOG_MAX - LOG_MIN + 1};
constexpr uint32_t kTunedOrders[kTunedOrderCount] = {{{", ".join(str(o) for o, _ in ORDERS)}}};
constexpr uint32_t kTunedTarget[kTunedOrderCount][kTunedSizeCount] = {{
    {rows_t}}};
constexpr uint32_t kTunedOps[kTunedOrderCount][kTunedSizeCount] = {{
    {rows_o}}};

}}  // namespace zxt
"""
    Path(path).write_text(text)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=str(HERE / "csrc" / "allpairs_tuned.h"))
    ap.add_argument("--report", default=None, help="also write the timings as markdown here")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    lib = _library()
    device = f"an {ARCH} device"
    table, lines = {}, []
    for order, group in ORDERS:
        table[order] = {}
        last = DEFAULT
        for lg in range(LOG_MIN, LOG_MAX + 1):
            n = 2 ** lg
            if float(n) * n * order > MAX_PAIRS:
                table[order][lg] = last
                continue
            best, times = tune_cell(lib, group, n, args.reps)
            table[order][lg] = last = best
            gain = times[DEFAULT] / times[best]
            lines.append(f"| {order} | {n} | {times[DEFAULT]:.3f} | {best[0]}, {best[1]} | {times[best]:.3f} | {gain:.2f}x |")
            print(f"|K|={order} N={n}: default {times[DEFAULT]:.3f} ms, best {best} {times[best]:.3f} ms",
                  file=sys.stderr, flush=True)
    write_header(args.out, table, device)
    print(json.dumps({"out": args.out, "cells": len(lines)}))
    if args.report:
        head = ["| \\|K\\| | N | default ms | tuned (target, ops) | tuned ms | gain |", "|---|---|---|---|---|---|"]
        Path(args.report).write_text("\n".join(head + lines) + "\n")


if __name__ == "__main__":
    main()
