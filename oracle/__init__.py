"""ORACLE — test infrastructure only (tests/, smoke() and bench.py's cpu_baseline).

ctypes wrappers over oracle/liboracle.so (the C++ restatement in qvm_oracle.cpp, built
by oracle/Makefile) plus the dense FP64 models in dense_f64.py.  The product never
imports this package.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.is_file():
            build()
        _LIB = C.CDLL(str(path))
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        u32 = C.c_uint32
        sig = {
            "orc_threads": ([C.c_int], C.c_int),
            "orc_haar": ([u32, u32, f32p], None),
            "orc_fold": ([f32p, u32, f32p, u32, u32], None),
            "orc_images": ([f32p, u32, f32p, u32, f32p], None),
            "orc_allpairs": ([f32p, u32, f32p, u32, f64p], None),
            "orc_near_brute": ([f32p, u32, f32p, u32, f32p, f64p], None),
            "orc_near_grid": ([f32p, u32, f32p, u32, f32p, C.c_double, u32, u32, f64p, f64p], C.c_int),
            "orc_near_grid64": ([f32p, u32, f32p, f64p, u32, f32p, C.c_double, u32, u32, f64p], C.c_int),
            "orc_nearest": ([f32p, u32, f32p, u32, C.c_float, u32p, f32p], None),
            "orc_nearest_grid": ([f32p, u32, f32p, u32, C.c_float, C.c_double, u32, u32, u32p, f32p], None),
            "orc_degrees": ([f32p, u32, f32p], None),
            "orc_median": ([f32p, u32], C.c_float),
            "orc_bb_step": ([f32p, u32, f32p, f32p, f64p, f32p, u32p, C.c_double, u32, C.c_float, f64p], None),
            "orc_mean_radius": ([C.c_uint64], C.c_double),
        }
        for name, (args, res) in sig.items():
            fn = getattr(_LIB, name)
            fn.argtypes, fn.restype = args, res
    return _LIB


def _q(q):
    return np.ascontiguousarray(q, dtype=np.float32).reshape(-1, 4)


def threads(n=0):
    return lib().orc_threads(int(n))


def haar(n, seed):
    out = np.empty((n, 4), np.float32)
    lib().orc_haar(n, seed, out)
    return out


def fold(q, ops, group_id):
    q = _q(q).copy()
    ops = _q(ops)
    lib().orc_fold(q, len(q), ops, len(ops), group_id)
    return q


def images(q, ops):
    q, ops = _q(q), _q(ops)
    out = np.empty((len(ops), len(q), 4), np.float32)
    lib().orc_images(q, len(q), ops, len(ops), out)
    return out


def allpairs(q, ops):
    """Rows (n, 6): fx, fy, fz, curvature, energy, nearest chord^2."""
    q, ops = _q(q), _q(ops)
    out = np.empty((len(q), 6), np.float64)
    lib().orc_allpairs(q, len(q), ops, len(ops), out)
    return out


def near_brute(q, ops, cut):
    q, ops = _q(q), _q(ops)
    out = np.empty((len(q), 6), np.float64)
    lib().orc_near_brute(q, len(q), ops, len(ops), np.ascontiguousarray(cut, np.float32), out)
    return out


def near_grid(q, ops, cut, radius, first=0, count=None):
    q, ops = _q(q), _q(ops)
    count = len(q) - first if count is None else count
    out = np.empty((count, 6), np.float64)
    stats = np.zeros(5, np.float64)
    lib().orc_near_grid(q, len(q), ops, len(ops), np.ascontiguousarray(cut, np.float32), float(radius), first, count,
                        out, stats)
    return out, dict(zip(("sites", "shell", "sdim", "candidates", "accepted"), stats.tolist()))


def near_grid64(q, ops, ops64, cut, radius, first=0, count=None):
    """orc_near_grid's pair set with FP64 pair terms (images from the FP64 operators):
    rows (count, 7) fx fy fz curvature energy nearest-chord^2 sum-of-|pair force|."""
    q, ops = _q(q), _q(ops)
    ops64 = np.ascontiguousarray(ops64, np.float64).reshape(-1, 4)
    count = len(q) - first if count is None else count
    out = np.empty((count, 7), np.float64)
    lib().orc_near_grid64(q, len(q), ops, ops64, len(ops), np.ascontiguousarray(cut, np.float32), float(radius),
                          first, count, out)
    return out


def nearest_grid(q, ops, cut, radius, first=0, count=None):
    """Nearest orbit point within the cutoff chord^2 `cut` for sources [first, first+count)
    through the near grid of `radius`: (rank, chord^2)."""
    q, ops = _q(q), _q(ops)
    count = len(q) - first if count is None else count
    rank = np.empty(count, np.uint32)
    chord = np.empty(count, np.float32)
    lib().orc_nearest_grid(q, len(q), ops, len(ops), float(cut), float(radius), first, count, rank, chord)
    return rank, chord


def nearest(q, ops, cut=0.0):
    q, ops = _q(q), _q(ops)
    rank = np.empty(len(q), np.uint32)
    chord = np.empty(len(q), np.float32)
    lib().orc_nearest(q, len(q), ops, len(ops), float(cut), rank, chord)
    return rank, chord


def degrees(chord):
    chord = np.ascontiguousarray(chord, np.float32)
    out = np.empty_like(chord)
    lib().orc_degrees(chord, len(chord), out)
    return out


def median(curvature):
    c = np.ascontiguousarray(curvature, np.float32)
    return float(lib().orc_median(c, len(c)))


def mean_radius(count):
    return float(lib().orc_mean_radius(int(count)))


class BB:
    """The BB step state of one sample (relax.wgsl), stepping q in place."""

    def __init__(self, n):
        self.prev = np.zeros(6 * n, np.float32)
        self.state = np.zeros(2, np.uint32)
        self.last = None

    def step(self, q, rows, radius, fixed=0, step_scale=1.0):
        n = len(q)
        force4 = np.ascontiguousarray(rows[:, :4], np.float32)
        nearest = np.ascontiguousarray(rows[:, 5], np.float32)
        energy = np.ascontiguousarray(rows[:, 4], np.float64)
        out = np.zeros(8, np.float64)
        lib().orc_bb_step(q, n, force4, nearest, energy, self.prev, self.state, float(radius), int(fixed),
                          float(step_scale), out)
        self.last = dict(zip(("sy", "yy", "ss", "ff", "energy", "median", "alpha", "step_max"), out.tolist()))
        return self.last


def relax(q, ops, group_id, steps, delta=0.0):
    """Reference relaxation loop (engine.zig:896-937): forces, BB step, fold; then a final
    force pass.  Returns (q, rows of the final positions, last step's sums)."""
    q = _q(q).copy()
    ops = _q(ops)
    n, order = len(q), len(ops)
    radius = mean_radius(n * order)
    bb = BB(n)
    if delta:
        from .dense_f64 import near_cut
        cut, _ = near_cut(n, order, delta)

    def forces(x):
        return near_brute(x, ops, cut) if delta else allpairs(x, ops)

    for _ in range(steps):
        r = forces(q)
        bb.step(q, r, radius)
        q = fold(q, ops, group_id)
    return q, forces(q), bb.last
