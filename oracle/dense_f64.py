"""ORACLE — test infrastructure only.  Dense NumPy FP64 models of the QVM path.

Restated from the closed forms the reference's own tests use as their oracle:
  rows (12 columns) ........ tests/test_pair_kernels.py:27-69
  near-field energy/sep .... tests/test_near_field.py:23-40
  measures ................. tests/test_measures.py:28-43
  mean radius .............. src/pairs.zig:205-215
Nothing in the product imports this module.
"""
import math

import numpy as np

PI = math.pi


def multiply(a, b):
    """Quaternion product, scalar first (broadcasting)."""
    a = np.asarray(a)
    b = np.asarray(b)
    w = a[..., :1] * b[..., :1] - (a[..., 1:] * b[..., 1:]).sum(-1, keepdims=True)
    v = a[..., :1] * b[..., 1:] + b[..., :1] * a[..., 1:] + np.cross(a[..., 1:], b[..., 1:])
    return np.concatenate([w, v], axis=-1)


def mean_radius(count):
    target = PI ** 2 / count
    lo, hi = 0.0, PI
    for _ in range(60):
        middle = 0.5 * (lo + hi)
        if PI * (middle - math.sin(middle)) < target:
            lo = middle
        else:
            hi = middle
    return 0.25 * (lo + hi)


def near_cut(n, order, delta, taper=0.5):
    """(chord^2, chord, taper start, 1/taper width) as f32, and the radius r (near.zig:95, :176)."""
    r = min(1.5, delta * mean_radius(n * order))
    chord = 2 * math.sin(r / 2)
    cut = np.array([chord * chord, chord, taper * r, 1 / max(1e-9, (1 - taper) * r)], np.float32)
    return cut, r


def rows(q, ops):
    """Columns: 0-2 Riesz 1..3, 3 log, 4 Green, 5 Quey energy, 6 nearest chord^2,
    8-10 force, 11 curvature (the reference's Row, in FP64)."""
    q = np.asarray(q, dtype=np.float64)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    ops = np.asarray(ops, dtype=np.float64)
    ops = ops / np.linalg.norm(ops, axis=-1, keepdims=True)
    n = len(q)
    out = np.zeros((n, 12), np.float64)
    out[:, 6] = 4
    eye = np.eye(n, dtype=bool)
    for k, op in enumerate(ops):
        y = multiply(q, op)
        y /= np.linalg.norm(y, axis=-1, keepdims=True)
        minus = q[:, None, :] - y[None, :, :]
        plus = q[:, None, :] + y[None, :, :]
        dm2 = (minus * minus).sum(-1)
        dp2 = (plus * plus).sum(-1)
        mask = eye if k == 0 else np.zeros((n, n), bool)
        safe_m = np.where(mask, 1, dm2)
        dm = np.sqrt(safe_m)
        dp = np.sqrt(dp2)
        im = 1 / dm
        ip = 1 / dp
        keep = ~mask
        out[:, 0] += np.where(keep, im + ip, 0).sum(-1)
        out[:, 1] += np.where(keep, im * im + ip * ip, 0).sum(-1)
        out[:, 2] += np.where(keep, im ** 3 + ip ** 3, 0).sum(-1)
        out[:, 3] += np.where(keep, -.5 * np.log(safe_m * dp2) + .5, 0).sum(-1)
        angle = 2 * np.arctan2(dm, dp)
        cot = (dp2 - safe_m) / (2 * dm * dp)
        out[:, 4] += np.where(keep, ((PI - 2 * angle) * cot - 1) / (4 * PI ** 2), 0).sum(-1)
        out[:, 6] = np.minimum(out[:, 6], np.where(mask, 4, np.minimum(dm2, dp2)).min(-1))
        dm = np.sqrt(np.maximum(dm2, 1e-30))
        dp = np.sqrt(np.maximum(dp2, 1e-30))
        angle = np.maximum(2 * np.arctan2(dm, dp), 1e-12)
        anti = np.maximum(2 * np.arctan2(dp, dm), 1e-12)
        source_keep = ~eye
        out[:, 5] += np.where(source_keep, 1 / angle + 1 / anti, 0).sum(-1)
        mag = 1 / (angle * angle + 1e-9) - 1 / (anti * anti + 1e-9)
        direction = (q[:, None, :1] * y[None, :, 1:] - y[None, :, :1] * q[:, None, 1:]
                     + np.cross(q[:, None, 1:], y[None, :, 1:]))
        force = -mag[..., None] * direction / (.5 * dm * dp + 1e-9)[..., None]
        out[:, 8:11] += np.where(source_keep[..., None], force, 0).sum(1)
        curv = 1 / (angle ** 3 + 1e-9) + 1 / (anti ** 3 + 1e-9)
        out[:, 11] += np.where(source_keep, curv, 0).sum(-1)
    out[:, :3] += np.array([.5, .25, .125])
    out[:, 3] += -np.log(2) + .25
    out[:, 4] += -3 / (8 * PI ** 2)
    return out


def near_rows(q, ops, delta, taper=0.5):
    """Windowed near-field rows in FP64: energy, nearest chord^2 and force (columns
    energy, nearest, fx, fy, fz, curvature) with the C1 window of pairs.wgsl:349-361."""
    q = np.asarray(q, np.float64)
    ops = np.asarray(ops, np.float64)
    n, order = len(q), len(ops)
    r = min(1.5, delta * mean_radius(n * order))
    cut = (2 * math.sin(r / 2)) ** 2
    images = np.concatenate([multiply(q, np.broadcast_to(op, q.shape)) for op in ops])
    images /= np.linalg.norm(images, axis=1, keepdims=True)
    source = np.tile(np.arange(n), order)
    out = np.zeros((n, 6))
    for i in range(n):
        dm2 = ((images - q[i]) ** 2).sum(1)
        dp2 = ((images + q[i]) ** 2).sum(1)
        near = np.minimum(dm2, dp2)
        far = np.maximum(dm2, dp2)
        own = source == i
        own[i] = False
        pick = (near <= cut) & (source != i)
        a = 2 * np.arctan(np.sqrt(near[pick] / far[pick]))
        b = PI - a
        s = np.clip((a - taper * r) / ((1 - taper) * r), 0, 1)
        window = 1 - s * s * (3 - 2 * s)
        slope = -6 * s * (1 - s) / ((1 - taper) * r)
        energy = PI / (a * b)
        sign = np.where(dm2[pick] <= dp2[pick], 1.0, -1.0)
        y = images[pick]
        x = q[i]
        direction = x[0] * y[:, 1:] - y[:, :1] * x[1:] + np.cross(x[1:], y[:, 1:])
        # The kernels' regularisers (pairs.wgsl:252-258), as tests/test_pair_kernels.py:58-63.
        along = sign * (1 / (a ** 2 + 1e-9) - 1 / (b ** 2 + 1e-9)) / (np.sin(a) + 1e-9)
        force = -along[:, None] * direction
        force = window[:, None] * force + (sign * energy * slope / np.sin(a))[:, None] * direction
        curvature = window * (1 / (a ** 3 + 1e-9) + 1 / (b ** 3 + 1e-9))
        out[i, 0] = (energy * window).sum()
        out[i, 1] = min(near[pick].min(initial=4.0), near[own].min(initial=4.0))
        out[i, 2:5] = force.sum(0)
        out[i, 5] = curvature.sum()
    return out


def near_energy(q, ops, delta):
    """stats['energy'] and separation of the near field (test_near_field.py:23-40)."""
    n, order = len(q), len(ops)
    r = near_rows(q, ops, delta)
    return r[:, 0].sum() / (n * order * max(n - 1, 1)), 4 * math.degrees(math.asin(math.sqrt(r[:, 1].min()) / 2))


def measures(q, ops):
    """measure() normalisations (pairs.zig:142-173, test_measures.py:28-43)."""
    r = rows(q, ops)
    order = len(ops)
    n = len(q)
    count = 2 * order * n
    sums = 2 * order * r[:, :5].sum(0)
    cgreen = 3 ** (5 / 3) / (5 * (2 * PI ** 2) ** (1 / 3) * (4 * PI) ** (2 / 3))
    delta = (sums[3] + count * .25 + count * math.log(count) / 3) / count
    return {
        'riesz_1': sums[0] / (count * count * 8 / (3 * PI)),
        'riesz_2': sums[1] / (count * count),
        'riesz_3': sums[2] / (count * count * 2 / (3 * PI) * math.log(count)),
        'green_ratio': sums[4] / (-cgreen * count ** (4 / 3)),
        'log_ratio': delta / (-5 / 12),
        'quey_energy': r[:, 5].sum() / (n * order * max(n - 1, 1)),
        'separation_deg': 4 * math.asin(math.sqrt(r[:, 6].min()) / 2) * 180 / PI,
        'force_rms': float(np.sqrt(np.square(r[:, 8:11]).sum(axis=1).mean())),
    }


def log_ratio(q, ops):
    """measure()'s log_ratio alone, in FP64 through one Gram matrix per operator: for unit x,
    y with c = x.y, |x-y|^2 |x+y|^2 = 4(1 - c^2) (tests/test_log_scale.py:19-35 uses the same
    identity).  Agrees with measures()['log_ratio'] to about 1e-11 at a fraction of its cost."""
    q = np.asarray(q, dtype=np.float64)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    ops = np.asarray(ops, dtype=np.float64)
    ops = ops / np.linalg.norm(ops, axis=-1, keepdims=True)
    n = len(q)
    total = 0.0
    for k, op in enumerate(ops):
        y = multiply(q, op)
        y /= np.linalg.norm(y, axis=-1, keepdims=True)
        c = q @ y.T
        product = 4 * (1 - c * c)
        if k == 0:
            np.fill_diagonal(product, 1.0)
        terms = -.5 * np.log(product) + .5
        if k == 0:
            np.fill_diagonal(terms, 0.0)
        total += terms.sum()
    total += n * (-math.log(2) + .25)
    count = 2 * len(ops) * n
    delta = total / n + .25 + math.log(count) / 3
    return delta / (-5 / 12)


def nearest_orbit(q, ops):
    """FP64 nearest orbit point of every source over the full signed orbit (the source's
    own identity image excluded): (rank j*2*order + 2*k + sign, chord^2)."""
    q = np.asarray(q, np.float64)
    ops = np.asarray(ops, np.float64)
    n, order = len(q), len(ops)
    images = np.stack([multiply(q, np.broadcast_to(op, q.shape)) for op in ops], axis=1)  # n, order, 4
    images /= np.linalg.norm(images, axis=-1, keepdims=True)
    signed = np.stack([images, -images], axis=2).reshape(n * order * 2, 4)  # rank order
    ranks = np.empty(n, np.int64)
    chords = np.empty(n)
    for i in range(n):
        d = ((signed - q[i]) ** 2).sum(1)
        d[i * 2 * order] = np.inf
        ranks[i] = int(np.argmin(d))
        chords[i] = d[ranks[i]]
    return ranks, chords
