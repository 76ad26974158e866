"""CPU pins of the reformulated pair arithmetic the gfx950 kernels use
(zxtalori_amd/csrc/qvm_device.h pair_terms), against the reference's quey_pair and
near_pair formulas (shaders/pairs.wgsl:246-262, :349-361) evaluated in FP64.

The kernel replaces sqrt, atan and four divisions by one rsq and one rcp:
t = near*rsq(near*far), sin a = t*far/2 (near+far = 4 for unit quaternions),
1/sin a = (near+far)/(2 sqrt(near far)), atan(t) = t*P(t^2), and the reciprocals of
a2*b2*sine, a3*b3, a*b from one reciprocal of their product.  Evaluated here in FP64
with the kernel's coefficients, the rewrite must agree with the reference form to the
polynomial's bound; the kernel's FP32 rounding is checked on the GPU against the oracle.
"""
import math
import re

import numpy as np

from tests.conftest import ROOT

SRC = (ROOT / "zxtalori_amd" / "csrc" / "qvm_device.h").read_text()


def poly_coefficients():
    body = SRC[SRC.index("atan01"):SRC.index("return t * p;")]
    return [float(x) for x in re.findall(r"(-?\d+\.\d+)f", body)]


def atan_poly(t):
    c = poly_coefficients()  # c[0] is the leading coefficient, then fma steps in order
    z = t * t
    p = c[0] * z + c[1]
    for k in c[2:]:
        p = p * z + k
    return t * p


def test_atan_polynomial_bound():
    t = np.linspace(0, 1, 200001)
    approx = atan_poly(np.float32(t).astype(np.float64))
    exact = np.arctan(np.float32(t).astype(np.float64))
    rel = np.abs(approx - exact) / np.maximum(exact, 1e-30)
    assert rel[1:].max() < 6e-8  # Remez bound 1.5e-8 plus the f32 rounding of the coefficients
    assert approx[0] == 0.0


def reference_pair(x, y, dm2, dp2, taper_start=None, inv_width=None, atan=math.atan, exact_sin=False):
    near = max(min(dm2, dp2), 1e-30)
    far = max(max(dm2, dp2), 1e-30)
    t = math.sqrt(near / far)
    a = max(2 * atan(t), 1e-12)
    b = math.pi - a
    a2, b2, a3, b3 = a * a + 1e-9, b * b + 1e-9, a ** 3 + 1e-9, b ** 3 + 1e-9
    sine = 2 * t / (1 + t * t) + 1e-9
    sign = 1.0 if dm2 <= dp2 else -1.0
    d = x[0] * y[1:] - y[0] * x[1:] + np.cross(x[1:], y[1:])
    along = sign * (b2 - a2) / (a2 * b2 * sine)
    force = -along * d
    curv = (a3 + b3) / (a3 * b3)
    energy = math.pi / (a * b)
    if taper_start is not None:
        s = min(max((a - taper_start) * inv_width, 0.0), 1.0)
        if s > 0:
            w = 1 - s * s * (3 - 2 * s)
            slope = -6 * s * (1 - s) * inv_width
            # exact_sin: sin(2 atan t) = 2t/(1+t^2) with the exact t, as the kernel takes it.
            sin_a = 2 * t / (1 + t * t) if exact_sin else math.sin(a)
            extra = sign * energy * slope / max(sin_a, 1e-9)
            force = w * force + extra * d
            curv *= w
            energy *= w
    return force, curv, energy


def kernel_pair(x, y, dm2, dp2, taper_start=None, inv_width=None):
    nr = max(min(dm2, dp2), 1e-30)
    fr = max(max(dm2, dp2), 1e-30)
    R = 1 / math.sqrt(nr * fr)
    t = nr * R
    a = max(2 * atan_poly(t), 1e-12)
    b = math.pi - a
    aa, bb = a * a, b * b
    a2, b2 = aa + 1e-9, bb + 1e-9
    a3, b3 = aa * a + 1e-9, bb * b + 1e-9
    sine = 0.5 * t * fr + 1e-9
    P1, P2, P3 = a2 * b2 * sine, a3 * b3, a * b
    D = 1 / (P1 * P2 * P3)
    iP1, iP2, iP3 = D * P2 * P3, D * P1 * P3, D * P1 * P2
    diff = (b2 - a2) if dm2 <= dp2 else (a2 - b2)
    coef = -diff * iP1
    curv = (a3 + b3) * iP2
    energy = math.pi * iP3
    if taper_start is not None:
        u = min(max((a - taper_start) * inv_width, 0.0), 1.0)
        w = (u * u) * (2 * u - 3) + 1
        slope = -6 * inv_width * (u * (1 - u))
        inv_sin = 0.5 * (nr + fr) * R
        e_slope = (energy if dm2 <= dp2 else -energy) * slope
        coef = w * coef + e_slope * inv_sin
        curv *= w
        energy *= w
    d = np.array([x[0] * y[1] - y[0] * x[1] + (x[2] * y[3] - x[3] * y[2]),
                  x[0] * y[2] - y[0] * x[2] + (x[3] * y[1] - x[1] * y[3]),
                  x[0] * y[3] - y[0] * x[3] + (x[1] * y[2] - x[2] * y[1])])
    return coef * d, curv, energy


def sample_pairs(count, seed=3):
    rng = np.random.default_rng(seed)
    for trial in range(count):
        x = rng.normal(size=4)
        x /= np.linalg.norm(x)
        # S3 angles from 1e-4 to pi: both signs of the nearer image.
        ang = 10 ** rng.uniform(-4, math.log10(math.pi))
        axis = rng.normal(size=4)
        axis -= axis.dot(x) * x
        axis /= np.linalg.norm(axis)
        y = math.cos(ang) * x + math.sin(ang) * axis
        y /= np.linalg.norm(y)
        window = (0.05, 1 / 0.05) if trial % 2 else (None, None)
        yield x, y, ((x - y) ** 2).sum(), ((x + y) ** 2).sum(), window


def test_pair_rewrite_is_the_reference_algebra():
    """With the same arctangent, the rewrite is the reference formula up to FP64 rounding."""
    for x, y, dm2, dp2, window in sample_pairs(4000):
        fr, cr, er = reference_pair(x, y, dm2, dp2, *window, atan=atan_poly, exact_sin=True)
        fk, ck, ek = kernel_pair(x, y, dm2, dp2, *window)
        # Scales: the unwindowed terms, so zeros of the window or of the force do not
        # turn rounding into a large relative error.
        f0, c0, e0 = reference_pair(x, y, dm2, dp2, atan=atan_poly)
        scale_f = abs(f0).max() + abs(e0) * 20 * np.abs(x).sum() * 2
        assert abs(ek - er) <= 1e-12 * e0
        assert abs(ck - cr) <= 1e-12 * c0
        assert np.abs(fk - fr).max() <= 1e-11 * scale_f


def test_pair_rewrite_error_against_exact_atan():
    """End to end, the polynomial's error moves energy and curvature by at most a few
    parts in 1e8 of their scale, and forces by the matching derivative."""
    for x, y, dm2, dp2, window in sample_pairs(4000):
        fr, cr, er = reference_pair(x, y, dm2, dp2, *window)
        fk, ck, ek = kernel_pair(x, y, dm2, dp2, *window)
        f0, c0, e0 = reference_pair(x, y, dm2, dp2)
        a = 2 * math.atan(math.sqrt(min(dm2, dp2) / max(dm2, dp2)))
        assert abs(ek - er) <= 1e-7 * e0
        assert abs(ck - cr) <= 2e-7 * c0
        # d(force)/da ~ 2/a^3 (+ window terms ~ E W'/sin a): a relative error of 3e-8 in a.
        scale_f = 2 / max(a, 1e-12) ** 3 * 3e-8 * a + abs(e0) * 20 * 20 * 3e-8 * a / max(math.sin(a), 1e-9)
        assert np.abs(fk - fr).max() <= 4 * scale_f + 1e-12 * np.abs(f0).max()


def test_duplicate_pair_stays_finite():
    x = np.array([0.9, 0.1, -0.3, 0.3])
    x /= np.linalg.norm(x)
    f, c, e = kernel_pair(x, x, 0.0, 4.0)
    assert np.isfinite(f).all() and np.isfinite(c) and np.isfinite(e)
    assert (f == 0).all()


# ---- the near-field form of the pair (near_pair_terms, near_force_of) ---------------
def series_coefficients():
    """The plain series coefficients c_k of atan2x_small, whose table holds 2 * c_k."""
    body = SRC[SRC.index("__device__ __forceinline__ float atan2x_small"):SRC.index("struct NearSum")]
    table = body[body.index("c[6]"):body.index("};")]
    found = re.findall(r"2\.0f \* (-?\d+\.\d+)f", table)
    assert len(found) == 6, table
    return [float(x) for x in found]


def atan_series(t, terms):
    c = series_coefficients()
    z = t * t
    p = c[terms]
    for k in range(terms - 1, -1, -1):
        p = p * z + c[k]
    return t * p


def test_atan_series_within_its_cutoff():
    """Each series length is used only for z = t^2 up to the bound atan_terms grants it;
    there its error against the exact arctangent stays below 1e-8 relative."""
    c = series_coefficients()
    assert c[:2] == [1.0, -0.33333334] and len(c) == 6
    for terms in range(2, 6):
        zmax = (5e-9 * (2 * terms + 3)) ** (1 / (terms + 1))
        t = np.linspace(1e-6, math.sqrt(zmax), 20001)
        approx = atan_series(t, terms)
        rel = np.abs(approx - np.arctan(t)) / np.arctan(t)
        assert rel.max() < 1e-8, (terms, rel.max())
        # one term fewer would not do over the same range
        if terms > 2:
            assert (np.abs(atan_series(t, terms - 1) - np.arctan(t)) / np.arctan(t)).max() > 5e-9


def test_doubled_series_is_twice_the_series_bit_for_bit():
    """Scaling every coefficient by 2 scales every rounded Horner step and the final product
    by 2 exactly (no overflow or underflow at these magnitudes), so the kernel's 2 atan(t)
    is bit for bit twice the plain series."""
    c = np.array(series_coefficients(), np.float32)
    t = np.linspace(1e-6, 0.6, 200001, dtype=np.float32)
    z = t * t
    for terms in range(2, 6):
        p, q = c[terms], np.float32(2) * c[terms]
        for k in range(terms - 1, -1, -1):
            p = (z * p + c[k]).astype(np.float32)
            q = (z * q + np.float32(2) * c[k]).astype(np.float32)
        np.testing.assert_array_equal((t * q).astype(np.float32), np.float32(2) * (t * p).astype(np.float32))


def direction_M(x, e):
    """-M(x) e as near_force_of applies it (e = sum of coef (x - y))."""
    x0, x1, x2, x3 = x
    return np.array([x1 * e[0] - x0 * e[1] + x3 * e[2] - x2 * e[3],
                     x2 * e[0] - x3 * e[1] - x0 * e[2] + x1 * e[3],
                     x3 * e[0] + x2 * e[1] - x1 * e[2] - x0 * e[3]])


def test_direction_is_linear_in_the_difference():
    """The reference's direction x0 y_v - y0 x_v + x_v x y_v equals -M(x)(x - y) for unit x
    (M(x) x = 0), so summing coef (x - y) and applying -M(x) once gives the same force."""
    rng = np.random.default_rng(11)
    for _ in range(2000):
        x = rng.normal(size=4)
        x /= np.linalg.norm(x)
        y = x + rng.normal(size=4) * 10 ** rng.uniform(-4, 0)
        y /= np.linalg.norm(y)
        d = x[0] * y[1:] - y[0] * x[1:] + np.cross(x[1:], y[1:])
        np.testing.assert_allclose(direction_M(x, x - y), d, rtol=0, atol=1e-14)
    # and the kernel's M is the one in the header (signs as near_force_of writes them)
    body = SRC[SRC.index("inline void near_force_of"):]
    body = body[:body.index("}")]
    assert body.count("x1 * g[0] - x0 * g[1] + x3 * g[2] - x2 * g[3]") == 1
    assert body.count("x2 * g[0] - x3 * g[1] - x0 * g[2] + x1 * g[3]") == 1
    assert body.count("x3 * g[0] + x2 * g[1] - x1 * g[2] - x0 * g[3]") == 1


def kernel_near_pair(x, y, taper_start, inv_width, terms):
    """near_pair_terms in FP64: returns (force contribution via -M(x) of coef (x - y), curv, energy)."""
    e = x - y
    dm2 = float(e @ e)
    nr = min(max(dm2, 1e-30), 4.0)
    fr = 4.0 - dm2
    R = 1 / math.sqrt(nr * fr)
    t = nr * R
    a = max(2 * (atan_series(t, terms) if terms else atan_poly(t)), 1e-12)
    b = math.pi - a
    aa, bb = a * a, b * b
    a2, b2 = aa + 1e-9, bb
    a3, b3 = aa * a + 1e-9, bb * b + 1e-9
    sine2 = t * fr + 2e-9  # twice the regularised sine
    P1, P2, P3 = a2 * b2 * sine2, a3 * b3, a * b
    D = 1 / (P1 * P2 * P3)
    E12 = D * P3
    iP1, iP2, iP3 = E12 * P2, E12 * P1, D * P1 * P2  # the pair energy is pi iP3
    u = min(max((a - taper_start) * inv_width, 0.0), 1.0)
    uu = u * u
    w = uu * (2 * u - 3) + 1
    slope_pi = -6 * math.pi * inv_width  # per source: pi times the slope's factor
    coef = w * (a2 - b2) * iP1 + iP3 * (slope_pi * (u - uu)) * R  # half the pair's coefficient
    # the kernel sums iP3 w and multiplies the source's sum by pi once, and doubles g
    return direction_M(x, 2 * coef * e), (a3 + b3) * iP2 * w, math.pi * (iP3 * w)


def test_near_pair_form_is_the_reference_pair():
    """Within a cutoff (near side, r up to 1.5 rad) the kernel's near-field pair equals the
    reference's windowed pair (with the same arctangent) to FP64 rounding, and the series
    arctangent moves it by less than the series bound."""
    rng = np.random.default_rng(5)
    for trial in range(3000):
        r = 10 ** rng.uniform(-2, math.log10(1.5))
        x = rng.normal(size=4)
        x /= np.linalg.norm(x)
        ang = r * rng.uniform(1e-3, 1.0)
        axis = rng.normal(size=4)
        axis -= axis.dot(x) * x
        axis /= np.linalg.norm(axis)
        y = math.cos(ang) * x + math.sin(ang) * axis
        dm2, dp2 = ((x - y) ** 2).sum(), ((x + y) ** 2).sum()
        window = (0.5 * r, 1 / (0.5 * r))
        z = math.tan(r / 2) ** 2 * 1.0001
        terms = next((k for k in range(2, 6) if z ** (k + 1) / (2 * k + 3) <= 5e-9), 0)
        fk, ck, ek = kernel_near_pair(x, y, *window, terms)
        atan = (lambda t: atan_series(t, terms)) if terms else atan_poly
        fr, cr, er = reference_pair(x, y, dm2, dp2, *window, atan=atan, exact_sin=True)
        f0, c0, e0 = reference_pair(x, y, dm2, dp2, atan=atan)
        scale_f = abs(f0).max() + abs(e0) * 20 * np.abs(x).sum() * 2 / r
        assert abs(ek - er) <= 1e-11 * e0
        assert abs(ck - cr) <= 1e-11 * c0
        assert np.abs(fk - fr).max() <= 1e-10 * scale_f
        # against the exact arctangent: a relative error delta in a (1e-8 for the series,
        # 6e-8 for the minimax polynomial in FP32 coefficients) moves E W by at most
        # E delta (a |E'/E| + a |W'|) <= E delta (2 + 6 a / width)
        fx, cx, ex = reference_pair(x, y, dm2, dp2, *window, exact_sin=True)
        delta = 1e-8 if terms else 6e-8
        a = 2 * math.atan(math.sqrt(dm2 / dp2))
        assert abs(ek - ex) <= e0 * delta * (2 + 6 * a * window[1])
