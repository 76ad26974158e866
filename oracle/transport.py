"""ORACLE — test infrastructure only: FP64 numpy restatement of the T/O/I FCC transports
(src/grids.zig:10-77, shaders/grids.wgsl:1-174 in the reference; their published record
is tests/fixtures/reference-small.npz and transports-h8-h12.npz, CPU float64).

Each source lattice (FCC points of an octahedron for T, a truncated cube for O, a cube
for I) is carried onto the group's fundamental zone in homochoric coordinates by a
volume-preserving map, then to quaternions.  This module computes every map from its
definition, with nested Gauss-Legendre quadrature and Newton/bisection solves and no
fitted tables, so it is independent of the device code's Chebyshev tables:

* T, O: per 1/48 chamber (sorted |x| <= |y| <= |z|), a Knothe-Rosenblatt map: the source
  chamber's volume fraction below azimuth phi equals the target's below phi'; along the
  azimuth, the volume fraction below polar angle theta equals the target's below theta';
  the radius keeps its fraction of the boundary radius.  The target boundary in
  homochoric space is rho^3 = 1.5 (atan(1/c) - c / (1 + c^2)) with c = 1 / (Rodrigues
  radius of the zone face in that direction).
* I: the cube goes to the homochoric ball by the cubochoric map (Rosca, Morawiec, De Graef,
  Modelling Simul. Mater. Sci. Eng. 22 (2014) 075013); each of the ball's 120 I_h chambers
  then maps onto the zone's chamber the same way (area fraction on the sphere, volume
  fraction in the zone).
* Homochoric to quaternion: |h|^3 = (3/4)(w - sin w), q = (cos w/2, sin w/2 h/|h|).
"""
import math

import numpy as np
from numpy.polynomial.legendre import leggauss

PI = math.pi
S2 = math.sqrt(2.0)
R_BALL = (3 * PI / 4) ** (1 / 3)
_X, _W = leggauss(64)
CHUNK = 1024


def cut_index(h):
    """The odd lattice sum nearest (1 + sqrt 2)(h + 1/2), the lower on a tie (grids.zig:10-15)."""
    return min(range(1, 10 * h + 5, 2), key=lambda x: (abs(x - (1 + S2) * (h + 0.5)), x))


def lattice(group, h):
    """Source points in the reference's order (i, then j, then k ascending) and beta."""
    cut = cut_index(h)
    bound = 2 * h if group == "T" else h
    ii = np.arange(-bound, bound + 1)
    idx = np.array(np.meshgrid(ii, ii, ii, indexing="ij")).reshape(3, -1).T
    keep = idx.sum(1) % 2 == 0
    if group == "T":
        keep &= np.abs(idx).sum(1) <= 2 * h
    elif group == "O":
        keep &= np.abs(idx).sum(1) < cut
    step = ((PI ** 2 / 16) ** (1 / 3) if group == "T" else PI ** (2 / 3)) / (2 * h + 1)
    return idx[keep] * step, cut / (h + 0.5) - 1


def rho3(c):
    s = np.maximum(c, 1e-300)
    return np.maximum(0.0, 1.5 * (np.arctan(1.0 / s) - s / (1.0 + s * s)))


def quad(f, a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    t = a[..., None] + 0.5 * (b - a)[..., None] * (_X + 1)
    return 0.5 * (b - a) * (_W * f(t)).sum(-1)


def newton(f, df, x, lo, hi, target, iterations=60):
    """Bracketed Newton for f(x) = target on [lo, hi], f increasing."""
    lo = np.broadcast_to(np.asarray(lo, float), x.shape).copy()
    hi = np.broadcast_to(np.asarray(hi, float), x.shape).copy()
    for _ in range(iterations):
        r = f(x) - target
        lo = np.where(r < 0, x, lo)
        hi = np.where(r > 0, x, hi)
        step = x - r / np.maximum(df(x), 1e-300)
        x_new = np.where((step > lo) & (step < hi), step, 0.5 * (lo + hi))
        done = np.all(np.abs(x_new - x) <= 1e-15 * np.maximum(1.0, np.abs(x)))
        x = x_new
        if done:
            break
    return x


_START = {}


def _start(name, cdf, lo, hi, u):
    """A starting azimuth for the CDF inversion: linear interpolation in a 4097-point table
    (errors ~1e-8, so a few Newton steps on the exact CDF finish the solve)."""
    if name not in _START:
        grid = np.linspace(lo, hi, 4097)
        values = np.concatenate([cdf(part) for part in np.array_split(grid, 16)])
        _START[name] = (grid, values / values[-1])
    grid, values = _START[name]
    return np.interp(u, values, grid)


def direction(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def polar_of_sorted(cu):
    a = np.abs(cu)
    order = np.argsort(a, axis=1, kind="stable")
    v = np.take_along_axis(a, order, 1)
    theta = np.arctan2(np.hypot(v[:, 0], v[:, 1]), v[:, 2])
    phi = np.arctan2(v[:, 1], v[:, 0])
    return np.linalg.norm(v, axis=1), theta, phi, order


def unsort(d, order, cu):
    out = np.zeros_like(d)
    np.put_along_axis(out, order, d, 1)
    return out * np.where(cu >= 0, 1.0, -1.0)


def chamber_theta_max(phi):
    """Polar limit of the chamber x <= y <= z at azimuth phi (the plane y = z)."""
    return np.arctan(1.0 / np.sin(phi))


# ---- T: octahedron |x|+|y|+|z| <= (pi^2/16)^(1/3) onto the zone |r|_1 <= 1 (Rodrigues)
def _t_r3(theta, phi):
    return rho3(np.cos(theta) + (np.cos(phi) + np.sin(phi)) * np.sin(theta))


def _t_inner(theta, phi):
    return quad(lambda t: _t_r3(t, phi[..., None]) * np.sin(t), 0 * theta, theta)


def _t_density(phi):
    return _t_inner(chamber_theta_max(phi), phi)


def _t_cdf(phi):
    return quad(_t_density, PI / 4 + 0 * phi, phi)


def _t_map(cu):
    r, theta, phi, order = polar_of_sorted(cu)
    phi = np.clip(phi, PI / 4, PI / 2)
    cp, sp = np.cos(phi), np.sin(phi)
    face = np.cos(theta) + (cp + sp) * np.sin(theta)
    # Source: the flat chamber's azimuth and polar volume fractions (closed forms).
    u = np.clip(2 * S2 * np.sin(phi - PI / 4) / (cp + 2 * sp), 0, 1)
    y = np.clip((np.sin(theta) * (cp + 2 * sp) / face) ** 2, 0, 1)
    total = _t_cdf(np.array(PI / 2))
    phi2 = newton(lambda x: _t_cdf(x) / total, lambda x: _t_density(x) / total,
                  _start("t", _t_cdf, PI / 4, PI / 2, u), PI / 4, PI / 2, u)
    high = chamber_theta_max(phi2)
    area = _t_density(phi2)
    theta2 = newton(lambda x: _t_inner(x, phi2) / area, lambda x: _t_r3(x, phi2) * np.sin(x) / area,
                    2 * np.arcsin(np.sqrt(y) * np.sin(0.5 * high)), 0.0, high, y)
    r_src = (PI ** 2 / 16) ** (1 / 3) / face
    return unsort(direction(theta2, phi2) * (np.cbrt(_t_r3(theta2, phi2)) * r / r_src)[:, None], order, cu)


# ---- O: truncated cube (|x_i| <= 1, |x|+|y|+|z| <= 1 + beta, in units of pi^(2/3)/2)
# onto the zone |r_i| <= sqrt 2 - 1, |r|_1 <= 1
def _o_r3(theta, phi):
    return rho3(np.maximum((1 + S2) * np.cos(theta), np.cos(theta) + (np.cos(phi) + np.sin(phi)) * np.sin(theta)))


def _o_inner(theta, phi):
    split = np.arctan(S2 / (np.cos(phi) + np.sin(phi)))
    f = lambda t: _o_r3(t, phi[..., None]) * np.sin(t)
    a = np.minimum(theta, split)
    return quad(f, 0 * theta, a) + np.where(theta > split, quad(f, a, theta), 0)


def _o_density(phi):
    return _o_inner(chamber_theta_max(phi), phi)


def _o_cdf(phi):
    k = 3 * PI / 8  # the two zone faces meet on the chamber's far wall here
    a = np.minimum(phi, k)
    return quad(_o_density, PI / 4 + 0 * phi, a) + np.where(phi > k, quad(_o_density, a, phi), 0)


def _o_source_inner(theta, phi, beta):
    """int_0^theta R^3 sin for R = min(1 / cos, (1 + beta) / (cos + s sin)), s = cos phi + sin phi."""
    s = np.cos(phi) + np.sin(phi)
    split = np.arctan(beta / s)
    tt, ts = np.tan(theta), np.tan(split)
    g = lambda t: (-1 / (1 + s * t) + 0.5 / (1 + s * t) ** 2) / (s * s)
    return np.where(theta <= split, 0.5 * tt * tt, 0.5 * ts * ts + (1 + beta) ** 3 * (g(tt) - g(ts)))


def _o_source_cdf(phi, beta):
    b = math.atan(1 / (beta - 1))
    f = lambda p: _o_source_inner(chamber_theta_max(p), p, beta)
    a = np.minimum(phi, b)
    return quad(f, PI / 4 + 0 * phi, a) + np.where(phi > b, quad(f, a, phi), 0)


def _o_map(cu, beta):
    r, theta, phi, order = polar_of_sorted(cu)
    phi = np.clip(phi, PI / 4, PI / 2)
    s = np.cos(phi) + np.sin(phi)
    u = np.clip(_o_source_cdf(phi, beta) / _o_source_cdf(np.array(PI / 2), beta), 0, 1)
    y = np.clip(_o_source_inner(theta, phi, beta) / _o_source_inner(chamber_theta_max(phi), phi, beta), 0, 1)
    total = _o_cdf(np.array(PI / 2))
    phi2 = newton(lambda x: _o_cdf(x) / total, lambda x: _o_density(x) / total,
                  _start("o", _o_cdf, PI / 4, PI / 2, u), PI / 4, PI / 2, u)
    high = chamber_theta_max(phi2)
    area = _o_density(phi2)
    theta2 = newton(lambda x: _o_inner(x, phi2) / area, lambda x: _o_r3(x, phi2) * np.sin(x) / area,
                    2 * np.arcsin(np.sqrt(y) * np.sin(0.5 * high)), 0.0, high, y)
    r_src = PI ** (2 / 3) / 2 * np.minimum(1 / np.cos(theta), (1 + beta) / (np.cos(theta) + np.sin(theta) * s))
    return unsort(direction(theta2, phi2) * (np.cbrt(_o_r3(theta2, phi2)) * r / r_src)[:, None], order, cu)


# ---- I: cube of edge pi^(2/3) -> ball (cubochoric) -> zone (dodecahedron, faces at tan(pi/10))
def cube_to_ball(xyz):
    """Rosca-Morawiec-De Graef: the pyramid of the largest |coordinate| (the first on a tie)
    is flattened to a square, taken to a disc by the Rosca-Plonka equal-area map, and
    lifted onto the ball by the inverse Lambert projection."""
    xyz = np.asarray(xyz, float)
    axis = np.argmax(np.abs(xyz), axis=1)
    perm = np.array([[1, 2, 0], [2, 0, 1], [0, 1, 2]])[axis]  # the largest coordinate goes last
    s = np.take_along_axis(xyz, perm, 1) * (PI / 6) ** (1 / 6)
    pref = math.sqrt(6 / PI)
    prek = R_BALL * 2 ** 0.25 / (PI ** (5 / 6) / 6 ** (1 / 6) / 2)
    x, y, z = s[:, 0], s[:, 1], s[:, 2]
    ax, ay = np.abs(x), np.abs(y)
    on_axis = np.maximum(ax, ay) == 0
    xfirst = ay <= ax
    big = np.where(xfirst, x, y)
    small = np.where(xfirst, y, x)
    ratio = np.where(on_axis, 0.0, small / np.where(big == 0, 1.0, big))
    c, sn = np.cos(PI / 12 * ratio), np.sin(PI / 12 * ratio)
    q = prek * big / np.sqrt(S2 - c)
    along, across = (S2 * c - 1) * q, S2 * sn * q
    t1 = np.where(xfirst, along, across)
    t2 = np.where(xfirst, across, along)
    cc = t1 * t1 + t2 * t2
    zsafe = np.where(z == 0, 1.0, z)
    lift = np.sqrt(np.maximum(1 - PI * cc / (24 * zsafe ** 2), 0))
    out = np.stack([t1 * lift, t2 * lift, pref * z - math.sqrt(PI) * cc / math.sqrt(24) / zsafe], 1)
    out[on_axis] = np.stack([0 * z, 0 * z, pref * z], 1)[on_axis]
    back = np.argsort(perm, axis=1)
    return np.take_along_axis(out, back, 1)


def _quat_mul(a, b):
    return np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                     a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                     a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                     a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])


def _matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def icosahedral_ops():
    """The 120 elements of I x {1, -1} as matrices, I closed from rotations by 2 pi / 5
    about z and about (2, 0, 1) (the operator tables' frame)."""
    def turn(axis, k):
        a = np.asarray(axis, float) / np.linalg.norm(axis)
        return np.array([math.cos(PI / k), *(a * math.sin(PI / k))])
    gens = [turn([0, 0, 1], 5), turn([2, 0, 1], 5)]
    elements = [np.array([1.0, 0, 0, 0])]
    i = 0
    while i < len(elements):
        for g in gens:
            p = _quat_mul(elements[i], g)
            if not any(abs(np.dot(e, p)) > 1 - 1e-9 for e in elements):
                elements.append(p)
        i += 1
    mats = [_matrix(q) for q in elements]
    return np.array(mats + [-m for m in mats])


# The chamber: azimuth [0, pi/5] around the five-fold axis z, inside the bisector of z and
# the neighbouring five-fold axis at azimuth 0.
_I_WALL = np.array([0, 0, 1.0]) - np.array([2 / math.sqrt(5), 0, 1 / math.sqrt(5)])
_I_WALL /= np.linalg.norm(_I_WALL)
_TAN10 = math.tan(PI / 10)


def _i_theta_max(phi):
    return PI - np.arctan2(_I_WALL[2], _I_WALL[0] * np.cos(phi) + _I_WALL[1] * np.sin(phi))


def _i_r3(theta):
    return rho3(np.cos(theta) / _TAN10)


def _i_inner(theta):
    return quad(lambda t: _i_r3(t) * np.sin(t), 0 * theta, theta)


def _i_density(phi):
    return _i_inner(_i_theta_max(phi))


def _i_cdf(phi):
    return quad(_i_density, 0 * phi, phi)


def _cap_cdf(phi):
    return phi - np.arcsin(np.clip(-_I_WALL[0] * np.sin(phi) + _I_WALL[1] * np.cos(phi), -1, 1))


def _i_map(cu):
    ops = icosahedral_ops()
    seed = direction(np.array(0.15), np.array(PI / 10))
    sites = np.einsum("kji,j->ki", ops, seed)
    ho = cube_to_ball(cu)
    radius = np.linalg.norm(ho, axis=1)
    d = ho / radius[:, None]
    m = ops[np.argmax(d @ sites.T, axis=1)]
    f = np.einsum("nij,nj->ni", m, d)
    theta = np.arctan2(np.hypot(f[:, 0], f[:, 1]), f[:, 2])
    phi = np.clip(np.arctan2(f[:, 1], f[:, 0]), 0, PI / 5)
    y = np.clip((np.sin(0.5 * theta) / np.sin(0.5 * _i_theta_max(phi))) ** 2, 0, 1)
    u = np.clip((_cap_cdf(phi) - _cap_cdf(np.array(0.0))) / (_cap_cdf(np.array(PI / 5)) - _cap_cdf(np.array(0.0))), 0, 1)
    total = _i_cdf(np.array(PI / 5))
    phi2 = newton(lambda x: _i_cdf(x) / total, lambda x: _i_density(x) / total,
                  _start("i", _i_cdf, 0.0, PI / 5, u), 0.0, PI / 5, u)
    high = _i_theta_max(phi2)
    area = _i_density(phi2)
    theta2 = newton(lambda x: _i_inner(x) / area, lambda x: _i_r3(x) * np.sin(x) / area,
                    2 * np.arcsin(np.sqrt(y) * np.sin(0.5 * high)), 0.0, high, y)
    mapped = direction(theta2, phi2) * (radius * np.cbrt(_i_r3(theta2)) / R_BALL)[:, None]
    return np.einsum("nji,nj->ni", m, mapped)


def homochoric_to_quaternion(ho):
    rho = np.linalg.norm(ho, axis=1)
    target = (4 / 3) * rho ** 3  # w - sin w
    w = newton(lambda x: x - np.sin(x), lambda x: 1 - np.cos(x), np.cbrt(6 * target), 0.0, PI, target, 80)
    scale = np.sin(0.5 * w) / np.where(rho > 0, rho, 1)
    q = np.concatenate([np.cos(0.5 * w)[:, None], ho * scale[:, None]], 1)
    q[rho == 0] = [1, 0, 0, 0]
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def transport(group, h, subset=None):
    """The transport grid of T, O or I at semi-edge h, float64 (N, 4), in lattice order
    (only the lattice points `subset` indexes, when given)."""
    cu, beta = lattice(group, h)
    if subset is not None:
        cu = cu[np.asarray(subset)]
    out = np.empty((len(cu), 4))
    for start in range(0, len(cu), CHUNK):
        part = cu[start:start + CHUNK]
        zero = (part == 0).all(1)
        safe = np.where(zero[:, None], 1.0, part)
        with np.errstate(divide="ignore", invalid="ignore"):
            if group == "T":
                ho = _t_map(safe)
            elif group == "O":
                ho = _o_map(safe, beta)
            else:
                ho = _i_map(safe)
        q = homochoric_to_quaternion(ho)
        q[zero] = [1, 0, 0, 0]
        out[start:start + CHUNK] = q
    return out
