// FCC transport grids of T, O and I (src/grids.zig:10-77 and shaders/grids.wgsl:1-174 in
// the reference), in FP64, shared by the host (tables) and the device (one thread per
// lattice point).
//
// A source lattice (FCC points of an octahedron for T, of a truncated cube for O, of a
// cube for I) is carried onto the group's fundamental zone in homochoric coordinates by a
// volume-preserving map, then to quaternions.  The zone boundary in direction u is
// rho_max^3 = 1.5 (atan(1/c) - c / (1 + c^2)), c = 1 / (Rodrigues radius of the zone face
// along u), which is the homochoric radius (3/4 (w - sin w))^(1/3) at tan(w/2) = 1/c.
//  * T and O: per 1/48 chamber (sorted |x| <= |y| <= |z|, azimuth phi in [pi/4, pi/2],
//    polar angle up to the plane y = z), a Knothe-Rosenblatt map: the source chamber's
//    volume fraction below azimuth phi equals the zone chamber's below phi', the volume
//    fraction below polar angle theta along the azimuth equals the zone's below theta', and
//    the radius keeps its fraction of the boundary radius.
//  * I: the cube goes onto the homochoric ball by the cubochoric map (Rosca, Morawiec,
//    De Graef, Modelling Simul. Mater. Sci. Eng. 22 (2014) 075013); each of the ball's 120
//    chambers under I x {1,-1} maps onto the zone's corresponding chamber the same way, from
//    the sphere's area fractions to the zone's volume fractions.
// The zone-side azimuth distributions are Chebyshev series (64 terms per smooth piece) that
// the host fits to nested Gauss-Legendre quadrature; every other fraction is a closed form
// or a 32-point Gauss-Legendre integral evaluated per point.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace zxt::fcc {

#define ZXT_HD __host__ __device__ inline

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrt2 = 1.41421356237309504880;
constexpr int kChebTerms = 64;
constexpr int kGauss = 32;

// A Chebyshev series on [lo, hi].
struct Cheb {
  double lo, hi;
  double c[kChebTerms];
};

struct Tables {
  double gx[kGauss], gw[kGauss];  // Gauss-Legendre nodes and weights on [-1, 1]
  // Zone-side azimuth CDF (normalised to 1 over the chamber) and its derivative.
  Cheb t_cdf, t_pdf;               // T: [pi/4, pi/2]
  Cheb o_cdf[2], o_pdf[2];         // O: [pi/4, 3pi/8], [3pi/8, pi/2]
  double o_knee;                   // O: the CDF at 3pi/8
  Cheb i_cdf, i_pdf;               // I: [0, pi/5]
  // I: the group and inversion as matrices, and the chamber seed's images (sites[k] =
  // ops[k]^T seed: a direction nearest sites[k] is carried into the chamber by ops[k]).
  double i_ops[120][9];
  double i_sites[120][3];
  // Cubochoric constants: (pi/6)^(1/6), sqrt(6/pi), R 2^(1/4) / (pi^(5/6) 6^(-1/6) / 2) with
  // R = (3 pi / 4)^(1/3), and sqrt(pi / 24).
  double cb_scale, cb_pref, cb_prek, cb_lift;
};

struct Job {
  uint32_t group;   // 8 T, 9 O, 10 I
  uint32_t count;   // lattice points
  uint32_t side;    // 2 * bound + 1 (rows are (i, j) pairs, i-major)
  int32_t bound;    // |i|, |j|, |k| <= bound: 2h for T, h for O and I
  int32_t h;        // semi-edge
  int32_t cut;      // O: |i|+|j|+|k| < cut
  double step;      // lattice spacing
  double beta;      // O: truncation plane |x|+|y|+|z| <= 1 + beta (cube half-edge 1)
  double src_total; // O: the source chamber's azimuth integral over [pi/4, pi/2]
};

ZXT_HD double clamp(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Row (i, j) of the source lattice (grids.zig:40-60): FCC points (i + j + k even) with
// |i|+|j|+|k| <= 2h (T), |k| <= h and |i|+|j|+|k| < cut (O), |k| <= h (I); k runs from k0
// to kmax in steps of 2.  Returns the row's point count.
ZXT_HD uint32_t row_span(const Job& job, int i, int j, int& k0, int& kmax) {
  const int a = (i < 0 ? -i : i) + (j < 0 ? -j : j);
  if (job.group == 8)
    kmax = 2 * job.h - a;
  else if (job.group == 9)
    kmax = job.cut - 1 - a < job.h ? job.cut - 1 - a : job.h;
  else
    kmax = job.h;
  k0 = ((kmax + i + j) & 1) ? 1 - kmax : -kmax;
  if (kmax < 0 || k0 > kmax) return 0;
  return uint32_t((kmax - k0) / 2 + 1);
}

ZXT_HD double cheb(const Cheb& s, double x) {
  const double t = (2.0 * x - s.lo - s.hi) / (s.hi - s.lo);
  double b1 = 0.0, b2 = 0.0;
  for (int k = kChebTerms - 1; k >= 1; --k) {
    const double b0 = 2.0 * t * b1 - b2 + s.c[k];
    b2 = b1;
    b1 = b0;
  }
  return t * b1 - b2 + s.c[0];
}

// int_a^b f by an n-point Gauss-Legendre rule.
template <typename F>
ZXT_HD double gauss(const double* x, const double* w, int n, double a, double b, F f) {
  const double half = 0.5 * (b - a), mid = 0.5 * (b + a);
  double sum = 0.0;
  for (int k = 0; k < n; ++k) sum += w[k] * f(half * x[k] + mid);
  return half * sum;
}

// Homochoric radius cubed of the point at Rodrigues radius 1/c.
ZXT_HD double rho3(double c) {
  const double s = c > 1e-300 ? c : 1e-300;
  const double r = 1.5 * (atan(1.0 / s) - s / (1.0 + s * s));
  return r > 0.0 ? r : 0.0;
}

// Polar limit of the T/O chamber x <= y <= z at azimuth phi: the plane y = z.
ZXT_HD double chamber_polar(double phi) { return atan(1.0 / sin(phi)); }

// Zone boundaries.  T: |r|_1 <= 1.  O: also |r_i| <= sqrt 2 - 1 (the two meet on the
// chamber's far wall at phi = 3pi/8 and, inside it, at polar angle o_split).
ZXT_HD double t_r3(double th, double ph) { return rho3(cos(th) + (cos(ph) + sin(ph)) * sin(th)); }
ZXT_HD double o_r3(double th, double ph) {
  const double a = (1.0 + kSqrt2) * cos(th), b = cos(th) + (cos(ph) + sin(ph)) * sin(th);
  return rho3(a > b ? a : b);
}
ZXT_HD double o_split(double ph) { return atan(kSqrt2 / (cos(ph) + sin(ph))); }

// I: the zone's chamber is azimuth [0, pi/5] about the five-fold axis z, inside the
// bisector of z and the five-fold axis (2/sqrt5, 0, 1/sqrt5); its face is z . r <= tan(pi/10).
struct IWall {
  static constexpr double x = -0.85065080835203993218;  // unit(z - (2, 0, 1)/sqrt 5)
  static constexpr double z = 0.52573111211913360603;
};
constexpr double kInvTan10 = 3.07768353717525340219;  // 1 / tan(pi/10)
ZXT_HD double i_polar(double ph) { return kPi - atan2(IWall::z, IWall::x * cos(ph)); }
ZXT_HD double i_r3(double th) { return rho3(cos(th) * kInvTan10); }
// Spherical area per unit azimuth of the chamber, 1 - cos(i_polar), integrated.
ZXT_HD double i_cap(double ph) { return ph - asin(clamp(-IWall::x * sin(ph), -1.0, 1.0)); }

// O source: the truncated cube chamber, radius min(1/cos, (1+beta)/(cos + s sin)),
// s = cos phi + sin phi; int_0^theta R^3 sin in closed form (t = tan theta).
ZXT_HD double o_source_inner(double th, double ph, double beta) {
  const double s = cos(ph) + sin(ph);
  const double split = atan(beta / s);
  const double tt = tan(th), ts = beta / s;
  if (th <= split) return 0.5 * tt * tt;
  const double qt = 1.0 + s * tt, qs = 1.0 + s * ts;
  const double g = (-1.0 / qt + 0.5 / (qt * qt)) - (-1.0 / qs + 0.5 / (qs * qs));
  return 0.5 * ts * ts + (1.0 + beta) * (1.0 + beta) * (1.0 + beta) * g / (s * s);
}
// Source azimuth integral of the O chamber from pi/4 to phi (kink where the truncation
// plane leaves the far wall, at phi_b = atan(1 / (beta - 1))).
ZXT_HD double o_source_cdf(const double* x, const double* w, int n, double ph, double beta) {
  const double knee = atan(1.0 / (beta - 1.0));
  auto f = [&](double p) { return o_source_inner(chamber_polar(p), p, beta); };
  const double a = ph < knee ? ph : knee;
  double v = gauss(x, w, n, kPi / 4, a, f);
  if (ph > knee) v += gauss(x, w, n, knee, ph, f);
  return v;
}

// Bracketed Newton for the increasing function f(x) = target on [lo, hi]; fd returns
// (f(x), f'(x)).
template <typename FD>
ZXT_HD double solve(double lo, double hi, double x, double target, FD fd) {
  for (int it = 0; it < 60; ++it) {
    double d;
    const double r = fd(x, d) - target;
    if (r < 0)
      lo = x;
    else if (r > 0)
      hi = x;
    else
      return x;
    double next = x - r / (d > 1e-300 ? d : 1e-300);
    if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
    if (fabs(next - x) <= 1e-15 * (fabs(x) > 1.0 ? fabs(x) : 1.0)) return next;
    x = next;
  }
  return x;
}

// The azimuth where a normalised CDF series reaches u.
ZXT_HD double invert(const Cheb& cdf, const Cheb& pdf, double u) {
  const double x0 = cdf.lo + clamp(u, 0.0, 1.0) * (cdf.hi - cdf.lo);
  return solve(cdf.lo, cdf.hi, x0, u, [&](double x, double& d) {
    d = cheb(pdf, x);
    return cheb(cdf, x);
  });
}

// The polar angle in [0, high] where int_0^theta f / int_0^high f = y; f has a kink at
// `knee` when 0 < knee < high.
template <typename F>
ZXT_HD double polar_solve(const Tables& tb, double y, double high, double knee, F f) {
  auto integral = [&](double th) {
    if (knee > 0.0 && th > knee)
      return gauss(tb.gx, tb.gw, kGauss, 0.0, knee, f) + gauss(tb.gx, tb.gw, kGauss, knee, th, f);
    return gauss(tb.gx, tb.gw, kGauss, 0.0, th, f);
  };
  const double total = integral(high);
  const double x0 = 2.0 * asin(clamp(sqrt(y) * sin(0.5 * high), 0.0, 1.0));
  return solve(0.0, high, clamp(x0, 0.0, high), y, [&](double th, double& d) {
    d = f(th) / total;
    return integral(th) / total;
  });
}

// |v| sorted ascending (stable), with the original axis of each.
struct Sorted {
  double v[3];
  int axis[3];
};
ZXT_HD Sorted sort_abs(const double* cu) {
  Sorted s{{fabs(cu[0]), fabs(cu[1]), fabs(cu[2])}, {0, 1, 2}};
  auto swap = [&](int a, int b) {
    if (s.v[a] > s.v[b]) {
      const double t = s.v[a];
      s.v[a] = s.v[b];
      s.v[b] = t;
      const int k = s.axis[a];
      s.axis[a] = s.axis[b];
      s.axis[b] = k;
    }
  };
  swap(0, 1);
  swap(1, 2);
  swap(0, 1);
  return s;
}
// Places a chamber vector back on the source point's axes and signs.
ZXT_HD void unsort(const Sorted& s, const double* d, const double* cu, double* out) {
  for (int k = 0; k < 3; ++k) {
    const int a = s.axis[k];
    out[a] = cu[a] >= 0.0 ? d[k] : -d[k];
  }
}

ZXT_HD void direction(double th, double ph, double r, double* d) {
  d[0] = r * sin(th) * cos(ph);
  d[1] = r * sin(th) * sin(ph);
  d[2] = r * cos(th);
}

// ---- the three maps ------------------------------------------------------------------
constexpr double kTEdge = 0.85125548036918854;  // (pi^2 / 16)^(1/3): T source |x|_1 bound
constexpr double kOEdge = 1.07251469855551277;  // pi^(2/3) / 2: O and I cube half-edge
constexpr double kBall = 1.33067003949146893;   // (3 pi / 4)^(1/3): homochoric ball radius

ZXT_HD void t_map(const Tables& tb, const double* cu, double* ho) {
  const Sorted s = sort_abs(cu);
  const double r = sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
  const double th = atan2(sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1]), s.v[2]);
  const double ph = clamp(atan2(s.v[1], s.v[0]), kPi / 4, kPi / 2);
  const double cp = cos(ph), sp = sin(ph);
  const double face = cos(th) + (cp + sp) * sin(th);
  // The flat chamber's azimuth and polar volume fractions.
  const double u = clamp(2.0 * kSqrt2 * sin(ph - kPi / 4) / (cp + 2.0 * sp), 0.0, 1.0);
  const double ys = sin(th) * (cp + 2.0 * sp) / face;
  const double y = clamp(ys * ys, 0.0, 1.0);
  const double ph2 = invert(tb.t_cdf, tb.t_pdf, u);
  const double th2 = polar_solve(tb, y, chamber_polar(ph2), 0.0, [&](double t) { return t_r3(t, ph2) * sin(t); });
  double d[3];
  direction(th2, ph2, cbrt(t_r3(th2, ph2)) * r * face / kTEdge, d);
  unsort(s, d, cu, ho);
}

ZXT_HD void o_map(const Tables& tb, const Job& job, const double* cu, double* ho) {
  const Sorted s = sort_abs(cu);
  const double r = sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2]);
  const double th = atan2(sqrt(s.v[0] * s.v[0] + s.v[1] * s.v[1]), s.v[2]);
  const double ph = clamp(atan2(s.v[1], s.v[0]), kPi / 4, kPi / 2);
  const double beta = job.beta;
  const double u = clamp(o_source_cdf(tb.gx, tb.gw, kGauss, ph, beta) / job.src_total, 0.0, 1.0);
  const double y = clamp(o_source_inner(th, ph, beta) / o_source_inner(chamber_polar(ph), ph, beta), 0.0, 1.0);
  const int piece = u <= tb.o_knee ? 0 : 1;
  const double ph2 = invert(tb.o_cdf[piece], tb.o_pdf[piece], u);
  const double th2 =
      polar_solve(tb, y, chamber_polar(ph2), o_split(ph2), [&](double t) { return o_r3(t, ph2) * sin(t); });
  const double sum = cos(ph) + sin(ph);
  const double a = 1.0 / cos(th), b = (1.0 + beta) / (cos(th) + sin(th) * sum);
  const double r_src = kOEdge * (a < b ? a : b);
  double d[3];
  direction(th2, ph2, cbrt(o_r3(th2, ph2)) * r / r_src, d);
  unsort(s, d, cu, ho);
}

// The cubochoric cube -> ball map: the pyramid of the largest |coordinate| (the first on a
// tie) is flattened, taken to a disc by the Rosca-Plonka equal-area square map and lifted
// onto the ball by the inverse Lambert azimuthal projection.
ZXT_HD void cube_to_ball(const Tables& tb, const double* xyz, double* out) {
  int top = 0;
  if (fabs(xyz[1]) > fabs(xyz[top])) top = 1;
  if (fabs(xyz[2]) > fabs(xyz[top])) top = 2;
  // The largest coordinate plays z; the other two keep their cyclic order.
  const int ax = (top + 1) % 3, ay = (top + 2) % 3;
  const double x = xyz[ax] * tb.cb_scale, y = xyz[ay] * tb.cb_scale, z = xyz[top] * tb.cb_scale;
  const double pref = tb.cb_pref, prek = tb.cb_prek;
  double l[3];
  if (x == 0.0 && y == 0.0) {
    l[0] = 0.0;
    l[1] = 0.0;
    l[2] = pref * z;
  } else {
    const bool xfirst = fabs(y) <= fabs(x);
    const double big = xfirst ? x : y, small = xfirst ? y : x;
    const double ang = kPi / 12.0 * small / big;
    const double c = cos(ang), sn = sin(ang);
    const double q = prek * big / sqrt(kSqrt2 - c);
    const double along = (kSqrt2 * c - 1.0) * q, across = kSqrt2 * sn * q;
    const double t1 = xfirst ? along : across, t2 = xfirst ? across : along;
    const double cc = t1 * t1 + t2 * t2;
    const double lift = sqrt(fmax(1.0 - kPi * cc / (24.0 * z * z), 0.0));
    l[0] = t1 * lift;
    l[1] = t2 * lift;
    l[2] = pref * z - tb.cb_lift * cc / z;
  }
  out[ax] = l[0];
  out[ay] = l[1];
  out[top] = l[2];
}

ZXT_HD void i_map(const Tables& tb, const double* cu, double* ho) {
  double ball[3];
  cube_to_ball(tb, cu, ball);
  const double radius = sqrt(ball[0] * ball[0] + ball[1] * ball[1] + ball[2] * ball[2]);
  const double d[3] = {ball[0] / radius, ball[1] / radius, ball[2] / radius};
  int best = 0;
  double score = -2.0;
  for (int k = 0; k < 120; ++k) {
    const double sc = d[0] * tb.i_sites[k][0] + d[1] * tb.i_sites[k][1] + d[2] * tb.i_sites[k][2];
    if (sc > score) {
      score = sc;
      best = k;
    }
  }
  const double* m = tb.i_ops[best];
  const double f[3] = {m[0] * d[0] + m[1] * d[1] + m[2] * d[2], m[3] * d[0] + m[4] * d[1] + m[5] * d[2],
                       m[6] * d[0] + m[7] * d[1] + m[8] * d[2]};
  const double th = atan2(sqrt(f[0] * f[0] + f[1] * f[1]), f[2]);
  const double ph = clamp(atan2(f[1], f[0]), 0.0, kPi / 5);
  const double ys = sin(0.5 * th) / sin(0.5 * i_polar(ph));
  const double y = clamp(ys * ys, 0.0, 1.0);
  const double c0 = i_cap(0.0);
  const double u = clamp((i_cap(ph) - c0) / (i_cap(kPi / 5) - c0), 0.0, 1.0);
  const double ph2 = invert(tb.i_cdf, tb.i_pdf, u);
  const double th2 = polar_solve(tb, y, i_polar(ph2), 0.0, [&](double t) { return i_r3(t) * sin(t); });
  double g[3];
  direction(th2, ph2, radius * cbrt(i_r3(th2)) / kBall, g);
  for (int k = 0; k < 3; ++k) ho[k] = m[k] * g[0] + m[3 + k] * g[1] + m[6 + k] * g[2];
}

// Homochoric vector -> unit quaternion: |h|^3 = (3/4)(w - sin w), q = (cos w/2, sin w/2 h/|h|).
ZXT_HD double omega_minus_sine(double w) {
  if (w < 0.5) {  // the series avoids the cancellation of w - sin w
    const double w2 = w * w;
    double term = w * w2 / 6.0, sum = 0.0;
    for (int k = 1; k < 12; ++k) {
      sum += term;
      term *= -w2 / double((2 * k + 2) * (2 * k + 3));
    }
    return sum;
  }
  return w - sin(w);
}
ZXT_HD void homochoric_to_quaternion(const double* ho, double* q) {
  const double rho = sqrt(ho[0] * ho[0] + ho[1] * ho[1] + ho[2] * ho[2]);
  if (rho == 0.0) {
    q[0] = 1.0;
    q[1] = q[2] = q[3] = 0.0;
    return;
  }
  const double target = (4.0 / 3.0) * rho * rho * rho;
  const double w = solve(0.0, kPi, fmin(cbrt(6.0 * target), kPi), target, [&](double x, double& d) {
    const double h = sin(0.5 * x);
    d = 2.0 * h * h;
    return omega_minus_sine(x);
  });
  const double s = sin(0.5 * w) / rho;
  q[0] = cos(0.5 * w);
  q[1] = ho[0] * s;
  q[2] = ho[1] * s;
  q[3] = ho[2] * s;
  const double n = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) q[k] *= n;
}

// One lattice point (i, j, k) of the job to its quaternion.
ZXT_HD void transport_point(const Tables& tb, const Job& job, int i, int j, int k, double* q) {
  const double cu[3] = {i * job.step, j * job.step, k * job.step};
  if (i == 0 && j == 0 && k == 0) {
    q[0] = 1.0;
    q[1] = q[2] = q[3] = 0.0;
    return;
  }
  double ho[3];
  if (job.group == 8)
    t_map(tb, cu, ho);
  else if (job.group == 9)
    o_map(tb, job, cu, ho);
  else
    i_map(tb, cu, ho);
  homochoric_to_quaternion(ho, q);
}

#undef ZXT_HD

}  // namespace zxt::fcc
