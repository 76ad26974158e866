// Device arithmetic of the QVM path for gfx950.
//
// Quaternions are float4 with x the scalar part (scalar first, right action q*g), as
// the reference stores them (src/groups.zig:1).
//
// Two kinds of arithmetic live here:
//  * exact-order functions (qmul, dot4, normalize, image, chord2): explicit fma, correctly
//    rounded sqrt and division, built with -ffp-contract=off.  The CPU oracle restates the
//    same order, so images, chords, pair sets and nearest-neighbour ranks agree bit for bit.
//  * the pair kernel (pair_terms): the reference's quey_pair and near_pair
//    (shaders/pairs.wgsl:246-262, :349-361) rewritten for the CDNA VALU.  One v_rsq and one
//    v_rcp per pair replace the sqrt, atan and four divisions: t = tan(a/2) = near*rsq(near*far),
//    sin a = 2t/(1+t^2) = t*far/2 (near+far = 4 for unit quaternions), 1/sin a =
//    (near+far)/(2 sqrt(near far)), atan(t) = t*P(t^2) on [0,1] (fit_atan.py), and the three
//    reciprocals 1/(a2 b2 sine), 1/(a3 b3), 1/(a b) from one reciprocal of their product.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include <cstdint>

namespace zxt {

constexpr float kPiF = 3.14159265358979323846f;
constexpr double kPi64 = 3.14159265358979323846;

__host__ __device__ inline float4 qmul(const float4 a, const float4 b) {
  float4 r;
  r.x = fmaf(-a.w, b.w, fmaf(-a.z, b.z, fmaf(-a.y, b.y, a.x * b.x)));
  r.y = fmaf(-a.w, b.z, fmaf(a.z, b.w, fmaf(b.x, a.y, a.x * b.y)));
  r.z = fmaf(-a.y, b.w, fmaf(a.w, b.y, fmaf(b.x, a.z, a.x * b.z)));
  r.w = fmaf(-a.z, b.y, fmaf(a.y, b.z, fmaf(b.x, a.w, a.x * b.w)));
  return r;
}
__host__ __device__ inline float dot4(const float4 a, const float4 b) {
  return fmaf(a.w, b.w, fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)));
}
__host__ __device__ inline float4 normalize4(const float4 v) {
  float s = sqrtf(dot4(v, v));
  return make_float4(v.x / s, v.y / s, v.z / s, v.w / s);
}
__host__ __device__ inline float4 neg4(const float4 v) { return make_float4(-v.x, -v.y, -v.z, -v.w); }
// Image of q by operator k: the identity image stays bit-identical to q (pairs.wgsl:287).
__host__ __device__ inline float4 image(const float4 q, const float4 op, uint32_t k) {
  return k == 0 ? q : normalize4(qmul(q, op));
}
__host__ __device__ inline float chord2(const float4 a, const float4 b) {
  float4 d = make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
  return dot4(d, d);
}
__host__ __device__ inline float chord2p(const float4 a, const float4 b) {
  float4 d = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
  return dot4(d, d);
}
// Order-preserving u32 key of a float (shell.wgsl:17) and its inverse.
__host__ __device__ inline uint32_t ordered_key(float x) {
  uint32_t b = __float_as_uint(x);
  return (b >> 31) ? ~b : (b ^ 0x80000000u);
}
__host__ __device__ inline float unordered_key(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k ^ 0x80000000u) : ~k);
}

// atan(t) / t = P(t^2) for t in [0, 1]: Remez minimax of relative error, degree 8
// (fit_atan.py; bound 1.5e-8 before FP32 evaluation).
__device__ __forceinline__ float atan01(float t) {
  float z = t * t;
  float p = fmaf(z, 0.0028498897f, -0.01606863f);
  p = fmaf(z, p, 0.04269152f);
  p = fmaf(z, p, -0.07504295f);
  p = fmaf(z, p, 0.10640934f);
  p = fmaf(z, p, -0.14203644f);
  p = fmaf(z, p, 0.1999262f);
  p = fmaf(z, p, -0.33333072f);
  p = fmaf(z, p, 1.0f);
  return t * p;
}

struct PairSum {
  float fx, fy, fz, curvature, energy;
};

// One Quey pair of source x and orbit point y (dm2 = |x-y|^2, dp2 = |x+y|^2), added to s.
// WINDOW applies the near field's C1 fade: W(s) = 1 - s^2(3-2s) over the outer half of
// the cutoff and its exact force term E W'(a) / sin a (pairs.wgsl:349-361).  At s = 0 the
// window is 1 and its slope 0, so it needs no branch.
template <bool WINDOW>
__device__ __forceinline__ void pair_terms(const float4 x, const float4 y, float dm2, float dp2, float taper_start,
                                           float inv_width, PairSum& s) {
  const float nr = fmaxf(fminf(dm2, dp2), 1e-30f);
  const float fr = fmaxf(fmaxf(dm2, dp2), 1e-30f);
  const float R = __builtin_amdgcn_rsqf(nr * fr);
  const float t = nr * R;
  const float a = fmaxf(2.0f * atan01(t), 1e-12f);
  const float b = kPiF - a;
  const float aa = a * a, bb = b * b;
  const float a2 = aa + 1e-9f, b2 = bb + 1e-9f;
  const float a3 = fmaf(aa, a, 1e-9f), b3 = fmaf(bb, b, 1e-9f);
  const float sine = fmaf(0.5f * t, fr, 1e-9f);
  const float P1 = a2 * b2 * sine, P2 = a3 * b3, P3 = a * b;
  const float P12 = P1 * P2;
  const float D = __builtin_amdgcn_rcpf(P12 * P3);
  const float iP1 = D * (P2 * P3), iP2 = D * (P1 * P3), iP3 = D * P12;
  const float diff = (dm2 <= dp2) ? (b2 - a2) : (a2 - b2);
  float coef = -diff * iP1;
  float curvature = (a3 + b3) * iP2;
  float energy = kPiF * iP3;
  if (WINDOW) {
    const float u = fminf(fmaxf((a - taper_start) * inv_width, 0.0f), 1.0f);
    const float w = fmaf(u * u, fmaf(2.0f, u, -3.0f), 1.0f);
    const float slope = -6.0f * inv_width * (u * (1.0f - u));
    const float inv_sin = 0.5f * (nr + fr) * R;
    const float e_slope = (dm2 <= dp2 ? energy : -energy) * slope;
    coef = fmaf(w, coef, e_slope * inv_sin);
    curvature *= w;
    energy *= w;
  }
  const float d1 = fmaf(x.x, y.y, -y.x * x.y) + fmaf(x.z, y.w, -x.w * y.z);
  const float d2 = fmaf(x.x, y.z, -y.x * x.z) + fmaf(x.w, y.y, -x.y * y.w);
  const float d3 = fmaf(x.x, y.w, -y.x * x.w) + fmaf(x.y, y.z, -x.z * y.y);
  s.fx = fmaf(coef, d1, s.fx);
  s.fy = fmaf(coef, d2, s.fy);
  s.fz = fmaf(coef, d3, s.fz);
  s.curvature += curvature;
  s.energy += energy;
}

// 2 atan(t) for small t: t times twice the Taylor series 1 - z/3 + z^2/5 - ... in z = t^2
// up to z^TERMS.  Within a near-field cutoff z <= tan(r/2)^2, and the launcher picks the
// fewest terms whose truncation, z^(TERMS+1)/(2 TERMS + 3), stays below 5e-9 there
// (atan_terms); TERMS = 0 selects the minimax polynomial on [0, 1].  The factor 2 rides on
// the coefficients: scaling by 2 commutes with every rounding of the Horner steps and of
// the last product, so the result is bit for bit 2 * (t * p) of the plain series, one
// multiplication fewer per pair.
template <int TERMS>
__device__ __forceinline__ float atan2x_small(float t) {
  if (TERMS == 0) return 2.0f * atan01(t);
  // Horner from the highest kept term: coefficients 2 (-1)^k / (2k + 1), k = TERMS..1.
  const float z = t * t;
  constexpr float c[6] = {2.0f * 1.0f,         2.0f * -0.33333334f, 2.0f * 0.2f,
                          2.0f * -0.14285715f, 2.0f * 0.11111111f,  2.0f * -0.09090909f};
  float p = c[TERMS];
#pragma unroll
  for (int k = TERMS - 1; k >= 0; --k) p = fmaf(z, p, c[k]);
  return t * p;
}

// Per-lane sums of near-field pairs.  The force direction of a pair is linear in the
// site, M(x) y = x0 y_v - y0 x_v + x_v x y_v with M(x) x = 0, so it is also -M(x) (x - y):
// the kernel sums g += coef (x - y) / 2 over the pairs and applies -2 M(x) once per source
// (near_force_of of 2 g), 4 FMAs per pair instead of 15.
struct NearSum {
  float gx, gy, gz, gw, curvature, energy;
};

// The near-field pair for a site y within the cutoff, from e = x - y and dm2 = |e|^2:
// dm2 <= chord2 < 2, so y is the nearer of +-y (sign +) and dp2 = 4 - dm2 for unit
// quaternions; then near + far = 4 and 1/sin a = 2 rsq(near far).  Otherwise the
// arithmetic of pair_terms<true> (pairs.wgsl:246-262, :349-361), with changes of form:
// the window coordinate (a - taper_start) / width is one fused multiply-add,
// a * inv_width + u_offset with u_offset = -taper_start * inv_width (one rounding instead
// of two); the pair energy pi / (a b) is summed without its factor pi, which the caller
// applies once per source in FP64, and rides with the window's slope factor -6 on the
// per-source constant slope_pi = -6 pi inv_width (the slope, times pi); the slope's
// u (1 - u) is u - u^2 from the u^2 the window already has; the regularised sine is taken
// at twice its value, t far + 2e-9 = 2 (sin a + 1e-9), which halves the force coefficient
// exactly (a power of two scales every rounding alike), and the caller doubles the force
// sum once per source; and the reciprocals of P1 = a2 b2 sine and P2 = a3 b3 come from the
// one reciprocal of P1 P2 P3 through 1 / (P1 P2) = P3 / (P1 P2 P3).  Each saves one
// instruction per pair; the rounding differs from the reference's form in the last bit.
// g accumulates coef (x - y) / 2.
template <int TERMS>
__device__ __forceinline__ void near_pair_terms(const float4 e, float dm2, float inv_width, float u_offset,
                                                float slope_pi, NearSum& s) {
  const float nr = __builtin_amdgcn_fmed3f(dm2, 1e-30f, 4.0f);
  const float fr = 4.0f - dm2;
  const float R = __builtin_amdgcn_rsqf(nr * fr);
  const float t = nr * R;
  const float a = fmaxf(atan2x_small<TERMS>(t), 1e-12f);
  const float b = kPiF - a;
  const float aa = a * a, bb = b * b;
  // Within a cutoff of at most 1.5 rad, b >= pi - 1.5, so b^2 >= 2.6 and b^2 + 1e-9 rounds
  // to b^2 in FP32 (half an ulp of 2.6 is 1.2e-7): the regulariser is dropped exactly.
  const float a2 = aa + 1e-9f, b2 = bb;
  const float a3 = fmaf(aa, a, 1e-9f), b3 = fmaf(bb, b, 1e-9f);
  const float sine2 = fmaf(t, fr, 2e-9f);  // 2 (sin a + 1e-9)
  const float P1 = a2 * b2 * sine2, P2 = a3 * b3, P3 = a * b;
  const float P12 = P1 * P2;
  const float D = __builtin_amdgcn_rcpf(P12 * P3);
  const float E12 = D * P3;                                 // 1 / (P1 P2)
  const float iP1 = E12 * P2, iP2 = E12 * P1, iP3 = D * P12;  // energy / pi = iP3
  const float u = fminf(fmaxf(fmaf(a, inv_width, u_offset), 0.0f), 1.0f);
  const float uu = u * u;
  const float w = fmaf(uu, fmaf(2.0f, u, -3.0f), 1.0f);
  const float slope_half = slope_pi * (u - uu);  // pi times the window's slope
  const float coef = fmaf(w, (a2 - b2) * iP1, iP3 * slope_half * R);
  s.gx = fmaf(coef, e.x, s.gx);
  s.gy = fmaf(coef, e.y, s.gy);
  s.gz = fmaf(coef, e.z, s.gz);
  s.gw = fmaf(coef, e.w, s.gw);
  s.curvature = fmaf((a3 + b3) * iP2, w, s.curvature);
  s.energy = fmaf(iP3, w, s.energy);  // times pi once per source
}

// The force of a source x from the summed g = sum coef (x - y): -M(x) g, in FP64.
__host__ __device__ inline void near_force_of(const float4 x, const double g[4], double f[3]) {
  const double x0 = x.x, x1 = x.y, x2 = x.z, x3 = x.w;
  f[0] = x1 * g[0] - x0 * g[1] + x3 * g[2] - x2 * g[3];
  f[1] = x2 * g[0] - x3 * g[1] - x0 * g[2] + x1 * g[3];
  f[2] = x3 * g[0] + x2 * g[1] - x1 * g[2] - x0 * g[3];
}

// Terms of atan2x_small for a cutoff chord^2: the fewest with truncation error below 5e-9
// for z up to chord2 / (4 - chord2) = tan(r/2)^2 (0: the minimax polynomial).
inline int atan_terms(float chord2) {
  const double z = double(chord2) / (4.0 - double(chord2)) * 1.0001;
  for (int k = 2; k <= 5; ++k)
    if (std::pow(z, k + 1) / (2 * k + 3) <= 5e-9) return k;
  return 0;
}

}  // namespace zxt
