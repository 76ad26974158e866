// ORACLE — test infrastructure only.  A CPU restatement (C++17, OpenMP) of the
// reference's QVM relaxation and spacing path, used by tests/, by smoke() and by
// bench.py's cpu_baseline leg as the checker.  Nothing in the product loads it.
//
// What it restates (reference file:line):
//   haar hash and Shoemake draw ......... shaders/haar.wgsl:3-9
//   fold (rescale, normalise, zone) ..... shaders/state.wgsl:1-51
//   quey_pair ........................... shaders/pairs.wgsl:246-262
//   near_pair window .................... shaders/pairs.wgsl:349-361
//   all-pairs rows (quey_partial) ....... shaders/pairs.wgsl:263-326
//   near-field rows (near_force) ........ shaders/pairs.wgsl:363-427, src/near.zig:90-188
//   zone planes, rim and shell .......... src/covering.zig:382-397, shaders/shell.wgsl:19-39
//   BB step (stats, median, alpha, move)  shaders/relax.wgsl:29-165, src/pairs.zig:339-363
//   nearest degrees ..................... shaders/pairs.wgsl:428-434
//
// Precision: per-pair values in FP32 with the reference's formulas (libm atanf,
// sqrtf, sinf), sums in FP64.  The image, chord and fold arithmetic follow an exact
// operation order (explicit fma, correctly rounded sqrt and division, no contraction:
// build with -ffp-contract=off) that the GPU kernels reproduce bit for bit, so pair
// sets and nearest-neighbour indices compare exactly.
//
// Parity pins (tests/test_oracle.py): the operator fixture pair-operators.npz, the
// dense FP64 models restated in oracle/dense_f64.py, finite differences of the
// energy, antipodal invariance, brute-force pair sets, an exact median by sort.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

constexpr float kPi = 3.14159265358979323846f;

struct Q {
  float w, x, y, z;
};

inline Q qmul(const Q& a, const Q& b) {
  Q r;
  r.w = std::fmaf(-a.z, b.z, std::fmaf(-a.y, b.y, std::fmaf(-a.x, b.x, a.w * b.w)));
  r.x = std::fmaf(-a.z, b.y, std::fmaf(a.y, b.z, std::fmaf(b.w, a.x, a.w * b.x)));
  r.y = std::fmaf(-a.x, b.z, std::fmaf(a.z, b.x, std::fmaf(b.w, a.y, a.w * b.y)));
  r.z = std::fmaf(-a.y, b.x, std::fmaf(a.x, b.y, std::fmaf(b.w, a.z, a.w * b.z)));
  return r;
}
inline float dot4(const Q& a, const Q& b) {
  return std::fmaf(a.z, b.z, std::fmaf(a.y, b.y, std::fmaf(a.x, b.x, a.w * b.w)));
}
inline Q normalize(const Q& v) {
  float s = std::sqrt(dot4(v, v));
  return {v.w / s, v.x / s, v.y / s, v.z / s};
}
inline Q neg(const Q& v) { return {-v.w, -v.x, -v.y, -v.z}; }
inline Q image(const Q& q, const Q& op, uint32_t k) { return k == 0 ? q : normalize(qmul(q, op)); }
inline float chord2(const Q& a, const Q& b) {
  Q d{a.w - b.w, a.x - b.x, a.y - b.y, a.z - b.z};
  return dot4(d, d);
}
inline float chord2p(const Q& a, const Q& b) {
  Q d{a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z};
  return dot4(d, d);
}
inline Q load(const float* p, size_t i) { return {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]}; }
inline void store(float* p, size_t i, const Q& v) {
  p[4 * i] = v.w;
  p[4 * i + 1] = v.x;
  p[4 * i + 2] = v.y;
  p[4 * i + 3] = v.z;
}

// ---- haar.wgsl:3-9 -------------------------------------------------------------
inline uint32_t hash32(uint32_t v) {
  uint32_t x = v;
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  return x ^ (x >> 16);
}
inline float random_unit(uint32_t v) { return (float(hash32(v) >> 8) + 0.5f) / 16777216.0f; }

// ---- state.wgsl:7-30 ----------------------------------------------------------
bool in_zone(const Q& p, uint32_t group, const float* ops, uint32_t order) {
  float ax = std::fabs(p.x), ay = std::fabs(p.y), az = std::fabs(p.z), w = p.w;
  switch (group) {
    case 0: return az <= w;
    case 1: return az <= 0.5773502691896258f * w;
    case 2: return az <= 0.4142135623730951f * w;
    case 3: return az <= 0.2679491924311227f * w;
    case 4: return std::max(std::max(ax, ay), az) <= w;
    case 5: return az <= 0.5773502691896258f * w && std::max(ax, 0.5f * ax + 0.8660254037844386f * ay) <= w;
    case 6: return az <= 0.4142135623730951f * w && std::max(std::max(ax, ay), 0.7071067811865476f * (ax + ay)) <= w;
    case 7: {
      float m = std::max(ax, ay), n = std::min(ax, ay);
      return az <= 0.2679491924311227f * w && std::max(m, 0.8660254037844386f * m + 0.5f * n) <= w;
    }
    case 8: return ax + ay + az <= w;
    case 9: return std::max(std::max(ax, ay), az) <= 0.4142135623730951f * w && ax + ay + az <= w;
    case 10: {
      if (w > 0.9510566f) return true;
      if (w < 0.9256147f) return false;
      for (uint32_t k = 1; k < order; ++k) {
        Q op = load(ops, k);
        if (std::fabs(std::fabs(op.w) - 0.8090169943749475f) < 1e-5f && std::fabs(dot4(p, op)) > w) return false;
      }
      return true;
    }
    default: return true;  // C1: the whole hemisphere w >= 0
  }
}

Q fold_one(Q a, uint32_t group, const float* ops, uint32_t order) {
  float comps[4] = {a.w, a.x, a.y, a.z};
  float scale = std::max(std::max(std::fabs(a.w), std::fabs(a.x)), std::max(std::fabs(a.y), std::fabs(a.z)));
  uint32_t sb;
  std::memcpy(&sb, &scale, 4);
  int maximum = int((sb >> 23) & 255u);
  float out[4];
  for (int c = 0; c < 4; ++c) {
    uint32_t bits;
    std::memcpy(&bits, &comps[c], 4);
    int exponent = int((bits >> 23) & 255u);
    int scaled = exponent + 127 - maximum;
    uint32_t shifted = (bits & 0x807fffffu) | (uint32_t(std::max(scaled, 0)) << 23);
    bool keep = exponent > 0 && scaled > 0;
    uint32_t r = keep ? shifted : (bits & 0x80000000u);
    std::memcpy(&out[c], &r, 4);
  }
  Q p = normalize({out[0], out[1], out[2], out[3]});
  if (p.w < 0.0f) p = neg(p);
  if (in_zone(p, group, ops, order)) return p;
  Q best = p;
  float w = p.w;
  for (uint32_t k = 0; k < order; ++k) {
    Q c = qmul(p, load(ops, k));
    if (std::fabs(c.w) > w) {
      best = c.w >= 0.0f ? c : neg(c);
      w = std::fabs(c.w);
    }
  }
  return normalize(best);
}

// ---- pairs.wgsl:246-262 and :349-361 --------------------------------------------
struct Pair {
  float f[3];
  float curvature, energy, angle;
};

Pair quey_pair(const Q& x, const Q& y, float dm2, float dp2) {
  float nearv = std::max(std::min(dm2, dp2), 1e-30f);
  float farv = std::max(std::max(dm2, dp2), 1e-30f);
  float t = std::sqrt(nearv / farv);
  float a = std::max(2.0f * std::atan(t), 1e-12f);
  float b = kPi - a;
  float a2 = a * a + 1e-9f, b2 = b * b + 1e-9f, a3 = a * a * a + 1e-9f, b3 = b * b * b + 1e-9f;
  float sine = 2.0f * t / (1.0f + t * t) + 1e-9f;
  float sign = dm2 <= dp2 ? 1.0f : -1.0f;
  float d1 = x.w * y.x - y.w * x.x + (x.y * y.z - x.z * y.y);
  float d2 = x.w * y.y - y.w * x.y + (x.z * y.x - x.x * y.z);
  float d3 = x.w * y.z - y.w * x.z + (x.x * y.y - x.y * y.x);
  float along = sign * (b2 - a2) / (a2 * b2 * sine);
  Pair p;
  p.f[0] = -along * d1;
  p.f[1] = -along * d2;
  p.f[2] = -along * d3;
  p.curvature = (a3 + b3) / (a3 * b3);
  p.energy = kPi / (a * b);
  p.angle = a;
  return p;
}

struct Cut {
  float chord2, chord, taper_start, inv_width;
};

Pair near_pair(const Q& x, const Q& y, float dm2, float dp2, const Cut& c) {
  Pair p = quey_pair(x, y, dm2, dp2);
  float s = std::min(std::max((p.angle - c.taper_start) * c.inv_width, 0.0f), 1.0f);
  if (s <= 0.0f) return p;
  float window = 1.0f - s * s * (3.0f - 2.0f * s);
  float slope = -6.0f * s * (1.0f - s) * c.inv_width;
  float sign = dm2 <= dp2 ? 1.0f : -1.0f;
  float d1 = x.w * y.x - y.w * x.x + (x.y * y.z - x.z * y.y);
  float d2 = x.w * y.y - y.w * x.y + (x.z * y.x - x.x * y.z);
  float d3 = x.w * y.z - y.w * x.z + (x.x * y.y - x.y * y.x);
  float extra = sign * p.energy * slope / std::max(std::sin(p.angle), 1e-9f);
  p.f[0] = window * p.f[0] + extra * d1;
  p.f[1] = window * p.f[1] + extra * d2;
  p.f[2] = window * p.f[2] + extra * d3;
  p.curvature *= window;
  p.energy *= window;
  return p;
}

// Row: fx, fy, fz, curvature, energy, nearest chord^2 (FP64 sums of FP32 pair terms).
inline void add_pair(double* row, const Pair& p) {
  row[0] += p.f[0];
  row[1] += p.f[1];
  row[2] += p.f[2];
  row[3] += p.curvature;
  row[4] += p.energy;
}

// ---- covering.zig:382-397 --------------------------------------------------------
std::vector<Q> zone_planes(const float* ops, uint32_t order) {
  std::vector<Q> planes;
  for (uint32_t i = 0; i < order; ++i) {
    for (double s : {1.0, -1.0}) {
      if (i == 0 && s == 1.0) continue;
      double n[4] = {1.0 - s * ops[4 * i], -s * ops[4 * i + 1], -s * ops[4 * i + 2], -s * ops[4 * i + 3]};
      double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + n[3] * n[3]);
      planes.push_back({float(n[0] / len), float(n[1] / len), float(n[2] / len), float(n[3] / len)});
    }
  }
  return planes;
}

}  // namespace

extern "C" {

int orc_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

void orc_haar(uint32_t n, uint32_t seed, float* q) {
  uint32_t hs = hash32(seed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) {
    uint32_t base = hash32(uint32_t(i) ^ hs);
    float u = random_unit(base);
    float a = 6.28318530718f * random_unit(base + 1u);
    float b = 6.28318530718f * random_unit(base + 2u);
    float su = std::sqrt(u), sv = std::sqrt(1.0f - u);
    store(q, size_t(i), {su * std::cos(b), sv * std::sin(a), sv * std::cos(a), su * std::sin(b)});
  }
}

void orc_fold(float* q, uint32_t n, const float* ops, uint32_t order, uint32_t group) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < int64_t(n); ++i) store(q, size_t(i), fold_one(load(q, size_t(i)), group, ops, order));
}

// Image of q[j] by operator k (identity image unrounded), as every kernel builds it.
void orc_images(const float* q, uint32_t n, const float* ops, uint32_t order, float* out) {
  for (uint32_t k = 0; k < order; ++k)
    for (uint32_t j = 0; j < n; ++j) store(out, size_t(k) * n + j, image(load(q, j), load(ops, k), k));
}

// All-pairs rows (quey_partial + quey_reduce): own orbit excluded from force and energy;
// nearest over every orbit point but the source itself.
void orc_allpairs(const float* q, uint32_t n, const float* ops, uint32_t order, double* rows) {
  std::vector<Q> img(size_t(n) * order);
  for (uint32_t k = 0; k < order; ++k)
    for (uint32_t j = 0; j < n; ++j) img[size_t(k) * n + j] = image(load(q, j), load(ops, k), k);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t ii = 0; ii < int64_t(n); ++ii) {
    uint32_t i = uint32_t(ii);
    Q x = load(q, i);
    double row[6] = {0, 0, 0, 0, 0, 4.0};
    float nearest = 4.0f;
    for (uint32_t k = 0; k < order; ++k) {
      for (uint32_t j = 0; j < n; ++j) {
        const Q& y = img[size_t(k) * n + j];
        float dm2 = chord2(x, y), dp2 = chord2p(x, y);
        if (!(j == i && k == 0)) nearest = std::min(nearest, std::min(dm2, dp2));
        if (j != i) add_pair(row, quey_pair(x, y, dm2, dp2));
      }
    }
    row[5] = nearest;
    std::memcpy(rows + 6 * i, row, sizeof(row));
  }
}

// Near-field rows by brute force over the whole signed orbit: every orbit point y with
// chord^2(x, y) <= cut counts once; the source's identity image is skipped and its other
// images bound only the nearest chord (pairs.wgsl:404-408).
void orc_near_brute(const float* q, uint32_t n, const float* ops, uint32_t order, const float* cut4, double* rows) {
  Cut c{cut4[0], cut4[1], cut4[2], cut4[3]};
  std::vector<Q> img(size_t(n) * order);
  for (uint32_t k = 0; k < order; ++k)
    for (uint32_t j = 0; j < n; ++j) img[size_t(k) * n + j] = image(load(q, j), load(ops, k), k);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t ii = 0; ii < int64_t(n); ++ii) {
    uint32_t i = uint32_t(ii);
    Q x = load(q, i);
    double row[6] = {0, 0, 0, 0, 0, 4.0};
    float nearest = 4.0f;
    for (uint32_t k = 0; k < order; ++k) {
      for (uint32_t j = 0; j < n; ++j) {
        for (int s = 0; s < 2; ++s) {
          Q y = s ? neg(img[size_t(k) * n + j]) : img[size_t(k) * n + j];
          float dm2 = chord2(x, y);
          if (!(dm2 <= c.chord2)) continue;
          if (j == i) {
            if (!(k == 0 && s == 0)) nearest = std::min(nearest, dm2);
            continue;
          }
          nearest = std::min(nearest, dm2);
          add_pair(row, near_pair(x, y, dm2, chord2p(x, y), c));
        }
      }
    }
    row[5] = nearest;
    std::memcpy(rows + 6 * i, row, sizeof(row));
  }
}

}  // extern "C"

namespace {

// The near field's data structures (src/near.zig:90-188, shaders/shell.wgsl:19-39): zone
// planes, rim sources, shell images within sin(r+1e-4) of the zone, a uniform grid of the
// stereographic chart p = q.yzw/(1+q.x), and per source a scan of the cells its cutoff
// cap (a ball in the chart) can reach.
struct NearGrid {
  uint32_t n = 0, order = 0;
  int sdim = 1;
  double cpu = 1, c0 = 1;
  float lo[3] = {0, 0, 0};
  std::vector<Q> site;          // sorted by cell, in scatter order within a cell
  std::vector<uint32_t> code;   // source | 0x80000000 for the identity image
  std::vector<uint32_t> rank;   // source*2*order + 2*operator + sign
  std::vector<uint32_t> off;    // sdim^3 + 1
  int cell1(double v, int axis) const {
    int cidx = int(std::floor((v - lo[axis]) * cpu));
    return std::min(std::max(cidx, 0), sdim - 1);
  }
  // Calls f(k) for every site k in the cell rows the cap of x crosses.  Returns the count.
  template <typename F>
  uint64_t visit(const Q& x, F&& f) const {
    double denom = c0 + x.w;
    double ctr[3] = {0, 0, 0}, rho = 0;
    bool all = denom < 1e-3;
    if (!all) {
      ctr[0] = x.x / denom;
      ctr[1] = x.y / denom;
      ctr[2] = x.z / denom;
      rho = std::sqrt(std::max(0.0, 1.0 - c0 * c0)) / denom * 1.001 + 1e-6;
    }
    uint64_t seen = 0;
    int zlo = all ? 0 : cell1(ctr[2] - rho, 2), zhi = all ? sdim - 1 : cell1(ctr[2] + rho, 2);
    int ylo = all ? 0 : cell1(ctr[1] - rho, 1), yhi = all ? sdim - 1 : cell1(ctr[1] + rho, 1);
    for (int z = zlo; z <= zhi; ++z) {
      for (int y = ylo; y <= yhi; ++y) {
        int x0 = 0, x1 = sdim - 1;
        if (!all) {
          double cy0 = (y == 0) ? -1e30 : lo[1] + y / cpu, cy1 = (y == sdim - 1) ? 1e30 : lo[1] + (y + 1) / cpu;
          double cz0 = (z == 0) ? -1e30 : lo[2] + z / cpu, cz1 = (z == sdim - 1) ? 1e30 : lo[2] + (z + 1) / cpu;
          double gy = std::max(0.0, std::max(cy0 - ctr[1], ctr[1] - cy1));
          double gz = std::max(0.0, std::max(cz0 - ctr[2], ctr[2] - cz1));
          double rest = rho * rho - gy * gy - gz * gz;
          if (rest < 0) continue;
          double span = std::sqrt(rest);
          x0 = cell1(ctr[0] - span, 0);
          x1 = cell1(ctr[0] + span, 0);
        }
        size_t rowbase = size_t(sdim) * (y + size_t(sdim) * z);
        for (uint32_t k = off[rowbase + x0]; k < off[rowbase + x1 + 1]; ++k) {
          ++seen;
          f(k);
        }
      }
    }
    return seen;
  }
};

NearGrid build_grid(const float* q, uint32_t n, const float* ops, uint32_t order, float chord2, double radius) {
  NearGrid g;
  g.n = n;
  g.order = order;
  std::vector<Q> planes = zone_planes(ops, order);
  float band = float(std::sin(radius + 1e-4));
  // Sites: core (identity images) then shell images of rim sources.
  std::vector<Q> site;
  std::vector<uint32_t> code, rank;
  site.reserve(size_t(n) * 2);
  code.reserve(size_t(n) * 2);
  rank.reserve(size_t(n) * 2);
  for (uint32_t i = 0; i < n; ++i) {
    site.push_back(load(q, i));
    code.push_back(i | 0x80000000u);
    rank.push_back(i * 2 * order);
  }
  int nt = omp_get_max_threads();
  std::vector<std::vector<Q>> tsite(nt);
  std::vector<std::vector<uint32_t>> tcode(nt), trank(nt);
#pragma omp parallel
  {
    int t = omp_get_thread_num();
#pragma omp for schedule(static)
    for (int64_t ii = 0; ii < int64_t(n); ++ii) {
      Q x = load(q, size_t(ii));
      bool rim = false;
      for (const Q& pl : planes)
        if (dot4(x, pl) <= band + 2e-6f) {
          rim = true;
          break;
        }
      if (!rim) continue;
      for (uint32_t k = 0; k < order; ++k) {
        Q base = image(x, load(ops, k), k);
        for (uint32_t s = 0; s < 2; ++s) {
          if (k == 0 && s == 0) continue;
          Q y = s ? neg(base) : base;
          if (y.w <= -0.9999999f) continue;
          bool inside = true;
          for (const Q& pl : planes)
            if (dot4(y, pl) < -band - 2e-6f) {
              inside = false;
              break;
            }
          if (!inside) continue;
          tsite[t].push_back(y);
          tcode[t].push_back(uint32_t(ii));
          trank[t].push_back(uint32_t(ii) * 2 * order + 2 * k + s);
        }
      }
    }
  }
  for (int t = 0; t < nt; ++t) {
    site.insert(site.end(), tsite[t].begin(), tsite[t].end());
    code.insert(code.end(), tcode[t].begin(), tcode[t].end());
    rank.insert(rank.end(), trank[t].begin(), trank[t].end());
  }
  size_t S = site.size();
  std::vector<float> px(S), py(S), pz(S);
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  for (size_t s = 0; s < S; ++s) {
    float d = 1.0f + site[s].w;
    px[s] = site[s].x / d;
    py[s] = site[s].y / d;
    pz[s] = site[s].z / d;
    lo[0] = std::min(lo[0], px[s]);
    hi[0] = std::max(hi[0], px[s]);
    lo[1] = std::min(lo[1], py[s]);
    hi[1] = std::max(hi[1], py[s]);
    lo[2] = std::min(lo[2], pz[s]);
    hi[2] = std::max(hi[2], pz[s]);
  }
  float width = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6f}) * 1.0001f;
  g.sdim = int(std::min(256.0, std::max(1.0, std::round(std::cbrt(double(S) / 2.0)))));
  g.cpu = g.sdim / double(width);
  for (int a = 0; a < 3; ++a) g.lo[a] = lo[a];
  size_t cells = size_t(g.sdim) * g.sdim * g.sdim;
  g.off.assign(cells + 1, 0);
  std::vector<uint32_t> scell(S), perm(S);
  for (size_t s = 0; s < S; ++s) {
    scell[s] = uint32_t(g.cell1(px[s], 0) + g.sdim * (g.cell1(py[s], 1) + size_t(g.sdim) * g.cell1(pz[s], 2)));
    g.off[scell[s] + 1]++;
  }
  for (size_t k = 0; k < cells; ++k) g.off[k + 1] += g.off[k];
  {
    std::vector<uint32_t> cur(g.off.begin(), g.off.end() - 1);
    for (size_t s = 0; s < S; ++s) perm[cur[scell[s]]++] = uint32_t(s);
  }
  g.site.resize(S);
  g.code.resize(S);
  g.rank.resize(S);
  for (size_t k = 0; k < S; ++k) {
    g.site[k] = site[perm[k]];
    g.code[k] = code[perm[k]];
    g.rank[k] = rank[perm[k]];
  }
  // The cap {y : x.y >= c0} is a ball in the chart: centre v/(c0+w), radius sqrt(1-c0^2)/(c0+w).
  g.c0 = 1.0 - double(chord2) * (1.0 + 1e-4) / 2.0 - 1e-6;
  return g;
}

// The reference's near-field pair (pairs.wgsl:246-262, :349-361) in FP64 with exact
// atan and sin: the error model the FP32 kernels are measured against.
struct D4 {
  double w, x, y, z;
};
inline D4 d_normalize(D4 v) {
  double s = std::sqrt(v.w * v.w + v.x * v.x + v.y * v.y + v.z * v.z);
  return {v.w / s, v.x / s, v.y / s, v.z / s};
}
inline D4 d_mul(const D4& a, const D4& b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
void near_pair64(const D4& x, const D4& y, double radius, double* row) {
  const double kPi64 = 3.14159265358979323846;
  double dm2 = 0, dp2 = 0;
  for (int c = 0; c < 4; ++c) {
    double xc = (&x.w)[c], yc = (&y.w)[c];
    dm2 += (xc - yc) * (xc - yc);
    dp2 += (xc + yc) * (xc + yc);
  }
  double nearv = std::min(dm2, dp2), farv = std::max(dm2, dp2);
  double a = 2.0 * std::atan(std::sqrt(nearv / farv));
  double b = kPi64 - a;
  double taper = 0.5 * radius, width = 0.5 * radius;
  double s = std::min(std::max((a - taper) / width, 0.0), 1.0);
  double window = 1.0 - s * s * (3.0 - 2.0 * s);
  double slope = -6.0 * s * (1.0 - s) / width;
  double energy = kPi64 / (a * b);
  double sign = dm2 <= dp2 ? 1.0 : -1.0;
  double d[3] = {x.w * y.x - y.w * x.x + (x.y * y.z - x.z * y.y), x.w * y.y - y.w * x.y + (x.z * y.x - x.x * y.z),
                 x.w * y.z - y.w * x.z + (x.x * y.y - x.y * y.x)};
  double along = sign * (1.0 / (a * a + 1e-9) - 1.0 / (b * b + 1e-9)) / (std::sin(a) + 1e-9);
  double extra = sign * energy * slope / std::sin(a);
  double f2 = 0;
  for (int c = 0; c < 3; ++c) {
    double f = -along * d[c] * window + extra * d[c];
    row[c] += f;
    f2 += f * f;
  }
  row[3] += window * (1.0 / (a * a * a + 1e-9) + 1.0 / (b * b * b + 1e-9));
  row[4] += energy * window;
  row[6] += std::sqrt(f2);
}

}  // namespace

extern "C" {

// Near-field rows through the reference's data structures (NearGrid above).  Rows are
// written for sources [first, first+count).  stats receives {sites, shell, sdim,
// candidates, accepted}.  This is the CPU baseline's timed path.
int orc_near_grid(const float* q, uint32_t n, const float* ops, uint32_t order, const float* cut4, double radius,
                  uint32_t first, uint32_t count, double* rows, double* stats) {
  Cut c{cut4[0], cut4[1], cut4[2], cut4[3]};
  NearGrid g = build_grid(q, n, ops, order, c.chord2, radius);
  double cand_total = 0, acc_total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : cand_total, acc_total)
  for (int64_t ii = 0; ii < int64_t(count); ++ii) {
    uint32_t i = first + uint32_t(ii);
    Q x = load(q, i);
    double row[6] = {0, 0, 0, 0, 0, 4.0};
    float nearest = 4.0f;
    double accepted = 0;
    cand_total += double(g.visit(x, [&](uint32_t k) {
      const Q& s = g.site[k];
      float dm2 = chord2(x, s);
      if (!(dm2 <= c.chord2)) return;
      uint32_t cd = g.code[k];
      if ((cd & 0x7fffffffu) == i) {
        if (cd == i) nearest = std::min(nearest, dm2);
        return;
      }
      nearest = std::min(nearest, dm2);
      accepted += 1;
      add_pair(row, near_pair(x, s, dm2, chord2p(x, s), c));
    }));
    acc_total += accepted;
    row[5] = nearest;
    std::memcpy(rows + 6 * size_t(ii), row, sizeof(row));
  }
  if (stats) {
    stats[0] = double(g.site.size());
    stats[1] = double(g.site.size() - n);
    stats[2] = g.sdim;
    stats[3] = cand_total;
    stats[4] = acc_total;
  }
  return 0;
}

// The same pair set as orc_near_grid (the FP32 cutoff test) with every pair term in FP64
// (near_pair64), on the kernels' geometry in exact arithmetic: the source and identity
// images are the given FP32 positions, other images normalize(q_j g_k) in FP64 with the
// FP64 operators ops64 (the reference's f64 literals).  Rows of 7: orc_near_grid's
// columns with column 5 the FP64 nearest chord^2 over the accepted pairs, and column 6
// the sum of the pair forces' magnitudes (the scale that cancels in a relaxed row).
int orc_near_grid64(const float* q, uint32_t n, const float* ops, const double* ops64, uint32_t order,
                    const float* cut4, double radius, uint32_t first, uint32_t count, double* rows) {
  NearGrid g = build_grid(q, n, ops, order, cut4[0], radius);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t ii = 0; ii < int64_t(count); ++ii) {
    uint32_t i = first + uint32_t(ii);
    Q x = load(q, i);
    const D4 x64{x.w, x.x, x.y, x.z};
    double row[7] = {0, 0, 0, 0, 0, 4.0, 0};
    g.visit(x, [&](uint32_t k) {
      const Q& s = g.site[k];
      float dm2 = chord2(x, s);
      if (!(dm2 <= cut4[0])) return;
      if ((g.code[k] & 0x7fffffffu) == i) return;
      uint32_t r = g.rank[k], j = r / (2 * order), op = (r % (2 * order)) / 2, sg = r % 2;
      Q qj = load(q, j);
      D4 y{qj.w, qj.x, qj.y, qj.z};
      if (op != 0) y = d_normalize(d_mul(y, {ops64[4 * op], ops64[4 * op + 1], ops64[4 * op + 2], ops64[4 * op + 3]}));
      if (sg) y = {-y.w, -y.x, -y.y, -y.z};
      double c2 = 0;
      for (int c = 0; c < 4; ++c) c2 += ((&x64.w)[c] - (&y.w)[c]) * ((&x64.w)[c] - (&y.w)[c]);
      row[5] = std::min(row[5], c2);
      near_pair64(x64, y, radius, row);
    });
    std::memcpy(rows + 7 * size_t(ii), row, sizeof(row));
  }
  return 0;
}

// Nearest orbit point within the cutoff for sources [first, first+count) through the near
// grid: chord^2 in FP32, the source's own identity image excluded, ties to the lowest rank
// source*2*order + 2*operator + sign (the spacing pass, pairs.zig:318-335); none within
// the cutoff gives rank 0xffffffff and chord 4.
void orc_nearest_grid(const float* q, uint32_t n, const float* ops, uint32_t order, float cut, double radius,
                      uint32_t first, uint32_t count, uint32_t* rank_out, float* chord_out) {
  NearGrid g = build_grid(q, n, ops, order, cut, radius);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t ii = 0; ii < int64_t(count); ++ii) {
    uint32_t i = first + uint32_t(ii);
    Q x = load(q, i);
    float best = 4.0f;
    uint32_t brank = 0xffffffffu;
    const uint32_t self = i * 2 * order;
    g.visit(x, [&](uint32_t k) {
      float dm2 = chord2(x, g.site[k]);
      uint32_t r = g.rank[k];
      if (!(dm2 <= cut) || r == self) return;
      if (dm2 < best || (dm2 == best && r < brank)) {
        best = dm2;
        brank = r;
      }
    });
    rank_out[ii] = brank;
    chord_out[ii] = best;
  }
}

// Nearest orbit point of each source by brute force: chord^2 in FP32 over every
// (source j, operator k, sign) but the source's own identity image, the minimum with
// ties broken by the lowest rank j*2*order + 2*k + sign.  With cut > 0 only points with
// chord^2 <= cut count (the near-field candidate set); none gives rank 0xffffffff, 4.0.
void orc_nearest(const float* q, uint32_t n, const float* ops, uint32_t order, float cut, uint32_t* rank_out,
                 float* chord_out) {
  std::vector<Q> img(size_t(n) * order);
  for (uint32_t k = 0; k < order; ++k)
    for (uint32_t j = 0; j < n; ++j) img[size_t(k) * n + j] = image(load(q, j), load(ops, k), k);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t ii = 0; ii < int64_t(n); ++ii) {
    uint32_t i = uint32_t(ii);
    Q x = load(q, i);
    float best = 4.0f;
    uint32_t brank = 0xffffffffu;
    for (uint32_t j = 0; j < n; ++j) {
      for (uint32_t k = 0; k < order; ++k) {
        const Q& y = img[size_t(k) * n + j];
        float d[2] = {chord2(x, y), chord2p(x, y)};
        for (uint32_t s = 0; s < 2; ++s) {
          if (j == i && k == 0 && s == 0) continue;
          if (cut > 0 && !(d[s] <= cut)) continue;
          uint32_t r = j * 2 * order + 2 * k + s;
          if (d[s] < best || (d[s] == best && r < brank)) {
            best = d[s];
            brank = r;
          }
        }
      }
    }
    rank_out[i] = brank;
    chord_out[i] = best;
  }
}

// SO(3) degrees from a chord^2, in FP32 as nearest_degrees (pairs.wgsl:428-434).
void orc_degrees(const float* chord, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i)
    out[i] = 4.0f * std::asin(std::min(1.0f, std::sqrt(chord[i]) * 0.5f)) * (180.0f / kPi);
}

// Exact median of max(c, 0) over both middle ranks, clamped below at 1e-12
// (relax.wgsl:79-134 selects it by radix passes; here by sorting).
float orc_median(const float* curvature, uint32_t n) {
  std::vector<float> v(n);
  for (uint32_t i = 0; i < n; ++i) v[i] = curvature[i] > 0.0f ? curvature[i] : 0.0f;
  std::sort(v.begin(), v.end());
  return std::max(1e-12f, 0.5f * (v[(n - 1) / 2] + v[n / 2]));
}

// One BB step (relax.wgsl relax_stats..relax_move) on rows of force (fx,fy,fz,curvature),
// energy and nearest.  Sums of s.y, y.y, s.s, f.f and energy in FP64.  prev holds per
// source the previous force (xyz) at [6i..6i+2] and move at [6i+3..6i+5].  state:
// [0] iteration, [1] flags (1 non-finite force, 2 duplicate).  out: sy, yy, ss, ff,
// energy, median, alpha, largest move.  fixed is index+1 of a held identity, or 0.
void orc_bb_step(float* q, uint32_t n, const float* force4, const float* nearest, const double* energy, float* prev,
                 uint32_t* state, double radius, uint32_t fixed, float step_scale, double* out) {
  bool first = state[0] == 0;
  double sy = 0, yy = 0, ss = 0, ff = 0, e = 0;
  uint32_t flags = state[1];
  for (uint32_t i = 0; i < n; ++i) {
    double f[3] = {force4[4 * i], force4[4 * i + 1], force4[4 * i + 2]};
    bool finite = true;
    for (double v : f) finite = finite && std::isfinite(v) && std::fabs(v) <= 3.4e38;
    if (!finite) flags |= 1u;
    if (!(nearest[i] > 0.0f)) flags |= 2u;
    ff += f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    e += energy[i];
    if (!first && i + 1 != fixed) {
      double s[3] = {prev[6 * i + 3], prev[6 * i + 4], prev[6 * i + 5]};
      double y[3] = {f[0] - prev[6 * i], f[1] - prev[6 * i + 1], f[2] - prev[6 * i + 2]};
      sy += s[0] * y[0] + s[1] * y[1] + s[2] * y[2];
      yy += y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
      ss += s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    }
  }
  std::vector<float> curv(n);
  for (uint32_t i = 0; i < n; ++i) curv[i] = force4[4 * i + 3];
  float median = orc_median(curv.data(), n);
  double alpha0 = (0.8 / 3.4) * radius * radius * radius;
  double alpha = alpha0;
  if (!first) {
    double a1 = std::fabs(yy) > 1e-20 ? std::fabs(sy / yy) : alpha0;
    double a2 = std::fabs(sy) > 1e-20 ? std::fabs(ss / sy) : alpha0;
    alpha = std::min(std::max(state[0] % 2 == 1 ? a1 : a2, 1e-12), std::max(1.0, 10.0 * radius));
  }
  state[0] += 1;
  state[1] = flags;
  float alpha32 = float(alpha), r32 = float(radius);
  float step_max = 0.0f;
  if (flags == 0) {
    for (uint32_t i = 0; i < n; ++i) {
      if (i + 1 == fixed) {
        for (int c = 0; c < 6; ++c) prev[6 * i + c] = 0.0f;
        store(q, i, {1.0f, 0.0f, 0.0f, 0.0f});
        continue;
      }
      float f[3] = {force4[4 * i], force4[4 * i + 1], force4[4 * i + 2]}, cw = force4[4 * i + 3];
      float scale = alpha32 * step_scale / std::max(1e-6f, cw / median + 1e-2f);
      float size = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) * scale;
      float clip = std::min(1.0f, 0.5f * r32 / std::max(size, 1e-12f));
      float d[3] = {f[0] * (scale * clip), f[1] * (scale * clip), f[2] * (scale * clip)};
      for (int c = 0; c < 3; ++c) {
        prev[6 * i + c] = f[c];
        prev[6 * i + 3 + c] = d[c];
      }
      step_max = std::max(step_max, size * clip);
      float angle = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      float sinc = angle > 1e-4f ? std::sin(angle) / angle : 1.0f - angle * angle / 6.0f;
      Q rot{std::cos(angle), d[0] * sinc, d[1] * sinc, d[2] * sinc};
      Q moved = normalize(qmul(rot, load(q, i)));
      if (moved.w < 0.0f) moved = neg(moved);
      store(q, i, moved);
    }
  }
  out[0] = sy;
  out[1] = yy;
  out[2] = ss;
  out[3] = ff;
  out[4] = e;
  out[5] = median;
  out[6] = alpha;
  out[7] = step_max;
}

// meanRadius (pairs.zig:205-215).
double orc_mean_radius(uint64_t count) {
  const double pi = 3.14159265358979323846;
  double target = pi * pi / double(count), lo = 0.0, hi = pi;
  for (int k = 0; k < 60; ++k) {
    double m = 0.5 * (lo + hi);
    if (pi * (m - std::sin(m)) < target)
      lo = m;
    else
      hi = m;
  }
  return 0.25 * (lo + hi);
}

}  // extern "C"
