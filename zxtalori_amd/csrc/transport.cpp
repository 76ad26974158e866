// FCC transport grids (src/grids.zig:10-77 in the reference): lattice rows and the map
// tables, on the host in FP64.
#include "transport.h"

#include <array>
#include <cmath>

#include "engine.h"
#include "hopf.h"

namespace zxt::transport {
namespace {

using fcc::kPi;

[[noreturn]] void fail(const char* what) { throw EngineError{what}; }

// Gauss-Legendre nodes and weights on [-1, 1]: Newton on P_n from the Chebyshev guesses.
void gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[n - 1 - i] = z;  // ascending
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// The Chebyshev series of f on [lo, hi] from its values at the 64 Chebyshev points.
template <typename F>
fcc::Cheb fit(double lo, double hi, F f) {
  constexpr int n = fcc::kChebTerms;
  fcc::Cheb s{lo, hi, {}};
  double v[n];
  for (int m = 0; m < n; ++m) {
    const double t = std::cos(kPi * (m + 0.5) / n);
    v[m] = f(0.5 * (hi + lo) + 0.5 * (hi - lo) * t);
  }
  for (int j = 0; j < n; ++j) {
    double sum = 0.0;
    for (int m = 0; m < n; ++m) sum += v[m] * std::cos(kPi * j * (m + 0.5) / n);
    s.c[j] = (j == 0 ? 1.0 : 2.0) * sum / n;
  }
  return s;
}

struct Rule {
  double x[64], w[64];
  Rule() { gauss_legendre(64, x, w); }
  template <typename F>
  double operator()(double a, double b, F f) const {
    return fcc::gauss(x, w, 64, a, b, f);
  }
};

// Fits the normalised CDF and density of `pdf` over [lo, hi], split at `knee` (lo < knee
// < hi) or unsplit (knee <= lo); returns the total.
template <typename P>
double fit_distribution(const Rule& g, double lo, double knee, double hi, P pdf, fcc::Cheb* cdf_out,
                        fcc::Cheb* pdf_out) {
  auto cdf = [&](double phi) {
    if (knee > lo && phi > knee) return g(lo, knee, pdf) + g(knee, phi, pdf);
    return g(lo, phi, pdf);
  };
  const double total = cdf(hi);
  if (knee > lo) {
    cdf_out[0] = fit(lo, knee, [&](double p) { return cdf(p) / total; });
    pdf_out[0] = fit(lo, knee, [&](double p) { return pdf(p) / total; });
    cdf_out[1] = fit(knee, hi, [&](double p) { return cdf(p) / total; });
    pdf_out[1] = fit(knee, hi, [&](double p) { return pdf(p) / total; });
  } else {
    cdf_out[0] = fit(lo, hi, [&](double p) { return cdf(p) / total; });
    pdf_out[0] = fit(lo, hi, [&](double p) { return pdf(p) / total; });
  }
  return total;
}

fcc::Tables build() {
  fcc::Tables tb{};
  gauss_legendre(fcc::kGauss, tb.gx, tb.gw);
  const Rule g;
  // T: zone volume per unit azimuth, int_0^polar rho^3 sin.
  fit_distribution(g, kPi / 4, 0.0, kPi / 2,
                   [&](double ph) {
                     return g(0.0, fcc::chamber_polar(ph), [&](double t) { return fcc::t_r3(t, ph) * std::sin(t); });
                   },
                   &tb.t_cdf, &tb.t_pdf);
  // O: the integrand has a kink at o_split inside the chamber for phi < 3pi/8.
  auto o_pdf = [&](double ph) {
    const double high = fcc::chamber_polar(ph), split = fcc::o_split(ph);
    auto f = [&](double t) { return fcc::o_r3(t, ph) * std::sin(t); };
    if (split < high) return g(0.0, split, f) + g(split, high, f);
    return g(0.0, high, f);
  };
  const double o_total = fit_distribution(g, kPi / 4, 3 * kPi / 8, kPi / 2, o_pdf, tb.o_cdf, tb.o_pdf);
  tb.o_knee = g(kPi / 4, 3 * kPi / 8, o_pdf) / o_total;
  // I: one wall bounds the chamber's polar angle.
  fit_distribution(g, 0.0, 0.0, kPi / 5,
                   [&](double ph) {
                     return g(0.0, fcc::i_polar(ph), [&](double t) { return fcc::i_r3(t) * std::sin(t); });
                   },
                   &tb.i_cdf, &tb.i_pdf);
  // I x {1, -1}: rotations in the operator tables' frame, then their negatives.
  const std::vector<std::array<double, 4>> rot = hopf::elements(10);
  const double seed[3] = {std::sin(0.15) * std::cos(kPi / 10), std::sin(0.15) * std::sin(kPi / 10), std::cos(0.15)};
  for (size_t k = 0; k < 120; ++k) {
    const auto& q = rot[k % 60];
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double m[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z),     2 * (x * z + w * y),
                         2 * (x * y + w * z),     1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                         2 * (x * z - w * y),     2 * (y * z + w * x),     1 - 2 * (x * x + y * y)};
    const double sign = k < 60 ? 1.0 : -1.0;
    for (int e = 0; e < 9; ++e) tb.i_ops[k][e] = sign * m[e];
    for (int c = 0; c < 3; ++c)
      tb.i_sites[k][c] = tb.i_ops[k][c] * seed[0] + tb.i_ops[k][3 + c] * seed[1] + tb.i_ops[k][6 + c] * seed[2];
  }
  const double ball = std::cbrt(3.0 * kPi / 4.0);
  tb.cb_scale = std::pow(kPi / 6.0, 1.0 / 6.0);
  tb.cb_pref = std::sqrt(6.0 / kPi);
  tb.cb_prek = ball * std::pow(2.0, 0.25) / (std::pow(kPi, 5.0 / 6.0) / std::pow(6.0, 1.0 / 6.0) / 2.0);
  tb.cb_lift = std::sqrt(kPi / 24.0);
  return tb;
}

// The odd lattice sum nearest (1 + sqrt 2)(h + 1/2), the lower on a tie (grids.zig:10-15).
int32_t cut_index(uint32_t h) {
  const double wanted = (1.0 + std::sqrt(2.0)) * (double(h) + 0.5);
  int32_t lower = int32_t(std::floor(wanted));
  const int32_t odd = lower % 2 != 0 ? lower : lower - 1;
  return wanted - double(odd) <= double(odd + 2) - wanted ? odd : odd + 2;
}

}  // namespace

const fcc::Tables& tables() {
  static const fcc::Tables tb = build();
  return tb;
}

uint64_t count(uint32_t group, uint32_t h) {
  if (group < 8 || group > 10) fail("TransportNeedsTOrOOrI");
  if (h < 1 || h > kMaxSemiEdge) fail("InvalidGridResolution");
  const Plan p = plan(group, h);
  return p.job.count;
}

Plan plan(uint32_t group, uint32_t h) {
  if (group < 8 || group > 10) fail("TransportNeedsTOrOOrI");
  if (h < 1 || h > kMaxSemiEdge) fail("InvalidGridResolution");
  Plan p;
  fcc::Job& job = p.job;
  job.group = group;
  job.h = int32_t(h);
  job.cut = cut_index(h);
  job.bound = group == 8 ? int32_t(2 * h) : int32_t(h);
  job.side = uint32_t(2 * job.bound + 1);
  job.step = (group == 8 ? std::cbrt(kPi * kPi / 16.0) : std::pow(kPi, 2.0 / 3.0)) / double(2 * h + 1);
  job.beta = double(job.cut) / (double(h) + 0.5) - 1.0;
  const fcc::Tables& tb = tables();
  job.src_total = group == 9 ? fcc::o_source_cdf(tb.gx, tb.gw, fcc::kGauss, kPi / 2, job.beta) : 1.0;
  p.offsets.resize(size_t(job.side) * job.side + 1);
  uint64_t total = 0;
  for (int32_t i = -job.bound; i <= job.bound; ++i)
    for (int32_t j = -job.bound; j <= job.bound; ++j) {
      p.offsets[size_t(i + job.bound) * job.side + size_t(j + job.bound)] = uint32_t(total);
      int32_t k0, kmax;
      total += fcc::row_span(job, i, j, k0, kmax);
      if (total > hopf::kMaxCount) fail("GridTooLarge");
    }
  p.offsets.back() = uint32_t(total);
  job.count = uint32_t(total);
  return p;
}

const char* description(uint32_t group) {
  switch (group) {
    case 8: return "FCC octahedron transport";
    case 9: return "FCC truncated cube transport";
    default: return "FCC cube transport";
  }
}

}  // namespace zxt::transport
