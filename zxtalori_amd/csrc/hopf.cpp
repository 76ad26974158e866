// Inverse Hopf product grids (src/hopf.zig in the reference, after SO3K_Grids
// so3k/product.py), on the host in FP64.
//
// The map pi(q) = R(q)^-1 z sends SO(3) onto the sphere, and pi(q g) = R(g)^-1 pi(q), so
// SO(3)/K is a bundle over the sphere modulo K whose fibres are the left turns R_z(t) q.
// A grid is a K-invariant sphere mesh (a subdivided octahedron or icosahedron mapped to
// the sphere by an equal-area map, or equal-area latitude rings for the six-fold root)
// times n equally spaced fibre angles.  Each K-orbit of mesh nodes contributes one node
// v, the one lying in the sector of the pole z, with turns R_z((2 pi j + phase) / n) s(v),
// j < n / (stabiliser order of v), where s(v) takes v to z and odd mesh layers are
// staggered by half a turn step.  A group H other than the three roots uses the root P
// that contains R H R^-1 and keeps one point q c R per H-orbit, c over the cosets of
// R H R^-1 in P.
//
// Point order follows the reference grid row by row (its mesh vertex and face order,
// node first occurrence, orbit, fibre turn, coset), so grids compare row by row.
#include "hopf.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "engine.h"

namespace zxt::hopf {
namespace {

using Quat = std::array<double, 4>;  // scalar first
using Vec = std::array<double, 3>;
using Mat = std::array<Vec, 3>;

constexpr double kPi = 3.14159265358979323846;
enum Id : uint32_t { C2, C3, C4, C6, D2, D3, D4, D6, T, O, I, C1 };
constexpr uint32_t kOrder[] = {2, 3, 4, 6, 4, 6, 8, 12, 12, 24, 60, 1};

[[noreturn]] void fail(const char* what) { throw EngineError{what}; }

// Runs body(begin, end) over [0, n) split into contiguous ranges, one per host thread (at
// most 16); every index is computed by the same code whichever thread takes it, so the
// result does not depend on the split.  Bodies must not throw.
template <class Body>
void parallel_ranges(size_t n, const Body& body) {
  const unsigned threads = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (threads == 1 || n < 8192) {
    body(size_t(0), n);
    return;
  }
  const size_t chunk = (n + threads - 1) / threads;
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (unsigned t = 0; t < threads; ++t) {
    const size_t a = size_t(t) * chunk, b = std::min(n, a + chunk);
    if (a >= b) continue;
    try {
      pool.emplace_back([&body, a, b] { body(a, b); });
    } catch (...) {
      body(a, b);  // no thread to be had: this range on the calling thread
    }
  }
  for (std::thread& t : pool) t.join();
}

Quat operator*(const Quat& a, const Quat& b) {
  return {a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
          a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
          a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
          a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]};
}
Quat conj(const Quat& q) { return {q[0], -q[1], -q[2], -q[3]}; }
double dot(const Quat& a, const Quat& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }
Quat unit(const Quat& q) {
  const double s = 1.0 / std::sqrt(dot(q, q));
  return {q[0] * s, q[1] * s, q[2] * s, q[3] * s};
}
double dot(const Vec& a, const Vec& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec unit(const Vec& v) {
  const double s = 1.0 / std::sqrt(dot(v, v));
  return {v[0] * s, v[1] * s, v[2] * s};
}
Vec operator-(const Vec& a, const Vec& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
// m v and m^T v.
Vec apply(const Mat& m, const Vec& v) { return {dot(m[0], v), dot(m[1], v), dot(m[2], v)}; }
Vec apply_t(const Mat& m, const Vec& v) {
  return {m[0][0] * v[0] + m[1][0] * v[1] + m[2][0] * v[2], m[0][1] * v[0] + m[1][1] * v[1] + m[2][1] * v[2],
          m[0][2] * v[0] + m[1][2] * v[1] + m[2][2] * v[2]};
}
// The rotation matrix of a unit quaternion.
Mat rotation(const Quat& q) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  return {{{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
           {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
           {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}}};
}
// The rotation by 2 pi / k about `axis`.
Quat turn(const Vec& axis, double k) {
  const Vec a = unit(axis);
  const double s = std::sin(kPi / k);
  return {std::cos(kPi / k), a[0] * s, a[1] * s, a[2] * s};
}

// R with R H R^-1 inside the root (so3k.product.embedding; hopf.zig:38-44).
Quat embedding(uint32_t group) {
  switch (group) {
    case C2:
    case D2:
    case T: return {0.9619383577839175, 0, 0.2732665289126717, 0};
    case C3:
    case D3: return {0.9459753748917731, 0.17451645846277514, 0.2687317799958305, -0.04957646865452023};
    default: return {1, 0, 0, 0};
  }
}

uint32_t fibre_count(uint32_t root, uint32_t f) {
  switch (root) {
    case O: return f < 6 ? 4 * f : 4 * f - 12 * ((2 * f + 9) / 18);
    case I: return 5 * f;
    default: return 6 * ((4 * f + 5) / 6);
  }
}

// Nodes on ring j of the six-fold grid, j = f being the equator.
uint32_t ring_size(uint32_t f, uint32_t j) {
  const uint32_t mult = j == f ? 12 : 6;
  const double x = 4.0 * double(f) * std::sin(kPi * double(j) / double(2 * f));
  return std::max(mult, mult * uint32_t(std::floor(x / double(mult) + 0.5)));
}

uint32_t node_count(uint32_t root, uint32_t f) {
  if (root == O) return 4 * f * f + 2;
  if (root == I) return 10 * f * f + 2;
  uint32_t total = 2 + ring_size(f, f);
  for (uint32_t j = 1; j < f; ++j) total += 2 * ring_size(f, j);
  return total;
}

// ---- the polyhedra ----------------------------------------------------------------
// Vertex and face order of the reference grid (its sorted vertices and hull simplices),
// which fixes the node order (hopf.zig:216-241).
constexpr uint8_t kOctaFaces[8][3] = {{5, 4, 3}, {5, 4, 0}, {2, 4, 3}, {2, 4, 0},
                                      {1, 5, 0}, {1, 5, 3}, {1, 2, 0}, {1, 2, 3}};
constexpr uint8_t kIcoFaces[20][3] = {{5, 3, 0},   {1, 3, 0},  {4, 5, 0},   {9, 5, 3},   {7, 1, 3},
                                      {7, 9, 11},  {7, 9, 3},  {10, 4, 5},  {10, 9, 5},  {10, 8, 11},
                                      {10, 9, 11}, {10, 8, 4}, {6, 7, 1},   {6, 7, 11},  {6, 8, 11},
                                      {2, 8, 4},   {2, 1, 0},  {2, 4, 0},   {2, 6, 1},   {2, 6, 8}};
// Icosahedron vertices off the poles: height sign (times 1/sqrt 5) and azimuth in degrees;
// entries 5 and 6 are the poles -z and +z.
constexpr double kIcoRing[12][2] = {{-1, 180}, {1, -144}, {1, 144}, {-1, -108}, {-1, 108}, {0, 0},
                                    {0, 0},    {1, -72},  {1, 72},  {-1, -36},  {-1, 36},  {1, 0}};

std::vector<Vec> polyhedron(bool ico) {
  if (!ico) return {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  std::vector<Vec> v(12);
  const double r = 2 / std::sqrt(5.0);
  for (int k = 0; k < 12; ++k) {
    const double a = kIcoRing[k][1] * kPi / 180;
    v[k] = {r * std::cos(a), r * std::sin(a), kIcoRing[k][0] / std::sqrt(5.0)};
  }
  v[5] = {0, 0, -1};
  v[6] = {0, 0, 1};
  return v;
}

// ---- the equal-area map from the flat faces to the sphere ----------------------------
// One chamber of the polyhedron's full symmetry group (rotations and their negatives) is a
// triangle with the pole z as a vertex, azimuths [lo, hi] and a far wall w . x = 0; it lies
// on one flat face.  A Knothe-Rosenblatt map matches first the azimuth marginals of flat
// area and of spherical area, then, along each azimuth, the flat area fraction (s / s_max)^2
// of the fan from the pole with the cap fraction sin^2(theta / 2) / sin^2(theta_max / 2).
// Points elsewhere are carried into the chamber by the symmetry whose image of the
// chamber seed is nearest, mapped, and carried back (hopf.zig:243-365).
class AreaMap {
 public:
  explicit AreaMap(bool ico) {
    const Vec seed = ico ? Vec{std::sin(0.2) * std::cos(0.1 * kPi), std::sin(0.2) * std::sin(0.1 * kPi), std::cos(0.2)}
                         : unit(Vec{1, 2, 3});
    lo_ = ico ? 0 : kPi / 4;
    hi_ = ico ? kPi / 5 : kPi / 2;
    const std::vector<Vec> v = polyhedron(ico);
    face_ = ico ? unit(Vec{v[6][0] + v[11][0] + v[8][0], v[6][1] + v[11][1] + v[8][1], v[6][2] + v[11][2] + v[8][2]})
                : unit(Vec{1, 1, 1});
    const std::vector<Quat> rot = elements(ico ? I : O);
    ops_.resize(2 * rot.size());
    for (size_t i = 0; i < rot.size(); ++i) {
      Mat m = rotation(rot[i]);
      // Entries that are integers in exact arithmetic come out of the products within
      // 1e-15; snapped, the chamber walls through them stay exact.
      for (Vec& row : m)
        for (double& x : row)
          if (std::fabs(x - std::round(x)) < 1e-10) x = std::round(x);
      ops_[i] = m;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) ops_[i + rot.size()][r][c] = -m[r][c];
    }
    // The far wall is the bisector of the seed and one of its images that bounds the polar
    // angle first along the middle azimuth (walls through the pole bound the azimuth).
    const double mid = 0.5 * (lo_ + hi_);
    double best = -1;
    seeds_.resize(ops_.size());
    for (size_t i = 0; i < ops_.size(); ++i) {
      seeds_[i] = apply_t(ops_[i], seed);
      Vec h = seed - seeds_[i];
      if (std::sqrt(dot(h, h)) <= 1e-9) continue;
      h = unit(h);
      if (std::fabs(h[2]) < 1e-12) continue;
      const double psi = std::atan2(h[2], h[0] * std::cos(mid) + h[1] * std::sin(mid));
      if (psi > best) {
        best = psi;
        wall_ = h;
      }
    }
    flat_lo_ = flat_cdf(lo_);
    flat_span_ = flat_cdf(hi_) - flat_lo_;
    cap_lo_ = cap_cdf(lo_);
    cap_span_ = cap_cdf(hi_) - cap_lo_;
  }

  Vec operator()(const Vec& v) const {
    size_t g = 0;
    double nearest = dot(v, seeds_[0]);
    for (size_t k = 1; k < seeds_.size(); ++k) {
      const double s = dot(v, seeds_[k]);
      if (s > nearest) {
        nearest = s;
        g = k;
      }
    }
    const Vec w = apply(ops_[g], v);
    const double phi = std::clamp(std::atan2(w[1], w[0]), lo_, hi_);
    const double theta = std::atan2(std::sqrt(w[0] * w[0] + w[1] * w[1]), w[2]);
    // Distance from the pole along the flat face at azimuth phi is proportional to
    // tan(theta) / (n_z + a tan(theta)).
    const double a = face_[0] * std::cos(phi) + face_[1] * std::sin(phi);
    const double t = std::tan(theta), tm = std::tan(polar_limit(phi));
    const double ratio = (t / (face_[2] + a * t)) / (tm / (face_[2] + a * tm));
    const double phi2 = cap_azimuth(std::clamp((flat_cdf(phi) - flat_lo_) / flat_span_, 0.0, 1.0));
    const double theta2 =
        2 * std::asin(std::sqrt(std::clamp(ratio * ratio, 0.0, 1.0)) * std::sin(0.5 * polar_limit(phi2)));
    const Vec out{std::sin(theta2) * std::cos(phi2), std::sin(theta2) * std::sin(phi2), std::cos(theta2)};
    return unit(apply_t(ops_[g], out));
  }

 private:
  // Polar angle of the far wall at azimuth phi.
  double polar_limit(double phi) const {
    return kPi - std::atan2(wall_[2], wall_[0] * std::cos(phi) + wall_[1] * std::sin(phi));
  }
  // Flat area per unit azimuth is proportional to 1 / (P cos phi + Q sin phi)^2, whose
  // antiderivative is (P sin - Q cos) / (P cos + Q sin).
  double flat_cdf(double phi) const {
    const double p = wall_[2] * face_[0] - face_[2] * wall_[0];
    const double q = wall_[2] * face_[1] - face_[2] * wall_[1];
    return (p * std::sin(phi) - q * std::cos(phi)) / (p * std::cos(phi) + q * std::sin(phi));
  }
  // Spherical area per unit azimuth, 1 - cos(polar_limit(phi)), integrated.
  double cap_cdf(double phi) const {
    return phi - std::asin(std::clamp(-wall_[0] * std::sin(phi) + wall_[1] * std::cos(phi), -1.0, 1.0));
  }
  // The azimuth where the normalised cap_cdf reaches `fraction`: Newton steps kept inside a
  // shrinking bracket (bisection when a step leaves it).
  double cap_azimuth(double fraction) const {
    double lo = lo_, hi = hi_, x = lo_ + fraction * (hi_ - lo_);
    for (int it = 0; it < 100; ++it) {
      const double r = (cap_cdf(x) - cap_lo_) / cap_span_ - fraction;
      if (r < 0)
        lo = x;
      else
        hi = x;
      const double slope = (1 - std::cos(polar_limit(x))) / cap_span_;
      double next = x - r / slope;
      if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
      if (std::fabs(next - x) < 1e-15) return next;
      x = next;
    }
    return x;
  }

  std::vector<Mat> ops_;
  std::vector<Vec> seeds_;
  double lo_, hi_;
  Vec face_, wall_{};
  double flat_lo_, flat_span_, cap_lo_, cap_span_;
};

struct Mesh {
  std::vector<Vec> nodes;
  std::vector<uint32_t> layers;  // lattice rows from the nearest vertex (rings from a pole)
};

// Frequency-f subdivision of every face, each shared node once in order of first
// occurrence, mapped to the sphere (hopf.zig:376-422).  A node is identified by its
// nonzero barycentric weights on polyhedron vertices.
Mesh polyhedral_mesh(uint32_t root, uint32_t f) {
  const bool ico = root == I;
  const std::vector<Vec> v = polyhedron(ico);
  const AreaMap area(ico);
  const size_t nfaces = ico ? 20 : 8;
  Mesh m;
  const uint32_t total = node_count(root, f);
  m.nodes.reserve(total);
  m.layers.reserve(total);
  std::unordered_set<uint64_t> seen;
  seen.reserve(2 * size_t(total));
  const double scale = 1.0 / double(f);
  // First the nodes in order of first occurrence (face, i, j), then their positions on the
  // sphere (the area map, the costly part) on all host threads.
  std::vector<std::array<uint32_t, 3>> site;
  site.reserve(total);
  for (uint32_t fi = 0; fi < nfaces; ++fi) {
    const uint8_t* face = ico ? kIcoFaces[fi] : kOctaFaces[fi];
    for (uint32_t i = 0; i <= f; ++i) {
      for (uint32_t j = 0; j + i <= f; ++j) {
        const uint32_t w[3] = {i, j, f - i - j};
        // Key: the (vertex, weight) pairs with nonzero weight, by vertex, 14 bits each.
        uint32_t pairs[3];
        int np = 0;
        for (int k = 0; k < 3; ++k)
          if (w[k]) pairs[np++] = uint32_t(face[k]) << 10 | w[k];
        std::sort(pairs, pairs + np);
        uint64_t key = uint64_t(np);
        for (int k = 0; k < np; ++k) key = key << 14 | pairs[k];
        if (!seen.insert(key).second) continue;
        if (site.size() == total) fail("BaseMeshCount");
        site.push_back({fi, i, j});
        m.layers.push_back(f - std::max(w[0], std::max(w[1], w[2])));
      }
    }
  }
  if (site.size() != total) fail("BaseMeshCount");
  m.nodes.resize(total);
  parallel_ranges(total, [&](size_t a, size_t b) {
    for (size_t k = a; k < b; ++k) {
      const uint8_t* face = ico ? kIcoFaces[site[k][0]] : kOctaFaces[site[k][0]];
      const uint32_t i = site[k][1], j = site[k][2];
      const uint32_t w[3] = {i, j, f - i - j};
      Vec x{0, 0, 0};
      for (int c = 0; c < 3; ++c) {
        const double bw = double(w[c]) * scale;
        for (int e = 0; e < 3; ++e) x[e] += bw * v[face[c]][e];
      }
      m.nodes[k] = area(unit(x));
    }
  });
  return m;
}

// Six-fold invariant equal-area rings: the poles, rings 1..f-1 in both hemispheres and the
// equator, each ring at the middle height of a band of area 4 pi / |V| per node;
// alternate rings are staggered by half a step, the equator holds the two-fold axes
// (hopf.zig:424-466).
Mesh ring_mesh(uint32_t f) {
  const uint32_t total = node_count(D6, f);
  Mesh m;
  m.nodes.reserve(total);
  m.layers.reserve(total);
  m.nodes.push_back({0, 0, 1});
  m.nodes.push_back({0, 0, -1});
  m.layers.push_back(0);
  m.layers.push_back(0);
  const double size = double(total);
  double top = 1 - 2 / size;
  for (uint32_t j = 1; j < f; ++j) {
    const uint32_t n = ring_size(f, j);
    const double width = 2 * double(n) / size;
    const double z = top - 0.5 * width;
    const double radius = std::sqrt(1 - z * z);
    for (const double height : {z, -z}) {
      for (uint32_t i = 0; i < n; ++i) {
        const double phi = 2 * kPi * (double(i) + 0.5 * double(j % 2)) / double(n);
        m.nodes.push_back({radius * std::cos(phi), radius * std::sin(phi), height});
        m.layers.push_back(j);
      }
    }
    top -= width;
  }
  const uint32_t equator = ring_size(f, f);
  for (uint32_t i = 0; i < equator; ++i) {
    const double phi = 2 * kPi * double(i) / double(equator);
    m.nodes.push_back({std::cos(phi), std::sin(phi), 0});
    m.layers.push_back(f);
  }
  if (m.nodes.size() != total) fail("BaseMeshCount");
  return m;
}

// Node lookup by position on a 1e-7 lattice, probing the 26 neighbouring cells when the
// point's own cell is empty.
class NodeIndex {
 public:
  explicit NodeIndex(const std::vector<Vec>& nodes) {
    table_.reserve(2 * nodes.size());
    for (uint32_t i = 0; i < nodes.size(); ++i) table_.emplace(cell(nodes[i], 0, 0, 0), i);
  }
  // Index of the node at p, or -1.
  long find(const Vec& p) const {
    auto it = table_.find(cell(p, 0, 0, 0));
    if (it != table_.end()) return it->second;
    for (int o = 0; o < 27; ++o) {
      it = table_.find(cell(p, o % 3 - 1, o / 3 % 3 - 1, o / 9 - 1));
      if (it != table_.end()) return it->second;
    }
    return -1;
  }

 private:
  using Cell = std::array<int32_t, 3>;
  struct CellHash {
    size_t operator()(const Cell& c) const {
      return size_t(uint64_t(uint32_t(c[0])) * 0x9E3779B97F4A7C15ull ^ uint64_t(uint32_t(c[1])) << 21 ^
                    uint64_t(uint32_t(c[2])) * 0xC2B2AE3D27D4EB4Full);
    }
  };
  static Cell cell(const Vec& p, int dx, int dy, int dz) {
    return {int32_t(std::round(p[0] * 1e7)) + dx, int32_t(std::round(p[1] * 1e7)) + dy,
            int32_t(std::round(p[2] * 1e7)) + dz};
  }
  std::unordered_map<Cell, uint32_t, CellHash> table_;
};

// The sector at the pole z in which each orbit keeps its node: the Voronoi cell of z among
// the polyhedron's vertices (for the six-fold root the northern hemisphere), azimuths
// 0 .. 2 pi / m for the m-fold z axis (hopf.zig:487-513).
struct Sector {
  std::vector<Vec> poles;
  double wedge;
  explicit Sector(uint32_t root) {
    if (root == O) {
      poles = polyhedron(false);
      wedge = 2 * kPi / 4;
    } else if (root == I) {
      poles = polyhedron(true);
      wedge = 2 * kPi / 5;
    } else {
      poles = {{0, 0, 1}, {0, 0, -1}};
      wedge = 2 * kPi / 6;
    }
  }
  bool contains(const Vec& w) const {
    for (const Vec& p : poles)
      if (dot(w, p) > w[2] + 1e-11) return false;
    if (std::sqrt(w[0] * w[0] + w[1] * w[1]) < 1e-11) return true;
    const double az = std::atan2(w[1], w[0]);
    return az >= -1e-11 && az <= wedge + 1e-11;
  }
};

// On the sector's edges an orbit keeps the node with the largest x, then y (to 1e-11).
bool preferred(const Vec& a, const Vec& b) {
  const double ax = std::round(a[0] * 1e11), bx = std::round(b[0] * 1e11);
  if (ax != bx) return ax > bx;
  return std::round(a[1] * 1e11) > std::round(b[1] * 1e11);
}

// A rotation taking the unit vector v to z (a half turn about x for v = -z).
Quat section(const Vec& v) {
  const Quat q{1 + v[2], v[1], -v[0], 0};
  if (std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) < 1e-12) return {0, 1, 0, 0};
  return unit(q);
}

}  // namespace

std::vector<std::array<double, 4>> elements(uint32_t group) {
  const Vec z{0, 0, 1}, x{1, 0, 0}, d{1, 1, 1};
  std::vector<Quat> gens;
  switch (group) {
    case C2: gens = {turn(z, 2)}; break;
    case C3: gens = {turn(z, 3)}; break;
    case C4: gens = {turn(z, 4)}; break;
    case C6: gens = {turn(z, 6)}; break;
    case D2: gens = {turn(z, 2), turn(x, 2)}; break;
    case D3: gens = {turn(z, 3), turn(x, 2)}; break;
    case D4: gens = {turn(z, 4), turn(x, 2)}; break;
    case D6: gens = {turn(z, 6), turn(x, 2)}; break;
    case T: gens = {turn(z, 2), turn(x, 2), turn(d, 3)}; break;
    case O: gens = {turn(z, 4), turn(d, 3)}; break;
    case I: gens = {turn(z, 5), turn({2, 0, 1}, 5)}; break;
    default: break;  // C1
  }
  std::vector<Quat> e{{1, 0, 0, 0}};
  for (size_t i = 0; i < e.size(); ++i) {
    for (const Quat& g : gens) {
      const Quat p = e[i] * g;
      const bool known =
          std::any_of(e.begin(), e.end(), [&](const Quat& k) { return std::fabs(dot(k, p)) > 1 - 1e-9; });
      if (!known) e.push_back(p);
    }
  }
  if (e.size() != kOrder[group]) fail("GroupClosure");
  for (Quat& q : e) {
    q = unit(q);
    if (q[0] < 0) q = {-q[0], -q[1], -q[2], -q[3]};
  }
  std::stable_sort(e.begin(), e.end(), [](const Quat& a, const Quat& b) { return a[0] > b[0]; });
  return e;
}

uint32_t root_of(uint32_t group) {
  switch (group) {
    case C4:
    case D4:
    case O: return O;
    case C6:
    case D6: return D6;
    default: return I;
  }
}

uint64_t raw_count(uint32_t group, uint32_t f) {
  const uint32_t root = root_of(group);
  return uint64_t(fibre_count(root, f)) * node_count(root, f) / kOrder[group];
}

uint32_t count(uint32_t group, uint32_t f) {
  if (f < 1 || f > kMaxFrequency) fail("InvalidGridResolution");
  const uint64_t n = raw_count(group, f);
  if (n > kMaxCount) fail("GridTooLarge");
  return uint32_t(n);
}

uint32_t nearest_frequency(uint32_t group, uint32_t target) {
  uint32_t best = 1;
  double score = INFINITY;
  for (uint32_t f = 1; f <= kMaxFrequency; ++f) {
    const uint64_t n = raw_count(group, f);
    if (n > kMaxCount) continue;
    const double s = std::fabs(std::log(double(n) / double(std::max<uint32_t>(target, 1))));
    if (s < score) {
      score = s;
      best = f;
    }
  }
  return best;
}

const char* description(uint32_t group) {
  switch (root_of(group)) {
    case I: return "Inverse Hopf product grid (icosahedral root)";
    case O: return "Inverse Hopf product grid (octahedral root)";
    default: return "Inverse Hopf product grid (hexagonal root)";
  }
}

void points(uint32_t group, uint32_t f, float* out) {
  if (group > C1) fail("UnknownGroup");
  const uint32_t total = count(group, f);
  const uint32_t root = root_of(group);
  const Mesh mesh = root == D6 ? ring_mesh(f) : polyhedral_mesh(root, f);
  const std::vector<Quat> parent = elements(root);
  std::vector<Mat> mats(parent.size());
  for (size_t i = 0; i < parent.size(); ++i) mats[i] = rotation(parent[i]);
  const std::vector<Quat> sub = elements(group);

  // Cosets c H' of H' = R H R^-1 in the root, each by its first element in root order.
  const Quat r = embedding(group), r_inv = conj(r);
  std::vector<Quat> cosets;
  std::vector<bool> covered(parent.size(), false);
  for (size_t i = 0; i < parent.size(); ++i) {
    if (covered[i]) continue;
    cosets.push_back(parent[i]);
    for (const Quat& h : sub) {
      const Quat member = parent[i] * ((r * h) * r_inv);
      for (size_t k = 0; k < parent.size(); ++k)
        if (std::fabs(dot(parent[k], member)) > 1 - 1e-9) covered[k] = true;
    }
  }
  if (cosets.size() * sub.size() != parent.size()) fail("GroupNotInRoot");
  // q c R = R_z(t) (s c R): the per-coset tails c R, then s (c R) once per orbit.
  std::vector<Quat> tails(cosets.size());
  for (size_t k = 0; k < cosets.size(); ++k) tails[k] = cosets[k] * r;

  // Fibre turns R_z(t) of both phases, t = (2 pi j + pi phase) / n.
  const uint32_t n = fibre_count(root, f);
  std::vector<Quat> spins(2 * size_t(n));
  for (uint32_t phase = 0; phase < 2; ++phase)
    for (uint32_t j = 0; j < n; ++j) {
      const double t = (2 * kPi * double(j) + kPi * double(phase)) / double(n);
      spins[phase * n + j] = {std::cos(0.5 * t), 0, 0, std::sin(0.5 * t)};
    }

  const NodeIndex index(mesh.nodes);
  const Sector sector(root);
  std::vector<bool> seen(mesh.nodes.size(), false);
  // One record per orbit in node order: its section, spin phase and turns; the rows are
  // written afterwards on all host threads, each orbit at its offset.
  struct Orbit {
    Quat s;
    uint32_t phase, turns;
    size_t first;
  };
  std::vector<Orbit> orbits;
  size_t rows = 0;
  for (size_t i = 0; i < mesh.nodes.size(); ++i) {
    if (seen[i]) continue;
    const Vec& v = mesh.nodes[i];
    // Walk the orbit: mark its nodes, count the images landing on v (the stabiliser
    // order) and keep the preferred image in the sector.
    uint32_t stabiliser = 0;
    bool have = false;
    Vec rep{};
    for (const Mat& m : mats) {
      const Vec image = apply_t(m, v);
      const long j = index.find(image);
      if (j < 0) fail("BaseMeshNotInvariant");
      const Vec d = mesh.nodes[size_t(j)] - image;
      if (dot(d, d) > 1e-18) fail("BaseMeshNotInvariant");
      seen[size_t(j)] = true;
      if (size_t(j) == i) ++stabiliser;
      if (sector.contains(image) && (!have || preferred(image, rep))) {
        rep = image;
        have = true;
      }
    }
    if (!have) fail("OrbitOutsideSector");
    if (n % stabiliser != 0) fail("FibreNotDivisible");
    const uint32_t turns = n / stabiliser;
    if (rows + size_t(turns) * cosets.size() > total) fail("ProductCount");
    orbits.push_back({section(rep), mesh.layers[i] % 2, turns, rows});
    rows += size_t(turns) * cosets.size();
  }
  if (rows != total) fail("ProductCount");
  parallel_ranges(orbits.size(), [&](size_t a, size_t b) {
    std::vector<Quat> heads(cosets.size());
    for (size_t o = a; o < b; ++o) {
      const Orbit& orb = orbits[o];
      for (size_t k = 0; k < cosets.size(); ++k) heads[k] = orb.s * tails[k];
      const Quat* row = &spins[size_t(orb.phase) * n];
      float* dst = out + orb.first * 4;
      for (uint32_t j = 0; j < orb.turns; ++j)
        for (const Quat& h : heads) {
          const Quat q = row[j] * h;
          dst[0] = float(q[0]);
          dst[1] = float(q[1]);
          dst[2] = float(q[2]);
          dst[3] = float(q[3]);
          dst += 4;
        }
    }
  });
}

}  // namespace zxt::hopf
