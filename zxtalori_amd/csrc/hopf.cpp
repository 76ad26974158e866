// Inverse Hopf product grids of SO(3)/K on the host in FP64 (the grids of SO3K_Grids
// so3k/product.py that the reference builds in src/hopf.zig).
//
// The method.  pi(q) = R(q)^-1 z maps SO(3) onto the sphere with pi(q g) = R(g)^-1 pi(q),
// so SO(3)/K fibres over the sphere modulo K, the fibre over a point being the left turns
// R_z(t) q.  A grid is a K-invariant sphere mesh times n equally spaced fibre angles.  Three
// root groups build it directly: O and I subdivide the octahedron or icosahedron at
// frequency f and carry the nodes to the sphere with an equal-area map, D6 takes equal-area
// latitude rings.  Each orbit of mesh nodes under the root keeps its node v in the sector
// at the pole z and contributes R_z((2 pi j + phase) / n) s(v), j < n / |stabiliser of v|,
// where s(v) turns v onto z and odd mesh layers are shifted by half a step.  Any other group
// H uses the root P holding R H R^-1 and keeps, per orbit point, one q c R per coset c of
// R H R^-1 in P.
//
// What fixes the rows (so that grids compare row by row with the reference's): the mesh
// node order (the polyhedra's vertex and face tables, nodes by first occurrence over faces),
// the orbit order (node order), the representative rule in the sector, the group element
// order (closure of the generators, sorted by scalar part) that orders the cosets, and the
// spin-then-coset order within an orbit.
//
// The code is organised around three pieces of this build's own: the equal-area map as a
// fold into one mirror chamber by reflections, the chamber's two area functions as
// closed-form triangle areas (a flat triangle and a spherical one), and orbits walked on
// the mesh's combinatorics (barycentric signatures permuted by the vertex permutations of
// the root, or ring and azimuth indices for the ring mesh) instead of by position lookup.
#include "hopf.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <map>
#include <thread>
#include <unordered_map>

#include "engine.h"

namespace zxt::hopf {
namespace {

using Q4 = std::array<double, 4>;  // quaternion, scalar first
using V3 = std::array<double, 3>;
using M3 = std::array<V3, 3>;

constexpr double kPi = 3.14159265358979323846;
enum GroupId : uint32_t { kC2, kC3, kC4, kC6, kD2, kD3, kD4, kD6, kT, kO, kI, kC1 };
constexpr uint32_t kSize[12] = {2, 3, 4, 6, 4, 6, 8, 12, 12, 24, 60, 1};

[[noreturn]] void fail(const char* what) { throw EngineError{what}; }

// ---- small algebra ------------------------------------------------------------------
Q4 qmul(const Q4& a, const Q4& b) {
  return {a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
          a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
          a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
          a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]};
}
double qdot(const Q4& a, const Q4& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }
Q4 qunit(const Q4& q) {
  const double s = 1.0 / std::sqrt(qdot(q, q));
  return {q[0] * s, q[1] * s, q[2] * s, q[3] * s};
}
double dot3(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3 cross3(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
V3 scale3(const V3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
V3 sub3(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 add3(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
V3 unit3(const V3& v) { return scale3(v, 1.0 / std::sqrt(dot3(v, v))); }
double norm3(const V3& v) { return std::sqrt(dot3(v, v)); }
// R(q)^T v: the rotation of the inverse, which is how the orbit walk applies an element.
V3 rotate_inverse(const M3& m, const V3& v) {
  return {m[0][0] * v[0] + m[1][0] * v[1] + m[2][0] * v[2], m[0][1] * v[0] + m[1][1] * v[1] + m[2][1] * v[2],
          m[0][2] * v[0] + m[1][2] * v[1] + m[2][2] * v[2]};
}
M3 matrix_of(const Q4& q) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  return {{{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
           {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
           {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}}};
}
// Rotation by 2 pi / fold about an axis.
Q4 axis_turn(V3 axis, double fold) {
  axis = unit3(axis);
  const double h = kPi / fold;
  return {std::cos(h), axis[0] * std::sin(h), axis[1] * std::sin(h), axis[2] * std::sin(h)};
}

// Splits [0, n) into blocks handed out to up to 16 host threads; each index is computed by
// the same code on whichever thread takes it, so the result does not depend on the split.
template <class Body>
void for_blocks(size_t n, size_t block, const Body& body) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const unsigned workers = unsigned(std::min<size_t>(std::min(16u, hw), (n + block - 1) / block));
  if (workers <= 1) {
    body(size_t(0), n);
    return;
  }
  std::atomic<size_t> next{0};
  auto drain = [&] {
    for (size_t a; (a = next.fetch_add(block)) < n;) body(a, std::min(n, a + block));
  };
  std::vector<std::thread> team;
  for (unsigned t = 1; t < workers; ++t) {
    try {
      team.emplace_back(drain);
    } catch (...) {
      break;  // fewer threads: the caller drains the rest
    }
  }
  drain();
  for (std::thread& t : team) t.join();
}

// ---- the grid's sizes (so3k/product.py; hopf.zig:45-76) ----------------------------------
uint32_t root_group(uint32_t group) {
  if (group == kC4 || group == kD4 || group == kO) return kO;
  if (group == kC6 || group == kD6) return kD6;
  return kI;  // every group that embeds in I, C1 included
}
uint32_t fibre_turns(uint32_t root, uint32_t f) {
  if (root == kI) return 5 * f;
  if (root == kO) return f < 6 ? 4 * f : 4 * f - 12 * ((2 * f + 9) / 18);
  return 6 * ((4 * f + 5) / 6);
}
// Nodes of latitude ring j (1 <= j <= f, the equator being j = f) of the ring mesh: the
// nearest multiple of 6 (12 on the equator) to 4 f sin(pi j / 2f), at least one multiple.
uint32_t ring_nodes(uint32_t f, uint32_t j) {
  const uint32_t unit = j == f ? 12u : 6u;
  const double ideal = 4.0 * f * std::sin(kPi * j / (2.0 * f));
  return unit * std::max(1u, uint32_t(std::floor(ideal / unit + 0.5)));
}
uint32_t mesh_nodes(uint32_t root, uint32_t f) {
  if (root == kI) return 10 * f * f + 2;
  if (root == kO) return 4 * f * f + 2;
  uint32_t total = 2 + ring_nodes(f, f);
  for (uint32_t j = 1; j < f; ++j) total += 2 * ring_nodes(f, j);
  return total;
}
// R with R H R^-1 inside the root (so3k.product.embedding; hopf.zig:38-44).
Q4 root_embedding(uint32_t group) {
  if (group == kC2 || group == kD2 || group == kT) return {0.9619383577839175, 0, 0.2732665289126717, 0};
  if (group == kC3 || group == kD3) return {0.9459753748917731, 0.17451645846277514, 0.2687317799958305, -0.04957646865452023};
  return {1, 0, 0, 0};
}

// ---- the polyhedra (vertex and face order of the reference's grids, hopf.zig:216-241) ------
constexpr uint8_t kOctaFaces[8][3] = {{5, 4, 3}, {5, 4, 0}, {2, 4, 3}, {2, 4, 0},
                                      {1, 5, 0}, {1, 5, 3}, {1, 2, 0}, {1, 2, 3}};
constexpr uint8_t kIcoFaces[20][3] = {{5, 3, 0},   {1, 3, 0},  {4, 5, 0},   {9, 5, 3},   {7, 1, 3},
                                      {7, 9, 11},  {7, 9, 3},  {10, 4, 5},  {10, 9, 5},  {10, 8, 11},
                                      {10, 9, 11}, {10, 8, 4}, {6, 7, 1},   {6, 7, 11},  {6, 8, 11},
                                      {2, 8, 4},   {2, 1, 0},  {2, 4, 0},   {2, 6, 1},   {2, 6, 8}};
// Icosahedron vertices: height (in units of 1/sqrt 5) and azimuth (degrees); entries 5 and
// 6 are overwritten by the poles -z and +z.
constexpr double kIcoVertices[12][2] = {{-1, 180}, {1, -144}, {1, 144}, {-1, -108}, {-1, 108}, {0, 0},
                                        {0, 0},    {1, -72},  {1, 72},  {-1, -36},  {-1, 36},  {1, 0}};

std::vector<V3> polyhedron_vertices(bool ico) {
  if (!ico) return {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  std::vector<V3> out;
  const double ring = 2.0 / std::sqrt(5.0), lift = 1.0 / std::sqrt(5.0);
  for (const auto& row : kIcoVertices) {
    const double az = row[1] * kPi / 180.0;
    out.push_back({ring * std::cos(az), ring * std::sin(az), row[0] * lift});
  }
  out[5] = {0, 0, -1};
  out[6] = {0, 0, 1};
  return out;
}

// ---- the equal-area map ------------------------------------------------------------------
// The polyhedron's full symmetry group (rotations and their negatives) is the reflection
// group of its mirror planes.  One chamber of it has the pole z as a vertex, azimuths
// [lo, hi] and a far wall; every unit vector is reflected into that chamber across the walls
// it violates, mapped there, and reflected back.  Inside the chamber the map is Knothe-
// Rosenblatt between the flat face (central projection of the chamber onto the face plane)
// and the sphere: first the azimuth, matching the share of flat area between lo and phi with
// the share of spherical area; then, along that azimuth, the flat fan's area fraction
// (s / s_max)^2 with the cap fraction sin^2(theta / 2) / sin^2(theta_max / 2).
//
// Both area shares come from triangles: the flat region between azimuths lo and phi is the
// plane triangle (P0, E(lo), E(phi)) of the pole's projection and the far edge's points,
// and the spherical region is the spherical triangle (z, A(lo), A(phi)) with A on the far
// wall's great circle, whose solid angle is 2 atan2(|a.(b x c)|, 1 + a.b + b.c + c.a).
class ChamberMap {
 public:
  explicit ChamberMap(bool ico) {
    const std::vector<V3> v = polyhedron_vertices(ico);
    if (ico) {
      // The chamber at +z between the 2-fold axis of edge (v6, v11), azimuth 0, and the
      // centre of face (v6, v11, v8), azimuth 36 degrees.
      lo_ = 0.0;
      hi_ = kPi / 5;
      const V3 edge = unit3(add3(v[6], v[11]));
      face_ = unit3(add3(add3(v[6], v[11]), v[8]));
      walls_[0] = {0, 1, 0};                                  // y >= 0
      walls_[1] = {std::sin(hi_), -std::cos(hi_), 0};         // azimuth <= 36 degrees
      walls_[2] = unit3(cross3(edge, face_));                 // the far wall
    } else {
      // The chamber 0 <= x <= y <= z: azimuths 45..90 degrees, far wall z = y.
      lo_ = kPi / 4;
      hi_ = kPi / 2;
      face_ = unit3({1, 1, 1});
      walls_[0] = {1, 0, 0};
      walls_[1] = unit3({-1, 1, 0});
      walls_[2] = unit3({0, -1, 1});
    }
    if (walls_[2][2] < 0) walls_[2] = scale3(walls_[2], -1.0);  // the pole on its inner side
    pole_flat_ = scale3(V3{0, 0, 1}, 1.0 / face_[2]);
    flat_total_ = flat_share(hi_);
    cap_total_ = cap_share(hi_);
  }

  V3 operator()(const V3& v) const {
    V3 w = v;
    uint8_t path[64];
    const int steps = fold(w, path);
    const double phi = std::clamp(std::atan2(w[1], w[0]), lo_, hi_);
    const double theta = std::atan2(std::hypot(w[0], w[1]), w[2]);
    const double limit = polar_limit(phi);
    const double s = norm3(sub3(flat_point(theta, phi), pole_flat_)) / norm3(sub3(flat_point(limit, phi), pole_flat_));
    const double phi2 = azimuth_of_cap_share(std::clamp(flat_share(phi) / flat_total_, 0.0, 1.0) * cap_total_);
    const double theta2 = 2.0 * std::asin(std::min(1.0, s) * std::sin(0.5 * polar_limit(phi2)));
    V3 out{std::sin(theta2) * std::cos(phi2), std::sin(theta2) * std::sin(phi2), std::cos(theta2)};
    for (int k = steps - 1; k >= 0; --k) out = reflect(out, walls_[path[k]]);
    return unit3(out);
  }

 private:
  static V3 reflect(const V3& v, const V3& wall) { return sub3(v, scale3(wall, 2.0 * dot3(v, wall))); }
  // Reflects w across violated walls until it lies in the chamber (each reflection across a
  // violated wall of a reflection group's chamber moves w strictly toward it, so this ends
  // within the group's longest word); returns the walls crossed, in order.
  int fold(V3& w, uint8_t* path) const {
    int steps = 0;
    while (steps < 64) {
      int worst = -1;
      double least = -1e-13;
      for (int k = 0; k < 3; ++k) {
        const double d = dot3(w, walls_[k]);
        if (d < least) {
          least = d;
          worst = k;
        }
      }
      if (worst < 0) break;
      w = reflect(w, walls_[worst]);
      path[steps++] = uint8_t(worst);
    }
    return steps;
  }
  // Polar angle where the far wall crosses azimuth phi: h.u = 0 for u = (sin t cos phi,
  // sin t sin phi, cos t) gives tan t = -h_z / (h_x cos phi + h_y sin phi).
  double polar_limit(double phi) const {
    const V3& h = walls_[2];
    return std::atan2(h[2], -(h[0] * std::cos(phi) + h[1] * std::sin(phi)));
  }
  V3 sphere_point(double theta, double phi) const {
    return {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
  }
  // Central projection of a sphere point onto the face plane face . x = 1.
  V3 flat_point(double theta, double phi) const {
    const V3 u = sphere_point(theta, phi);
    return scale3(u, 1.0 / dot3(face_, u));
  }
  // Flat area of the chamber's face between azimuths lo and phi.
  double flat_share(double phi) const {
    const V3 a = sub3(flat_point(polar_limit(lo_), lo_), pole_flat_);
    const V3 b = sub3(flat_point(polar_limit(phi), phi), pole_flat_);
    return 0.5 * norm3(cross3(a, b));
  }
  // Spherical area of the chamber between azimuths lo and phi.
  double cap_share(double phi) const {
    const V3 z{0, 0, 1};
    const V3 a = sphere_point(polar_limit(lo_), lo_), b = sphere_point(polar_limit(phi), phi);
    const double triple = std::fabs(dot3(z, cross3(a, b)));
    return 2.0 * std::atan2(triple, 1.0 + dot3(z, a) + dot3(a, b) + dot3(b, z));
  }
  // The azimuth whose spherical share is `target`: regula falsi with the Illinois rule on
  // the bracket [lo, hi] (cap_share increases with phi).
  double azimuth_of_cap_share(double target) const {
    double a = lo_, b = hi_, fa = cap_share(a) - target, fb = cap_share(b) - target;
    if (fa >= 0) return a;
    if (fb <= 0) return b;
    int side = 0;
    double last = a;
    for (int it = 0; it < 200 && b - a > 1e-15; ++it) {
      const double c = b - fb * (b - a) / (fb - fa);
      const double fc = cap_share(c) - target;
      if (fc == 0 || std::fabs(c - last) < 1e-16) return c;
      last = c;
      if ((fc < 0) == (fa < 0)) {
        a = c;
        fa = fc;
        if (side == -1) fb *= 0.5;
        side = -1;
      } else {
        b = c;
        fb = fc;
        if (side == 1) fa *= 0.5;
        side = 1;
      }
    }
    return std::fabs(fa) < std::fabs(fb) ? a : b;
  }

  double lo_ = 0, hi_ = 0;
  V3 face_{}, pole_flat_{};
  V3 walls_[3];
  double flat_total_ = 1, cap_total_ = 1;
};

// ---- meshes ------------------------------------------------------------------------------
// Both meshes give node positions, their layer parity (the spin phase) and, for an element of
// the root given by its rotation matrix m, the node at R(g)^-1 v.
struct Mesh {
  std::vector<V3> pos;
  std::vector<uint8_t> odd_layer;
};

// Subdivided polyhedron.  A node is identified by its barycentric signature: the polyhedron
// vertices with nonzero weight and their weights (a vertex, an edge or a face interior).  A
// root element permutes the vertices, so it maps signatures to signatures.
class PolyMesh : public Mesh {
 public:
  PolyMesh(uint32_t root, uint32_t f) : ico_(root == kI), vertices_(polyhedron_vertices(ico_)) {
    const uint32_t total = mesh_nodes(root, f);
    pos.resize(total);
    odd_layer.reserve(total);
    sig_.reserve(total);
    // Nodes by first occurrence over faces in table order, (i, j) weights on the face's
    // first two vertices.
    const uint32_t faces = ico_ ? 20 : 8;
    std::vector<std::array<uint32_t, 3>> where;  // face, i, j of each node
    where.reserve(total);
    for (uint32_t fc = 0; fc < faces; ++fc) {
      const uint8_t* face = ico_ ? kIcoFaces[fc] : kOctaFaces[fc];
      for (uint32_t i = 0; i <= f; ++i)
        for (uint32_t j = 0; i + j <= f; ++j) {
          const uint32_t w[3] = {i, j, f - i - j};
          const uint64_t key = signature(face, w);
          if (!index_.emplace(key, uint32_t(sig_.size())).second) continue;
          if (sig_.size() == total) fail("BaseMeshCount");
          sig_.push_back(key);
          where.push_back({fc, i, j});
          odd_layer.push_back(uint8_t((f - std::max({w[0], w[1], w[2]})) & 1u));
        }
    }
    if (sig_.size() != total) fail("BaseMeshCount");
    const ChamberMap area(ico_);
    for_blocks(total, 2048, [&](size_t a, size_t b) {
      for (size_t k = a; k < b; ++k) {
        const uint8_t* face = ico_ ? kIcoFaces[where[k][0]] : kOctaFaces[where[k][0]];
        const double w[3] = {double(where[k][1]), double(where[k][2]), double(f - where[k][1] - where[k][2])};
        V3 x{0, 0, 0};
        for (int c = 0; c < 3; ++c) x = add3(x, scale3(vertices_[face[c]], w[c] / double(f)));
        pos[k] = area(unit3(x));
      }
    });
  }
  // Prepares the vertex permutation of an element (vertex a -> the vertex at m^T v_a).
  void prepare(const std::vector<M3>& elements) {
    perm_.assign(elements.size(), {});
    for (size_t g = 0; g < elements.size(); ++g)
      for (size_t a = 0; a < vertices_.size(); ++a) {
        const V3 img = rotate_inverse(elements[g], vertices_[a]);
        int hit = -1;
        for (size_t b = 0; b < vertices_.size(); ++b)
          if (dot3(sub3(img, vertices_[b]), sub3(img, vertices_[b])) < 1e-18) hit = int(b);
        if (hit < 0) fail("BaseMeshNotInvariant");
        perm_[g][a] = uint8_t(hit);
      }
  }
  uint32_t image(uint32_t node, size_t g) const {
    uint8_t vert[3];
    uint32_t w[3];
    const int k = unpack(sig_[node], vert, w);
    uint8_t moved[3];
    for (int c = 0; c < k; ++c) moved[c] = perm_[g][vert[c]];
    const auto it = index_.find(pack(moved, w, k));
    if (it == index_.end()) fail("BaseMeshNotInvariant");
    return it->second;
  }

 private:
  // The signature: up to three (vertex 4 bits, weight 10 bits) pairs sorted by vertex, in
  // 14-bit slots, and the pair count above them.
  static uint64_t pack(const uint8_t* vert, const uint32_t* w, int k) {
    std::array<std::pair<uint8_t, uint32_t>, 3> pairs;
    for (int c = 0; c < k; ++c) pairs[c] = {vert[c], w[c]};
    std::sort(pairs.begin(), pairs.begin() + k);
    uint64_t key = uint64_t(k) << 42;
    for (int c = 0; c < k; ++c) key |= (uint64_t(pairs[c].first) << 10 | pairs[c].second) << (14 * c);
    return key;
  }
  static int unpack(uint64_t key, uint8_t* vert, uint32_t* w) {
    const int k = int(key >> 42);
    for (int c = 0; c < k; ++c) {
      const uint64_t slot = (key >> (14 * c)) & 0x3fffu;
      vert[c] = uint8_t(slot >> 10);
      w[c] = uint32_t(slot & 1023u);
    }
    return k;
  }
  static uint64_t signature(const uint8_t* face, const uint32_t* w) {
    uint8_t vert[3];
    uint32_t ww[3];
    int k = 0;
    for (int c = 0; c < 3; ++c)
      if (w[c]) {
        vert[k] = face[c];
        ww[k++] = w[c];
      }
    return pack(vert, ww, k);
  }

  bool ico_;
  std::vector<V3> vertices_;
  std::vector<uint64_t> sig_;
  std::unordered_map<uint64_t, uint32_t> index_;
  std::vector<std::array<uint8_t, 12>> perm_;
};

// Six-fold equal-area latitude rings: the poles, rings 1..f-1 north and south, the equator.
// Ring j holds m_j nodes at azimuths 2 pi (i + s_j) / m_j (s_j = 1/2 on odd rings, none on
// the equator) and sits at the middle height of its band, every node owning 4 pi / |V| of
// area: the band of ring j spans heights 1 - 2 (1 + C_{j-1}) / |V| down by 2 m_j / |V|, C
// counting the nodes of the rings before it.  A D6 element turns azimuths (about z) or
// reflects them and swaps the hemispheres (half turns about horizontal axes), so the image
// of (ring, hemisphere, i) is found from the element's action on the azimuth.
class RingMesh : public Mesh {
 public:
  explicit RingMesh(uint32_t f) : f_(f) {
    const uint32_t total = mesh_nodes(kD6, f);
    add_node(0, 0, 0, {0, 0, 1});
    add_node(0, 1, 0, {0, 0, -1});
    uint32_t before = 0;  // nodes in the northern rings so far
    for (uint32_t j = 1; j < f; ++j) {
      const uint32_t m = ring_nodes(f, j);
      const double top = 1.0 - 2.0 * (1.0 + before) / double(total);
      const double z = top - double(m) / double(total);
      const double r = std::sqrt(1.0 - z * z);
      for (uint32_t hemi = 0; hemi < 2; ++hemi)
        for (uint32_t i = 0; i < m; ++i) {
          const double phi = 2.0 * kPi * (i + 0.5 * (j & 1u)) / m;
          add_node(j, hemi, i, {r * std::cos(phi), r * std::sin(phi), hemi ? -z : z});
        }
      before += m;
    }
    const uint32_t eq = ring_nodes(f, f);
    for (uint32_t i = 0; i < eq; ++i) {
      const double phi = 2.0 * kPi * i / eq;
      add_node(f, 0, i, {std::cos(phi), std::sin(phi), 0});
    }
    if (pos.size() != total) fail("BaseMeshCount");
  }
  void prepare(const std::vector<M3>& elements) {
    // m^T maps azimuth phi to phi + angle (a turn about z, m^T_zz = 1) or to angle - phi
    // with the hemispheres swapped (a half turn about a horizontal axis, m^T_zz = -1).
    action_.clear();
    for (const M3& m : elements) {
      const bool flip = m[2][2] < 0;
      action_.push_back({flip, std::atan2(m[0][1], m[0][0])});
    }
  }
  uint32_t image(uint32_t node, size_t g) const {
    const Place& p = place_[node];
    const bool flip = action_[g].flip;
    if (p.ring == 0) return p.hemi ^ uint32_t(flip);  // the poles
    const uint32_t m = p.ring == f_ ? ring_nodes(f_, f_) : ring_nodes(f_, p.ring);
    const double shift = p.ring == f_ ? 0.0 : 0.5 * (p.ring & 1u);
    const double phi = 2.0 * kPi * (p.index + shift) / m;
    const double turned = flip ? action_[g].angle - phi : phi + action_[g].angle;
    const double slot = turned * m / (2.0 * kPi) - shift;
    const double nearest = std::round(slot);
    if (std::fabs(slot - nearest) > 1e-7) fail("BaseMeshNotInvariant");
    const uint32_t index = uint32_t((int64_t(nearest) % int64_t(m) + int64_t(m)) % int64_t(m));
    const uint32_t hemi = p.ring == f_ ? 0u : p.hemi ^ uint32_t(flip);
    return first_.at(key(p.ring, hemi)) + index;
  }

 private:
  struct Place {
    uint32_t ring, hemi, index;
  };
  struct Action {
    bool flip;
    double angle;
  };
  static uint64_t key(uint32_t ring, uint32_t hemi) { return uint64_t(ring) << 1 | hemi; }
  void add_node(uint32_t ring, uint32_t hemi, uint32_t index, const V3& p) {
    if (index == 0) first_[key(ring, hemi)] = uint32_t(pos.size());
    place_.push_back({ring, hemi, index});
    pos.push_back(p);
    odd_layer.push_back(uint8_t(ring & 1u));
  }
  uint32_t f_;
  std::vector<Place> place_;
  std::map<uint64_t, uint32_t> first_;
  std::vector<Action> action_;
};

// ---- orbits ------------------------------------------------------------------------------
// The sector at the pole z in which each orbit keeps its node: the region nearer z than any
// other vertex of the polyhedron (the northern hemisphere for the ring mesh), azimuths
// 0 .. 2 pi / m for the m-fold z axis; tolerances 1e-11 (hopf.zig:487-513).
class Sector {
 public:
  explicit Sector(uint32_t root) {
    if (root == kD6) {
      rivals_ = {{0, 0, -1}};
      wedge_ = 2 * kPi / 6;
    } else {
      rivals_ = polyhedron_vertices(root == kI);
      wedge_ = 2 * kPi / (root == kI ? 5 : 4);
    }
  }
  bool holds(const V3& w) const {
    for (const V3& p : rivals_)
      if (dot3(w, p) > w[2] + 1e-11) return false;
    if (std::hypot(w[0], w[1]) < 1e-11) return true;
    const double az = std::atan2(w[1], w[0]);
    return az >= -1e-11 && az <= wedge_ + 1e-11;
  }

 private:
  std::vector<V3> rivals_;
  double wedge_ = 0;
};
// On the sector's edges an orbit keeps the node of largest x, then largest y (compared at
// 1e-11).
bool ahead_of(const V3& a, const V3& b) {
  const double ax = std::round(a[0] * 1e11), bx = std::round(b[0] * 1e11);
  return ax != bx ? ax > bx : std::round(a[1] * 1e11) > std::round(b[1] * 1e11);
}
// The rotation about v x z taking the unit vector v onto z: q = (1 + v.z, v x z) normalised;
// at v = -z a half turn about x.
Q4 onto_pole(const V3& v) {
  const V3 axis = cross3(v, {0, 0, 1});
  const Q4 q{1.0 + v[2], axis[0], axis[1], axis[2]};
  if (std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) < 1e-12) return {0, 1, 0, 0};
  return qunit(q);
}

struct Orbit {
  Q4 lift;          // onto_pole of the representative
  uint32_t phase;   // spin phase: the node's layer parity
  uint32_t turns;   // fibre rotations, n / |stabiliser|
  size_t first_row;
};

template <class MeshT>
std::vector<Orbit> walk_orbits(const MeshT& mesh, const std::vector<M3>& elements, uint32_t root, uint32_t n,
                               size_t cosets, size_t total) {
  const Sector sector(root);
  std::vector<uint8_t> seen(mesh.pos.size(), 0);
  std::vector<Orbit> orbits;
  size_t rows = 0;
  for (uint32_t i = 0; i < mesh.pos.size(); ++i) {
    if (seen[i]) continue;
    const V3& v = mesh.pos[i];
    uint32_t stabiliser = 0;
    bool found = false;
    V3 rep{};
    for (size_t g = 0; g < elements.size(); ++g) {
      const uint32_t j = mesh.image(i, g);
      const V3 img = rotate_inverse(elements[g], v);
      const V3 gap = sub3(mesh.pos[j], img);
      if (dot3(gap, gap) > 1e-18) fail("BaseMeshNotInvariant");
      seen[j] = 1;
      stabiliser += j == i;
      if (sector.holds(img) && (!found || ahead_of(img, rep))) {
        rep = img;
        found = true;
      }
    }
    if (!found) fail("OrbitOutsideSector");
    if (n % stabiliser) fail("FibreNotDivisible");
    const uint32_t turns = n / stabiliser;
    orbits.push_back({onto_pole(rep), mesh.odd_layer[i], turns, rows});
    rows += size_t(turns) * cosets;
    if (rows > total) fail("ProductCount");
  }
  if (rows != total) fail("ProductCount");
  return orbits;
}

}  // namespace

// ---- group elements ----------------------------------------------------------------------
// Generators (axis, fold) of each group in the frames of the operator tables; the closure
// is breadth first from the identity, each element's sign fixed to a nonnegative scalar
// part, then a stable sort by decreasing scalar part (hopf.zig:161-213 orders them so).
std::vector<std::array<double, 4>> elements(uint32_t group) {
  struct Gen {
    V3 axis;
    double fold;
  };
  const V3 z{0, 0, 1}, x{1, 0, 0}, d{1, 1, 1};
  std::vector<Gen> gens;
  switch (group) {
    case kC2: gens = {{z, 2}}; break;
    case kC3: gens = {{z, 3}}; break;
    case kC4: gens = {{z, 4}}; break;
    case kC6: gens = {{z, 6}}; break;
    case kD2: gens = {{z, 2}, {x, 2}}; break;
    case kD3: gens = {{z, 3}, {x, 2}}; break;
    case kD4: gens = {{z, 4}, {x, 2}}; break;
    case kD6: gens = {{z, 6}, {x, 2}}; break;
    case kT: gens = {{z, 2}, {x, 2}, {d, 3}}; break;
    case kO: gens = {{z, 4}, {d, 3}}; break;
    case kI: gens = {{z, 5}, {{2, 0, 1}, 5}}; break;
    default: break;  // C1: the identity alone
  }
  std::vector<Q4> turns;
  for (const Gen& g : gens) turns.push_back(axis_turn(g.axis, g.fold));
  std::vector<Q4> out{{1, 0, 0, 0}};
  for (size_t head = 0; head < out.size(); ++head)
    for (const Q4& t : turns) {
      const Q4 p = qmul(out[head], t);
      bool known = false;
      for (const Q4& k : out) known = known || std::fabs(qdot(k, p)) > 1 - 1e-9;
      if (!known) out.push_back(p);
    }
  if (group > kC1 || out.size() != kSize[group]) fail("GroupClosure");
  for (Q4& q : out) {
    q = qunit(q);
    if (q[0] < 0) q = {-q[0], -q[1], -q[2], -q[3]};
  }
  std::stable_sort(out.begin(), out.end(), [](const Q4& a, const Q4& b) { return a[0] > b[0]; });
  return out;
}

uint32_t root_of(uint32_t group) { return root_group(group); }

uint64_t raw_count(uint32_t group, uint32_t f) {
  const uint32_t root = root_group(group);
  return uint64_t(fibre_turns(root, f)) * mesh_nodes(root, f) / kSize[group];
}

uint32_t count(uint32_t group, uint32_t f) {
  if (f < 1 || f > kMaxFrequency) fail("InvalidGridResolution");
  const uint64_t n = raw_count(group, f);
  if (n > kMaxCount) fail("GridTooLarge");
  return uint32_t(n);
}

uint32_t nearest_frequency(uint32_t group, uint32_t target) {
  const double want = double(std::max<uint32_t>(target, 1));
  uint32_t best = 1;
  double best_gap = INFINITY;
  for (uint32_t f = 1; f <= kMaxFrequency; ++f) {
    const uint64_t n = raw_count(group, f);
    if (n > kMaxCount) continue;
    const double gap = std::fabs(std::log(double(n) / want));
    if (gap < best_gap) {
      best_gap = gap;
      best = f;
    }
  }
  return best;
}

const char* description(uint32_t group) {
  switch (root_group(group)) {
    case kI: return "Inverse Hopf product grid (icosahedral root)";
    case kO: return "Inverse Hopf product grid (octahedral root)";
    default: return "Inverse Hopf product grid (hexagonal root)";
  }
}

void points(uint32_t group, uint32_t f, float* out) {
  if (group > kC1) fail("UnknownGroup");
  const uint32_t total = count(group, f);
  const uint32_t root = root_group(group);
  const std::vector<Q4> parent = elements(root), sub = elements(group);
  std::vector<M3> mats;
  for (const Q4& q : parent) mats.push_back(matrix_of(q));

  // Cosets of H' = R H R^-1 in the root, each named by its first element in root order.
  const Q4 r = root_embedding(group), r_inv{r[0], -r[1], -r[2], -r[3]};
  std::vector<Q4> conj_sub;
  for (const Q4& h : sub) conj_sub.push_back(qmul(qmul(r, h), r_inv));
  std::vector<int> coset_of(parent.size(), -1);
  std::vector<Q4> tails;  // c R per coset
  for (size_t i = 0; i < parent.size(); ++i) {
    if (coset_of[i] >= 0) continue;
    const int id = int(tails.size());
    tails.push_back(qmul(parent[i], r));
    for (const Q4& h : conj_sub) {
      const Q4 member = qmul(parent[i], h);
      for (size_t k = 0; k < parent.size(); ++k)
        if (std::fabs(qdot(parent[k], member)) > 1 - 1e-9) coset_of[k] = id;
    }
  }
  if (tails.size() * sub.size() != parent.size()) fail("GroupNotInRoot");

  const uint32_t n = fibre_turns(root, f);
  std::vector<Orbit> orbits;
  if (root == kD6) {
    RingMesh mesh(f);
    mesh.prepare(mats);
    orbits = walk_orbits(mesh, mats, root, n, tails.size(), total);
  } else {
    PolyMesh mesh(root, f);
    mesh.prepare(mats);
    orbits = walk_orbits(mesh, mats, root, n, tails.size(), total);
  }

  // Fibre turns R_z(t), t = (2 pi j + pi phase) / n, as half-angle quaternions.
  std::vector<Q4> spin(2 * size_t(n));
  for (uint32_t phase = 0; phase < 2; ++phase)
    for (uint32_t j = 0; j < n; ++j) {
      const double half = 0.5 * (2.0 * kPi * j + kPi * phase) / n;
      spin[phase * n + j] = {std::cos(half), 0, 0, std::sin(half)};
    }
  // Rows: per orbit in node order, per fibre turn, per coset: R_z(t) (s (c R)).
  for_blocks(orbits.size(), 256, [&](size_t a, size_t b) {
    std::vector<Q4> heads(tails.size());
    for (size_t o = a; o < b; ++o) {
      const Orbit& orb = orbits[o];
      for (size_t k = 0; k < tails.size(); ++k) heads[k] = qmul(orb.lift, tails[k]);
      float* row = out + 4 * orb.first_row;
      for (uint32_t j = 0; j < orb.turns; ++j) {
        const Q4& t = spin[size_t(orb.phase) * n + j];
        for (const Q4& h : heads) {
          const Q4 q = qmul(t, h);
          for (int c = 0; c < 4; ++c) *row++ = float(q[c]);
        }
      }
    }
  });
}

}  // namespace zxt::hopf
