// Host engine: sample state, relaxation, measure and spacing drivers.
//
// Follows the reference's control flow (src/engine.zig:233-376, :886-1020;
// src/pairs.zig:220-413; src/near.zig:90-196) over HIP: every step records kernels on
// one stream, and the host waits only where the reference does (end of a relax call,
// measure, spacing, the near field's one-time sizing pass).
#include "engine.h"

#include <shared_mutex>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <unordered_map>

#include "groups_table.h"
#include "hopf.h"
#include "qvm_device.h"
#include "transport.h"

namespace zxt {

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr uint32_t kMaxCount = 10000000;      // hopf.max_count (src/hopf.zig:23)
constexpr double kFullMeasurePairs = 5e10;    // engine.zig:1123
constexpr uint32_t kSpacingAllPairs = 20000;  // engine.zig:1126
constexpr double kSpacingDelta = 4;           // engine.zig:1127
constexpr double kSpacingCellSites = 0.75;    // sites per cell of the spacing pass's grid
// and its shape: y/z cells 1.75 times that edge, 4 x-cells each (fewer rows per cap, trimmed
// closer; about as many cells).  Measured on MI355X, I, N = 1e7, whole pass, same box:
// cubic 3.47 ms; 1.75 x 4 3.32; 2.0 x 8 3.34; 1.5 x 4 3.37; 2.0 x 16 3.53; 2.5 x 16 3.62;
// 1.0 x 4 3.65 (more cells: the search is faster, the scans and source order slower).
constexpr uint32_t kSpacingXSplit = 4;
constexpr double kSpacingEdge = 1.75;
// Sites of one near field: the force kernel addresses them by 32-bit byte offsets
// (16 B each, 64 entries of slack past the capacity).
constexpr uint64_t kMaxSites = (uint64_t(1) << 28) - 64;
// Cell shape of the force passes' grid: y/z cells kForceEdge times the edge of a cube
// holding the chosen sites per cell, split into kForceXSplit x-cells each.  A cutoff ball
// then spans fewer rows, trimmed to finer x-cells: fewer, fuller 64-site chunks per
// accepted pair.  Measured on MI355X at I, N = 1e6, reach 20 (force kernel, same box):
// cubic cells 14.53 ms; edge 1.35 with 16 x-cells 13.32-13.40; 1.5/16 13.46; 1.2/16 13.53;
// 1.0/16 13.78; 2.0/8 14.14; 2.5/16 14.60; 32 x-cells no better than 16.
constexpr uint32_t kForceXSplit = 16;
constexpr double kForceEdge = 1.35;
const char* kVersion = "0.3.0-mi355x";

[[noreturn]] void fail(const std::string& what) { throw EngineError{what}; }

double nearest_degrees(double chord2) { return 4.0 * std::asin(std::min(1.0, std::sqrt(chord2) / 2.0)) * 180.0 / kPi; }
}  // namespace

double mean_radius(uint64_t count) {
  const double target = kPi * kPi / double(count);
  double lo = 0.0, hi = kPi;
  for (int k = 0; k < 60; ++k) {
    const double m = 0.5 * (lo + hi);
    if (kPi * (m - std::sin(m)) < target)
      lo = m;
    else
      hi = m;
  }
  return 0.25 * (lo + hi);
}

// ---- CaptureLock -----------------------------------------------------------------
namespace {
std::shared_mutex capture_mutex;
thread_local int capture_holds = 0;
}  // namespace

CaptureLock::CaptureLock(bool capture) : mode_(0) {
  if (capture_holds++ > 0) return;
  if (capture) {
    capture_mutex.lock_shared();
    mode_ = 1;
  } else {
    capture_mutex.lock();
    mode_ = 2;
  }
}
CaptureLock::~CaptureLock() {
  --capture_holds;
  if (mode_ == 1) capture_mutex.unlock_shared();
  if (mode_ == 2) capture_mutex.unlock();
}

// ---- DevBuf ----------------------------------------------------------------------
void DevBuf::release() {
  if (p) {
    CaptureLock lock(false);
    (void)hipFree(p);
  }
  p = nullptr;
  bytes = 0;
}
void DevBuf::ensure(size_t b) {
  if (p && b <= bytes) return;
  CaptureLock lock(false);
  release();
  const size_t want = std::max<size_t>(b, 256);
  if (hipMalloc(&p, want) != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    fail("OutOfDeviceMemory");
  }
  bytes = want;
}

void Engine::check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(std::string("GpuError: ") + what + ": " + hipGetErrorString(e));
}

// ---- NearField -------------------------------------------------------------------
NearBuffers NearField::view() const {
  NearBuffers b;
  b.planes = planes.as<float4>();
  b.nplanes = nplanes;
  b.band = band;
  b.reach = float(radius + 1e-4);
  b.capacity = capacity;
  b.sdim = sdim;
  b.xdiv = xdiv;
  b.info = info.as<NearInfo>();
  b.rim = rim.as<uint32_t>();
  b.uq = uq.as<float4>();
  b.ucode = ucode.as<uint32_t>();
  b.urank = urank.as<uint32_t>();
  b.ucell = ucell.as<uint32_t>();
  b.perm = perm.as<uint32_t>();
  b.sq = sq.as<float4>();
  b.scode = scode.as<uint32_t>();
  b.srank = srank.as<uint32_t>();
  b.offsets = offsets.as<uint32_t>();
  b.uslot = uslot.as<uint32_t>();
  b.order = order_buf.as<uint32_t>();
  b.opos = uq.as<float4>();  // the unsorted sites are spent once the build has sorted them
  b.scan_tmp = scan_tmp.as<uint32_t>();
  b.counters = counters.as<unsigned long long>();
  b.counts = counts.as<uint32_t>();
  b.crowded = crowded.as<uint32_t>();
  b.crowded_capacity = crowded.p ? crowded_capacity(capacity) : 0u;
  b.huge = huge.as<uint32_t>();
  b.huge_capacity = huge.p ? huge_capacity(capacity) : 0u;
  return b;
}

static void release_near(NearField& f) {
  for (DevBuf* b : {&f.planes, &f.info, &f.rim, &f.uq, &f.ucode, &f.urank, &f.ucell, &f.perm, &f.sq, &f.scode,
                    &f.srank, &f.offsets, &f.uslot, &f.order_buf,
                    &f.scan_tmp, &f.counters, &f.counts, &f.crowded, &f.huge})
    b->release();
  f.ready = false;
}

// Sizes the buffers from one sites pass (one wait), as near.zig:90-152 does.
void NearField::setup(Engine& e, const float4* q, uint32_t n_, uint32_t order_, uint32_t group_,
                      double mean_radius_) {
  release_near(*this);
  n = n_;
  order = order_;
  group = group_;
  radius = std::min(1.5, delta * mean_radius_);
  const double chord = 2.0 * std::sin(radius / 2.0);
  cut.chord2 = float(chord * chord);
  cut.taper_start = float(taper * radius);
  cut.inv_width = float(1.0 / std::max(1e-9, (1.0 - taper) * radius));
  // Every site the FP32 test chord^2 <= chord2 can accept lies within this cap.
  const double margin = double(cut.chord2) * (1.0 + 1e-5) + 1e-6;
  const double c0 = 1.0 - margin / 2.0;
  cut.c0 = float(c0);
  cut.s0 = float(std::sqrt(std::max(0.0, 1.0 - c0 * c0)));
  const double inner_chord = 2.0 * std::sin(0.5 * radius / 2.0);
  cut.inner_chord2 = float(inner_chord * inner_chord);
  const double inner_c0 = 1.0 - (double(cut.inner_chord2) * (1.0 + 1e-5) + 1e-6) / 2.0;
  cut.inner_c0 = float(inner_c0);
  cut.inner_s0 = float(std::sqrt(std::max(0.0, 1.0 - inner_c0 * inner_c0)));
  band = float(std::sin(radius + 1e-4));
  // x.g lies at S3 angle acos|g.w| from x (and -x.g at pi minus that) for every x, so the
  // source's own non-identity images enter the cutoff only when r reaches the smallest.
  double own_angle = kPi;
  for (uint32_t i = 1; i < order; ++i)
    own_angle = std::min(own_angle, std::acos(std::min(1.0, std::fabs(double(kOperators[kGroupOffset[group] + i][0])))));
  cut.own = radius + 1e-3 >= own_angle ? 1u : 0u;
  // Zone planes: the bisectors of the identity and every other signed image
  // (covering.zig:382-397), normalised in FP64.
  std::vector<float4> pl;
  const uint32_t off = kGroupOffset[group];
  for (uint32_t i = 0; i < order; ++i) {
    for (double s : {1.0, -1.0}) {
      if (i == 0 && s == 1.0) continue;
      const float* op = kOperators[off + i];
      double v[4] = {1.0 - s * op[0], -s * op[1], -s * op[2], -s * op[3]};
      const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
      pl.push_back(make_float4(float(v[0] / len), float(v[1] / len), float(v[2] / len), float(v[3] / len)));
    }
  }
  // Nearest planes first (by their distance from the identity, the normal's scalar part):
  // the rim and shell tests then stop as soon as the remaining planes are provably out of
  // reach (k_core_rim, shell_planes).  The tests are "any" and "all", so order only
  // changes when they stop.
  std::stable_sort(pl.begin(), pl.end(), [](const float4& a, const float4& b) { return a.x < b.x; });
  nplanes = uint32_t(pl.size());
  planes.ensure(pl.size() * sizeof(float4));
  Engine::check(hipMemcpyAsync(planes.p, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice, e.stream),
                "planes");
  info.ensure(sizeof(NearInfo));
  rim.ensure(size_t(n) * 4);
  uint64_t cap = std::max<uint64_t>(uint64_t(n) + 1024, uint64_t(n) * 4);
  bool sized = false;
  for (int attempt = 0; attempt < 5; ++attempt) {
    capacity = uint32_t(std::min<uint64_t>(cap, kMaxSites));
    uq.ensure(size_t(capacity) * 16);
    ucode.ensure(size_t(capacity) * 4);
    urank.ensure(size_t(capacity) * 4);
    sdim = 1;
    launch_near_sites(e.prof, e.stream, q, n, e.ops_device(), order, view());
    Engine::check(hipMemcpyAsync(&status, info.p, sizeof(NearInfo), hipMemcpyDeviceToHost, e.stream), "info");
    e.sync();
    if (!status.overflow) {
      sized = true;
      break;
    }
    cap = std::max<uint64_t>(cap * 2, uint64_t(status.sites) + status.sites / 4);
  }
  if (!sized) fail("NearShellTooLarge");
  const uint64_t count = status.sites;
  if (e.near_squeeze_pending) {
    // Core sites are written at their source's index, so the sources always fit.
    e.near_squeeze_pending = false;
    const int64_t want = std::max<int64_t>(int64_t(count) + e.near_squeeze, int64_t(n));
    capacity = uint32_t(std::min<int64_t>(want, int64_t(capacity)));
  } else {
    // Room for the shell to change as the sources move.
    capacity = uint32_t(std::min<uint64_t>(std::max<uint64_t>(capacity, count + count / 4 + 1024), kMaxSites));
  }
  // Sites per cell of the box.  A ball of `reach` mean radii holds about reach^3 sites:
  // thousands at reach 20, where coarse cells (8 sites) win by giving fewer, fuller chunks
  // of 64 sites per row at the cost of more candidates.  Measured on MI355X at I, N = 1e6
  // with cubic cells: the force pass at reach 20 takes 15.5 ms with 8 sites per cell
  // against 17.3 ms with 2, at reach 8 2.2 ms with 4 against 2.6 ms with 2.  Force passes
  // then reshape the cells (kForceEdge, kForceXSplit).  The spacing pass sets its own
  // (cell_sites): its first search covers half its reach, a few sites, best found in
  // fine cells (I, N = 1e7, the whole pass: 4.24 ms with 0.75 sites per cell, 4.33 with
  // 0.5, 4.40 with 1.0, 4.53 with 1.6, 4.83 with 3).
  const double reach = radius / mean_radius_;
  const double per_cell = cell_sites > 0 ? cell_sites : std::min(8.0, std::max(1.5, reach / 2.5));
  xdiv = x_split;
  const double max_dim = std::floor(320.0 / std::cbrt(double(xdiv)));  // at most 320^3 cells
  sdim = uint32_t(std::min(max_dim, std::max(1.0, std::round(std::cbrt(double(count) / per_cell) / cell_edge))));
  const size_t cells = cell_count(sdim, xdiv);
  uq.ensure(size_t(capacity) * 16);
  ucode.ensure(size_t(capacity) * 4);
  urank.ensure(size_t(capacity) * 4);
  ucell.ensure((std::max<size_t>(capacity, tile_visits(sdim, xdiv)) + 1) * 4);  // then the visits' scan
  perm.ensure(size_t(capacity) * 4);
  // The force kernel loads whole 64-site chunks from a row's start: 64 entries of slack.
  sq.ensure((size_t(capacity) + 64) * 16);
  scode.ensure((size_t(capacity) + 64) * 4);
  srank.ensure(size_t(capacity) * 4);
  offsets.ensure((cells + 1) * 4);
  uslot.ensure(size_t(capacity) * 4);
  order_buf.ensure(size_t(n) * 4);
  scan_tmp.ensure(scan_tmp_words(std::max<size_t>(cells, tile_visits(sdim, xdiv))) * 4);
  counters.ensure(4 * sizeof(unsigned long long));
  counts.ensure(size_t(n) * 8);
  crowded.ensure(size_t(crowded_capacity(capacity)) * 4);
  huge.ensure(size_t(huge_capacity(capacity)) * 4);
  ready = true;
}

void NearField::ready_for(Engine& e, const float4* q, uint32_t n_, uint32_t order_, uint32_t group_,
                          double mean_radius_) {
  if (!ready || n != n_ || order != order_ || group != group_) setup(e, q, n_, order_, group_, mean_radius_);
}

void NearField::build(Engine& e, const float4* q, uint32_t n_, uint32_t order_, uint32_t group_,
                      double mean_radius_, bool ordered) {
  ready_for(e, q, n_, order_, group_, mean_radius_);
  launch_near_build(e.prof, e.stream, q, n, e.ops_device(), order, view(), ordered);
}

bool NearField::overflowed(Engine& e) {
  read_status(e);
  e.sync();
  return took_overflow(e);
}

void NearField::read_status(Engine& e) {
  Engine::check(hipMemcpyAsync(&status, info.p, sizeof(NearInfo), hipMemcpyDeviceToHost, e.stream), "info");
  Engine::check(hipMemcpyAsync(e.last_near_counts, counters.p, sizeof(e.last_near_counts), hipMemcpyDeviceToHost,
                               e.stream),
                "counters");
}

bool NearField::took_overflow(Engine& e) {
  e.last_near = status;
  e.last_near_sdim = sdim;
  e.last_near_cut = cut;
  e.last_near_radius = radius;
  if (!status.overflow) return false;
  ready = false;
  ++e.near_resizes;
  return true;
}

// ---- RelaxState --------------------------------------------------------------------
void RelaxState::release() {
  for (DevBuf* b : {&force, &energy, &nearest, &prev, &state, &partials, &ap_partial, &ap_pnear, &nn_partial,
                    &nn_chord, &nn_rank, &nn_degrees})
    b->release();
  release_near(near);
  n = 0;
  current = false;
}

void RelaxState::set_near(uint32_t group, const double* delta) {
  if (has_near) {
    if (delta && near.group == group && near.delta == *delta) return;
    release_near(near);
    has_near = false;
    current = false;
  }
  if (delta) {
    has_near = true;
    near.delta = *delta;
    near.group = group;
    near.ready = false;
    current = false;
  }
}

void RelaxState::prepare(Engine& e, uint32_t n_) {
  if (n == n_ && state.p) return;
  const bool keep_near = has_near;
  const double delta = near.delta;
  const uint32_t group = near.group;
  release();
  has_near = keep_near;
  near.delta = delta;
  near.group = group;
  n = n_;
  prev.ensure(size_t(n) * 32);
  launch_fill_u32(e.prof, e.stream, prev.as<uint32_t>(), 0u, size_t(n) * 8);
  state.ensure(sizeof(BBState));
  launch_bb_reset(e.prof, e.stream, state.as<BBState>());
  partials.ensure(size_t(bb_partials(n)) * 5 * sizeof(double));
}

void RelaxState::ensure_rows(uint32_t n_) {
  force.ensure(size_t(n_) * 16);
  energy.ensure(size_t(n_) * 8);
  nearest.ensure(size_t(n_) * 4);
}

void RelaxState::record_forces(Engine& e, uint32_t n_, uint32_t order_) {
  ensure_rows(n_);
  float4* q = static_cast<float4*>(e.q_device());
  if (has_near) {
    near.build(e, q, n_, order_, e.group(), mean_radius(uint64_t(n_) * order_));
    launch_near_force(e.prof, e.stream, q, n_, near.view(), near.cut, rows());
  } else {
    launch_allpairs(e.prof, e.stream, single_job(allpairs_desc(e, n_, order_)));
  }
}

JobDesc RelaxState::allpairs_desc(Engine& e, uint32_t n_, uint32_t order_) {
  ensure_rows(n_);
  JobDesc d = describe_job(static_cast<float4*>(e.q_device()), e.ops_device(), n_, order_, e.group());
  d.plan = plan_allpairs(n_, order_, e.pair_budget_scale, allpairs_shape(n_, order_));
  ap_partial.ensure(size_t(d.plan.splits()) * n_ * 5 * sizeof(double));
  ap_pnear.ensure(size_t(d.plan.splits()) * n_ * sizeof(float));
  d.ap_partial = ap_partial.as<double>();
  d.ap_pnear = ap_pnear.as<float>();
  d.rows = rows();
  return d;
}

JobDesc RelaxState::step_desc(Engine& e, double radius, uint32_t fixed) {
  JobDesc d = describe_job(static_cast<float4*>(e.q_device()), e.ops_device(), n, e.order(), e.group());
  d.rows = rows();
  d.prev = prev.as<float4>();
  d.state = state.as<BBState>();
  d.bb_partials = partials.as<double>();
  d.radius = radius;
  d.step_scale = 1.0f;
  d.fixed = fixed;
  return d;
}

// ---- Engine ------------------------------------------------------------------------
Engine::Engine(hipStream_t shared) {
  CaptureLock lock(false);
  if (shared) {
    stream = shared;
  } else {
    check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
    owns_stream_ = true;
  }
  prof.set_stream(stream);
  ops_.ensure(sizeof(kOperators));
  check(hipMemcpy(ops_.p, kOperators, sizeof(kOperators), hipMemcpyHostToDevice), "operators");
  set_group(9);
}

Engine::~Engine() {
  (void)hipStreamSynchronize(stream);
  relax_state_.release();
  spacing_state_.release();
  measure_state_.release();
  q_.release();
  ops_.release();
  if (owns_stream_) {
    CaptureLock lock(false);
    (void)hipStreamDestroy(stream);
  }
}

void Engine::sync() { check(hipStreamSynchronize(stream), "wait"); }

void* Engine::q_device() const { return q_.p; }
uint32_t Engine::order() const { return kGroupOrder[group_]; }
const float* Engine::operations() const { return kOperators[kGroupOffset[group_]]; }
const float4* Engine::ops_device() const { return ops_.as<float4>() + kGroupOffset[group_]; }
uint32_t Engine::ratio_limit() const { return uint32_t(std::floor(std::sqrt(kFullMeasurePairs / order()))); }

void Engine::set_group(uint32_t group) {
  if (group >= kGroupCount) fail("UnknownGroup");
  group_ = group;
}

void Engine::fold(uint32_t n) { launch_fold(prof, stream, q_.as<float4>(), n, ops_device(), order(), group_); }

void Engine::fetch() {
  host_.resize(size_t(n_) * 4);
  check(hipMemcpyAsync(host_.data(), q_.p, size_t(n_) * 16, hipMemcpyDeviceToHost, stream), "fetch");
  sync();
}

void Engine::reset_run() {
  point_separation_iteration_ = -1;
  point_index_iteration_ = -1;
  point_covering_iteration_ = -1;
  cover_ = CoverResult{};
  opt_ = Optimizer{};
  relax_state_.release();
  spacing_state_.release();
  measure_state_.release();
  iteration_ = 0;
  step_deg_ = 0;
  force_rms_ = 0;
  measured_ = false;
  full_measures_ = false;
  cancelled.store(false);
  err.clear();
}

void Engine::adopt(uint32_t n) {
  n_ = n;
  fetch();
  fixed_index_ = fix_identity_ ? identity_index() : -1;
  reset_run();
}

long Engine::identity_index() const {
  for (uint32_t i = 0; i < n_; ++i) {
    const float* a = &host_[size_t(i) * 4];
    if (std::fabs(a[1]) + std::fabs(a[2]) + std::fabs(a[3]) < 1e-6f) return long(i);
  }
  return -1;
}

void Engine::set(uint32_t group, const float* values, size_t count4) {
  if (count4 == 0 || count4 % 4 != 0 || count4 > size_t(4) * kMaxCount) fail("InvalidCount");
  for (size_t i = 0; i < count4 / 4; ++i) {
    double norm = 0.0;
    for (int c = 0; c < 4; ++c) {
      const float x = values[4 * i + c];
      if (!std::isfinite(x)) fail("NonFiniteQuaternion");
      norm += double(x) * x;
    }
    if (norm < 1e-20) fail("ZeroQuaternion");
  }
  const uint32_t old = group_;
  set_group(group);
  try {
    const uint32_t n = uint32_t(count4 / 4);
    q_.ensure(size_t(n) * 16);
    check(hipMemcpyAsync(q_.p, values, size_t(n) * 16, hipMemcpyHostToDevice, stream), "upload");
    fold(n);
    adopt(n);
  } catch (...) {
    group_ = old;
    throw;
  }
  generator_ = "Imported quaternions";
  inserted_count_ = 0;
  grid_kind_.clear();
  h_ = 0;
  seed_ = 0;
  base_count_ = n_;
}

void Engine::generate_random(uint32_t group, uint32_t count, uint32_t seed) {
  if (count < 1 || count > kMaxCount) fail("InvalidCount");
  const uint32_t old = group_;
  set_group(group);
  try {
    q_.ensure(size_t(count) * 16);
    launch_haar(prof, stream, q_.as<float4>(), count, seed);
    fold(count);
    adopt(count);
  } catch (...) {
    group_ = old;
    throw;
  }
  grid_kind_.clear();
  h_ = 0;
  inserted_count_ = 0;
  generator_ = "Haar random";
  seed_ = seed;
  base_count_ = count;
}

void Engine::generate(uint32_t group, uint32_t h, uint32_t target, uint32_t seed, uint32_t kind) {
  if (target > hopf::kMaxCount) fail("GridTooLarge");
  if (kind == 2) {
    generate_transport(group, std::max<uint32_t>(1, h), seed);
    return;
  }
  const uint32_t f = h == 0 ? hopf::nearest_frequency(group, target == 0 ? 2048 : target) : h;
  const uint32_t n = hopf::count(group, f);
  // Written by the generator's threads, so no zero fill first.
  const std::unique_ptr<float[]> rows(new float[size_t(n) * 4]);
  hopf::points(group, f, rows.get());
  const uint32_t old = group_;
  set_group(group);
  try {
    q_.ensure(size_t(n) * 16);
    check(hipMemcpyAsync(q_.p, rows.get(), size_t(n) * 16, hipMemcpyHostToDevice, stream), "upload");
    // Grid points are orbit images: the fold picks each one's zone representative.
    fold(n);
    adopt(n);
  } catch (...) {
    group_ = old;
    throw;
  }
  generator_ = hopf::description(group);
  inserted_count_ = 0;
  grid_kind_ = "hopf";
  h_ = f;
  seed_ = seed;
  base_count_ = n;
}

// FCC transport of T, O or I at semi-edge h (engine.zig:247-282 with grids.zig:40-77):
// lattice order, not folded (the map lands in the zone).
void Engine::generate_transport(uint32_t group, uint32_t h, uint32_t seed) {
  const transport::Plan plan = transport::plan(group, h);
  const uint32_t n = plan.job.count;
  DevBuf tables, offsets;
  tables.ensure(sizeof(fcc::Tables));
  offsets.ensure(plan.offsets.size() * sizeof(uint32_t));
  check(hipMemcpyAsync(tables.p, &transport::tables(), sizeof(fcc::Tables), hipMemcpyHostToDevice, stream), "upload");
  check(hipMemcpyAsync(offsets.p, plan.offsets.data(), plan.offsets.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                       stream),
        "upload");
  const uint32_t old = group_;
  set_group(group);
  try {
    q_.ensure(size_t(n) * 16);
    launch_transport(prof, stream, tables.as<fcc::Tables>(), plan.job, offsets.as<uint32_t>(), q_.as<float4>());
    adopt(n);
  } catch (...) {
    group_ = old;
    throw;
  }
  generator_ = transport::description(group);
  inserted_count_ = 0;
  grid_kind_ = "transport";
  h_ = h;
  seed_ = seed;
  base_count_ = n;
}

void Engine::set_neighbourhood(double delta) {
  if (!(delta == 0 || (delta >= 2 && delta <= 1000))) fail("NeighbourhoodOutside2To1000");
  neighbourhood_ = delta;
}

void Engine::sync_neighbourhood() {
  relax_state_.set_near(group_, neighbourhood_ > 0 ? &neighbourhood_ : nullptr);
  relax_state_.near.x_split = kForceXSplit;
  relax_state_.near.cell_edge = kForceEdge;
  measure_neighbourhood_ = neighbourhood_;
}

PinnedVec<float>& Engine::separation_out() {
  point_separation_.resize(n_);
  point_separation_iteration_ = long(iteration_);
  return point_separation_;
}

Measures Engine::quick_measures(const PinnedVec<float4>& f, const PinnedVec<double>& energy,
                                const PinnedVec<float>& nearest) const {
  const uint32_t n = n_;
  double nearv = 4.0, e = 0.0, sq = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    nearv = std::min(nearv, double(nearest[i]));
    e += energy[i];
    sq += double(f[i].x) * f[i].x + double(f[i].y) * f[i].y + double(f[i].z) * f[i].z;
  }
  if (!(nearv > 0)) fail("DuplicateOrientations");
  Measures m;
  const double nd = n;
  m.count_extended = 2.0 * order() * nd;
  m.quey_energy = e / (nd * order() * std::max(nd - 1.0, 1.0));
  m.force_rms = std::sqrt(sq / nd);
  m.separation_deg = nearest_degrees(nearv);
  return m;
}

// One force pass of the current positions, read back (retried when the near field's
// shell outgrew its buffers); the forces are kept for the next step.
void Engine::read_forces(RelaxState& rs) {
  for (int attempt = 0; attempt < 4; ++attempt) {
    rs.prepare(*this, n_);
    rs.current = false;
    rs.record_forces(*this, n_, order());
    download(rows_force_, rs.force, n_);
    download(rows_energy_, rs.energy, n_);
    download(rows_nearest_, rs.nearest, n_);
    sync();
    if (rs.has_near && rs.near.overflowed(*this)) continue;
    rs.current = true;
    return;
  }
  fail("NearShellTooLarge");
}

// Measures of the current positions (pairs.zig:302-316).
Measures Engine::relax_measure(RelaxState& rs, bool keep_separation) {
  read_forces(rs);
  Measures m = quick_measures(rows_force_, rows_energy_, rows_nearest_);
  if (keep_separation) {
    auto& out = separation_out();
    for (uint32_t i = 0; i < n_; ++i) out[i] = float(nearest_degrees(rows_nearest_[i]));
  }
  return m;
}

void Engine::force_rows(float* force4, double* energy, float* nearest) {
  if (n_ < 2) fail("TooFewOrientations");
  sync_neighbourhood();
  read_forces(relax_state_);
  std::memcpy(force4, rows_force_.data(), size_t(n_) * 16);
  std::memcpy(energy, rows_energy_.data(), size_t(n_) * 8);
  std::memcpy(nearest, rows_nearest_.data(), size_t(n_) * 4);
}

// Riesz, Green and log ratios from the FP64 row sums, normalised as pairs.zig:142-173.
void Engine::full_measure(Measures& m) {
  const uint32_t ord = order();
  const AllPairsPlan plan = plan_allpairs(n_, ord, pair_budget_scale);
  measure_partial_.ensure(size_t(plan.splits()) * n_ * 5 * sizeof(double));
  measure_out_.ensure(size_t(n_) * 5 * sizeof(double));
  launch_measure(prof, stream, q_.as<float4>(), n_, ops_device(), ord, plan, measure_partial_.as<double>(),
                 measure_out_.as<double>());
  std::vector<double> rows;
  download(rows, measure_out_, size_t(n_) * 5);
  sync();
  // The omitted identity self pair still has a real antipode (pairs.wgsl:231-236).
  const double self[5] = {0.5, 0.25, 0.125, -0.4431471805599453, -3.0 / (8.0 * kPi * kPi)};
  double sum[5] = {0, 0, 0, 0, 0};
  for (uint32_t i = 0; i < n_; ++i)
    for (int c = 0; c < 5; ++c) sum[c] += rows[5 * size_t(i) + c] + self[c];
  const double extension = 2.0 * ord;
  const double total = double(n_) * extension;
  for (double& v : sum) v *= extension;
  const double c_green =
      std::pow(3.0, 5.0 / 3.0) / (5.0 * std::pow(2.0 * kPi * kPi, 1.0 / 3.0) * std::pow(4.0 * kPi, 2.0 / 3.0));
  const double delta_log = (sum[3] + 0.25 * total + std::log(total) * total / 3.0) / total;
  m.riesz_1 = sum[0] / (total * total * (8.0 / (3.0 * kPi)));
  m.riesz_2 = sum[1] / (total * total);
  m.riesz_3 = sum[2] / ((2.0 / (3.0 * kPi)) * total * total * std::log(total));
  m.green_ratio = sum[4] / (-c_green * std::pow(total, 4.0 / 3.0));
  m.log_ratio = delta_log / (-5.0 / 12.0);
}

void Engine::measure() {
  if (n_ == 0) fail("NoOrientations");
  if (n_ > ratio_limit()) {
    sync_neighbourhood();
    measures_ = relax_measure(relax_state_, true);
    full_measures_ = false;
  } else {
    // All-pairs energy and separation, then per-point separations over the relaxation's
    // neighbourhood (engine.zig:317-324).
    measure_state_.set_near(group_, nullptr);
    measures_ = relax_measure(measure_state_, true);
    full_measure(measures_);
    full_measures_ = true;
    sync_neighbourhood();
    if (neighbourhood_ > 0) relax_measure(relax_state_, true);
    measure_neighbourhood_ = 0;
  }
  measured_ = true;
  measure_iteration_ = iteration_;
}

double Engine::nearest_pass(RelaxState& rs) {
  for (int attempt = 0; attempt < 4; ++attempt) {
    rs.current = false;
    rs.nn_rank.ensure(size_t(n_) * 4);
    rs.nn_degrees.ensure(size_t(n_) * 4 + 4);
    float* degrees = rs.nn_degrees.as<float>();
    uint32_t* least_bits = reinterpret_cast<uint32_t*>(degrees + n_);
    float4* q = q_.as<float4>();
    if (rs.has_near) {
      // The nearest orbit point is a minimum over keys: the sites' order within a cell
      // does not matter.
      rs.near.build(*this, q, n_, order(), group_, mean_radius(uint64_t(n_) * order()), false);
      launch_near_nearest(prof, stream, n_, order(), rs.near.view(), rs.near.cut, degrees, rs.nn_rank.as<uint32_t>(),
                          least_bits);
    } else {
      rs.nn_chord.ensure(size_t(n_) * 4);
      rs.nn_partial.ensure(size_t(order()) * n_ * 8);
      launch_allpairs_nearest(prof, stream, q, n_, ops_device(), order(), rs.nn_partial.as<uint64_t>(),
                              rs.nn_chord.as<float>(), rs.nn_rank.as<uint32_t>());
      launch_degrees(prof, stream, rs.nn_chord.as<float>(), n_, degrees, least_bits);
    }
    auto& out = separation_out();
    uint32_t least_key = 0;
    check(hipMemcpyAsync(out.data(), degrees, size_t(n_) * 4, hipMemcpyDeviceToHost, stream), "download");
    check(hipMemcpyAsync(&least_key, least_bits, 4, hipMemcpyDeviceToHost, stream), "download");
    download(point_index_, rs.nn_rank, n_);
    sync();
    if (rs.has_near && rs.near.overflowed(*this)) continue;
    point_index_iteration_ = long(iteration_);
    float least;
    std::memcpy(&least, &least_key, 4);
    if (!(least > 0)) fail("DuplicateOrientations");
    return least;
  }
  fail("NearShellTooLarge");
}

void Engine::spacing() {
  if (n_ < 2) fail("TooFewOrientations");
  const double delta = kSpacingDelta;
  spacing_state_.set_near(group_, n_ > kSpacingAllPairs ? &delta : nullptr);
  spacing_state_.near.cell_sites = kSpacingCellSites;
  spacing_state_.near.x_split = kSpacingXSplit;
  spacing_state_.near.cell_edge = kSpacingEdge;
  const double least = nearest_pass(spacing_state_);
  if (measured_ && measure_iteration_ == iteration_) {
    measures_.separation_deg = least;
    return;
  }
  measures_ = Measures{};
  measures_.count_extended = 2.0 * order() * double(n_);
  measures_.separation_deg = least;
  measured_ = true;
  measure_iteration_ = iteration_;
  full_measures_ = false;
}

void Engine::begin_relax() {
  if (n_ < 2) fail("TooFewOrientations");
  cancelled.store(false);
  opt_.active = false;  // a relaxed sample leaves the optimiser's best behind
  sync_neighbourhood();
  relax_radius_ = mean_radius(uint64_t(n_) * order());
}

// The forces of one BB step (if the last pass's are not current): a near-field pass
// (recorded here) or, for all pairs, the job to add to a multi-sample launch.
// BB state and row buffers of the next step, so that its step descriptor is complete
// before its force pass is recorded.
void Engine::prepare_step() {
  if (fix_identity_ && fixed_index_ < 0) fail("IdentityMissingEnableAddIdentity");
  relax_state_.prepare(*this, n_);
  relax_state_.ensure_rows(n_);
}

bool Engine::step_forces(JobDesc& allpairs) {
  prepare_step();
  RelaxState& rs = relax_state_;
  if (rs.current) return false;
  if (rs.has_near) {
    rs.record_forces(*this, n_, order());
    return false;
  }
  allpairs = rs.allpairs_desc(*this, n_, order());
  return true;
}

bool Engine::near_job(NearJob& job) {
  prepare_step();
  RelaxState& rs = relax_state_;
  if (rs.current) return false;
  const uint32_t ord = order();
  float4* q = static_cast<float4*>(q_device());
  rs.near.ready_for(*this, q, n_, ord, group_, mean_radius(uint64_t(n_) * ord));
  job = NearJob{};
  job.q = q;
  job.ops = ops_device();
  job.n = n_;
  job.order = ord;
  job.b = rs.near.view();
  job.cut = rs.near.cut;
  job.key = near_force_key(job.cut);
  job.rows = rs.rows();
  return true;
}

JobDesc Engine::step_job() { return relax_state_.step_desc(*this, relax_radius_, fixed_slot()); }

void Engine::stepped() {
  relax_state_.current = false;
  ++iteration_;
}

void Engine::relax(uint32_t steps) {
  begin_relax();
  bool recorded = false;
  const uint32_t count = std::min<uint32_t>(steps, 1000);
  // Steps are recorded without reading anything back, but every sync_every steps (about
  // 2e10 pair evaluations, some tens of ms) the host waits for the device, so zo_cancel
  // from another thread stops a long call between steps (engine.zig:896-915).
  const double all = double(n_) * double(order());
  const double neighbours =
      neighbourhood_ > 0 ? std::min(all, 0.94 * neighbourhood_ * neighbourhood_ * neighbourhood_ + 64.0) : all;
  const uint32_t sync_every = uint32_t(std::max(1.0, std::floor(2e10 / (double(n_) * neighbours))));
  for (uint32_t index = 0; index < count; ++index) {
    if (cancelled.load()) break;
    JobDesc ap;
    if (step_forces(ap)) launch_allpairs(prof, stream, single_job(ap));
    const JobSet one = single_job(step_job());
    launch_bb_step(prof, stream, one);  // move and fold included
    stepped();
    recorded = true;
    if ((index + 1) % sync_every == 0 && index + 1 < count) sync();
  }
  finish_relax(recorded);
}

// One wait ends a batch of steps: last step, positions, energy and separation
// (pairs.zig:377-399); these forces start the next step.
void Engine::finish_relax(bool recorded) {
  for (int attempt = 0; attempt < 4; ++attempt) {
    finish_prepare();
    relax_state_.record_forces(*this, n_, order());
    finish_reads(recorded);
    sync();
    if (finish_check(recorded)) return;
    recorded = false;  // the last step's result was read; the redone pass reads none
  }
  fail("NearShellTooLarge");
}

void Engine::finish_prepare() {
  measured_ = false;
  relax_state_.prepare(*this, n_);
  relax_state_.current = false;
}

void Engine::finish_reads(bool recorded) {
  RelaxState& rs = relax_state_;
  if (recorded)
    check(hipMemcpyAsync(&finish_state_, rs.state.p, sizeof(BBState), hipMemcpyDeviceToHost, stream), "state");
  host_.resize(size_t(n_) * 4);
  check(hipMemcpyAsync(host_.data(), q_.p, size_t(n_) * 16, hipMemcpyDeviceToHost, stream), "positions");
  download(rows_force_, rs.force, n_);
  download(rows_energy_, rs.energy, n_);
  download(rows_nearest_, rs.nearest, n_);
  if (rs.has_near) rs.near.read_status(*this);
}

bool Engine::finish_check(bool recorded) {
  RelaxState& rs = relax_state_;
  if (recorded) {
    const BBState& st = finish_state_;
    if (st.flags & 1u) fail("NonFiniteForce");
    if (st.flags & 2u) fail("DuplicateOrientations");
    float step_max;
    std::memcpy(&step_max, &st.step_bits, 4);
    step_deg_ = double(step_max) * 360.0 / kPi;
    force_rms_ = std::sqrt(st.ff / double(n_));
  }
  if (rs.has_near && rs.near.took_overflow(*this)) return false;
  auto& sep = separation_out();
  measures_ = quick_measures(rows_force_, rows_energy_, rows_nearest_);
  for (uint32_t i = 0; i < n_; ++i) sep[i] = float(nearest_degrees(rows_nearest_[i]));
  full_measures_ = false;
  measured_ = true;
  measure_iteration_ = iteration_;
  rs.current = true;
  return true;
}

size_t Engine::device_bytes() const {
  size_t b = q_.bytes + ops_.bytes + measure_partial_.bytes + measure_out_.bytes;
  for (const RelaxState* rs : {&relax_state_, &spacing_state_, &measure_state_}) {
    for (const DevBuf* d : {&rs->force, &rs->energy, &rs->nearest, &rs->prev, &rs->state, &rs->partials,
                            &rs->ap_partial, &rs->ap_pnear, &rs->nn_partial, &rs->nn_chord, &rs->nn_rank, &rs->nn_degrees})
      b += d->bytes;
    const NearField& f = rs->near;
    for (const DevBuf* d : {&f.planes, &f.info, &f.rim, &f.uq, &f.ucode, &f.urank, &f.ucell, &f.perm, &f.sq, &f.scode,
                            &f.srank, &f.offsets, &f.uslot,
                            &f.order_buf, &f.scan_tmp, &f.counters, &f.counts, &f.crowded, &f.huge})
      b += d->bytes;
  }
  return b;
}

void Engine::identity(bool add, bool fixed) {
  if (n_ == 0) fail("NoOrientations");
  long index = identity_index();
  if (add && index < 0) {
    const std::string generator = generator_, grid = grid_kind_;
    const uint32_t base = base_count_, seed = seed_, h = h_, inserted = inserted_count_;
    std::vector<float> values(host_.begin(), host_.end());
    values.insert(values.end(), {1.0f, 0.0f, 0.0f, 0.0f});
    set(group_, values.data(), values.size());
    generator_ = generator;
    grid_kind_ = grid;
    h_ = h;
    inserted_count_ = inserted;
    base_count_ = base;
    seed_ = seed;
    index = long(n_) - 1;
  }
  if (fixed && index < 0) fail("IdentityMissingEnableAddIdentity");
  if (fix_identity_ != fixed) relax_state_.release();
  fix_identity_ = fixed;
  fixed_index_ = fixed ? index : -1;
  if (fixed) {
    const float one[4] = {1.0f, 0.0f, 0.0f, 0.0f};
    std::memcpy(&host_[size_t(index) * 4], one, 16);
    cover_ = CoverResult{};
    point_covering_iteration_ = -1;
    check(hipMemcpyAsync(static_cast<char*>(q_.p) + size_t(index) * 16, one, 16, hipMemcpyHostToDevice, stream),
          "identity");
    sync();
    relax_state_.current = false;
  }
}

void Engine::force_passes(uint32_t passes) {
  if (n_ < 2) fail("TooFewOrientations");
  sync_neighbourhood();
  RelaxState& rs = relax_state_;
  rs.prepare(*this, n_);
  for (int attempt = 0; attempt < 4; ++attempt) {
    for (uint32_t k = 0; k < passes; ++k) rs.record_forces(*this, n_, order());
    sync();
    if (rs.has_near && rs.near.overflowed(*this)) continue;
    rs.current = true;
    return;
  }
  fail("NearShellTooLarge");
}

const float* Engine::point_values(uint32_t kind) const {
  if (kind == 0) {
    if (point_separation_iteration_ < 0 || point_separation_.size() != n_) return nullptr;
    return point_separation_.data();
  }
  if (kind == 1 || kind == 2) {
    const PinnedVec<float>& v = kind == 1 ? point_covering_ : point_faces_;
    if (point_covering_iteration_ < 0 || v.size() != n_) return nullptr;
    return v.data();
  }
  return nullptr;
}

// ---- covering radius ----------------------------------------------------------------
// Per source, the farthest vertex of its Voronoi cell among all signed orbit images
// (k_cover); the sample's covering radius is the largest.  Clouds of at most 131,072
// images take every image as a candidate; larger ones the near-field site grid built at
// the cutoff c_max, which grows (x1.6) until every cell is certified within it.
void Engine::cover() {
  if (n_ < 4) fail("TooFewOrientations");
  if (cover_.valid && cover_.iteration == long(iteration_) && point_covering_iteration_ == long(iteration_)) {
    cover_.cached = true;
    return;
  }
  cover_measure(false, false);
}

void Engine::cover_measure(bool vertices, bool tets, bool codes) {
  if (n_ < 4) fail("TooFewOrientations");
  const uint32_t ord = order();
  const double density = double(n_) * ord;
  const double spacing = std::cbrt(kPi * kPi / density);  // S3 angle per cloud point
  const bool all_images = density <= 131072.0;
  const double limit = all_images ? 3.0 : 1.2;
  double c_max = std::min(limit, 5.0 * spacing);
  const float c_start = float(std::min(c_max, 2.4 * spacing));
  cover_theta_.ensure(size_t(n_) * 4);
  cover_witness_.ensure(size_t(n_) * 16);
  cover_flags_.ensure(size_t(n_) * 4);
  cover_faces_.ensure(size_t(n_) * 4);
  const uint32_t vcap = vertices ? uint32_t(std::min<uint64_t>(uint64_t(n_) * 64, 1ull << 26)) : 0;
  const uint32_t tcap = tets ? uint32_t(std::min<uint64_t>(uint64_t(n_) * 64, 1ull << 25)) : 0;
  if (vertices) {
    cover_vq_.ensure(size_t(vcap) * 16);
    cover_vr_.ensure(size_t(vcap) * 4);
  }
  if (tets) cover_tets_.ensure(size_t(tcap) * 16);
  if (codes) cover_wcode_.ensure(size_t(n_) * 16);
  if (vertices || tets) cover_vcount_.ensure(8);
  float4* q = q_.as<float4>();
  uint32_t rounds = 0, rebuilt = 0;
  while (rounds < 8) {
    ++rounds;
    NearBuffers view;
    const NearBuffers* grid = nullptr;
    NearCut cut{};
    if (!all_images) {
      const double mr = mean_radius(uint64_t(n_) * ord);
      const double delta = c_max / mr;
      if (cover_near_.delta != delta) {
        cover_near_.delta = delta;
        cover_near_.ready = false;
      }
      cover_near_.build(*this, q, n_, ord, group_, mr);
      view = cover_near_.view();
      if (view.xdiv != 1) fail("GpuError: the covering walk takes cubic cells");
      grid = &view;
      cut = cover_near_.cut;
    }
    if (vertices || tets) launch_fill_u32(prof, stream, cover_vcount_.as<uint32_t>(), 0u, 2);
    launch_cover(prof, stream, q, n_, ops_device(), ord, grid, cut, c_start, float(c_max), cover_theta_.as<float>(),
                 cover_witness_.as<float4>(), cover_flags_.as<uint32_t>(), cover_faces_.as<float>(),
                 vertices ? cover_vq_.as<float4>() : nullptr, vertices ? cover_vr_.as<float>() : nullptr,
                 vertices ? cover_vcount_.as<uint32_t>() : nullptr, vcap, tets ? cover_tets_.as<uint4>() : nullptr,
                 tets ? cover_vcount_.as<uint32_t>() + 1 : nullptr, tcap, codes ? cover_wcode_.as<uint4>() : nullptr);
    if (codes) {
      download(cover_wcode_host_, cover_wcode_, n_);
      cover_witnesses_.resize(n_);
      check(hipMemcpyAsync(cover_witnesses_.data(), cover_witness_.p, size_t(n_) * 16, hipMemcpyDeviceToHost, stream),
            "download");
    }
    download(point_covering_, cover_theta_, n_);
    download(point_faces_, cover_faces_, n_);
    download(cover_flags_host_, cover_flags_, n_);
    sync();
    if (!all_images && cover_near_.overflowed(*this)) {
      if (++rebuilt > 4) fail("NearShellTooLarge");
      --rounds;
      continue;
    }
    uint32_t open = 0;
    for (uint32_t i = 0; i < n_; ++i) {
      if (cover_flags_host_[i] == 2) fail("CoveringNeighbourOverflow");
      open += cover_flags_host_[i] != 0 ? 1u : 0u;
    }
    if (open) {
      if (c_max >= limit) fail("CoveringDomainNotEnclosed");
      c_max = std::min(limit, c_max * 1.6);
      continue;
    }
    uint32_t worst = 0;
    for (uint32_t i = 1; i < n_; ++i)
      if (point_covering_[i] > point_covering_[worst]) worst = i;
    CoverResult r;
    r.valid = true;
    r.iteration = long(iteration_);
    r.radius_deg = double(point_covering_[worst]) * 360.0 / kPi;
    r.domain_deg = c_max * 360.0 / kPi;
    check(hipMemcpy(r.witness, cover_witness_.as<float4>() + worst, 16, hipMemcpyDeviceToHost), "download");
    r.rounds = rounds;
    r.probes = n_ * rounds;
    r.sites = all_images ? uint32_t(std::min<double>(4e9, 2.0 * density)) : cover_near_.status.sites;
    for (uint32_t i = 0; i < n_; ++i) point_covering_[i] = float(double(point_covering_[i]) * 360.0 / kPi);
    cover_ = r;
    point_covering_iteration_ = long(iteration_);
    if (vertices || tets) {
      uint32_t counts[2] = {0, 0};
      check(hipMemcpy(counts, cover_vcount_.p, 8, hipMemcpyDeviceToHost), "download");
      if (vertices) {
        const uint32_t count = std::min(counts[0], vcap);
        download(cover_vq_host_, cover_vq_, count);
        download(cover_vr_host_, cover_vr_, count);
      }
      if (tets) download(cover_tets_host_, cover_tets_, std::min(counts[1], tcap));
      sync();
    }
    return;
  }
  fail("CoveringRetryLimit");
}

// ---- exact counts (engine.zig:199-232, insert.zig, dedup.zig) --------------------------
namespace {

using Q4 = std::array<double, 4>;

Q4 mul4(const Q4& a, const Q4& b) {
  return {a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3], a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
          a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1], a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]};
}
double dot4d(const Q4& a, const Q4& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }
Q4 unit4(Q4 q) {
  const double s = 1.0 / std::sqrt(dot4d(q, q));
  for (double& x : q) x *= s;
  return q;
}

// Points of S3 hashed by their vector part on a uniform grid; both signs of every point.
class SphereHash {
 public:
  explicit SphereHash(double cell) : cell_(std::max(cell, 1e-9)) {}
  // Room for `count` add() calls (each stores q and -q) without rehashing.
  void reserve(size_t count) {
    cells_.reserve(2 * count);
    points_.reserve(2 * count);
  }
  void add(const Q4& q) {
    for (double sign : {1.0, -1.0}) {
      const Q4 p{sign * q[0], sign * q[1], sign * q[2], sign * q[3]};
      cells_[key(p[1], p[2], p[3])].push_back(uint32_t(points_.size()));
      points_.push_back(p);
    }
  }
  // Whether some stored point has dot(q, p) > threshold.
  bool near(const Q4& q, double threshold) const {
    const double chord = std::sqrt(std::max(0.0, 2.0 - 2.0 * threshold)) * (1 + 1e-9) + 1e-12;
    const auto lo = key(q[1] - chord, q[2] - chord, q[3] - chord), hi = key(q[1] + chord, q[2] + chord, q[3] + chord);
    for (int64_t x = lo[0]; x <= hi[0]; ++x)
      for (int64_t y = lo[1]; y <= hi[1]; ++y)
        for (int64_t z = lo[2]; z <= hi[2]; ++z) {
          auto it = cells_.find({x, y, z});
          if (it == cells_.end()) continue;
          for (uint32_t i : it->second)
            if (dot4d(q, points_[i]) > threshold) return true;
        }
    return false;
  }
  // Index of a stored point within `tolerance` (4D distance) of q, or -1.
  long within(const Q4& q, double tolerance) const {
    const auto lo = key(q[1] - tolerance, q[2] - tolerance, q[3] - tolerance);
    const auto hi = key(q[1] + tolerance, q[2] + tolerance, q[3] + tolerance);
    for (int64_t x = lo[0]; x <= hi[0]; ++x)
      for (int64_t y = lo[1]; y <= hi[1]; ++y)
        for (int64_t z = lo[2]; z <= hi[2]; ++z) {
          auto it = cells_.find({x, y, z});
          if (it == cells_.end()) continue;
          for (uint32_t i : it->second) {
            double d = 0.0;
            for (int k = 0; k < 4; ++k) d += (points_[i][k] - q[k]) * (points_[i][k] - q[k]);
            if (d < tolerance * tolerance) return long(i);
          }
        }
    return -1;
  }

 private:
  using Key = std::array<int64_t, 3>;
  struct KeyHash {
    size_t operator()(const Key& k) const {
      return size_t(uint64_t(k[0]) * 0x9E3779B97F4A7C15ull ^ uint64_t(k[1]) * 0xC2B2AE3D27D4EB4Full ^ uint64_t(k[2]));
    }
  };
  Key key(double x, double y, double z) const {
    return {int64_t(std::floor(x / cell_)), int64_t(std::floor(y / cell_)), int64_t(std::floor(z / cell_))};
  }
  double cell_;
  std::vector<Q4> points_;
  std::unordered_map<Key, std::vector<uint32_t>, KeyHash> cells_;
};

}  // namespace

// Up to `wanted` centres of the largest empty caps, greedily (insert.zig:150-240): every
// Voronoi vertex of the sample (k_cover), carried into the zone, largest first; a centre is
// taken when its cap holds no orbit image of the sample or of a centre already taken (3e-7
// of slack on the cap's height), and an identical centre only once.
std::vector<float> Engine::cap_centres(uint32_t wanted) {
  cover_measure(true, false);
  const uint32_t ord = order();
  std::vector<Q4> ops(ord);
  for (uint32_t k = 0; k < ord; ++k) {
    const float* o = kOperators[kGroupOffset[group_] + k];
    ops[k] = unit4({o[0], o[1], o[2], o[3]});
  }
  // Candidates largest radius first, ties in vertex order (a stable sort's order), produced
  // in growing sorted prefixes: only the first few thousand of the millions of vertices are
  // ever visited, and only those are carried into the zone.
  const size_t total = cover_vq_host_.size();
  if (total == 0) fail("NoEmptyDelaunayCentres");
  std::vector<uint32_t> rank(total);
  for (size_t i = 0; i < total; ++i) rank[i] = uint32_t(i);
  const float* vr = cover_vr_host_.data();
  const auto before = [vr](uint32_t a, uint32_t b) { return vr[a] > vr[b] || (vr[a] == vr[b] && a < b); };
  size_t sorted = 0;
  const auto extend = [&](size_t upto) {
    upto = std::min(total, upto);
    if (upto <= sorted) return;
    if (upto < total) std::nth_element(rank.begin() + long(sorted), rank.begin() + long(upto), rank.end(), before);
    std::sort(rank.begin() + long(sorted), rank.begin() + long(upto), before);
    sorted = upto;
  };
  extend(size_t(wanted) * 16 + 4096);
  const double largest = double(vr[rank[0]]);
  SphereHash occupied(2.0 * std::sin(0.5 * largest));
  occupied.reserve(size_t(n_) + wanted);
  for (uint32_t i = 0; i < n_; ++i) occupied.add(unit4({host_[4 * i], host_[4 * i + 1], host_[4 * i + 2], host_[4 * i + 3]}));
  std::vector<float> out;
  for (size_t at = 0; at < total && out.size() / 4 < wanted; ++at) {
    if (at == sorted) extend(2 * sorted);
    const uint32_t i = rank[at];
    const float4 v = cover_vq_host_[i];
    const Q4 c = unit4({v.x, v.y, v.z, v.w});
    // The zone's image: c g^-1 with the largest |scalar part| = |c . g|.
    uint32_t best = 0;
    double top = -1.0;
    for (uint32_t k = 0; k < ord; ++k) {
      const double w = std::fabs(dot4d(c, ops[k]));
      if (w > top) {
        top = w;
        best = k;
      }
    }
    const Q4& g = ops[best];
    Q4 r = mul4(c, {g[0], -g[1], -g[2], -g[3]});
    if (r[0] < 0)
      for (double& x : r) x = -x;
    const Q4 q = unit4(r);
    if (occupied.near(q, 1.0 - 1e-12)) continue;  // the same centre from another cell
    const double height = std::cos(double(vr[i]));
    bool empty = true;
    for (const Q4& h : ops) {
      if (occupied.near(mul4(q, {h[0], -h[1], -h[2], -h[3]}), height + 3e-7)) {
        empty = false;
        break;
      }
    }
    if (!empty) continue;
    occupied.add(q);
    for (double x : q) out.push_back(float(x));
  }
  if (out.empty()) fail("NoEmptyDelaunayCentres");
  return out;
}

// Keeps the first of every group of rows equal up to the orbit (2e-6), in place; the first
// `unique` rows are known distinct (dedup.zig:5-60).  Returns the count kept.
uint32_t Engine::orbit_compact(std::vector<float>& q, uint32_t unique) const {
  const double tolerance = 2e-6;
  const uint32_t ord = order();
  std::vector<Q4> ops(ord);
  for (uint32_t k = 0; k < ord; ++k) {
    const float* o = kOperators[kGroupOffset[group_] + k];
    ops[k] = {o[0], o[1], o[2], o[3]};
  }
  SphereHash kept(2.0 * tolerance);
  uint32_t count = 0;
  const uint32_t rows = uint32_t(q.size() / 4);
  kept.reserve(rows);
  for (uint32_t i = 0; i < rows; ++i) {
    const Q4 src{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    bool duplicate = false;
    if (i >= unique) {
      for (const Q4& g : ops) {
        if (kept.within(unit4(mul4(src, g)), tolerance) >= 0) {
          duplicate = true;
          break;
        }
      }
    }
    if (duplicate) continue;
    for (int k = 0; k < 4; ++k) q[4 * count + k] = float(src[k]);
    kept.add(src);
    ++count;
  }
  return count;
}

void Engine::generate_exact(uint32_t group, uint32_t target, bool add_identity, uint32_t kind) {
  if (target < 32 || target > 1000000) fail("ExactCountRange32To1000000");
  const uint32_t budget = target - (add_identity ? 1u : 0u);
  uint32_t h = 1;
  if (kind == 2) {
    if (group < 8 || group > 10) fail("TransportNeedsTOrOOrI");
    while (h < 64 && transport::count(group, h + 1) <= budget) ++h;
  } else {
    uint32_t best = 0;
    uint64_t most = 0;
    for (uint32_t f = 1; f <= hopf::kMaxFrequency; ++f) {
      const uint64_t n = hopf::raw_count(group, f);
      if (n <= budget && n > most) {
        most = n;
        best = f;
      }
    }
    if (!best) fail("BaseGridExceedsTarget");
    h = best;
  }
  generate(group, h, target, 0, kind);
  const uint32_t base = n_;
  if (add_identity) identity(true, fix_identity_);
  const std::string generator = generator_, grid = grid_kind_;
  for (int round = 0; round < 8 && n_ != target; ++round) {
    if (n_ > target) fail("BaseGridExceedsTarget");
    const std::vector<float> additions = cap_centres(target - n_);
    std::vector<float> values(host_.begin(), host_.end());
    values.insert(values.end(), additions.begin(), additions.end());
    const uint32_t keep = orbit_compact(values, n_);
    if (keep <= n_) fail("NoDistinctDelaunayCentres");
    values.resize(size_t(keep) * 4);
    set(group, values.data(), values.size());
  }
  if (n_ != target) fail("InsufficientDelaunayCentres");
  generator_ = generator;
  grid_kind_ = grid;
  h_ = h;
  base_count_ = base;
  inserted_count_ = target - base;
}

// The Delaunay tetrahedra around the sample's orbit points (engine.zig:352-358): a covering
// measurement that also lists, for every certified cell, the source and the three
// neighbours whose bisectors meet at each vertex.  Vertices are stereographic (x, y, z)/(1 + w)
// of the sample's points (first, in order) and of the other orbit images that close its
// cells; each tetrahedron once.
void Engine::build_mesh() {
  if (n_ < 4) fail("TooFewOrientations");
  cover_measure(false, true);
  const uint32_t ord = order(), two = 2 * ord;
  std::vector<std::array<uint32_t, 4>> tets;
  tets.reserve(cover_tets_host_.size());
  for (size_t i = 0; i < cover_tets_host_.size(); ++i) {
    const uint4 t = cover_tets_host_[i];
    if (t.y == ~0u || t.z == ~0u || t.w == ~0u) continue;
    std::array<uint32_t, 4> k{t.x, t.y, t.z, t.w};
    std::sort(k.begin(), k.end());
    tets.push_back(k);
  }
  std::sort(tets.begin(), tets.end());
  tets.erase(std::unique(tets.begin(), tets.end()), tets.end());
  std::unordered_map<uint32_t, uint32_t> index;
  mesh_vertices_.clear();
  auto vertex = [&](uint32_t code) -> uint32_t {
    const uint32_t j = code / two;
    if (code % two == 0) return j;
    auto it = index.find(code);
    if (it != index.end()) return it->second;
    const uint32_t at = n_ + uint32_t(index.size());
    index.emplace(code, at);
    return at;
  };
  std::vector<uint32_t> flat;
  flat.reserve(tets.size() * 4);
  for (const auto& t : tets)
    for (uint32_t c : t) flat.push_back(vertex(c));
  std::vector<uint32_t> codes(index.size());
  for (const auto& kv : index) codes[kv.second - n_] = kv.first;
  mesh_vertices_.resize(size_t(n_ + codes.size()) * 4);
  auto put = [&](uint32_t at, const float4 q) {
    const double w = 1.0 + double(q.x);
    mesh_vertices_[4 * at + 0] = float(q.y / w);
    mesh_vertices_[4 * at + 1] = float(q.z / w);
    mesh_vertices_[4 * at + 2] = float(q.w / w);
    mesh_vertices_[4 * at + 3] = 0.0f;
  };
  const float4* ops = reinterpret_cast<const float4*>(operations());
  for (uint32_t j = 0; j < n_; ++j) put(j, make_float4(host_[4 * j], host_[4 * j + 1], host_[4 * j + 2], host_[4 * j + 3]));
  for (size_t i = 0; i < codes.size(); ++i) {
    const uint32_t c = codes[i], j = c / two, k = (c % two) / 2;
    const float4 src = make_float4(host_[4 * j], host_[4 * j + 1], host_[4 * j + 2], host_[4 * j + 3]);
    float4 y = image(src, ops[k], k);
    if (c & 1u) y = neg4(y);
    put(n_ + uint32_t(i), y);
  }
  mesh_tets_.swap(flat);
}

// ---- covering optimisation -----------------------------------------------------------
void Engine::set_kicks(uint32_t patience, float progress, float size, float trust) {
  if (patience < 1 || patience > 1000000) fail("InvalidKickPatience");
  if (!(progress >= 0.0f && progress < 1.0f)) fail("InvalidKickProgress");
  if (!(size >= 0.0f && size <= 1.0f)) fail("InvalidKickSize");
  if (!(trust > 0.0f && trust <= 1.0f)) fail("InvalidKickTrust");
  kicks_ = Kicks{patience, progress, size, trust};
}

// Rows `index` of the sample set to `rows` (4 floats each), folded into the zone first
// when asked (on the GPU, those rows only), device and host copies both.
void Engine::write_rows(const std::vector<uint32_t>& index, const std::vector<float>& rows, bool fold) {
  const size_t m = index.size();
  if (!m) return;
  opt_rows_.ensure(m * 16);
  check(hipMemcpyAsync(opt_rows_.p, rows.data(), m * 16, hipMemcpyHostToDevice, stream), "upload");
  if (fold) launch_fold(prof, stream, opt_rows_.as<float4>(), uint32_t(m), ops_device(), order(), group_);
  for (size_t i = 0; i < m; ++i)
    check(hipMemcpyAsync(q_.as<float4>() + index[i], opt_rows_.as<float4>() + i, 16, hipMemcpyDeviceToDevice, stream),
          "rows");
  std::vector<float> folded(m * 4);
  check(hipMemcpyAsync(folded.data(), opt_rows_.p, m * 16, hipMemcpyDeviceToHost, stream), "download");
  sync();
  for (size_t i = 0; i < m; ++i) std::memcpy(&host_[size_t(index[i]) * 4], &folded[4 * i], 16);
}

// The whole sample replaced by values (already in the zone), keeping the run's records.
void Engine::write_sample(const std::vector<float>& values) {
  check(hipMemcpyAsync(q_.p, values.data(), values.size() * 4, hipMemcpyHostToDevice, stream), "upload");
  sync();
  std::memcpy(host_.data(), values.data(), values.size() * 4);
}

void Engine::optimize_covering(uint32_t steps, uint32_t mode, bool finish, float floor_deg) {
  if (n_ < 4) fail("TooFewOrientations");
  if ((mode & 3u) > 1 || (mode & ~0xff001fu) != 0) fail("InvalidCoveringMode");
  if (!(floor_deg >= 0.0f && floor_deg < 180.0f)) fail("InvalidSeparationFloor");
  if (fix_identity_ && fixed_index_ < 0) fail("IdentityMissingEnableAddIdentity");
  cancelled.store(false);
  const uint32_t ord = order(), two = 2 * ord;
  const float4* ops = reinterpret_cast<const float4*>(operations());
  // Moves pull a cap's points this fraction of the way to its centre; "face" (mode 1)
  // moves half as far per step.
  const double pull = (mode & 3u) == 1 ? 0.5 : 1.0;
  auto measure = [&]() { cover_measure(false, false, true); };
  auto current_ok = [&]() {
    return cover_.valid && cover_.iteration == long(iteration_) && cover_wcode_host_.size() == n_;
  };
  if (!opt_.active) {
    if (!current_ok()) measure();
    opt_ = Optimizer{};
    opt_.active = true;
    opt_.best = opt_.reference = cover_.radius_deg;
    opt_.best_q.assign(host_.begin(), host_.end());
  } else if (!current_ok()) {
    measure();
  }
  // The separation floor is checked with a spacing pass (SO(3) degrees).
  auto separation_ok = [&]() {
    if (floor_deg <= 0.0f) return true;
    const double delta = kSpacingDelta;
    spacing_state_.set_near(group_, n_ > kSpacingAllPairs ? &delta : nullptr);
    spacing_state_.near.cell_sites = kSpacingCellSites;
    spacing_state_.near.x_split = kSpacingXSplit;
    spacing_state_.near.cell_edge = kSpacingEdge;
    return nearest_pass(spacing_state_) >= double(floor_deg);
  };
  uint64_t rng = 0x9E3779B97F4A7C15ull ^ (uint64_t(opt_.kicks + 1) << 17);
  auto uniform = [&]() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return double(rng >> 11) * (1.0 / 9007199254740992.0);
  };
  for (uint32_t s = 0; s < steps; ++s) {
    if (cancelled.load()) break;
    uint32_t worst = 0;
    for (uint32_t i = 1; i < n_; ++i)
      if (point_covering_[i] > point_covering_[worst]) worst = i;
    const uint4 codes = cover_wcode_host_[worst];
    const float4 wv = cover_witnesses_[worst];
    const double v[4] = {wv.x, wv.y, wv.z, wv.w};
    std::vector<uint32_t> index;
    std::vector<float> rows, old_rows;
    for (uint32_t code : {codes.x, codes.y, codes.z, codes.w}) {
      if (code == ~0u) continue;
      const uint32_t j = code / two, k = (code % two) / 2;
      const bool neg = (code & 1u) != 0;
      if (fix_identity_ && long(j) == fixed_index_) continue;
      if (std::find(index.begin(), index.end(), j) != index.end()) continue;
      const float4 qj = make_float4(host_[4 * j], host_[4 * j + 1], host_[4 * j + 2], host_[4 * j + 3]);
      float4 y = image(qj, ops[k], k);
      if (neg) y = neg4(y);
      double yd[4] = {y.x, y.y, y.z, y.w};
      const double side = yd[0] * v[0] + yd[1] * v[1] + yd[2] * v[2] + yd[3] * v[3] < 0 ? -1.0 : 1.0;
      double nyv[4], norm = 0.0;
      for (int c = 0; c < 4; ++c) {
        nyv[c] = yd[c] + opt_.alpha * pull * (side * v[c] - yd[c]);
        norm += nyv[c] * nyv[c];
      }
      norm = 1.0 / std::sqrt(norm);
      // q_j' = (+-y') g_k^-1
      const float4 yn = make_float4(float(nyv[0] * norm), float(nyv[1] * norm), float(nyv[2] * norm),
                                    float(nyv[3] * norm));
      const float4 g = ops[k];
      float4 qn = qmul(neg ? neg4(yn) : yn, make_float4(g.x, -g.y, -g.z, -g.w));
      index.push_back(j);
      rows.insert(rows.end(), {qn.x, qn.y, qn.z, qn.w});
      old_rows.insert(old_rows.end(), {qj.x, qj.y, qj.z, qj.w});
    }
    ++opt_.steps;
    if (index.empty()) {
      ++opt_.rejected;
      continue;
    }
    const CoverResult before = cover_;
    std::vector<float> before_radius(point_covering_.begin(), point_covering_.end());
    std::vector<uint4> before_codes(cover_wcode_host_.begin(), cover_wcode_host_.end());
    std::vector<float4> before_witness = cover_witnesses_;
    write_rows(index, rows, true);
    ++iteration_;
    measure();
    const bool better = cover_.radius_deg < before.radius_deg - 1e-12 && separation_ok();
    if (better) {
      ++opt_.improved;
      opt_.alpha = std::min(0.6, opt_.alpha * 1.25);
      if (cover_.radius_deg < opt_.best - 1e-12) {
        opt_.best = cover_.radius_deg;
        opt_.best_q.assign(host_.begin(), host_.end());
      }
      if ((opt_.reference - cover_.radius_deg) >= double(kicks_.progress) * opt_.reference) {
        opt_.reference = cover_.radius_deg;
        opt_.since_gain = 0;
      } else {
        ++opt_.since_gain;
      }
    } else {
      ++opt_.rejected;
      ++opt_.since_gain;
      write_rows(index, old_rows, false);
      --iteration_;
      cover_ = before;
      std::copy(before_radius.begin(), before_radius.end(), point_covering_.data());
      std::copy(before_codes.begin(), before_codes.end(), cover_wcode_host_.data());
      cover_witnesses_ = before_witness;
      point_covering_iteration_ = long(iteration_);
      opt_.alpha = opt_.alpha * 0.5 < 1e-3 ? 0.3 : opt_.alpha * 0.5;
    }
    if (opt_.since_gain >= kicks_.patience) {
      // Plateau: from the best sample, every free point turned by a random rotation of RMS
      // angle `size` times the best radius; steps restart at `trust`.
      std::vector<float> values = opt_.best_q;
      const double angle = double(kicks_.size) * opt_.best * kPi / 360.0;
      for (uint32_t i = 0; i < n_ && angle > 0.0; ++i) {
        if (fix_identity_ && long(i) == fixed_index_) continue;
        double a[3], len = 0.0;
        for (double& x : a) {
          x = 2.0 * uniform() - 1.0;
          len += x * x;
        }
        len = std::sqrt(std::max(len, 1e-30));
        const double t = angle * std::sqrt(3.0) * uniform();
        const float4 r = make_float4(float(std::cos(0.5 * t)), float(std::sin(0.5 * t) * a[0] / len),
                                     float(std::sin(0.5 * t) * a[1] / len), float(std::sin(0.5 * t) * a[2] / len));
        const float4 q = make_float4(values[4 * i], values[4 * i + 1], values[4 * i + 2], values[4 * i + 3]);
        const float4 out = normalize4(qmul(q, r));
        values[4 * i] = out.x;
        values[4 * i + 1] = out.y;
        values[4 * i + 2] = out.z;
        values[4 * i + 3] = out.w;
      }
      write_sample(values);
      fold(n_);
      fetch();
      ++iteration_;
      measure();
      ++opt_.kicks;
      opt_.since_gain = 0;
      opt_.reference = cover_.radius_deg;
      opt_.alpha = std::max(double(kicks_.trust) * 10.0, 1e-3);
    }
  }
  // The sample kept is the best one seen (kicked samples included only once a move from
  // them was accepted, so a floor the moves respect holds for it).
  if (finish && !std::equal(opt_.best_q.begin(), opt_.best_q.end(), host_.begin())) {
    write_sample(opt_.best_q);
    ++iteration_;
    measure();
  }
  measured_ = measured_ && measure_iteration_ == iteration_;
}

const uint32_t* Engine::point_indices() const {
  if (point_index_iteration_ < 0 || point_index_.size() != n_) return nullptr;
  return point_index_.data();
}

// ---- stats JSON (engine.zig:938-1020) ----------------------------------------------
namespace {
struct Json {
  std::string s = "{";
  bool first = true;
  void key(const char* k) {
    if (!first) s += ",";
    first = false;
    s += "\"";
    s += k;
    s += "\":";
  }
  void num(const char* k, double v) {
    key(k);
    if (!std::isfinite(v)) {
      s += "null";
      return;
    }
    char b[64];
    snprintf(b, sizeof(b), "%.17g", v);
    s += b;
  }
  void integer(const char* k, long long v) {
    key(k);
    s += std::to_string(v);
  }
  void null(const char* k) {
    key(k);
    s += "null";
  }
  void boolean(const char* k, bool v) {
    key(k);
    s += v ? "true" : "false";
  }
  void raw(const char* k, const char* text) {
    key(k);
    s += text;
  }
  void floats(const char* k, const float* v, int count) {
    key(k);
    s += "[";
    for (int i = 0; i < count; ++i) {
      char b[64];
      snprintf(b, sizeof(b), "%s%.9g", i ? "," : "", double(v[i]));
      s += b;
    }
    s += "]";
  }
  void str(const char* k, const std::string& v) {
    key(k);
    s += "\"";
    for (char c : v) {
      if (c == '"' || c == '\\') s += '\\';
      s += c;
    }
    s += "\"";
  }
};
}  // namespace

std::string Engine::stats() {
  const bool m = measured_;
  const bool current = m && measure_iteration_ == iteration_;
  const uint32_t ord = order();
  Json j;
  j.str("version", kVersion);
  j.str("group", kGroupNames[group_]);
  j.integer("order", ord);
  j.integer("count", n_);
  j.integer("count_extended", 2ll * ord * n_);
  j.str("generator", generator_);
  if (grid_kind_.empty())
    j.null("grid");
  else
    j.str("grid", grid_kind_);
  j.integer("h", h_);
  j.integer("seed", seed_);
  j.integer("iteration", iteration_);
  j.boolean("fixed_identity", fix_identity_);
  const long id = n_ ? identity_index() : -1;
  if (id >= 0)
    j.integer("identity_index", id);
  else
    j.null("identity_index");
  j.integer("base_count", base_count_);
  j.integer("inserted_count", inserted_count_);
  j.str("triangulation_backend", "hip");
  // Simplex bound (SO3K_Transports so3kt/measures/reference.py; covering_reference.zig:4-15):
  // sum_i c_i count^-(2i+1)/3 in SO(3) degrees, count = 2 |K| N.
  double bound = NAN;
  if (n_ > 0) {
    static const double c[8] = {1.889832101454757, 0.573930763512449, 0.581069217539205, 0.859789414116394,
                                1.490131881944242, 2.791922498626376, 5.481645765756282, 11.12163594341394};
    const double t = std::cbrt(1.0 / (2.0 * ord * double(n_)));
    double v = 0.0;
    for (int i = 7; i >= 0; --i) v = c[i] + t * t * v;
    bound = t * v * 360.0 / kPi;
  }
  const bool cover_current = cover_.valid && cover_.iteration == long(iteration_);
  j.num("simplex_bound_deg", bound);
  if (cover_.valid) {
    j.num("covering_excess", cover_.radius_deg / bound - 1.0);
  } else {
    j.null("covering_excess");
  }
  if (cover_current && current && measures_.separation_deg > 0)
    j.num("mesh_ratio", cover_.radius_deg / measures_.separation_deg);
  else
    j.null("mesh_ratio");
  if (cover_.valid)
    j.num("covering_radius_deg", cover_.radius_deg);
  else
    j.null("covering_radius_deg");
  j.boolean("covering_current", cover_current);
  j.boolean("covering_cached", cover_.valid && cover_.cached);
  if (cover_.valid) {
    j.integer("covering_unresolved_caps", 0);
    j.integer("covering_iteration", cover_.iteration);
    j.floats("covering_witness", cover_.witness, 4);
    j.num("covering_domain_bound_deg", cover_.domain_deg);
    j.integer("covering_shell_points", cover_.sites);
    j.integer("covering_fence_points", 0);
    j.null("covering_tetrahedra");
    j.integer("covering_certificate_probes", cover_.probes);
    j.integer("covering_exact_predicates", 0);
    j.integer("covering_rounds", cover_.rounds);
  } else {
    for (const char* k : {"covering_unresolved_caps", "covering_iteration", "covering_witness",
                          "covering_domain_bound_deg", "covering_shell_points", "covering_fence_points",
                          "covering_tetrahedra", "covering_certificate_probes", "covering_exact_predicates",
                          "covering_rounds"})
      j.null(k);
  }
  j.integer("covering_optimizer_steps", opt_.steps);
  j.integer("covering_optimizer_kicks", opt_.kicks);
  if (opt_.active)
    j.num("covering_optimizer_best_deg", opt_.best);
  else
    j.null("covering_optimizer_best_deg");
  j.integer("covering_optimizer_improved", opt_.improved);
  j.integer("covering_optimizer_rejected", opt_.rejected);
  // The reference's counters of its local re-tessellation and LP jobs (engine.zig:1000-1011)
  // have no counterpart in this optimiser, which measures the whole covering after every
  // move: present, so that readers of the record find every key, and null.
  for (const char* k : {"covering_optimizer_jobs", "covering_optimizer_skipped", "covering_optimizer_local_rounds",
                        "covering_optimizer_mismatches", "covering_optimizer_local_error",
                        "covering_optimizer_skip_reasons", "covering_optimizer_overflows",
                        "covering_optimizer_full_steps", "covering_optimizer_live_builds",
                        "covering_optimizer_live_failures", "covering_optimizer_phase_ms"})
    j.null(k);
  {
    char b[256];
    snprintf(b, sizeof(b), "{\"patience\":%u,\"progress\":%.9g,\"size\":%.9g,\"trust\":%.9g}", kicks_.patience,
             double(kicks_.progress), double(kicks_.size), double(kicks_.trust));
    j.raw("covering_kick_settings", b);
  }
  j.boolean("measured", m);
  j.boolean("measures_current", current);
  if (m)
    j.integer("measures_iteration", measure_iteration_);
  else
    j.null("measures_iteration");
  j.boolean("energy_ratios_current", full_measures_ && current);
  if (neighbourhood_ > 0)
    j.num("relax_neighbourhood", neighbourhood_);
  else
    j.null("relax_neighbourhood");
  if (point_separation_iteration_ >= 0)
    j.integer("point_separation_iteration", point_separation_iteration_);
  else
    j.null("point_separation_iteration");
  if (point_covering_iteration_ >= 0)
    j.integer("point_covering_iteration", point_covering_iteration_);
  else
    j.null("point_covering_iteration");
  if (m && measure_neighbourhood_ > 0)
    j.num("energy_neighbourhood", measure_neighbourhood_);
  else
    j.null("energy_neighbourhood");
  j.integer("ratio_limit", ratio_limit());
  std::string precision;
  if (!m)
    precision = "Not measured";
  else if (!current)
    precision = "Energy and separation from an earlier iteration";
  else if (!std::isfinite(measures_.quey_energy))
    precision = "Separation current; energy not measured";
  else if (full_measures_)
    precision = "All measures current";
  else if (n_ > ratio_limit())
    precision = "Energy and separation current; this sample exceeds the ratio limit";
  else
    precision = "Energy and separation current; refresh energy ratios";
  j.str("precision", precision);
  const double nan = std::numeric_limits<double>::quiet_NaN();
  j.num("energy", m ? measures_.quey_energy : nan);
  j.num("quey_energy", m ? measures_.quey_energy : nan);
  j.num("riesz_1", m ? measures_.riesz_1 : nan);
  j.num("riesz_2", m ? measures_.riesz_2 : nan);
  j.num("riesz_3", m ? measures_.riesz_3 : nan);
  j.num("green_ratio", m ? measures_.green_ratio : nan);
  j.num("log_ratio", m ? measures_.log_ratio : nan);
  j.num("separation_deg", m ? measures_.separation_deg : nan);
  j.num("step_deg", step_deg_);
  j.num("force_rms", m ? measures_.force_rms : force_rms_);
  int dev = 0;
  hipDeviceProp_t prop{};
  std::string adapter = "HIP";
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    adapter = std::string(prop.name) + " (" + prop.gcnArchName + ")";
  j.str("adapter", adapter);
  j.s += "}";
  text = j.s;
  return text;
}

}  // namespace zxt
