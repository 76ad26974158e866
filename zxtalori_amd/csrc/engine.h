// Host engine of the MI355X build: one sample, its device state and the relaxation,
// measure and spacing drivers (the roles of src/engine.zig, src/pairs.zig RelaxState and
// src/near.zig in the reference), over the HIP runtime with one stream.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"

namespace zxt {

// Graph captures and the host calls that can end another thread's capture (device and
// pinned host allocation and free, stream creation and destruction, synchronous copies)
// exclude each other process-wide: a capture holds this lock shared, such a call holds it
// exclusively.  A thread that already holds it (inside its own capture, or inside an
// enclosing call) does not take it again.
class CaptureLock {
 public:
  explicit CaptureLock(bool capture);
  ~CaptureLock();
  CaptureLock(const CaptureLock&) = delete;
  CaptureLock& operator=(const CaptureLock&) = delete;

 private:
  int mode_;  // 0 not taken, 1 shared, 2 exclusive
};

// A grown-only device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release();
  // Ensures at least `b` bytes; contents are not kept.  Throws on failure.
  void ensure(size_t b);
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct EngineError {
  std::string message;
};

// A grown-only host array in pinned memory (device-to-host copies at full PCIe rate), with
// the few vector operations the engine uses.  resize keeps the first min(old, new) items.
template <typename T>
class PinnedVec {
 public:
  PinnedVec() = default;
  PinnedVec(const PinnedVec&) = delete;
  PinnedVec& operator=(const PinnedVec&) = delete;
  ~PinnedVec() {
    if (p_) {
      CaptureLock lock(false);
      (void)hipHostFree(p_);
    }
  }
  void resize(size_t n) {
    if (n > cap_) {
      CaptureLock lock(false);
      void* q = nullptr;
      if (hipHostMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) throw EngineError{"OutOfHostMemory"};
      if (p_) {
        std::memcpy(q, p_, n_ * sizeof(T));
        (void)hipHostFree(p_);
      }
      p_ = static_cast<T*>(q);
      cap_ = n;
    }
    n_ = n;
  }
  size_t size() const { return n_; }
  size_t bytes() const { return cap_ * sizeof(T); }
  T* data() { return p_; }
  const T* data() const { return p_; }
  T& operator[](size_t i) { return p_[i]; }
  const T& operator[](size_t i) const { return p_[i]; }
  const T* begin() const { return p_; }
  const T* end() const { return p_ + n_; }

 private:
  T* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

// measure() results (pairs.zig:17-30).
struct Measures {
  double count_extended = NAN;
  double riesz_1 = NAN, riesz_2 = NAN, riesz_3 = NAN, green_ratio = NAN, log_ratio = NAN;
  double quey_energy = NAN;
  double separation_deg = NAN;
  double force_rms = NAN;
};

struct StepResult {
  double max_step_deg = 0, force_rms = 0, energy = 0;
};

class Engine;

// Near-field sites, cells and constants of one sample (near.zig:31-196).
struct NearField {
  double delta = 0;
  double taper = 0.5;
  double cell_sites = 0;  // sites per cell the grid aims for (0: chosen from the reach)
  uint32_t n = 0, order = 0, group = 0;
  bool ready = false;
  double radius = 0;
  NearCut cut{};
  uint32_t capacity = 0, sdim = 1, nplanes = 0;
  // Cell shape the owner asks for: y/z cells cell_edge times the edge of a cube holding
  // the chosen sites per cell, split into x_split x-cells each (1 and 1: cubic cells, as
  // the covering walk takes them); xdiv is the grid's x-cells per y/z cell.
  uint32_t x_split = 1, xdiv = 1;
  double cell_edge = 1.0;
  float band = 0;
  // The last pass's sites, shell and overflow flag (valid after a wait).
  NearInfo status{};
  DevBuf planes, info, rim, uq, ucode, urank, ucell, perm, sq, scode, srank, offsets, uslot, order_buf,
      scan_tmp, counters, counts, crowded, huge;
  NearBuffers view() const;
  void setup(Engine& e, const float4* q, uint32_t n, uint32_t order, uint32_t group, double mean_radius);
  // Records sites and sorts for the positions q (setting up first when needed); `ordered`
  // as launch_near_build takes it.
  void build(Engine& e, const float4* q, uint32_t n, uint32_t order, uint32_t group, double mean_radius,
             bool ordered = true);
  // Sets up (sizing pass, allocation, one wait) unless set up for this sample already.
  void ready_for(Engine& e, const float4* q, uint32_t n, uint32_t order, uint32_t group, double mean_radius);
  // After a wait: whether the last pass overflowed (then the next pass sizes again).
  bool overflowed(Engine& e);
  // overflowed() in two halves, for callers that wait once for many samples: the reads of
  // the pass's status and counters (enqueued), then, after the wait, the check.
  void read_status(Engine& e);
  bool took_overflow(Engine& e);
};

// Relaxation state (pairs.zig:220-413): rows, BB history, step state, near field.
struct RelaxState {
  uint32_t n = 0;
  uint32_t order = 0;
  bool current = false;
  bool has_near = false;
  NearField near;
  DevBuf force, energy, nearest, prev, state, partials, ap_partial, ap_pnear, nn_partial, nn_chord, nn_rank,
      nn_degrees;
  Rows rows() const { return Rows{force.as<float4>(), energy.as<double>(), nearest.as<float>()}; }
  void release();
  void set_near(uint32_t group, const double* delta);
  void prepare(Engine& e, uint32_t n);
  void ensure_rows(uint32_t n);
  void record_forces(Engine& e, uint32_t n, uint32_t order);
  // The job descriptors of the multi-sample kernels: an all-pairs force pass (buffers
  // ensured) and a BB step (after prepare).
  JobDesc allpairs_desc(Engine& e, uint32_t n, uint32_t order);
  JobDesc step_desc(Engine& e, double radius, uint32_t fixed);
};

class Batch;

class Engine {
  friend class Batch;

 public:
  // With a stream, launches go to it (a batch's); otherwise the engine owns one.
  explicit Engine(hipStream_t shared = nullptr);
  ~Engine();

  // Public operations (src/engine.zig).  They throw EngineError.
  void set(uint32_t group, const float* values, size_t count4);
  void generate_random(uint32_t group, uint32_t count, uint32_t seed);
  // Inverse Hopf grid at frequency h, or with h = 0 the frequency whose count is nearest
  // `target` (engine.zig:242-282), for kind 0 or 1; kind 2 is the FCC transport of T, O or I
  // at semi-edge max(1, h).
  void generate(uint32_t group, uint32_t h, uint32_t target, uint32_t seed, uint32_t kind);
  void generate_transport(uint32_t group, uint32_t h, uint32_t seed);
  // Covering radius of the sample (engine.zig:378-398, covering.zig:555-577): measured
  // again only when the sample changed since the last measurement.
  void cover();
  // Exactly `target` orientations (engine.zig:199-232): the largest base grid that fits,
  // then centres of the largest empty caps until the count is reached.
  void generate_exact(uint32_t group, uint32_t target, bool add_identity, uint32_t kind);
  // Covering optimisation: `steps` greedy cap moves (mode 0 "tet", 1 "face" = half steps);
  // `finish` leaves the best sample seen; the floor (SO(3) degrees) rejects moves that bring
  // any pair nearer than it.  set_kicks: plateau patience, relative gain, kick size, trust.
  void set_kicks(uint32_t patience, float progress, float size, float trust);
  void optimize_covering(uint32_t steps, uint32_t mode, bool finish, float floor_deg);
  // The Delaunay mesh around the sample (covering measured too): stereographic vertices
  // (4 floats each, the sample first) and tetrahedra (4 indices each).
  void build_mesh();
  const float* mesh_vertices() const { return mesh_vertices_.data(); }
  const uint32_t* mesh_tets() const { return mesh_tets_.data(); }
  uint32_t mesh_vertex_count() const { return uint32_t(mesh_vertices_.size() / 4); }
  uint32_t mesh_tet_count() const { return uint32_t(mesh_tets_.size() / 4); }
  std::vector<float> mesh_vertices_;
  std::vector<uint32_t> mesh_tets_;
  void measure();
  void spacing();
  void relax(uint32_t steps);
  void set_neighbourhood(double delta);
  void identity(bool add, bool fixed);
  void force_passes(uint32_t passes);
  void force_rows(float* force4, double* energy, float* nearest);
  size_t device_bytes() const;
  std::string stats();
  const float* point_values(uint32_t kind) const;
  const uint32_t* point_indices() const;

  uint32_t count() const { return n_; }
  uint32_t group() const { return group_; }
  uint32_t order() const;
  const float* operations() const;
  const float* quaternions() const { return host_.data(); }
  uint32_t ratio_limit() const;

  std::atomic<bool> cancelled{false};
  Prof prof;
  double pair_budget_scale = 1.0;
  hipStream_t stream = nullptr;
  std::string err;
  std::string text;
  NearInfo last_near{};
  uint32_t last_near_sdim = 0;
  unsigned long long last_near_counts[2] = {0, 0};  // accepted pairs, sites tested
  NearCut last_near_cut{};
  double last_near_radius = 0;
  // Development hook: the next near-field sizing leaves room for exactly its site count
  // plus near_squeeze sites (never fewer than the sources), instead of a quarter more, so
  // that the pass overflows when the shell grows; cleared once used.
  bool near_squeeze_pending = false;
  int32_t near_squeeze = 0;
  uint32_t near_resizes = 0;  // near-field passes redone after their shell overflowed

  // Sends this engine's launches to s (returns the previous stream); a batch runs the
  // near-field passes of its jobs on several streams this way.
  hipStream_t use_stream(hipStream_t s) {
    hipStream_t old = stream;
    stream = s;
    prof.set_stream(s);
    return old;
  }
  const float4* ops_device() const;
  void* q_device() const;
  // index + 1 of the held identity, or 0.
  uint32_t fixed_slot() const { return fixed_index_ >= 0 ? uint32_t(fixed_index_ + 1) : 0u; }
  void sync();
  template <typename V>
  void download(V& dst, const DevBuf& src, size_t count) {
    dst.resize(count);
    check(hipMemcpyAsync(dst.data(), src.p, count * sizeof(dst[0]), hipMemcpyDeviceToHost, stream), "download");
  }
  static void check(hipError_t e, const char* what);

 private:
  void set_group(uint32_t group);
  void fold(uint32_t n);
  void fetch();
  void adopt(uint32_t n);
  void reset_run();
  long identity_index() const;
  void sync_neighbourhood();
  Measures relax_measure(RelaxState& rs, bool keep_separation);
  void full_measure(Measures& m);
  DevBuf measure_partial_, measure_out_;
  // One force pass read back into rows_force_, rows_energy_, rows_nearest_.
  void read_forces(RelaxState& rs);
  PinnedVec<float4> rows_force_;
  PinnedVec<double> rows_energy_;
  PinnedVec<float> rows_nearest_;
  double nearest_pass(RelaxState& rs);
  Measures quick_measures(const PinnedVec<float4>& f, const PinnedVec<double>& energy,
                          const PinnedVec<float>& nearest) const;
  PinnedVec<float>& separation_out();
  // relax() in parts, so that a batch runs the same steps for many samples.
  void begin_relax();
  void prepare_step();
  bool step_forces(JobDesc& allpairs);
  // A near-field sample's force pass of one step as a job of a multi-sample launch (the
  // near field set up first); false when the last pass's forces are current.
  bool near_job(NearJob& job);
  JobDesc step_job();
  void stepped();
  void finish_relax(bool recorded);
  // finish_relax in parts, so that a batch records the final force passes of many samples
  // together and waits once: set-up before the pass, the reads after it (enqueued), and,
  // after the wait, the last step's result and the measures (false: the near field
  // overflowed, and finish_relax(false) redoes the pass).
  void finish_prepare();
  void finish_reads(bool recorded);
  bool finish_check(bool recorded);
  BBState finish_state_{};
  double relax_radius_ = 0;
  bool owns_stream_ = false;

  uint32_t group_ = 9;
  DevBuf ops_, q_;
  PinnedVec<float> host_;
  uint32_t n_ = 0;
  uint32_t iteration_ = 0;
  uint32_t seed_ = 0;
  uint32_t base_count_ = 0;
  std::string generator_ = "none";
  std::string grid_kind_;  // "hopf" or "transport", empty for samples that are not grids
  uint32_t h_ = 0;
  bool measured_ = false;
  Measures measures_;
  uint32_t measure_iteration_ = 0;
  bool full_measures_ = false;
  RelaxState relax_state_, spacing_state_, measure_state_;
  double neighbourhood_ = 0, measure_neighbourhood_ = 0;
  PinnedVec<float> point_separation_;
  // Covering: the last result, per-point cap radii (SO(3) degrees) and Voronoi face counts.
  struct CoverResult {
    bool valid = false, cached = false;
    long iteration = -1;
    double radius_deg = 0, domain_deg = 0;
    float witness[4] = {0, 0, 0, 0};
    uint32_t rounds = 0, probes = 0, sites = 0;
  };
  CoverResult cover_;
  NearField cover_near_;
  DevBuf cover_theta_, cover_witness_, cover_flags_, cover_faces_;
  PinnedVec<float> point_covering_, point_faces_;
  PinnedVec<uint32_t> cover_flags_host_;
  long point_covering_iteration_ = -1;
  DevBuf cover_vq_, cover_vr_, cover_vcount_;
  PinnedVec<float4> cover_vq_host_;
  PinnedVec<float> cover_vr_host_;
  uint32_t inserted_count_ = 0;
  DevBuf cover_tets_, cover_wcode_, opt_rows_;
  PinnedVec<uint4> cover_tets_host_, cover_wcode_host_;
  std::vector<float4> cover_witnesses_;
  // One covering measurement (with every certified cell's vertices, or its Delaunay
  // tetrahedra, when asked).
  void cover_measure(bool vertices, bool tets, bool codes = false);
  // Covering optimisation (engine.zig:399-420, covering.zig Live): greedy moves of the
  // points spanning the largest empty cap toward its centre, kept when the measured radius
  // drops; plateau kicks from the best sample.
  struct Kicks {
    uint32_t patience = 150;
    float progress = 5e-4f, size = 0.04f, trust = 0.03f;
  };
  struct Optimizer {
    bool active = false;
    uint32_t steps = 0, kicks = 0, improved = 0, rejected = 0, since_gain = 0;
    double best = 0.0, reference = 0.0, alpha = 0.3;
    std::vector<float> best_q;
  };
  Kicks kicks_;
  Optimizer opt_;
  void write_rows(const std::vector<uint32_t>& index, const std::vector<float>& rows, bool fold);
  void write_sample(const std::vector<float>& values);
  std::vector<float> cap_centres(uint32_t wanted);
  uint32_t orbit_compact(std::vector<float>& q, uint32_t unique) const;
  PinnedVec<uint32_t> point_index_;
  long point_separation_iteration_ = -1;
  long point_index_iteration_ = -1;
  double step_deg_ = 0, force_rms_ = 0;
  bool fix_identity_ = false;
  long fixed_index_ = -1;
};

double mean_radius(uint64_t count);

}  // namespace zxt
