// Launch interface between the host engine (engine.cpp) and the gfx950 kernels
// (kernels.hip).  Plain structs of device pointers; every launch goes to one stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "transport_math.h"

namespace zxt {

// Per-source force rows of one pass: (fx, fy, fz, curvature) in FP32, the energy sum in
// FP64 and the nearest chord^2 (over every orbit point but the source itself).
// The reference keeps these in a 48-byte Row (pairs.zig:11-15); the build stores only
// the columns the relaxation path reads.
struct Rows {
  float4* force = nullptr;
  double* energy = nullptr;
  float* nearest = nullptr;
};

// Near-field constants of one sample (near.zig:95, :176; pairs.wgsl:349-361).
struct NearCut {
  float chord2;       // accept a site when chord^2(x, y) <= chord2
  float taper_start;  // window starts falling at this S3 angle (0.5 r)
  float inv_width;    // 1 / (0.5 r)
  float c0;           // x.y >= c0 bounds every accepted site (with margin)
  float s0;           // sqrt(1 - c0^2)
  uint32_t own;       // the source's own non-identity images can lie within the cutoff
  // The spacing pass first searches the cap of half the reach: inner_c0, inner_s0 bound
  // every site whose chord^2 passes inner_chord2, as c0, s0 do for chord2.
  float inner_chord2;
  float inner_c0;
  float inner_s0;
};

// Device-resident status of one near-field pass.  lo/hi are order-preserving u32 keys of
// the sites' chart box; lower/cpu the box the cell grid uses this pass.
struct NearInfo {
  uint32_t sites;     // sites written (core + shell), may exceed capacity
  uint32_t overflow;  // set when sites > capacity
  uint32_t rim;       // rim sources listed
  uint32_t wmin_key;  // order-preserving key of the smallest source scalar part
  uint32_t crowded;   // cells listed for the crowded-cell sort (more than 16 sites)
  uint32_t huge;      // of those, cells of more than 1024 sites
  uint32_t lo[3];
  uint32_t hi[3];
  float lower[3];
  float cpu;          // cells per chart unit
};

// BB step state, device-resident (relax.wgsl:17-22 keeps the same in 16 words).
constexpr uint32_t kDigitBins = 2048;  // radix-select bins per middle rank (11-bit digits)
struct BBState {
  uint32_t iteration;  // steps begun
  uint32_t flags;      // 1 non-finite force, 2 duplicate orientation
  uint32_t step_bits;  // f32 bits of the largest move this step
  uint32_t step;       // the running step's index (iteration before it began)
  double sy, yy, ss, ff, energy;
  uint32_t rank[2];    // remaining rank of the two middle ranks after the first two digits
  uint32_t prefix[2];  // their first two digits
  float median;
  float alpha;
  double alpha64;
  // Radix-select histograms of the median key: the first digit (both middle ranks share
  // it), then per rank the second and third digits.
  uint32_t hist0[kDigitBins];
  uint32_t hist1[2][kDigitBins];
  uint32_t hist2[2][kDigitBins / 2];
};

// Scratch of an all-pairs pass: partial rows of `splits` (operator batch x site chunk).
struct AllPairsPlan {
  uint32_t ops_per_split = 1;
  uint32_t site_chunk = 256;
  uint32_t site_splits = 1;
  uint32_t op_splits = 1;
  __host__ __device__ uint32_t splits() const { return ops_per_split ? site_splits * op_splits : 0; }
};

// Per-launch timing with HIP events on the engine's stream (zo_profile).
class Prof {
 public:
  explicit Prof(hipStream_t s = nullptr) : stream_(s) {}
  ~Prof();
  void set_stream(hipStream_t s) { stream_ = s; }
  bool on() const { return on_; }
  void enable(bool v);
  void run(const char* name, const std::function<void()>& launch);
  // Waits for recorded launches and folds them into the totals.
  void collect();
  std::string report();
  void clear();
  // Totals in ms and launch counts, by kernel name.
  struct Total {
    std::string name;
    double ms;
    uint64_t count;
  };
  std::vector<Total> totals();

 private:
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  hipStream_t stream_;
  bool on_ = false;
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> pool_;
  std::vector<Total> totals_;
  hipEvent_t get_event();
};

// One sample's device state for the multi-sample kernels (all pairs, BB step, fold).
struct JobDesc {
  float4* q = nullptr;
  const float4* ops = nullptr;
  uint32_t n = 0, order = 0, group = 0;
  uint32_t fixed = 0;  // index + 1 of a held identity, or 0
  Rows rows;
  float4* prev = nullptr;        // 2n: previous force, previous move
  BBState* state = nullptr;
  double* bb_partials = nullptr;  // 5 per point block
  double* ap_partial = nullptr;   // 5 * splits * n
  float* ap_pnear = nullptr;      // splits * n
  AllPairsPlan plan;
  double radius = 0;
  float step_scale = 1.0f;
  uint32_t point_blocks = 0;   // ceil(n / 256)
  uint32_t median_blocks = 0;  // min(point_blocks, 256)
  uint32_t reserved = 0;       // keeps the struct free of padding (tables compare bytewise)
};
static_assert(sizeof(JobDesc) == 136, "JobDesc has no padding bytes");

// The jobs of one launch: a device table of JobDesc with per-stage prefix sums of each
// job's workgroups (grid block b belongs to the job j with start[j] <= b < start[j+1]),
// or a single job passed by value (table == nullptr).
struct JobSet {
  const JobDesc* table = nullptr;
  uint32_t count = 0;
  const uint32_t* point_start = nullptr;   // count + 1
  const uint32_t* ap_start = nullptr;      // count + 1
  const uint32_t* median_start = nullptr;  // count + 1
  uint32_t point_total = 0, ap_total = 0, median_total = 0;
  JobDesc one;
};
JobDesc describe_job(float4* q, const float4* ops, uint32_t n, uint32_t order, uint32_t group);
JobSet single_job(const JobDesc& d);

// ---- launchers (kernels.hip) ----------------------------------------------------
void launch_haar(Prof& p, hipStream_t s, float4* q, uint32_t n, uint32_t seed);
// FCC transport grid (job.count points, rows described by offsets) into out.
void launch_transport(Prof& p, hipStream_t s, const fcc::Tables* tables, const fcc::Job& job, const uint32_t* offsets,
                      float4* out);
void launch_fold(Prof& p, hipStream_t s, const JobSet& jobs);
void launch_fold(Prof& p, hipStream_t s, float4* q, uint32_t n, const float4* ops, uint32_t order, uint32_t group);
void launch_fill_u32(Prof& p, hipStream_t s, uint32_t* dst, uint32_t value, size_t count);

// Tile plan of an all-pairs pass: about `target` workgroups (site chunks x operator
// batches x 256-source blocks), operators batched `ops_per_split` to a workgroup (0: only
// when the operator count alone would exceed 4 x target).
struct AllPairsShape {
  uint32_t target = 2048;
  uint32_t ops_per_split = 0;
};
// The force pass's shape for n points and |K| = order: the tuned table (allpairs_tuned.h,
// written by zxtalori_amd/tune.py) unless set_allpairs_override holds a shape.
AllPairsShape allpairs_shape(uint32_t n, uint32_t order);
void set_allpairs_override(uint32_t target, uint32_t ops_per_split);
AllPairsPlan plan_allpairs(uint32_t n, uint32_t order, double budget_scale, AllPairsShape shape = {});
// All-pairs rows of every job in the set (each job's plan, partial rows and rows).
void launch_allpairs(Prof& p, hipStream_t s, const JobSet& jobs);
void launch_allpairs_nearest(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                             uint64_t* partial, float* chord_out, uint32_t* rank_out);
// zo_spacing's all-pairs path with the group's operators read from the compiled-in table
// and no profiling hooks (nothing allocated, host or device): nearest keys into partial
// (order x n), chords into chord (n), then ranks, SO(3) degrees and the least degree's
// bits (*least_bits).  Returns the launches' error state.
hipError_t launch_spacing_table(hipStream_t s, const float4* q, uint32_t n, uint32_t group, uint64_t* partial,
                                float* chord, uint32_t* rank_out, float* degrees_out, uint32_t* least_bits);
// The text zo_create_error returns (abi.cpp), for failures of calls without an engine.
void set_create_error(const char* text);
// FP64 Riesz 1-3, log and Green row sums (out: n x 5), all pairs, self pair excluded.
void launch_measure(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                    const AllPairsPlan& plan, double* partial, double* out);

// Near field.  Buffers are owned by the engine's NearField (engine.cpp).
struct NearBuffers {
  const float4* planes = nullptr;
  uint32_t nplanes = 0;
  float band = 0;        // sin(r + 1e-4): rim and shell margin (shell.wgsl:19-22)
  float reach = 0;       // r + 1e-4: the angle within which a site can matter
  uint32_t capacity = 0;
  uint32_t sdim = 1;             // cells along y and z
  uint32_t xdiv = 1;             // cells along x per y/z cell (x-cells xdiv times finer)
  NearInfo* info = nullptr;
  uint32_t* rim = nullptr;       // n
  float4* uq = nullptr;          // capacity, unsorted sites
  uint32_t* ucode = nullptr;     // capacity
  uint32_t* urank = nullptr;     // capacity
  uint32_t* ucell = nullptr;     // capacity + 1: each site's cell; then the core flags' scan
  uint32_t* perm = nullptr;      // capacity
  float4* sq = nullptr;          // capacity, sorted by cell
  uint32_t* scode = nullptr;     // capacity
  uint32_t* srank = nullptr;     // capacity
  uint32_t* offsets = nullptr;   // cell_count(sdim, xdiv) + 1
  uint32_t* uslot = nullptr;     // capacity: each site's arrival slot in its cell
  uint32_t* order = nullptr;     // n, sources in the tile-by-tile visit order of their identity images
  float4* opos = nullptr;        // n, their positions in that order, unordered builds only (shares uq)
  uint32_t* scan_tmp = nullptr;  // scan_tmp_words(max(cells, capacity))
  unsigned long long* counters = nullptr;  // [0] accepted pairs, [1] sites tested (force pass)
  uint32_t* counts = nullptr;    // 2n: per source, pairs and sites tested of the last force pass
  uint32_t* crowded = nullptr;   // crowded_capacity: cells with more sites than one thread sorts
  uint32_t crowded_capacity = 0;
  uint32_t pad1 = 0;
  uint32_t* huge = nullptr;      // huge_capacity: cells with more sites than one wave sorts
  uint32_t huge_capacity = 0;
  uint32_t pad2 = 0;
};
static_assert(sizeof(NearBuffers) == 200, "NearBuffers has no padding bytes");
// Cells of a near-field grid: sdim * xdiv along x, sdim along y and z.
__host__ __device__ inline size_t cell_count(uint32_t sdim, uint32_t xdiv) {
  return size_t(sdim) * xdiv * sdim * sdim;
}
// The sources are visited by tiles of cells: kTile^3 cubic cells, or with x-cells xdiv times
// finer, (kTile / 2 * xdiv) x (kTile / 2)^2 cells.  The padded grid of tiles has this many
// visits (at least cell_count).
constexpr uint32_t kTile = 8;
__host__ __device__ inline uint32_t tile_edge(uint32_t xdiv) { return xdiv > 1 ? kTile / 2 : kTile; }
__host__ __device__ inline uint32_t tile_visits(uint32_t sdim, uint32_t xdiv) {
  const uint32_t e = tile_edge(xdiv), nt = (sdim + e - 1) / e;
  return nt * nt * nt * e * e * e * xdiv;
}
// capacity / 17 + 1 bounds the cells holding more than 16 of `capacity` sites, and
// capacity / 1025 + 1 those holding more than 1024.
inline uint32_t crowded_capacity(uint32_t capacity) { return capacity / 17 + 1; }
inline uint32_t huge_capacity(uint32_t capacity) { return capacity / 1025 + 1; }
constexpr uint32_t kCrowdedBlocks = 1024;
constexpr uint32_t kHugeBlocks = 64;
size_t scan_tmp_words(size_t count);

// Writes the sites (core + shell) and their chart box only (sizing pass).
void launch_near_sites(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                       const NearBuffers& b);
// Builds the sites (core + shell), sorts them into the cell grid and orders the sources
// by the tile-by-tile visit order of their identity images.  Reads nothing back.  With
// `ordered`, the sites of every cell are in rank order (sums over them then run in the same
// order every pass); without it, in arrival order (enough for minima, such as spacing).
void launch_near_build(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                       const NearBuffers& b, bool ordered = true);
// Force rows (mode 0) or nearest orbit point with rank (mode 1) over the built sites.
void launch_near_force(Prof& p, hipStream_t s, const float4* q, uint32_t n, const NearBuffers& b, const NearCut& cut,
                       Rows out);
// Spacing over the built sites: per source the SO(3) degrees of its nearest orbit point
// and that point's rank, and the least degree's bits (*least_bits).
void launch_near_nearest(Prof& p, hipStream_t s, uint32_t n, uint32_t order, const NearBuffers& b, const NearCut& cut,
                         float* degrees_out, uint32_t* rank_out, uint32_t* least_bits);

// The force kernel variant a cutoff takes: own images in reach (8) plus arctangent terms.
uint32_t near_force_key(const NearCut& cut);

// One sample of a multi-sample near-field pass (a batch step): every build stage and the
// force pass run as one launch for all samples (one per force variant), each workgroup
// finding its sample by the stage's prefix sums of workgroups.  Same arithmetic per sample
// as launch_near_build + launch_near_force, so the rows are bit-identical.
struct NearJob {
  const float4* q = nullptr;
  const float4* ops = nullptr;
  uint32_t n = 0, order = 0;
  NearBuffers b;
  NearCut cut{};
  uint32_t key = 0;  // near_force_key(cut)
  Rows rows;
};
static_assert(sizeof(NearJob) == 288, "NearJob has no padding bytes");
enum NearStage : uint32_t {
  kNsFill,        // cell offsets cleared
  kNsRim,         // core sites and rim list
  kNsShell,       // shell sites
  kNsSites,       // per-site count, place, rank place
  kNsCellTiles,   // scan of the cell counts: tiles, then totals (one workgroup), then adds
  kNsCellAdd,
  kNsCrowded,     // crowded-cell sorts
  kNsHuge,
  kNsVisits,      // source visit counts and order
  kNsVisitTiles,  // scan of the visit counts
  kNsVisitAdd,
  kNsForce,       // one wave per source
  kNsReduce,      // pair and candidate counters
  kNearStages
};
uint32_t near_stage_blocks(const NearJob& j, uint32_t stage);
constexpr uint32_t kNearKeys = 16;
struct NearSet {
  const NearJob* table = nullptr;  // sorted by key
  uint32_t count = 0;
  const uint32_t* start = nullptr;  // kNearStages x (count + 1) workgroup prefix sums
  uint32_t total[kNearStages] = {};
  // Force launches: per variant, the first force workgroup and how many.
  uint32_t keys = 0;
  uint32_t key[kNearKeys] = {};
  uint32_t key_first[kNearKeys] = {};
  uint32_t key_blocks[kNearKeys] = {};
};
void launch_near_jobs(Prof& p, hipStream_t s, const NearSet& set);

// SO(3) degrees of nearest chords (out) and their minimum's bits (*least).
// Covering radius per source (S3 angle of its Voronoi cell's farthest vertex), the vertex,
// a flag (0 certified, 1 needs a cutoff beyond c_max, 2 too many neighbours) and the cell's
// face count; grid: the near-field buffers built at cutoff c_max, or null to take every
// signed image.  With vq, every vertex of every certified cell is appended (direction and
// S3 radius) up to vcap, *vcount counting them all (zero it first); with tets, the
// Delaunay tetrahedra at those vertices as four orbit codes j*2|K| + 2k + sign (~0 for a
// box face) up to tcap, *tcount counting them; with wcode, per source the codes of the four
// points spanning its cell's farthest vertex.
void launch_cover(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                  const NearBuffers* grid, const NearCut& cut, float c_start, float c_max, float* theta,
                  float4* witness, uint32_t* flags, float* faces, float4* vq = nullptr, float* vr = nullptr,
                  uint32_t* vcount = nullptr, uint32_t vcap = 0, uint4* tets = nullptr, uint32_t* tcount = nullptr,
                  uint32_t tcap = 0, uint4* wcode = nullptr);
void launch_degrees(Prof& p, hipStream_t s, const float* chord, uint32_t n, float* out, uint32_t* least);
// Issue-rate probe: wave64 instructions per second of non-packed FP32 FMA chains (kind 0)
// or transcendental chains (kind 1) over `blocks` workgroups of 256 (scratch: blocks x 256
// floats); -1 on failure.
double probe_issue_rate(hipStream_t s, uint32_t kind, uint32_t iters, float* scratch, uint32_t blocks);

// BB step (relax.wgsl) of every job in the set, resident: stats, sums, median, move.
void launch_bb_step(Prof& p, hipStream_t s, const JobSet& jobs);
uint32_t bb_partials(uint32_t n);
void launch_bb_reset(Prof& p, hipStream_t s, BBState* state);

}  // namespace zxt
