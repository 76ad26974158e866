// gfx950 kernels of the QVM relaxation and spacing path.
//
// Reference kernels they replace (shaders/*.wgsl), by subsystem:
//   haar, fold ................. haar.wgsl:5-10, state.wgsl:31-51
//   all-pairs rows ............. pairs.wgsl:263-326 (quey_partial, quey_reduce)
//   near-field sites ........... shell.wgsl:66-130 (rim, shell), near.wgsl:22-82 (sorts, images)
//   near-field rows ............ pairs.wgsl:363-427 (near_force, near_reduce)
//   nearest orbit point ........ pairs.wgsl:428-434 and the nearest column of the rows
//   BB step .................... relax.wgsl:29-165
//
// Design (MI355X: 256 CUs in 8 XCDs, wave64, 160 KB LDS per CU):
//  * All pairs: one thread per source, 256-source workgroups, images of a 256-site tile
//    built in LDS per operator (never stored), (operator batch x site chunk) splits in
//    blockIdx.y so small samples still launch thousands of workgroups; partial rows in
//    FP64, reduced in a fixed order.
//  * Near field: one wave per source.  A source's cutoff cap is a ball in the
//    stereographic chart (centre v/(c0+w), radius sqrt(1-c0^2)/(c0+w)); the wave lists
//    the cell rows that ball crosses, trims each row in x, and streams the rows' site runs
//    64 candidates at a time.  Candidates within the cutoff are compacted through an LDS
//    ring (ballot + mbcnt), so the ~100-op pair kernel always runs on full waves; the
//    accepted order is deterministic, and so are the sums.  Sources are visited tile by
//    tile (8^3 cells), and workgroups are remapped so each XCD takes a contiguous range
//    of them and its L2 holds their sites.
//  * Sorts are counting sorts (atomic counts, a three-kernel scan, atomic scatter, then an
//    in-cell sort by a unique key), so every pass is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>

#include "groups_table.h"
#include "kernels.h"
#include "qvm_device.h"
#include "transport_math.h"
#include "allpairs_tuned.h"

namespace zxt {

namespace {

constexpr uint32_t kBlock = 256;

inline uint32_t blocks_for(size_t n, uint32_t per) { return uint32_t((n + per - 1) / per); }

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Cross-lane steps by DPP (immediate lane patterns: no address registers, which the
// ds_bpermute shuffles would keep live across the loops that call these).  row_shr:n
// within each 16-lane row, then row_bcast:15 and row_bcast:31 carry row totals upward;
// lanes without a source lane take 0.
template <int CTRL, int ROWS>
__device__ inline uint32_t dpp_u32(uint32_t v) {
  return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS>
__device__ inline double dpp_f64(double v) {
  const uint64_t b = uint64_t(__double_as_longlong(v));
  const uint32_t lo = dpp_u32<CTRL, ROWS>(uint32_t(b)), hi = dpp_u32<CTRL, ROWS>(uint32_t(b >> 32));
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

// Inclusive prefix sum over the wave (lane i: v_0 + ... + v_i).
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  v += dpp_u32<0x111, 0xf>(v);  // row_shr:1
  v += dpp_u32<0x112, 0xf>(v);  // row_shr:2
  v += dpp_u32<0x114, 0xf>(v);  // row_shr:4
  v += dpp_u32<0x118, 0xf>(v);  // row_shr:8
  v += dpp_u32<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_u32<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// The wave's FP64 sum, in lane 63 (the same order every run).
__device__ inline double wave_sum_to_last(double v) {
  v += dpp_f64<0x111, 0xf>(v);
  v += dpp_f64<0x112, 0xf>(v);
  v += dpp_f64<0x114, 0xf>(v);
  v += dpp_f64<0x118, 0xf>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  return v;
}

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

__device__ inline uint32_t wave_min_u32(uint32_t v) {
  for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d));
  return v;
}

// Loads through a pointer of the global address space are global loads with a scalar
// base; through a generic pointer read from memory (a job table's), the compiler cannot
// tell where it points and emits flat loads, which take 64-bit VGPR addresses.
template <typename T>
using global_t = __attribute__((address_space(1))) T;
template <typename T>
__device__ inline global_t<T>* to_global(T* p) {
  return (global_t<T>*)p;
}
__device__ inline float4 load4(const global_t<const float4>* p) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = *(const global_t<const v4>*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// Workgroup b runs on XCD b % 8: give each XCD a contiguous range of virtual blocks.
__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t grid) {
  const uint32_t per = grid / 8;  // grid is a multiple of 8
  return (b % 8) * per + b / 8;
}
inline uint32_t xcd_grid(uint32_t blocks) { return (blocks + 7) / 8 * 8; }

__device__ inline uint32_t hash32(uint32_t v) {
  uint32_t x = v;
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  return x ^ (x >> 16);
}
__device__ inline float random_unit(uint32_t v) { return (float(hash32(v) >> 8) + 0.5f) / 16777216.0f; }

__device__ inline float3 chart(const float4 q) {
  const float d = 1.0f + q.x;
  return make_float3(q.y / d, q.z / d, q.w / d);
}

}  // namespace

// ---- Prof ---------------------------------------------------------------------------
Prof::~Prof() {
  for (auto& r : pending_) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : pool_) (void)hipEventDestroy(e);
}
hipEvent_t Prof::get_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
void Prof::enable(bool v) {
  if (!v) collect();
  on_ = v;
}
void Prof::run(const char* name, const std::function<void()>& launch) {
  if (!on_) {
    launch();
    return;
  }
  Rec r{name, get_event(), get_event()};
  if (r.a) (void)hipEventRecord(r.a, stream_);
  launch();
  if (r.b) (void)hipEventRecord(r.b, stream_);
  pending_.push_back(r);
  if (pending_.size() > 4096) collect();
}
void Prof::collect() {
  if (pending_.empty()) return;
  (void)hipStreamSynchronize(stream_);
  for (auto& r : pending_) {
    float ms = 0.0f;
    if (r.a && r.b) (void)hipEventElapsedTime(&ms, r.a, r.b);
    bool found = false;
    for (auto& t : totals_) {
      if (t.name == r.name) {
        t.ms += ms;
        t.count += 1;
        found = true;
        break;
      }
    }
    if (!found) totals_.push_back({r.name, double(ms), 1});
    if (r.a) pool_.push_back(r.a);
    if (r.b) pool_.push_back(r.b);
  }
  pending_.clear();
}
std::vector<Prof::Total> Prof::totals() {
  collect();
  return totals_;
}
void Prof::clear() {
  collect();
  totals_.clear();
}
std::string Prof::report() {
  collect();
  std::string out;
  char line[256];
  for (auto& t : totals_) {
    snprintf(line, sizeof(line), "%s\t%llu\t%.6f\n", t.name.c_str(), (unsigned long long)t.count, t.ms);
    out += line;
  }
  return out;
}

// ---- haar (haar.wgsl:5-10) ---------------------------------------------------------
__global__ void k_haar(float4* __restrict__ q, uint32_t n, uint32_t seed_hash) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t base = hash32(i ^ seed_hash);
  const float u = random_unit(base);
  const float a = 6.28318530718f * random_unit(base + 1u);
  const float b = 6.28318530718f * random_unit(base + 2u);
  const float su = sqrtf(u), sv = sqrtf(1.0f - u);
  q[i] = make_float4(su * cosf(b), sv * sinf(a), sv * cosf(a), su * sinf(b));
}

void launch_haar(Prof& p, hipStream_t s, float4* q, uint32_t n, uint32_t seed) {
  // hash(seed) on the host: the same integer arithmetic.
  uint32_t x = seed;
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  p.run("haar", [&] { hipLaunchKernelGGL(k_haar, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, q, n, x); });
}

// ---- FCC transports (grids.zig:40-77, grids.wgsl:165-174) ------------------------------
// One thread per lattice point, in FP64: the point's (i, j) row by binary search of the row
// offsets, then the volume-preserving map onto the zone (transport_math.h) and the
// homochoric-to-quaternion solve; the result is rounded once to FP32.
__global__ __launch_bounds__(256) void k_transport(const fcc::Tables* __restrict__ tb, fcc::Job job,
                                                   const uint32_t* __restrict__ offsets, float4* __restrict__ out) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= job.count) return;
  uint32_t lo = 0, hi = job.side * job.side;  // offsets[lo] <= n < offsets[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (offsets[mid] <= n)
      lo = mid;
    else
      hi = mid;
  }
  const int i = int(lo / job.side) - job.bound, j = int(lo % job.side) - job.bound;
  int k0, kmax;
  fcc::row_span(job, i, j, k0, kmax);
  double q[4];
  fcc::transport_point(*tb, job, i, j, k0 + 2 * int(n - offsets[lo]), q);
  out[n] = make_float4(float(q[0]), float(q[1]), float(q[2]), float(q[3]));
}

void launch_transport(Prof& p, hipStream_t s, const fcc::Tables* tables, const fcc::Job& job, const uint32_t* offsets,
                      float4* out) {
  p.run("transport", [&] {
    hipLaunchKernelGGL(k_transport, dim3(blocks_for(job.count, 256)), dim3(256), 0, s, tables, job, offsets, out);
  });
}

// ---- fold (state.wgsl:1-51) ---------------------------------------------------------
__device__ bool in_zone(const float4 p, uint32_t group, const float4* __restrict__ ops, uint32_t order) {
  const float ax = fabsf(p.y), ay = fabsf(p.z), az = fabsf(p.w), w = p.x;
  switch (group) {
    case 0: return az <= w;
    case 1: return az <= 0.5773502691896258f * w;
    case 2: return az <= 0.4142135623730951f * w;
    case 3: return az <= 0.2679491924311227f * w;
    case 4: return fmaxf(fmaxf(ax, ay), az) <= w;
    case 5: return az <= 0.5773502691896258f * w && fmaxf(ax, 0.5f * ax + 0.8660254037844386f * ay) <= w;
    case 6:
      return az <= 0.4142135623730951f * w && fmaxf(fmaxf(ax, ay), 0.7071067811865476f * (ax + ay)) <= w;
    case 7: {
      const float m = fmaxf(ax, ay), nn = fminf(ax, ay);
      return az <= 0.2679491924311227f * w && fmaxf(m, 0.8660254037844386f * m + 0.5f * nn) <= w;
    }
    case 8: return ax + ay + az <= w;
    case 9: return fmaxf(fmaxf(ax, ay), az) <= 0.4142135623730951f * w && ax + ay + az <= w;
    case 10: {
      if (w > 0.9510566f) return true;
      if (w < 0.9256147f) return false;
      for (uint32_t k = 1; k < order; ++k) {
        const float4 op = ops[k];
        if (fabsf(fabsf(op.x) - 0.8090169943749475f) < 1e-5f && fabsf(dot4(p, op)) > w) return false;
      }
      return true;
    }
    default: return true;  // C1
  }
}

__device__ inline uint32_t find_job(const uint32_t* __restrict__ start, uint32_t count, uint32_t b) {
  if (!start) return 0;
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
__device__ inline JobDesc load_job(const JobSet& s, uint32_t j) { return s.table ? s.table[j] : s.one; }
__device__ inline uint32_t job_start(const uint32_t* __restrict__ start, uint32_t j) { return start ? start[j] : 0u; }

// One point into the fundamental zone (state.wgsl:1-51): rescaled by exponent bits (no
// division can overflow), normalised, scalar part made nonnegative, then, outside the zone,
// the orbit point of largest scalar part.
__device__ float4 fold_point(const float4 a, uint32_t group, const float4* __restrict__ ops, uint32_t order) {
  const float scale = fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w)));
  const int maximum = int((__float_as_uint(scale) >> 23) & 255u);
  float c[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t bits = __float_as_uint(c[k]);
    const int exponent = int((bits >> 23) & 255u);
    const int scaled = exponent + 127 - maximum;
    const uint32_t shifted = (bits & 0x807fffffu) | (uint32_t(max(scaled, 0)) << 23);
    const bool keep = exponent > 0 && scaled > 0;
    c[k] = __uint_as_float(keep ? shifted : (bits & 0x80000000u));
  }
  float4 p = normalize4(make_float4(c[0], c[1], c[2], c[3]));
  if (p.x < 0.0f) p = neg4(p);
  if (in_zone(p, group, ops, order)) return p;
  float4 best = p;
  float w = p.x;
  for (uint32_t k = 0; k < order; ++k) {
    const float4 cq = qmul(p, ops[k]);
    if (fabsf(cq.x) > w) {
      best = cq.x >= 0.0f ? cq : neg4(cq);
      w = fabsf(cq.x);
    }
  }
  return normalize4(best);
}

__global__ void k_fold(JobSet s) {
  const uint32_t j = find_job(s.point_start, s.count, blockIdx.x);
  const JobDesc d = load_job(s, j);
  const uint32_t i = (blockIdx.x - job_start(s.point_start, j)) * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  d.q[i] = fold_point(d.q[i], d.group, d.ops, d.order);
}

JobDesc describe_job(float4* q, const float4* ops, uint32_t n, uint32_t order, uint32_t group) {
  JobDesc d;
  d.q = q;
  d.ops = ops;
  d.n = n;
  d.order = order;
  d.group = group;
  d.point_blocks = blocks_for(n, kBlock);
  d.median_blocks = std::max(1u, std::min(d.point_blocks, 256u));
  return d;
}

JobSet single_job(const JobDesc& d) {
  JobSet s;
  s.count = 1;
  s.one = d;
  s.point_total = d.point_blocks;
  s.ap_total = d.point_blocks * d.plan.splits();
  s.median_total = d.median_blocks;
  return s;
}

void launch_fold(Prof& p, hipStream_t s, const JobSet& jobs) {
  if (!jobs.point_total) return;
  p.run("fold", [&] { hipLaunchKernelGGL(k_fold, dim3(jobs.point_total), dim3(kBlock), 0, s, jobs); });
}

void launch_fold(Prof& p, hipStream_t s, float4* q, uint32_t n, const float4* ops, uint32_t order, uint32_t group) {
  launch_fold(p, s, single_job(describe_job(q, ops, n, order, group)));
}

__global__ void k_fill_u32(uint32_t* __restrict__ dst, uint32_t value, size_t count) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
    dst[i] = value;
}
void launch_fill_u32(Prof& p, hipStream_t s, uint32_t* dst, uint32_t value, size_t count) {
  const uint32_t g = uint32_t(std::min<size_t>(4096, (count + kBlock - 1) / kBlock + 1));
  p.run("fill", [&] { hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(kBlock), 0, s, dst, value, count); });
}

// ---- all pairs (pairs.wgsl:263-326) -------------------------------------------------
// A job's workgroups are (split, source block): split = operator batch x site chunk of
// its plan.  Partial row layout: 5 doubles (fx, fy, fz, curvature, energy) per (split,
// source), nearest chord^2 separately; the reduce sums the splits in a fixed order.
__global__ __launch_bounds__(256) void k_allpairs(JobSet s) {
  __shared__ float4 tile[kBlock];
  const uint32_t jb = find_job(s.ap_start, s.count, blockIdx.x);
  const JobDesc d = load_job(s, jb);
  const uint32_t local = blockIdx.x - job_start(s.ap_start, jb);
  const uint32_t split = local / d.point_blocks, sblock = local % d.point_blocks;
  const float4* __restrict__ q = d.q;
  const uint32_t n = d.n;
  const uint32_t k0 = (split / d.plan.site_splits) * d.plan.ops_per_split, k1 = min(d.order, k0 + d.plan.ops_per_split);
  const uint32_t j0 = (split % d.plan.site_splits) * d.plan.site_chunk, j1 = min(n, j0 + d.plan.site_chunk);
  const uint32_t i = sblock * kBlock + threadIdx.x;
  const bool live = i < n;
  const float4 x = live ? q[i] : make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  float nearest = 4.0f;
  for (uint32_t k = k0; k < k1; ++k) {
    const float4 op = d.ops[k];
    for (uint32_t start = j0; start < j1; start += kBlock) {
      const uint32_t j = start + threadIdx.x;
      if (j < j1) tile[threadIdx.x] = image(q[j], op, k);
      __syncthreads();
      if (live) {
        const uint32_t cnt = min(kBlock, j1 - start);
        for (uint32_t t0 = 0; t0 < cnt; t0 += 64) {
          PairSum ps{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
          const uint32_t t1 = min(cnt, t0 + 64);
          for (uint32_t t = t0; t < t1; ++t) {
            const float4 y = tile[t];
            const float dm2 = chord2(x, y), dp2 = chord2p(x, y);
            const uint32_t other = start + t;
            if (other != i || k != 0) nearest = fminf(nearest, fminf(dm2, dp2));
            if (other != i) pair_terms<false>(x, y, dm2, dp2, 0.0f, 0.0f, ps);
          }
          acc[0] += ps.fx;
          acc[1] += ps.fy;
          acc[2] += ps.fz;
          acc[3] += ps.curvature;
          acc[4] += ps.energy;
        }
      }
      __syncthreads();
    }
  }
  if (live) {
    const size_t o = size_t(split) * n + i;
#pragma unroll
    for (int c = 0; c < 5; ++c) d.ap_partial[5 * o + c] = acc[c];
    d.ap_pnear[o] = nearest;
  }
}

// The sum of a source's split partials.  Each workgroup takes kReduceSources sources of one
// source block (kReduceParts workgroups per block): lane l of group g adds splits g,
// g + kReduceParts, ... of source l (consecutive lanes read consecutive sources), then one
// lane per source adds the groups' sums in group order.  The order is fixed, so the rows
// are the same bits every run and in any mix; a thread per source summing every split in
// turn left a small sample (N = 1000, 96 splits) with 16 waves on the chip, each a long
// chain of dependent loads and adds.
constexpr uint32_t kReduceSources = 16, kReduceParts = kBlock / kReduceSources;
__global__ __launch_bounds__(256) void k_allpairs_reduce(JobSet s) {
  __shared__ double part[kReduceParts][5][kReduceSources];
  __shared__ float part_near[kReduceParts][kReduceSources];
  const uint32_t block = blockIdx.x / kReduceParts, sub = blockIdx.x % kReduceParts;
  const uint32_t jb = find_job(s.point_start, s.count, block);
  const JobDesc d = load_job(s, jb);
  const uint32_t l = threadIdx.x % kReduceSources, g = threadIdx.x / kReduceSources;
  const uint32_t i = (block - job_start(s.point_start, jb)) * kBlock + sub * kReduceSources + l;
  const bool live = i < d.n;
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  float nearest = 4.0f;
  const uint32_t splits = d.plan.splits();
  if (live)
    for (uint32_t sp = g; sp < splits; sp += kReduceParts) {
      const size_t o = size_t(sp) * d.n + i;
#pragma unroll
      for (int c = 0; c < 5; ++c) a[c] += d.ap_partial[5 * o + c];
      nearest = fminf(nearest, d.ap_pnear[o]);
    }
#pragma unroll
  for (int c = 0; c < 5; ++c) part[g][c][l] = a[c];
  part_near[g][l] = nearest;
  __syncthreads();
  if (g == 0 && live) {
    double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    float tn = 4.0f;
    for (uint32_t h = 0; h < kReduceParts; ++h) {
#pragma unroll
      for (int c = 0; c < 5; ++c) t[c] += part[h][c][l];
      tn = fminf(tn, part_near[h][l]);
    }
    d.rows.force[i] = make_float4(float(t[0]), float(t[1]), float(t[2]), float(t[3]));
    d.rows.energy[i] = t[4];
    d.rows.nearest[i] = tn;
  }
}

namespace {
std::atomic<uint64_t> g_allpairs_override{0};
}  // namespace

void set_allpairs_override(uint32_t target, uint32_t ops_per_split) {
  g_allpairs_override.store(uint64_t(target) << 32 | ops_per_split);
}

AllPairsShape allpairs_shape(uint32_t n, uint32_t order) {
  const uint64_t o = g_allpairs_override.load();
  if (o >> 32) return {uint32_t(o >> 32), uint32_t(o)};
  int row = -1;
  for (int r = 0; r < kTunedOrderCount; ++r)
    if (kTunedOrders[r] == order) row = r;
  if (row < 0) return {};
  int lg = int(std::lround(std::log2(double(std::max<uint32_t>(n, 1)))));
  lg = std::min(std::max(lg, kTunedLogMin), kTunedLogMin + kTunedSizeCount - 1);
  return {kTunedTarget[row][lg - kTunedLogMin], kTunedOps[row][lg - kTunedLogMin]};
}

AllPairsPlan plan_allpairs(uint32_t n, uint32_t order, double budget_scale, AllPairsShape shape) {
  AllPairsPlan p;
  const uint32_t sb = blocks_for(n, kBlock);
  const uint32_t target = std::max(shape.target, 1u);
  p.ops_per_split = 1;
  p.op_splits = order;
  if (shape.ops_per_split) {
    p.ops_per_split = std::min(order, shape.ops_per_split);
    p.op_splits = (order + p.ops_per_split - 1) / p.ops_per_split;
  } else if (uint64_t(sb) * order > 4 * uint64_t(target)) {
    p.ops_per_split = uint32_t((uint64_t(sb) * order + 4 * uint64_t(target) - 1) / (4 * uint64_t(target)));
    p.op_splits = (order + p.ops_per_split - 1) / p.ops_per_split;
  }
  uint32_t chunks = 1;
  if (uint64_t(sb) * p.op_splits < target) chunks = uint32_t((target + sb * p.op_splits - 1) / (sb * p.op_splits));
  // A smaller budget (zo_pair_budget) splits the sites finer: only the rounding changes.
  if (budget_scale > 0.0 && budget_scale < 1.0) chunks = uint32_t(std::min<double>(1e6, chunks / budget_scale));
  // Site chunks are whole tiles of kBlock sites, or, when a small sample has fewer tiles
  // than the target asks for, whole waves of 64: a workgroup's threads each walk a chunk
  // of sites in turn, and with few workgroups that walk is the pass's critical path.
  chunks = std::max(1u, chunks);
  if (chunks <= blocks_for(n, kBlock)) {
    p.site_chunk = blocks_for(n, chunks * kBlock) * kBlock;
  } else {
    chunks = std::min(chunks, blocks_for(n, 64));
    p.site_chunk = blocks_for(n, chunks * 64) * 64;
  }
  p.site_splits = blocks_for(n, p.site_chunk);
  if (budget_scale > 0.0 && budget_scale < 1.0) {
    p.ops_per_split = 1;
    p.op_splits = order;
  }
  return p;
}

void launch_allpairs(Prof& p, hipStream_t s, const JobSet& jobs) {
  if (!jobs.ap_total) return;
  p.run("allpairs", [&] { hipLaunchKernelGGL(k_allpairs, dim3(jobs.ap_total), dim3(kBlock), 0, s, jobs); });
  p.run("allpairs_reduce",
        [&] { hipLaunchKernelGGL(k_allpairs_reduce, dim3(jobs.point_total * kReduceParts), dim3(kBlock), 0, s, jobs); });
}

// Nearest orbit point over all pairs with its rank j*2*order + 2*k + sign; the minimum
// chord^2, ties to the lowest rank.  blockIdx.y is the operator.
// TABLE: the operators come from the compiled-in table (ops holds the group's first row
// index there) instead of a device copy; the same values, so the same bits.
template <bool TABLE>
__global__ __launch_bounds__(256) void k_allpairs_nearest(const float4* __restrict__ q, uint32_t n,
                                                          const float4* __restrict__ ops, uint32_t table_row,
                                                          uint32_t order, uint64_t* __restrict__ partial) {
  __shared__ float4 tile[kBlock];
  const uint32_t k = blockIdx.y;
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool live = i < n;
  const float4 x = live ? q[i] : make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  float4 op;
  if (TABLE) {
    const float* row = kOperators[table_row + k];
    op = make_float4(row[0], row[1], row[2], row[3]);
  } else {
    op = ops[k];
  }
  uint64_t best = ~uint64_t(0);
  for (uint32_t start = 0; start < n; start += kBlock) {
    const uint32_t j = start + threadIdx.x;
    if (j < n) tile[threadIdx.x] = image(q[j], op, k);
    __syncthreads();
    if (live) {
      const uint32_t cnt = min(kBlock, n - start);
      for (uint32_t t = 0; t < cnt; ++t) {
        const float4 y = tile[t];
        const uint32_t jj = start + t;
        const uint32_t r = jj * 2 * order + 2 * k;
        const float dm2 = chord2(x, y), dp2 = chord2p(x, y);
        const uint64_t km = (uint64_t(__float_as_uint(dm2)) << 32) | r;
        const uint64_t kp = (uint64_t(__float_as_uint(dp2)) << 32) | (r + 1);
        if (!(jj == i && k == 0) && km < best) best = km;
        if (kp < best) best = kp;
      }
    }
    __syncthreads();
  }
  if (live) partial[size_t(k) * n + i] = best;
}

__global__ void k_nearest_reduce(const uint64_t* __restrict__ partial, uint32_t n, uint32_t order,
                                 float* __restrict__ chord_out, uint32_t* __restrict__ rank_out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t best = ~uint64_t(0);
  for (uint32_t k = 0; k < order; ++k) {
    const uint64_t v = partial[size_t(k) * n + i];
    best = v < best ? v : best;
  }
  chord_out[i] = best == ~uint64_t(0) ? 4.0f : __uint_as_float(uint32_t(best >> 32));
  rank_out[i] = best == ~uint64_t(0) ? 0xffffffffu : uint32_t(best);
}

void launch_allpairs_nearest(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                             uint64_t* partial, float* chord_out, uint32_t* rank_out) {
  p.run("allpairs_nearest", [&] {
    hipLaunchKernelGGL(k_allpairs_nearest<false>, dim3(blocks_for(n, kBlock), order), dim3(kBlock), 0, s, q, n, ops,
                       0u, order, partial);
  });
  p.run("nearest_reduce", [&] {
    hipLaunchKernelGGL(k_nearest_reduce, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, partial, n, order,
                       chord_out, rank_out);
  });
}


// ---- full measure (pairs.wgsl:151-244) ----------------------------------------------
// Riesz s=1,2,3, Green and log sums over every orbit point but the source itself, in
// native FP64 (the reference carries the log in double-single arithmetic).  Images and
// chords are FP64 from the FP32 positions, normalised as the reference's dense model
// does.  Partial sums per (split, source): r1, r2, r3, log, green.
__device__ inline double4 qmul_d(const double4 a, const double4 b) {
  return make_double4(a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w, a.x * b.y + b.x * a.y + a.z * b.w - a.w * b.z,
                      a.x * b.z + b.x * a.z + a.w * b.y - a.y * b.w, a.x * b.w + b.x * a.w + a.y * b.z - a.z * b.y);
}
__device__ inline double4 normalize_d(const double4 v) {
  const double s = 1.0 / sqrt(v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w);
  return make_double4(v.x * s, v.y * s, v.z * s, v.w * s);
}
__device__ inline double4 to_d(const float4 v) { return make_double4(v.x, v.y, v.z, v.w); }

__global__ __launch_bounds__(256) void k_measure(const float4* __restrict__ q, uint32_t n,
                                                 const float4* __restrict__ ops, uint32_t order,
                                                 uint32_t ops_per_split, uint32_t site_chunk, uint32_t site_splits,
                                                 double* __restrict__ partial) {
  __shared__ double4 tile[kBlock];
  const uint32_t split = blockIdx.y;
  const uint32_t k0 = (split / site_splits) * ops_per_split, k1 = min(order, k0 + ops_per_split);
  const uint32_t j0 = (split % site_splits) * site_chunk, j1 = min(n, j0 + site_chunk);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const bool live = i < n;
  const double4 x = normalize_d(to_d(live ? q[i] : make_float4(1.0f, 0.0f, 0.0f, 0.0f)));
  double r1 = 0.0, r2 = 0.0, r3 = 0.0, lg = 0.0, green = 0.0;
  constexpr double kPiD = 3.14159265358979323846;
  constexpr double kInvFourPi2 = 1.0 / (4.0 * kPiD * kPiD);
  for (uint32_t k = k0; k < k1; ++k) {
    const double4 op = normalize_d(to_d(ops[k]));
    for (uint32_t start = j0; start < j1; start += kBlock) {
      const uint32_t j = start + threadIdx.x;
      if (j < j1) {
        const double4 y = to_d(q[j]);
        tile[threadIdx.x] = normalize_d(k == 0 ? y : qmul_d(y, op));
      }
      __syncthreads();
      if (live) {
        const uint32_t cnt = min(kBlock, j1 - start);
        for (uint32_t t = 0; t < cnt; ++t) {
          if (k == 0 && start + t == i) continue;
          const double4 y = tile[t];
          const double mx = x.x - y.x, my = x.y - y.y, mz = x.z - y.z, mw = x.w - y.w;
          const double px = x.x + y.x, py = x.y + y.y, pz = x.z + y.z, pw = x.w + y.w;
          const double dm2 = mx * mx + my * my + mz * mz + mw * mw;
          const double dp2 = px * px + py * py + pz * pz + pw * pw;
          const double im = rsqrt(dm2), ip = rsqrt(dp2);
          const double im2 = im * im, ip2 = ip * ip;
          r1 += im + ip;
          r2 += im2 + ip2;
          r3 += im2 * im + ip2 * ip;
          lg += 0.5 - 0.5 * log(dm2 * dp2);
          const double a = 2.0 * atan(sqrt(fmin(dm2, dp2) / fmax(dm2, dp2)));
          const double cot = fabs(dp2 - dm2) * 0.5 * im * ip;
          green += ((kPiD - 2.0 * a) * cot - 1.0) * kInvFourPi2;
        }
      }
      __syncthreads();
    }
  }
  if (live) {
    double* o = partial + 5 * (size_t(split) * n + i);
    o[0] = r1;
    o[1] = r2;
    o[2] = r3;
    o[3] = lg;
    o[4] = green;
  }
}

__global__ void k_measure_reduce(const double* __restrict__ partial, uint32_t n, uint32_t splits,
                                 double* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t s = 0; s < splits; ++s)
    for (int c = 0; c < 5; ++c) a[c] += partial[5 * (size_t(s) * n + i) + c];
  for (int c = 0; c < 5; ++c) out[5 * size_t(i) + c] = a[c];
}

void launch_measure(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                    const AllPairsPlan& plan, double* partial, double* out) {
  const uint32_t splits = plan.splits();
  p.run("measure", [&] {
    hipLaunchKernelGGL(k_measure, dim3(blocks_for(n, kBlock), splits), dim3(kBlock), 0, s, q, n, ops, order,
                       plan.ops_per_split, plan.site_chunk, plan.site_splits, partial);
  });
  p.run("measure_reduce", [&] {
    hipLaunchKernelGGL(k_measure_reduce, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, partial, n, splits, out);
  });
}

// ---- scan (exclusive, in place; data[count] receives the total) -----------------------
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kBlock * kScanItems;
size_t scan_tmp_words(size_t count) { return (count + kScanTile - 1) / kScanTile + 1; }

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
  const uint32_t t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (uint32_t d = 1; d < kBlock; d <<= 1) {
    uint32_t o = t >= d ? lds[t - d] : 0u;
    __syncthreads();
    lds[t] += o;
    __syncthreads();
  }
  total = lds[kBlock - 1];
  const uint32_t r = lds[t] - v;
  __syncthreads();
  return r;
}

__device__ __attribute__((always_inline)) inline void scan_tiles_body(
    uint32_t* __restrict__ data, size_t count, uint32_t* __restrict__ tile_sums, uint32_t block, uint32_t grid) {
  __shared__ uint32_t lds[kBlock];
  const size_t base = size_t(block) * kScanTile + size_t(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k) {
    v[k] = base + k < count ? data[base + k] : 0u;
    const uint32_t c = v[k];
    v[k] = sum;
    sum += c;
  }
  uint32_t total;
  const uint32_t off = block_excl_scan(sum, lds, total);
#pragma unroll
  for (uint32_t k = 0; k < kScanItems; ++k)
    if (base + k < count) data[base + k] = v[k] + off;
  if (threadIdx.x == 0) tile_sums[block] = total;
}
__global__ __launch_bounds__(256) void k_scan_tiles(
    uint32_t* __restrict__ data, size_t count, uint32_t* __restrict__ tile_sums) {
  scan_tiles_body(data, count, tile_sums, blockIdx.x, gridDim.x);
}

__device__ __attribute__((always_inline)) inline void scan_sums_body(
    uint32_t* __restrict__ tile_sums, uint32_t ntiles, uint32_t* __restrict__ total_out, uint32_t block,
    uint32_t grid) {
  __shared__ uint32_t lds[kBlock];
  const uint32_t per = (ntiles + kBlock - 1) / kBlock;
  const uint32_t b0 = threadIdx.x * per, b1 = min(ntiles, b0 + per);
  uint32_t sum = 0;
  for (uint32_t b = b0; b < b1; ++b) sum += tile_sums[b];
  uint32_t total;
  uint32_t at = block_excl_scan(sum, lds, total);
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t c = tile_sums[b];
    tile_sums[b] = at;
    at += c;
  }
  if (threadIdx.x == 0) *total_out = total;
}
__global__ __launch_bounds__(256) void k_scan_sums(
    uint32_t* __restrict__ tile_sums, uint32_t ntiles, uint32_t* __restrict__ total_out) {
  scan_sums_body(tile_sums, ntiles, total_out, blockIdx.x, gridDim.x);
}

__device__ __attribute__((always_inline)) inline void scan_add_body(
    uint32_t* __restrict__ data, size_t count, const uint32_t* __restrict__ tile_sums, uint32_t block,
    uint32_t grid) {
  for (size_t i = size_t(block) * blockDim.x + threadIdx.x; i < count; i += size_t(grid) * blockDim.x)
    data[i] += tile_sums[i / kScanTile];
}
__global__ void k_scan_add(uint32_t* __restrict__ data, size_t count, const uint32_t* __restrict__ tile_sums) {
  scan_add_body(data, count, tile_sums, blockIdx.x, gridDim.x);
}

// Exclusive scan of data[0..count) in place, data[count] = total.  tmp: scan_tmp_words(count).
static void launch_scan(Prof& p, hipStream_t s, uint32_t* data, size_t count, uint32_t* tmp) {
  const uint32_t ntiles = uint32_t((count + kScanTile - 1) / kScanTile);
  p.run("scan", [&] {
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, s, data, count, tmp);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, tmp, ntiles, data + count);
    const uint32_t g = uint32_t(std::min<size_t>(8192, (count + kBlock - 1) / kBlock + 1));
    hipLaunchKernelGGL(k_scan_add, dim3(g), dim3(kBlock), 0, s, data, count, tmp);
  });
}

// ---- near-field sites (shell.wgsl:19-130, near.wgsl:22-82) -----------------------------
__device__ inline void near_reset_body(NearInfo* info, uint32_t n) {
  info->sites = n;
  info->overflow = 0;
  info->rim = 0;
  info->wmin_key = 0xffffffffu;
  info->crowded = 0;
  info->huge = 0;
  for (int c = 0; c < 3; ++c) {
    info->lo[c] = 0xffffffffu;
    info->hi[c] = 0u;
  }
}
__global__ void k_near_reset(NearInfo* info, uint32_t n) {
  near_reset_body(info, n);
}

__device__ inline void box_reduce_commit(NearInfo* info, uint32_t lo[3], uint32_t hi[3]) {
  __shared__ uint32_t slo[3][kBlock / 64], shi[3][kBlock / 64];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uint32_t a = lo[c], b = hi[c];
    for (int d = 32; d > 0; d >>= 1) {
      a = min(a, __shfl_xor(a, d));
      b = max(b, __shfl_xor(b, d));
    }
    if (lane == 0) {
      slo[c][wave] = a;
      shi[c][wave] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    uint32_t a = 0xffffffffu, b = 0u;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
      a = min(a, slo[c][w]);
      b = max(b, shi[c][w]);
    }
    if (a != 0xffffffffu) atomicMin(&info->lo[c], a);
    if (b != 0u) atomicMax(&info->hi[c], b);
  }
}

// Slots for a workgroup's items from one atomic on counter: returns each thread's first
// slot for its mine items (in thread order).  Every thread of the workgroup calls it.
// Atomics on one address from every workgroup of the chip serialise (the counter is
// shared by all XCDs), so one per workgroup rather than one per wave.
__device__ inline uint32_t block_reserve(uint32_t* counter, uint32_t mine) {
  __shared__ uint32_t part[kBlock / 64 + 1];
  const uint32_t wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
  const uint32_t incl = wave_incl_scan(mine);
  if ((threadIdx.x & 63) == 63) part[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < waves; ++w) {
      const uint32_t t = part[w];
      part[w] = total;
      total += t;
    }
    part[kBlock / 64] = total ? atomicAdd(counter, total) : 0u;
  }
  __syncthreads();
  const uint32_t at = part[kBlock / 64] + part[wave] + incl - mine;
  __syncthreads();  // part is reused by the next call
  return at;
}

// |q - e|, the chord from the identity, without the cancellation of sqrt(2 - 2 q.w).
__device__ inline float identity_chord(const float4 q) {
  const float d0 = q.x - 1.0f;
  return sqrtf(fmaf(q.w, q.w, fmaf(q.z, q.z, fmaf(q.y, q.y, d0 * d0))));
}

__device__ inline void box_add(uint32_t lo[3], uint32_t hi[3], const float4 y) {
  const float3 p = chart(y);
  const uint32_t k0 = ordered_key(p.x), k1 = ordered_key(p.y), k2 = ordered_key(p.z);
  lo[0] = min(lo[0], k0);
  hi[0] = max(hi[0], k0);
  lo[1] = min(lo[1], k1);
  hi[1] = max(hi[1], k1);
  lo[2] = min(lo[2], k2);
  hi[2] = max(hi[2], k2);
}

// Core sites (every source's identity image, at its own index) and the rim list: the
// sources within the band of some zone plane (shell.wgsl:68-96).
// The workgroup's minimum (maximum, sum) of v into *dst with one atomic, as every
// workgroup's atomics on one address serialise.  Every thread of the workgroup calls.
__device__ inline void block_min_commit(uint32_t* dst, uint32_t v) {
  __shared__ uint32_t sm[kBlock / 64];
  for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) v = min(v, sm[w]);
    if (v != 0xffffffffu) atomicMin(dst, v);
  }
}
__device__ inline void block_max_commit(uint32_t* dst, uint32_t v) {
  __shared__ uint32_t sm[kBlock / 64];
  for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) v = max(v, sm[w]);
    if (v != 0u) atomicMax(dst, v);
  }
}
__device__ inline void block_sum_commit(unsigned long long* dst, unsigned long long v) {
  __shared__ unsigned long long sm[kBlock / 64];
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) v += sm[w];
    if (v != 0u) atomicAdd(dst, v);
  }
}

__device__ __attribute__((always_inline)) inline void core_rim_body(
    const float4* __restrict__ q, uint32_t n, const float4* __restrict__ planes, uint32_t nplanes, float band,
    uint32_t two_order, NearBuffers b, uint32_t block, uint32_t grid) {
  if (n == 0) return;  // uniform: no thread reaches the workgroup's barriers
  uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  uint32_t wmin = 0xffffffffu;
  const uint32_t stride = grid * blockDim.x;
  const uint32_t rounds = (n + stride - 1) / stride;  // at most 32 (core_rim_blocks)
  const uint32_t first = block * blockDim.x + threadIdx.x;
  uint32_t rim_mask = 0;  // bit r: source r * stride + first is a rim source
  // Rounds four at a time, with no branch around a round's loads and stores: a lane past n
  // works on source n - 1 (rewriting its bytes unchanged; the box and the minimum are
  // idempotent) and its rim bit is dropped.  Each batch's positions are loaded while the
  // batch before is processed.  (Stores count in the same in-order counter as loads, so a
  // branch around them made every round wait for the round before's stores.)
  const uint32_t last = n - 1;
  float4 cur[4], nxt[4];
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) cur[j] = q[min(j * stride + first, last)];
  for (uint32_t r0 = 0; r0 < rounds; r0 += 4) {
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) nxt[j] = q[min((r0 + 4 + j) * stride + first, last)];
    float de[4];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const uint32_t i = (r0 + j) * stride + first, ic = min(i, last);
      const float4 x = cur[j];
      b.uq[ic] = x;
      b.ucode[ic] = ic | 0x80000000u;
      b.urank[ic] = ic * two_order;
      box_add(lo, hi, x);
      wmin = min(wmin, ordered_key(x.x));
      de[j] = identity_chord(x);
    }
    // The rim test of the four sources in one walk over the planes (the walk's scalar loads
    // and loop control were most of the kernel when each source walked alone).  For each
    // source: x.p >= p.w - |x - e| (e the identity, |p| = 1) and the planes come nearest
    // first, so once that bound clears the band (with 1e-5 to spare for FP32 rounding), no
    // later plane can hold x within it; a source leaves the walk there or at the first
    // plane that holds it.
    uint32_t open = 0xfu, rim = 0u;  // bit j: source j still walking / a rim source
    for (uint32_t k = 0; k < nplanes && open; ++k) {
      const float4 pl = planes[k];
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const bool past = pl.x - de[j] > band + 2e-6f + 1e-5f;
        const bool held = dot4(cur[j], pl) <= band + 2e-6f;
        const uint32_t bit = 1u << j;
        if ((open & bit) && !past && held) rim |= bit;
        if (past || held) open &= ~bit;
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const uint32_t r = r0 + j;
      rim_mask |= uint32_t((rim >> j) & 1u && r * stride + first < n) << r;
    }
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) cur[j] = nxt[j];
  }
  // One append per workgroup for all its rounds.
  uint32_t at = block_reserve(&b.info->rim, uint32_t(__popc(rim_mask)));
  for (uint32_t m = rim_mask; m; m &= m - 1) b.rim[at++] = uint32_t(__ffs(m) - 1) * stride + first;
  box_reduce_commit(b.info, lo, hi);
  block_min_commit(&b.info->wmin_key, wmin);
}
__global__ __launch_bounds__(256) void k_core_rim(
    const float4* __restrict__ q, uint32_t n, const float4* __restrict__ planes, uint32_t nplanes, float band,
    uint32_t two_order, NearBuffers b) {
  core_rim_body(q, n, planes, nplanes, band, two_order, b, blockIdx.x, gridDim.x);
}

// Workgroups of k_core_rim: at least 512 (measured at I, N = 1e6: 0.06 ms against 0.17 ms
// with 2048), and enough that no thread takes more than 32 sources (its rim flags fit one
// word).
static uint32_t core_rim_blocks(uint32_t n) {
  return std::max<uint32_t>(512, blocks_for(n, kBlock * 32));
}

// Workgroups of k_shell (each takes rim sources round by round): few enough that the
// workgroups' atomics on the site counter and the box stay cheap, as every workgroup's
// atomics on one address serialise.
constexpr uint32_t kShellBlocks = 512;

// The scalar part of x op (the first component of image(x, op, k)): the images x op_k and
// -(x op_k) (images 2k and 2k + 1) have scalar parts w and -w.
__device__ inline float image_scalar(const float4 x, const float4 op) {
  return fmaf(-x.w, op.w, fmaf(-x.z, op.z, fmaf(-x.y, op.y, x.x * op.x)));
}

// Whether image img (1..2|K|-1: operator img/2, negated when odd) of x, whose scalar part
// already clears w_floor, is a shell site: the reference's test, within the band of every
// zone halfspace (shell.wgsl:19-22).  A site matters only within angle r of some source,
// so within acos(min source w) + r of the identity: its scalar part is at least w_floor, a
// cheap, exact necessary condition that spares most images this test.
__device__ inline bool shell_planes(const float4 x, uint32_t img, const float4* __restrict__ ops,
                                    const float4* __restrict__ planes, uint32_t nplanes, float band) {
  const uint32_t k = img >> 1;
  float4 y = image(x, ops[k], k);
  if (img & 1u) y = neg4(y);
  if (!(y.x > -0.9999999f)) return false;
  // Planes nearest first: once y.p >= p.w - |y - e| clears -band (with 1e-5 to spare),
  // every later plane holds y too.
  const float de = identity_chord(y);
  for (uint32_t pk = 0; pk < nplanes; ++pk) {
    const float4 pl = planes[pk];
    if (pl.x - de >= -band - 2e-6f + 1e-5f) break;
    if (dot4(y, pl) < -band - 2e-6f) return false;
  }
  return true;
}

// Of the images first + b for the set bits b of cand, those that are shell sites, as bits.
// Each lane walks its own candidates: the loop runs as often as the lane with the most.
__device__ inline uint32_t shell_keep(const float4 x, uint32_t first, uint32_t cand, const float4* __restrict__ ops,
                                      const float4* __restrict__ planes, uint32_t nplanes, float band) {
  uint32_t keep = 0;
  for (; cand; cand &= cand - 1) {
    const uint32_t bit = uint32_t(__builtin_ctz(cand));
    if (shell_planes(x, first + bit, ops, planes, nplanes, band)) keep |= 1u << bit;
  }
  return keep;
}

// Images (k, sign) != (identity, +) of the rim sources that are shell sites, appended
// after the core (shell.wgsl:97-130).  One thread per rim source, in two passes over its
// images: the scalar-part test for all of them at once (no divergence: a rim source has
// few images that pass it, but some lane of the wave almost always does, so testing the
// planes image by image ran the plane test for nearly every image), then the plane test
// for each lane's own candidates.  It counts its sites, the workgroup reserves room for
// all of its threads' sites with one atomic, then each thread writes its own.  The append
// order varies; the cell sort's in-cell key order makes the result deterministic.
__device__ __attribute__((always_inline)) inline void shell_body(
    const float4* __restrict__ q, uint32_t n, const float4* __restrict__ ops, uint32_t order,
    const float4* __restrict__ planes, uint32_t nplanes, float band, NearBuffers b, uint32_t block, uint32_t grid) {
  uint32_t lo[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, hi[3] = {0u, 0u, 0u};
  const uint32_t m = b.info->rim;
  const float w_floor =
      cosf(fminf(3.14159265f, acosf(fminf(1.0f, unordered_key(b.info->wmin_key))) + b.reach)) - 1e-6f;
  const uint32_t stride = grid * blockDim.x;
  const uint32_t rounds = (m + stride - 1) / stride;
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t t = r * stride + block * blockDim.x + threadIdx.x;
    const bool live = t < m;
    const uint32_t src = live ? b.rim[t] : 0u;
    const float4 x = live ? q[src] : make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    // The passing images, as bits of four words (image img is bit img & 31 of word
    // img >> 5; per <= 119 < 128): the writing pass recomputes only their positions, in
    // the same order, without the tests.
    uint32_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (live) {
      // Operators 16 at a time, one word each: images 2k and 2k + 1 are bits 2 (k & 15)
      // and 2 (k & 15) + 1.  The operators are uniform (scalar loads).
      for (uint32_t kb = 0; kb < order; kb += 16) {
        uint32_t word = 0;
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j)
          if (kb + j < order) {
            const float w = image_scalar(x, ops[kb + j]);
            word |= (!(w < w_floor) ? 1u : 0u) << (2 * j);
            word |= (!(-w < w_floor) ? 2u : 0u) << (2 * j);
          }
        if (kb == 0) m0 = word & ~1u;  // image 0 is x itself: the core site
        else if (kb == 16) m1 = word;
        else if (kb == 32) m2 = word;
        else m3 = word;
      }
      m0 = shell_keep(x, 0, m0, ops, planes, nplanes, band);
      m1 = shell_keep(x, 32, m1, ops, planes, nplanes, band);
      m2 = shell_keep(x, 64, m2, ops, planes, nplanes, band);
      m3 = shell_keep(x, 96, m3, ops, planes, nplanes, band);
    }
    const uint32_t mine = uint32_t(__popc(m0) + __popc(m1) + __popc(m2) + __popc(m3));
    const uint32_t base = block_reserve(&b.info->sites, mine);
    if (live && mine) {
      uint32_t slot = base;
      for (uint32_t word = 0; word < 4; ++word) {
        uint32_t bits = word == 0 ? m0 : word == 1 ? m1 : word == 2 ? m2 : m3;
        for (; bits; bits &= bits - 1) {
          const uint32_t img = word * 32 + uint32_t(__builtin_ctz(bits));
          const uint32_t k = img >> 1;
          float4 y = image(x, ops[k], k);
          if (img & 1u) y = neg4(y);
          if (slot < b.capacity) {
            b.uq[slot] = y;
            b.ucode[slot] = src;
            b.urank[slot] = src * (2 * order) + img;
            box_add(lo, hi, y);
          } else {
            b.info->overflow = 1u;
          }
          ++slot;
        }
      }
    }
  }
  box_reduce_commit(b.info, lo, hi);
}
__global__ __launch_bounds__(256) void k_shell(
    const float4* __restrict__ q, uint32_t n, const float4* __restrict__ ops, uint32_t order,
    const float4* __restrict__ planes, uint32_t nplanes, float band, NearBuffers b) {
  shell_body(q, n, ops, order, planes, nplanes, band, b, blockIdx.x, gridDim.x);
}

__device__ inline void box_finalize_body(NearInfo* info, uint32_t sdim) {
  float lo[3], hi[3];
  float width = 1e-6f;
  for (int c = 0; c < 3; ++c) {
    lo[c] = unordered_key(info->lo[c]);
    hi[c] = unordered_key(info->hi[c]);
    width = fmaxf(width, hi[c] - lo[c]);
  }
  width *= 1.0001f;
  for (int c = 0; c < 3; ++c) info->lower[c] = lo[c] - 1e-5f * width;
  info->cpu = float(sdim) / (width * 1.0001f);
}
__global__ void k_box_finalize(NearInfo* info, uint32_t sdim) {
  box_finalize_body(info, sdim);
}

__device__ inline uint32_t cell_axis(float v, float lower, float cpu, uint32_t dim) {
  const float c = floorf((v - lower) * cpu);
  return uint32_t(fminf(fmaxf(c, 0.0f), float(dim - 1)));
}

// Each site's cell and its arrival slot in that cell (the count before its own).
__device__ __attribute__((always_inline)) inline void site_count_body(NearBuffers b, uint32_t block, uint32_t grid) {
  const uint32_t S = min(b.info->sites, b.capacity);
  const float lx = b.info->lower[0], ly = b.info->lower[1], lz = b.info->lower[2], cpu = b.info->cpu;
  const uint32_t d = b.sdim, dx = b.sdim * b.xdiv;
  const float cpx = cpu * float(b.xdiv);  // x-cells per chart unit
  for (uint32_t s = block * blockDim.x + threadIdx.x; s < S; s += grid * blockDim.x) {
    const float3 p = chart(b.uq[s]);
    const uint32_t c = cell_axis(p.x, lx, cpx, dx) + dx * (cell_axis(p.y, ly, cpu, d) + d * cell_axis(p.z, lz, cpu, d));
    b.ucell[s] = c;
    b.uslot[s] = atomicAdd(&b.offsets[c], 1u);
  }
}
__global__ void k_site_count(NearBuffers b) {
  site_count_body(b, blockIdx.x, gridDim.x);
}

// Cells of up to kThreadCell sites are put in rank order by counting (k_site_rank_place);
// larger ones by sorting an index list (k_cell_sort_waves, k_cell_sort_block).
constexpr uint32_t kThreadCell = 16;

// With the counts scanned into offsets, each site takes its cell's start plus its slot.
// Unordered (spacing): the site itself (position and rank) goes there.  Ordered: a cell of at most kThreadCell
// sites gets the site's rank there (k_site_rank_place then moves each site to its place
// in rank order); a larger one the site's index (perm), for the crowded-cell sorts.
__device__ __attribute__((always_inline)) inline void site_place_body(
    NearBuffers b, uint32_t ordered, uint32_t block, uint32_t grid) {
  const uint32_t S = min(b.info->sites, b.capacity);
  for (uint32_t s = block * blockDim.x + threadIdx.x; s < S; s += grid * blockDim.x) {
    const uint32_t c = b.ucell[s];
    const uint32_t begin = b.offsets[c], dst = begin + b.uslot[s];
    if (!ordered) {  // no site codes: the spacing pass tells identity images by rank
      b.sq[dst] = b.uq[s];
      b.srank[dst] = b.urank[s];
    } else {
      b.perm[dst] = b.offsets[c + 1] - begin > kThreadCell ? s : b.urank[s];
    }
  }
}
__global__ void k_site_place(NearBuffers b, uint32_t ordered) {
  site_place_body(b, ordered, blockIdx.x, gridDim.x);
}

// Moves unsorted site s to sorted position k.
__device__ inline void move_site(const NearBuffers& b, uint32_t k, uint32_t s) {
  b.sq[k] = b.uq[s];
  b.scode[k] = b.ucode[s];
  b.srank[k] = b.urank[s];
}

// In-cell order by the site's unique rank, so sums run in the same order every pass
// (near.wgsl:70-82).  In a cell of at most kThreadCell sites each site's place is the count
// of its cell's ranks below its own (read from the ranks k_site_place laid out; the loads
// are independent, no sorting chain); larger cells are listed for k_cell_sort_waves (by
// the site that arrived first).
__device__ __attribute__((always_inline)) inline void site_rank_place_body(
    NearBuffers b, uint32_t block, uint32_t grid) {
  const uint32_t S = min(b.info->sites, b.capacity);
  for (uint32_t s = block * blockDim.x + threadIdx.x; s < S; s += grid * blockDim.x) {
    const uint32_t c = b.ucell[s];
    const uint32_t begin = b.offsets[c], m = b.offsets[c + 1] - begin;
    if (m > kThreadCell) {
      if (b.uslot[s] == 0) b.crowded[atomicAdd(&b.info->crowded, 1u)] = c;
      continue;
    }
    const uint32_t key = b.urank[s];
    uint32_t below = 0;
    for (uint32_t k = 0; k < m; ++k) below += b.perm[begin + k] < key ? 1u : 0u;
    move_site(b, begin + below, s);
  }
}
__global__ void k_site_rank_place(NearBuffers b) {
  site_rank_place_body(b, blockIdx.x, gridDim.x);
}

// Crowded cells (more than kThreadCell items: coarse cells, or a wide cutoff over a small
// sample crowding the chart's centre), one wave each: up to 64 items placed by rank (the
// count of smaller keys; keys are unique), larger cells sorted in the wave's LDS by a
// bitonic network padded to a power of two.  Cells of more than kWaveChunk items are
// listed for k_cell_sort_block.
constexpr uint32_t kWaveChunk = 1024;
struct SortLds {
  uint32_t k[kWaveChunk];
  uint32_t v[kWaveChunk];
};
__device__ __attribute__((always_inline)) inline void cell_sort_waves_body(
    const uint32_t* __restrict__ offsets, uint32_t* __restrict__ perm, const uint32_t* __restrict__ key,
    const uint32_t* __restrict__ crowded_list, const uint32_t* __restrict__ crowded, uint32_t list_capacity,
    uint32_t* __restrict__ huge_list, uint32_t* __restrict__ huge, NearBuffers sites, uint32_t block,
    uint32_t grid) {
  __shared__ SortLds lds[kBlock / 64];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  SortLds& L = lds[wave];
  const uint32_t total = min(*crowded, list_capacity);
  const uint32_t waves = grid * (kBlock / 64);
  for (uint32_t b = block * (kBlock / 64) + wave; b < total; b += waves) {
    const uint32_t c = crowded_list[b];
    const uint32_t begin = offsets[c], m = offsets[c + 1] - begin;
    if (m > kWaveChunk) {
      if (lane == 0) huge_list[atomicAdd(huge, 1u)] = c;
      continue;
    }
    if (m <= 64) {
      // Up to one item per lane: each lane's rank is the number of smaller keys (unique),
      // read back from LDS as broadcasts; no sorting network.
      const uint32_t v = lane < m ? perm[begin + lane] : 0u;
      const uint32_t k = lane < m ? (key ? key[v] : v) : 0xffffffffu;
      L.k[lane] = k;
      wave_sync();
      uint32_t rank = 0;
      for (uint32_t j = 0; j < m; ++j) rank += L.k[j] < k ? 1u : 0u;
      if (lane < m) {
        perm[begin + rank] = v;
        move_site(sites, begin + rank, v);
      }
      wave_sync();
      continue;
    }
    uint32_t width = 64;
    while (width < m) width <<= 1;
    for (uint32_t t = lane; t < width; t += 64) {
      const uint32_t v = t < m ? perm[begin + t] : 0xffffffffu;
      L.v[t] = v;
      L.k[t] = t < m ? (key ? key[v] : v) : 0xffffffffu;
    }
    wave_sync();
    for (uint32_t size = 2; size <= width; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t t = lane; t < width / 2; t += 64) {
          const uint32_t lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
          const bool up = (lo & size) == 0;
          const uint32_t klo = L.k[lo], khi = L.k[hi];
          if ((klo > khi) == up) {
            L.k[lo] = khi;
            L.k[hi] = klo;
            const uint32_t v = L.v[lo];
            L.v[lo] = L.v[hi];
            L.v[hi] = v;
          }
        }
        wave_sync();
      }
    }
    for (uint32_t t = lane; t < m; t += 64) {
      perm[begin + t] = L.v[t];
      move_site(sites, begin + t, L.v[t]);
    }
    wave_sync();
  }
}
__global__ __launch_bounds__(256) void k_cell_sort_waves(
    const uint32_t* __restrict__ offsets, uint32_t* __restrict__ perm, const uint32_t* __restrict__ key,
    const uint32_t* __restrict__ crowded_list, const uint32_t* __restrict__ crowded, uint32_t list_capacity,
    uint32_t* __restrict__ huge_list, uint32_t* __restrict__ huge, NearBuffers sites) {
  cell_sort_waves_body(offsets, perm, key, crowded_list, crowded, list_capacity, huge_list, huge, sites, blockIdx.x,
    gridDim.x);
}

// Cells of more than kWaveChunk items, one workgroup each: chunks of up to kBlockChunk
// sorted in LDS (bitonic), then, for cells of more than one chunk, every item placed at
// its rank (its index in its chunk plus, per other chunk, the count of smaller keys found
// by binary search) through tmp.
constexpr uint32_t kBlockChunk = 8192;
__device__ __attribute__((always_inline)) inline void cell_sort_block_body(
    const uint32_t* __restrict__ offsets, uint32_t* __restrict__ perm, const uint32_t* __restrict__ key,
    const uint32_t* __restrict__ huge_list, const uint32_t* __restrict__ huge, uint32_t list_capacity,
    uint32_t* __restrict__ tmp, NearBuffers sites, uint32_t block, uint32_t grid) {
  __shared__ uint32_t sk[kBlockChunk], sv[kBlockChunk];
  const uint32_t total = min(*huge, list_capacity);
  for (uint32_t b = block; b < total; b += grid) {
    const uint32_t c = huge_list[b];
    const uint32_t begin = offsets[c], m = offsets[c + 1] - begin;
    for (uint32_t cb = 0; cb < m; cb += kBlockChunk) {
      const uint32_t len = min(kBlockChunk, m - cb);
      uint32_t width = 256;
      while (width < len) width <<= 1;
      for (uint32_t t = threadIdx.x; t < width; t += blockDim.x) {
        const uint32_t v = t < len ? perm[begin + cb + t] : 0xffffffffu;
        sv[t] = v;
        sk[t] = t < len ? (key ? key[v] : v) : 0xffffffffu;
      }
      __syncthreads();
      for (uint32_t size = 2; size <= width; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
          for (uint32_t t = threadIdx.x; t < width / 2; t += blockDim.x) {
            const uint32_t lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
            const bool up = (lo & size) == 0;
            const uint32_t klo = sk[lo], khi = sk[hi];
            if ((klo > khi) == up) {
              sk[lo] = khi;
              sk[hi] = klo;
              const uint32_t v = sv[lo];
              sv[lo] = sv[hi];
              sv[hi] = v;
            }
          }
          __syncthreads();
        }
      }
      for (uint32_t t = threadIdx.x; t < len; t += blockDim.x) perm[begin + cb + t] = sv[t];
      __syncthreads();
    }
    if (m > kBlockChunk) {
      const uint32_t chunks = (m + kBlockChunk - 1) / kBlockChunk;
      for (uint32_t at = threadIdx.x; at < m; at += blockDim.x) {
        const uint32_t v = perm[begin + at];
        const uint32_t kv = key ? key[v] : v;
        const uint32_t own = at / kBlockChunk;
        uint32_t rank = at - own * kBlockChunk;
        for (uint32_t o = 0; o < chunks; ++o) {
          if (o == own) continue;
          uint32_t lo = o * kBlockChunk, hi = min(m, lo + kBlockChunk);  // first key >= kv in [lo, hi)
          const uint32_t base = lo;
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            const uint32_t u = perm[begin + mid];
            if ((key ? key[u] : u) < kv)
              lo = mid + 1;
            else
              hi = mid;
          }
          rank += lo - base;
        }
        tmp[begin + rank] = v;
      }
      __syncthreads();
      for (uint32_t at = threadIdx.x; at < m; at += blockDim.x) perm[begin + at] = tmp[begin + at];
      __syncthreads();
    }
    for (uint32_t at = threadIdx.x; at < m; at += blockDim.x) move_site(sites, begin + at, perm[begin + at]);
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void k_cell_sort_block(
    const uint32_t* __restrict__ offsets, uint32_t* __restrict__ perm, const uint32_t* __restrict__ key,
    const uint32_t* __restrict__ huge_list, const uint32_t* __restrict__ huge, uint32_t list_capacity,
    uint32_t* __restrict__ tmp, NearBuffers sites) {
  cell_sort_block_body(offsets, perm, key, huge_list, huge, list_capacity, tmp, sites, blockIdx.x, gridDim.x);
}

// The sources are visited tile by tile (tiles of tile_edge y/z cells a side, tiles in
// row-major order, cells row-major within a tile), each cell's identity images in their
// sorted order: waves of neighbouring sources read overlapping site runs, and each XCD
// takes a contiguous range of them (xcd_remap) whose neighbourhood its L2 holds.  Visit v
// of the padded tile grid (tile_visits entries) names the cell, or none past the grid's
// edge.
__device__ inline bool visit_cell(uint32_t v, uint32_t sdim, uint32_t xdiv, uint32_t& cell) {
  const uint32_t e = tile_edge(xdiv), ex = e * xdiv, nt = (sdim + e - 1) / e, per = ex * e * e;
  const uint32_t w = v % per, t = v / per;
  const uint32_t x = (t % nt) * ex + w % ex, y = (t / nt % nt) * e + w / ex % e, z = (t / (nt * nt)) * e + w / (ex * e);
  if (x >= sdim * xdiv || y >= sdim || z >= sdim) return false;
  cell = x + sdim * xdiv * (y + sdim * z);
  return true;
}

// ucell (free once the scatter has run) takes each visit's count of identity images: the
// sites whose code has the core bit or, with two_order = 2|K| (unordered builds, which
// keep no codes), whose rank src * 2|K| + image has image 0.
__device__ __attribute__((always_inline)) inline void core_count_body(
    NearBuffers b, uint32_t two_order, uint32_t block, uint32_t grid) {
  const uint32_t visits = tile_visits(b.sdim, b.xdiv);
  for (uint32_t v = block * blockDim.x + threadIdx.x; v < visits; v += grid * blockDim.x) {
    uint32_t cell, count = 0;
    if (visit_cell(v, b.sdim, b.xdiv, cell))
      for (uint32_t k = b.offsets[cell]; k < b.offsets[cell + 1]; ++k)
        count += two_order ? uint32_t(b.srank[k] % two_order == 0) : b.scode[k] >> 31;
    b.ucell[v] = count;
  }
}
__global__ void k_core_count(NearBuffers b, uint32_t two_order) {
  core_count_body(b, two_order, blockIdx.x, gridDim.x);
}

// With ucell scanned: each visited cell writes its sources from its offset on (and, for
// the spacing pass, two_order != 0, their positions).
__device__ __attribute__((always_inline)) inline void core_order_body(
    NearBuffers b, uint32_t two_order, uint32_t block, uint32_t grid) {
  const uint32_t visits = tile_visits(b.sdim, b.xdiv);
  for (uint32_t v = block * blockDim.x + threadIdx.x; v < visits; v += grid * blockDim.x) {
    uint32_t cell;
    if (!visit_cell(v, b.sdim, b.xdiv, cell)) continue;
    uint32_t at = b.ucell[v];
    for (uint32_t k = b.offsets[cell]; k < b.offsets[cell + 1]; ++k) {
      if (two_order) {
        const uint32_t rank = b.srank[k];
        if (rank % two_order == 0) {
          b.order[at] = rank / two_order;
          b.opos[at] = b.sq[k];  // the identity image: the source's position, bit for bit
          ++at;
        }
      } else {
        const uint32_t code = b.scode[k];
        if (code >> 31) b.order[at++] = code & 0x7fffffffu;
      }
    }
  }
}
__global__ void k_core_order(NearBuffers b, uint32_t two_order) {
  core_order_body(b, two_order, blockIdx.x, gridDim.x);
}

static uint32_t stride_grid(size_t work) { return uint32_t(std::min<size_t>(8192, (work + kBlock - 1) / kBlock + 1)); }

void launch_near_sites(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                       const NearBuffers& b) {
  p.run("near_reset", [&] { hipLaunchKernelGGL(k_near_reset, dim3(1), dim3(1), 0, s, b.info, n); });
  p.run("core_rim", [&] {
    hipLaunchKernelGGL(k_core_rim, dim3(std::min(core_rim_blocks(n), blocks_for(n, kBlock))), dim3(kBlock), 0, s, q, n,
                       b.planes, b.nplanes, b.band, 2 * order, b);
  });
  p.run("shell", [&] {
    hipLaunchKernelGGL(k_shell, dim3(kShellBlocks), dim3(kBlock), 0, s, q, n, ops, order, b.planes, b.nplanes, b.band, b);
  });
  p.run("box", [&] { hipLaunchKernelGGL(k_box_finalize, dim3(1), dim3(1), 0, s, b.info, b.sdim); });
}

void launch_near_build(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                       const NearBuffers& b, bool ordered) {
  const size_t cells = cell_count(b.sdim, b.xdiv);
  launch_fill_u32(p, s, b.offsets, 0u, cells + 1);
  launch_near_sites(p, s, q, n, ops, order, b);
  const uint32_t g = stride_grid(b.capacity);
  p.run("site_count", [&] { hipLaunchKernelGGL(k_site_count, dim3(g), dim3(kBlock), 0, s, b); });
  launch_scan(p, s, b.offsets, cells, b.scan_tmp);
  p.run("site_place", [&] { hipLaunchKernelGGL(k_site_place, dim3(g), dim3(kBlock), 0, s, b, ordered ? 1u : 0u); });
  if (ordered) p.run("site_sort", [&] {
    hipLaunchKernelGGL(k_site_rank_place, dim3(g), dim3(kBlock), 0, s, b);
    hipLaunchKernelGGL(k_cell_sort_waves, dim3(kCrowdedBlocks), dim3(kBlock), 0, s, b.offsets, b.perm, b.urank,
                       b.crowded, &b.info->crowded, b.crowded_capacity, b.huge, &b.info->huge, b);
    hipLaunchKernelGGL(k_cell_sort_block, dim3(kHugeBlocks), dim3(kBlock), 0, s, b.offsets, b.perm, b.urank, b.huge,
                       &b.info->huge, b.huge_capacity, b.ucell, b);
  });
  const uint32_t visits = tile_visits(b.sdim, b.xdiv), gv = stride_grid(visits);
  const uint32_t two_order = ordered ? 0u : 2 * order;
  p.run("core_order", [&] { hipLaunchKernelGGL(k_core_count, dim3(gv), dim3(kBlock), 0, s, b, two_order); });
  launch_scan(p, s, b.ucell, visits, b.scan_tmp);
  p.run("core_order", [&] { hipLaunchKernelGGL(k_core_order, dim3(gv), dim3(kBlock), 0, s, b, two_order); });
  // The order only schedules the waves: each source's row is its own.
}

// ---- near-field rows (pairs.wgsl:363-427) -------------------------------------------
// The queue of accepted sites of one wave: x - y and |x - y|^2 exactly as the candidate
// test computed them, one array per component (4-byte strides: a wave's 64 consecutive
// entries fall in distinct banks).  It is a stack: a chunk of candidates pushes its
// accepted sites on top, and the pair stage pops the top 64, so at most 63 + 64 entries
// are ever held and no index wraps.
struct WaveLds {
  float stack[5 * 128];  // entry e: x - y at e, e + 128, e + 256, e + 384; |x - y|^2 at e + 512
};

__device__ inline float row_gap(float c, int idx, uint32_t dim) {
  // Distance (cell units) from c to cell interval [idx, idx+1]; edge cells are unbounded
  // outward, so sites clamped into them are never pruned.
  const float lo = idx == 0 ? -3e38f : float(idx);
  const float hi = idx == int(dim) - 1 ? 3e38f : float(idx + 1);
  return fmaxf(0.0f, fmaxf(lo - c, c - hi));
}

// The wave's FP64 total of a per-lane value, as a wave-uniform value (the same order every
// run).
__device__ inline double wave_total(double v) {
  const uint64_t bits = uint64_t(__double_as_longlong(wave_sum_to_last(v)));
  const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(bits), 63);
  const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(bits >> 32), 63);
  return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

// Raw buffer view of a device array for wave loads: a uniform 32-bit byte offset in a
// scalar register plus a fixed per-lane offset (the last descriptor word is the gfx9
// format's; no range check beyond what the offsets cover).
__device__ inline __amdgpu_buffer_rsrc_t buffer_view(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(0xffffffffu), 0x00020000);
}
__device__ inline float4 buffer_load4(__amdgpu_buffer_rsrc_t r, uint32_t lane_bytes, uint32_t wave_bytes) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, lane_bytes, wave_bytes, 0);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}
__device__ inline uint32_t buffer_load1(__amdgpu_buffer_rsrc_t r, uint32_t lane_bytes, uint32_t wave_bytes) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, lane_bytes, wave_bytes, 0);
}

// Near-field force rows (pairs.wgsl:363-427): per source (one wave) the sum of the
// windowed Quey pair over every orbit point within the cutoff, outside its own orbit.
//
// The cutoff cap is a ball in the chart; up to 64 of the cell rows it crosses are set up
// at once, one per lane, each trimmed in x to the ball (to x-cells xdiv times finer than
// the y/z cells, so a row ends close to the ball).  The wave then walks them row by
// row (uniform): a row's sites are contiguous, so lane l tests the site row start + c + l,
// loaded at a uniform byte offset (a scalar register, advanced by 1 KiB per chunk) plus the
// lane's fixed offset.  A row's full 64-site chunks need no lane mask; its last, partial
// chunk loads past the row's end (the site arrays hold 64 entries beyond their capacity)
// and masks those lanes out of the test.  Accepted candidates are pushed on the wave's LDS
// stack (ballot + mbcnt) and each full 64 runs the pair arithmetic on a full wave; only a
// row's last chunk runs partly empty.  Per-lane FP32 sums go into per-lane FP64 sums every
// 64 rounds, and the wave's totals are taken once at the end.  Scalar work per chunk is
// kept to the loop, the push and the stack count: the scalar unit is shared by the CU's
// four SIMDs, and scalar row masks built for every chunk took it to 79 % of its issue
// rate, cancelling the vector instructions they saved.
//
// OWN: whether the source's own orbit can lie within the cutoff.  x.g is at S3 angle
// acos|g.w| from x for every x, so below the smallest such angle (0.63 rad for I) no own
// image is ever a candidate: then the candidate stage needs no site codes and no test at
// all beyond the cutoff, the source's own identity image (chord exactly 0: the site
// bit-identical to x, rebuilt from q before every pass) being dropped when its pair is
// evaluated.  A zero chord enters the pair arithmetic as chord^2 2 (S3 angle pi/2, beyond
// every window, so its terms vanish exactly), stays out of the nearest chord, and is
// counted; a second one (an exact duplicate of x) is restored at the end.
// TERMS: the arctangent series (atan2x_small).
// vb: the workgroup's index in the source order (4 sources each).
template <bool OWN, int TERMS>
__device__ __attribute__((always_inline)) inline void near_body(
    const float4* __restrict__ q_in, uint32_t n, const uint32_t* __restrict__ order_in,
    const float4* __restrict__ sq_in, const uint32_t* __restrict__ scode_in,
    const uint32_t* __restrict__ offsets_in, const NearInfo* __restrict__ info_in, uint32_t sdim, uint32_t xdiv,
    NearCut cut, Rows rows, uint32_t* __restrict__ counts, uint32_t vb) {
  const auto* __restrict__ q = to_global(q_in);
  const auto* __restrict__ order = to_global(order_in);
  const auto* __restrict__ offsets = to_global(offsets_in);
  const auto* __restrict__ info = to_global(info_in);
  const __amdgpu_buffer_rsrc_t sites = buffer_view(sq_in), codes = buffer_view(scode_in);
  __shared__ WaveLds lds[kBlock / 64];
  // The wave index is uniform: the source, its position and every per-source constant
  // below then live in scalar registers, leaving the VGPRs to the candidate and pair work.
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint32_t slot = vb * (kBlock / 64) + wave;
  if (slot >= n) return;  // the whole wave leaves; no workgroup barrier follows
  WaveLds& L = lds[wave];
  const uint32_t src = order[slot];
  const float4 x = load4(q + src);
  const float cpu = info->cpu;
  const float fdim = float(sdim), xd = float(xdiv), fdx = float(sdim * xdiv);
  const float denom = cut.c0 + x.x;
  const bool full = !(denom > 1e-3f);
  // Chart position in cell units: y and z in y/z cells, x in x-cells (xdiv per y/z cell).
  float cx = 0.5f * fdx, cy = 0.5f * fdim, cz = 0.5f * fdim, rc = 3e18f;
  if (!full) {
    const float inv = 1.0f / denom;
    cx = (x.y * inv - info->lower[0]) * (cpu * xd);
    cy = (x.z * inv - info->lower[1]) * cpu;
    cz = (x.w * inv - info->lower[2]) * cpu;
    rc = (cut.s0 * inv * 1.001f + 1e-6f) * cpu + 1e-3f;
  }
  auto clampc = [&](float v) -> int { return int(fminf(fmaxf(floorf(v), 0.0f), fdim - 1.0f)); };
  auto clampx = [&](float v) -> int { return int(fminf(fmaxf(floorf(v), 0.0f), fdx - 1.0f)); };
  const int zlo = clampc(cz - rc), zhi = clampc(cz + rc), ylo = clampc(cy - rc), yhi = clampc(cy + rc);
  const uint32_t ny = uint32_t(yhi - ylo + 1), nrows = ny * uint32_t(zhi - zlo + 1);
  const float rc2 = rc * rc;
  const uint32_t self_code = src | 0x80000000u;
  const uint32_t lane16 = lane * 16u, lane4 = lane * 4u;

  NearSum ps{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // per lane
  // Chords are sums of squares (>= 0), so their bits order as unsigned integers: one
  // integer min per update, no float canonicalisation.
  uint32_t nearest_bits = __float_as_uint(4.0f);
  uint32_t zeros = 0;  // zero chords met (!OWN): counted from ballots in uniform control flow
  const float u_offset = -cut.taper_start * cut.inv_width;
  const float slope_pi = -6.0f * kPiF * cut.inv_width;  // pi times the slope's factor (g is halved)
  float* top = &L.stack[0];  // the stack's top (uniform): 0 to 63 entries between chunks
  uint32_t rounds = 0;
  uint32_t candidates = 0;  // sites tested (uniform); accepted pairs are 64 rounds + the tail

  // The pair stage for the 64 stack entries from `top` (a uniform address).
  auto process = [&](const float* top) {
    float dm2 = top[lane + 512];
    if (!OWN) {
      const uint64_t zero = __builtin_amdgcn_ballot_w64(dm2 == 0.0f);
      zeros += uint32_t(__popcll(zero));
      // 2 exceeds every accepted chord (< 2), so it stands for "no pair" in the minimum.
      dm2 = __builtin_amdgcn_inverse_ballot_w64(zero) ? 2.0f : dm2;
      nearest_bits = min(nearest_bits, __float_as_uint(dm2));
    }
    const float4 d = make_float4(top[lane], top[lane + 128], top[lane + 256], top[lane + 384]);
    near_pair_terms<TERMS>(d, dm2, cut.inv_width, u_offset, slope_pi, ps);
  };
  // Every 64 rounds the lanes' FP32 sums (64 pairs each) go into their FP64 sums.
  auto flush = [&]() {
    acc[0] += double(ps.gx);
    acc[1] += double(ps.gy);
    acc[2] += double(ps.gz);
    acc[3] += double(ps.gw);
    acc[4] += double(ps.curvature);
    acc[5] += double(ps.energy);
    ps = NearSum{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  };
  // One chunk of 64 sites at byte offset `at` of the site array; `in_row`: its lanes that
  // lie in the row.
  auto chunk = [&](uint32_t at, uint64_t in_row) {
    const float4 y = buffer_load4(sites, lane16, at);
    const float4 d = make_float4(x.x - y.x, x.y - y.y, x.z - y.z, x.w - y.w);
    const float dm2 = dot4(d, d);  // chord2(x, y), the same operations
    uint64_t mask = __builtin_amdgcn_ballot_w64(dm2 <= cut.chord2) & in_row;
    if (OWN) {
      // Every orbit point but the source itself bounds the nearest chord; the source's own
      // images exert no force (pairs.wgsl:404-408).
      const uint32_t code = buffer_load1(codes, lane4, at >> 2);
      if (__builtin_amdgcn_inverse_ballot_w64(mask) && code != self_code)
        nearest_bits = min(nearest_bits, __float_as_uint(dm2));
      mask &= ~__builtin_amdgcn_ballot_w64((code & 0x7fffffffu) == src);
    }
    if (mask) {
      if (__builtin_amdgcn_inverse_ballot_w64(mask)) {
        const uint32_t e =
            __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
        top[e] = d.x;
        top[e + 128] = d.y;
        top[e + 256] = d.z;
        top[e + 384] = d.w;
        top[e + 512] = dm2;
      }
      top += __popcll(mask);
      wave_sync();
      if (top >= &L.stack[64]) {
        // The top 64 entries: the ones below stay where they are.
        top -= 64;
        process(top);
        rounds = __builtin_amdgcn_readfirstlane(rounds + 1);
        if ((rounds & 63u) == 0) flush();
        wave_sync();
      }
    }
  };

  for (uint32_t rb = 0; rb < nrows; rb += 64) {
    const uint32_t r = rb + lane;
    uint32_t start = 0, len = 0;
    if (r < nrows) {
      const int yy = ylo + int(r % ny), zz = zlo + int(r / ny);
      const float gy = row_gap(cy, yy, sdim), gz = row_gap(cz, zz, sdim);
      const float rest = rc2 - gy * gy - gz * gz;
      if (rest >= 0.0f) {
        const float span = full ? 3e18f : sqrtf(rest) * xd;  // in x-cells
        const int x0 = clampx(cx - span), x1 = clampx(cx + span);
        const uint32_t row = sdim * xdiv * (uint32_t(yy) + sdim * uint32_t(zz));
        start = offsets[row + x0];
        len = offsets[row + x1 + 1] - start;
      }
    }
    candidates += uint32_t(__builtin_amdgcn_readlane(int(wave_incl_scan(len)), 63));
    for (uint32_t rr = 0; rr < 64; ++rr) {
      const uint32_t n_row = uint32_t(__builtin_amdgcn_readlane(int(len), int(rr)));
      if (n_row == 0) continue;
      uint32_t at = uint32_t(__builtin_amdgcn_readlane(int(start), int(rr))) * 16u;
      const uint32_t full_end = at + (n_row & ~63u) * 16u;
      for (; at < full_end; at += 1024u) chunk(at, ~0ull);
      if (n_row & 63u) chunk(at, (1ull << (n_row & 63u)) - 1ull);
    }
  }
  // The tail (fewer than 64 entries), its zero chords counted in uniform control flow.
  const uint32_t qcount = uint32_t(top - &L.stack[0]);
  {
    const uint64_t active = (1ull << qcount) - 1ull;
    float dm2 = L.stack[lane + 512];
    if (!OWN) {
      const uint64_t zero = __builtin_amdgcn_ballot_w64(dm2 == 0.0f) & active;
      zeros += uint32_t(__popcll(zero));
      dm2 = __builtin_amdgcn_inverse_ballot_w64(zero) ? 2.0f : dm2;
    }
    if (lane < qcount) {
      if (!OWN) nearest_bits = min(nearest_bits, __float_as_uint(dm2));
      const float4 d = make_float4(L.stack[lane], L.stack[lane + 128], L.stack[lane + 256], L.stack[lane + 384]);
      near_pair_terms<TERMS>(d, dm2, cut.inv_width, u_offset, slope_pi, ps);
    }
  }
  flush();
  double total[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) total[i] = wave_total(acc[i]);
  uint32_t pairs = 64 * rounds + qcount;
  uint32_t nearest_key = wave_min_u32(nearest_bits);
  if (!OWN) {
    if (zeros > 1) {
      // Each duplicate of x is a pair at chord 0 (pairs.wgsl:404-408 keeps it): its terms,
      // clamped by the pair's regularisers, with no force (x - y = 0).
      NearSum z{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      near_pair_terms<TERMS>(make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, cut.inv_width, u_offset, slope_pi, z);
      total[4] += double(zeros - 1) * double(z.curvature);
      total[5] += double(zeros - 1) * double(z.energy);
      nearest_key = 0u;
    }
    if (nearest_key == __float_as_uint(2.0f)) nearest_key = __float_as_uint(4.0f);  // no pair at all
    pairs -= min(zeros, 1u);  // the identity image is no pair
  }
  // The pairs' force coefficients were summed halved and their energies without pi.
  const double g[4] = {2.0 * total[0], 2.0 * total[1], 2.0 * total[2], 2.0 * total[3]};
  const double curvature = total[4], e = total[5] * kPi64;
  const float nearest = __uint_as_float(nearest_key);
  if (lane == 0) {
    double f[3];
    near_force_of(x, g, f);
    rows.force[src] = make_float4(float(f[0]), float(f[1]), float(f[2]), float(curvature));
    rows.energy[src] = e;
    rows.nearest[src] = nearest;
    counts[2 * size_t(src)] = pairs;
    counts[2 * size_t(src) + 1] = candidates;
  }
}

template <bool OWN, int TERMS>
__global__ __launch_bounds__(256) void k_near(const float4* __restrict__ q, uint32_t n,
                                              const uint32_t* __restrict__ order, const float4* __restrict__ sq,
                                              const uint32_t* __restrict__ scode,
                                              const uint32_t* __restrict__ offsets, const NearInfo* __restrict__ info,
                                              uint32_t sdim, uint32_t xdiv, NearCut cut,
                                              uint32_t real_blocks, Rows rows, uint32_t* __restrict__ counts) {
  // Each XCD takes a contiguous range of the tile-by-tile source order, so its L2 holds
  // that range's sites: 420 MB of HBM traffic per pass at I, N = 1e6, against 1061 MB
  // with the workgroups dealt round-robin over the XCDs (at the same kernel time).
  const uint32_t vb = xcd_remap(blockIdx.x, gridDim.x);
  if (vb >= real_blocks) return;
  near_body<OWN, TERMS>(q, n, order, sq, scode, offsets, info, sdim, xdiv, cut, rows, counts, vb);
}

// Sums the per-source pair and candidate counts of a force pass into counters[0..1].
__device__ __attribute__((always_inline)) inline void count_reduce_body(
    const uint32_t* __restrict__ counts, uint32_t n, unsigned long long* __restrict__ counters, uint32_t block,
    uint32_t grid) {
  unsigned long long a = 0, b = 0;
  for (uint32_t i = block * blockDim.x + threadIdx.x; i < n; i += grid * blockDim.x) {
    a += counts[2 * size_t(i)];
    b += counts[2 * size_t(i) + 1];
  }
  block_sum_commit(&counters[0], a);
  block_sum_commit(&counters[1], b);
}
__global__ __launch_bounds__(256) void k_count_reduce(
    const uint32_t* __restrict__ counts, uint32_t n, unsigned long long* __restrict__ counters) {
  count_reduce_body(counts, n, counters, blockIdx.x, gridDim.x);
}

void launch_near_force(Prof& p, hipStream_t s, const float4* q, uint32_t n, const NearBuffers& b, const NearCut& cut,
                       Rows out) {
  const uint32_t real = blocks_for(n, kBlock / 64);
  p.run("near_force", [&] {
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(xcd_grid(real)), dim3(kBlock), 0, s, q, n, b.order, b.sq, b.scode, b.offsets,
                         b.info, b.sdim, b.xdiv, cut, real, out, b.counts);
    };
    const int terms = atan_terms(cut.chord2);
    if (cut.own) {
      switch (terms) {
        case 2: go(k_near<true, 2>); break;
        case 3: go(k_near<true, 3>); break;
        case 4: go(k_near<true, 4>); break;
        case 5: go(k_near<true, 5>); break;
        default: go(k_near<true, 0>); break;
      }
    } else {
      switch (terms) {
        case 2: go(k_near<false, 2>); break;
        case 3: go(k_near<false, 3>); break;
        case 4: go(k_near<false, 4>); break;
        case 5: go(k_near<false, 5>); break;
        default: go(k_near<false, 0>); break;
      }
    }
  });
  launch_fill_u32(p, s, reinterpret_cast<uint32_t*>(b.counters), 0u, 4);
  p.run("count_reduce", [&] {
    hipLaunchKernelGGL(k_count_reduce, dim3(std::min<uint32_t>(256, blocks_for(n, kBlock))), dim3(kBlock), 0, s,
                       b.counts, n, b.counters);
  });
}

// ---- near field of many samples in one launch per stage (a batch step) ------------
uint32_t near_force_key(const NearCut& cut) { return (cut.own ? 8u : 0u) + uint32_t(atan_terms(cut.chord2)); }

// Workgroups per job and stage.  The per-site and per-cell stages take 8 items per
// thread: a workgroup first finds its job (a search over the stage's prefix sums), so
// each should do more than one round of work (the site stages: 6.7 -> 4.6 ms per step of
// the 256-job mix).  The visit stages keep one visit per thread: each walks its cell's
// sites, and fewer threads leave too little parallelism (0.8 -> 1.3 ms per step).
uint32_t near_stage_blocks(const NearJob& j, uint32_t stage) {
  const size_t cells = cell_count(j.b.sdim, j.b.xdiv), visits = tile_visits(j.b.sdim, j.b.xdiv);
  const uint32_t per = kBlock * 8;
  switch (stage) {
    case kNsFill: return std::min<uint32_t>(4096, blocks_for(cells + 1, per));
    case kNsRim: return std::min(core_rim_blocks(j.n), blocks_for(j.n, kBlock));
    case kNsShell: return std::min<uint32_t>(kShellBlocks, blocks_for(j.n, kBlock));  // rim sources <= n
    case kNsSites: return std::min<uint32_t>(8192, blocks_for(j.b.capacity, per));
    case kNsCellTiles: return blocks_for(cells, kScanTile);
    case kNsCellAdd: return std::min<uint32_t>(8192, blocks_for(cells, per));
    case kNsCrowded: return std::min<uint32_t>(kCrowdedBlocks, blocks_for(j.b.crowded_capacity, kBlock / 64));
    case kNsHuge: return std::min<uint32_t>(kHugeBlocks, j.b.huge_capacity);
    case kNsVisits: return stride_grid(visits);
    case kNsVisitTiles: return blocks_for(visits, kScanTile);
    case kNsVisitAdd: return std::min<uint32_t>(8192, blocks_for(visits, per));
    case kNsForce: return blocks_for(j.n, kBlock / 64);
    case kNsReduce: return std::min<uint32_t>(256, blocks_for(j.n, kBlock));
    default: return 0;
  }
}

namespace {
// The job whose workgroup range of `stage` holds workgroup g (uniform), with the
// workgroup's index within the job's range and the range's size.
struct JobPart {
  const NearJob* job;
  uint32_t block, grid;
};
__device__ inline JobPart near_part(const NearSet& set, uint32_t stage, uint32_t g) {
  const uint32_t* st = set.start + size_t(stage) * (set.count + 1);
  uint32_t lo = 0, hi = set.count;  // st[lo] <= g < st[hi]; a job with no workgroups is never found
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (st[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return {set.table + lo, g - st[lo], st[lo + 1] - st[lo]};
}
}  // namespace

// Per job: the pass's status and the force pass's counters start again.
__global__ void k_jobs_reset(NearSet set) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= set.count) return;
  const NearJob& job = set.table[j];
  near_reset_body(job.b.info, job.n);
  job.b.counters[0] = 0ull;
  job.b.counters[1] = 0ull;
}
__global__ void k_jobs_fill_offsets(NearSet set) {
  const JobPart w = near_part(set, kNsFill, blockIdx.x);
  const size_t count = cell_count(w.job->b.sdim, w.job->b.xdiv) + 1;
  for (size_t i = size_t(w.block) * blockDim.x + threadIdx.x; i < count; i += size_t(w.grid) * blockDim.x)
    w.job->b.offsets[i] = 0u;
}
__global__ __launch_bounds__(256) void k_jobs_core_rim(NearSet set) {
  const JobPart w = near_part(set, kNsRim, blockIdx.x);
  const NearJob& j = *w.job;
  core_rim_body(j.q, j.n, j.b.planes, j.b.nplanes, j.b.band, 2 * j.order, j.b, w.block, w.grid);
}
__global__ __launch_bounds__(256) void k_jobs_shell(NearSet set) {
  const JobPart w = near_part(set, kNsShell, blockIdx.x);
  const NearJob& j = *w.job;
  shell_body(j.q, j.n, j.ops, j.order, j.b.planes, j.b.nplanes, j.b.band, j.b, w.block, w.grid);
}
__global__ void k_jobs_box(NearSet set) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < set.count) box_finalize_body(set.table[j].b.info, set.table[j].b.sdim);
}
// stage 0 count, 1 place (ordered), 2 rank place.
__global__ void k_jobs_sites(NearSet set, uint32_t stage) {
  const JobPart w = near_part(set, kNsSites, blockIdx.x);
  if (stage == 0)
    site_count_body(w.job->b, w.block, w.grid);
  else if (stage == 1)
    site_place_body(w.job->b, 1u, w.block, w.grid);
  else
    site_rank_place_body(w.job->b, w.block, w.grid);
}
// Scans of the cell counts (visits 0: offsets over the grid's cells) or of the visit counts
// (visits 1: ucell over tile_visits), each in its job's scan_tmp.
__device__ inline void scan_of(const NearJob& j, uint32_t visits, uint32_t*& data, size_t& count) {
  const uint32_t d = j.b.sdim;
  data = visits ? j.b.ucell : j.b.offsets;
  count = visits ? size_t(tile_visits(d, j.b.xdiv)) : cell_count(d, j.b.xdiv);
}
__global__ __launch_bounds__(256) void k_jobs_scan_tiles(NearSet set, uint32_t visits) {
  const JobPart w = near_part(set, visits ? kNsVisitTiles : kNsCellTiles, blockIdx.x);
  uint32_t* data;
  size_t count;
  scan_of(*w.job, visits, data, count);
  scan_tiles_body(data, count, w.job->b.scan_tmp, w.block, w.grid);
}
__global__ __launch_bounds__(256) void k_jobs_scan_sums(NearSet set, uint32_t visits) {
  const NearJob& j = set.table[blockIdx.x];
  uint32_t* data;
  size_t count;
  scan_of(j, visits, data, count);
  scan_sums_body(j.b.scan_tmp, uint32_t((count + kScanTile - 1) / kScanTile), data + count, 0u, 1u);
}
__global__ void k_jobs_scan_add(NearSet set, uint32_t visits) {
  const JobPart w = near_part(set, visits ? kNsVisitAdd : kNsCellAdd, blockIdx.x);
  uint32_t* data;
  size_t count;
  scan_of(*w.job, visits, data, count);
  scan_add_body(data, count, w.job->b.scan_tmp, w.block, w.grid);
}
__global__ __launch_bounds__(256) void k_jobs_sort_waves(NearSet set) {
  const JobPart w = near_part(set, kNsCrowded, blockIdx.x);
  const NearBuffers& b = w.job->b;
  cell_sort_waves_body(b.offsets, b.perm, b.urank, b.crowded, &b.info->crowded, b.crowded_capacity, b.huge,
                       &b.info->huge, b, w.block, w.grid);
}
__global__ __launch_bounds__(256) void k_jobs_sort_block(NearSet set) {
  const JobPart w = near_part(set, kNsHuge, blockIdx.x);
  const NearBuffers& b = w.job->b;
  cell_sort_block_body(b.offsets, b.perm, b.urank, b.huge, &b.info->huge, b.huge_capacity, b.ucell, b, w.block,
                       w.grid);
}
// stage 0 visit counts, 1 source order.
__global__ void k_jobs_visits(NearSet set, uint32_t stage) {
  const JobPart w = near_part(set, kNsVisits, blockIdx.x);
  if (stage == 0)
    core_count_body(w.job->b, 0u, w.block, w.grid);
  else
    core_order_body(w.job->b, 0u, w.block, w.grid);
}
// The force workgroups of the jobs of one variant: [first, first + blocks) of the stage.
template <bool OWN, int TERMS>
__global__ __launch_bounds__(256) void k_jobs_near(NearSet set, uint32_t first, uint32_t blocks) {
  const uint32_t vb = xcd_remap(blockIdx.x, gridDim.x);
  if (vb >= blocks) return;
  const JobPart w = near_part(set, kNsForce, first + vb);
  const NearJob& j = *w.job;
  near_body<OWN, TERMS>(j.q, j.n, j.b.order, j.b.sq, j.b.scode, j.b.offsets, j.b.info, j.b.sdim, j.b.xdiv, j.cut, j.rows,
                        j.b.counts, w.block);
}
__global__ __launch_bounds__(256) void k_jobs_count_reduce(NearSet set) {
  const JobPart w = near_part(set, kNsReduce, blockIdx.x);
  count_reduce_body(w.job->b.counts, w.job->n, w.job->b.counters, w.block, w.grid);
}

void launch_near_jobs(Prof& p, hipStream_t s, const NearSet& set) {
  if (!set.count) return;
  const uint32_t per_job = blocks_for(set.count, kBlock);
  auto stage = [&](const char* name, uint32_t st, auto kernel, auto... args) {
    if (!set.total[st]) return;
    p.run(name, [&] { hipLaunchKernelGGL(kernel, dim3(set.total[st]), dim3(kBlock), 0, s, set, args...); });
  };
  p.run("near_reset", [&] { hipLaunchKernelGGL(k_jobs_reset, dim3(per_job), dim3(kBlock), 0, s, set); });
  stage("fill", kNsFill, k_jobs_fill_offsets);
  stage("core_rim", kNsRim, k_jobs_core_rim);
  stage("shell", kNsShell, k_jobs_shell);
  p.run("box", [&] { hipLaunchKernelGGL(k_jobs_box, dim3(per_job), dim3(kBlock), 0, s, set); });
  stage("site_count", kNsSites, k_jobs_sites, 0u);
  stage("scan", kNsCellTiles, k_jobs_scan_tiles, 0u);
  p.run("scan", [&] { hipLaunchKernelGGL(k_jobs_scan_sums, dim3(set.count), dim3(kBlock), 0, s, set, 0u); });
  stage("scan", kNsCellAdd, k_jobs_scan_add, 0u);
  stage("site_place", kNsSites, k_jobs_sites, 1u);
  stage("site_sort", kNsSites, k_jobs_sites, 2u);
  stage("site_sort", kNsCrowded, k_jobs_sort_waves);
  stage("site_sort", kNsHuge, k_jobs_sort_block);
  stage("core_order", kNsVisits, k_jobs_visits, 0u);
  stage("scan", kNsVisitTiles, k_jobs_scan_tiles, 1u);
  p.run("scan", [&] { hipLaunchKernelGGL(k_jobs_scan_sums, dim3(set.count), dim3(kBlock), 0, s, set, 1u); });
  stage("scan", kNsVisitAdd, k_jobs_scan_add, 1u);
  stage("core_order", kNsVisits, k_jobs_visits, 1u);
  for (uint32_t k = 0; k < set.keys; ++k) {
    const uint32_t first = set.key_first[k], blocks = set.key_blocks[k];
    p.run("near_force", [&] {
      auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(xcd_grid(blocks)), dim3(kBlock), 0, s, set, first, blocks);
      };
      switch (set.key[k]) {
        case 8 + 2: go(k_jobs_near<true, 2>); break;
        case 8 + 3: go(k_jobs_near<true, 3>); break;
        case 8 + 4: go(k_jobs_near<true, 4>); break;
        case 8 + 5: go(k_jobs_near<true, 5>); break;
        case 8 + 0: go(k_jobs_near<true, 0>); break;
        case 2: go(k_jobs_near<false, 2>); break;
        case 3: go(k_jobs_near<false, 3>); break;
        case 4: go(k_jobs_near<false, 4>); break;
        case 5: go(k_jobs_near<false, 5>); break;
        default: go(k_jobs_near<false, 0>); break;
      }
    });
  }
  stage("count_reduce", kNsReduce, k_jobs_count_reduce);
}

// Nearest orbit point within the cutoff for spacing (pairs.zig:318-335): a small cap (4
// mean radii, a few dozen candidates in a handful of short cell rows), so a group of
// kSpacingLanes lanes owns a source: the group sets up up to kSpacingLanes rows at once,
// one per lane, then walks them (group-uniform), its lanes testing consecutive sites of a
// row; the key (chord^2 bits, rank) makes the minimum take the lowest rank on ties.  The
// source's own identity image is skipped by its rank.
//
// The search runs first over the cap of half the reach (an eighth of the volume, about 7
// orbit points on average): when its least chord^2 is within that cap's cutoff, every site
// that could be nearer (or tie) lies in the cells it scanned, so it is the minimum over the
// whole cutoff, the reference's result.  Only the rare sources with nothing that near are
// searched again over the full cap.
constexpr uint32_t kSpacingLanes = 8;
__global__ __launch_bounds__(256) void k_spacing(uint32_t n, const uint32_t* __restrict__ order,
                                                 const float4* __restrict__ sq,
                                                 const uint32_t* __restrict__ srank,
                                                 const uint32_t* __restrict__ offsets,
                                                 const NearInfo* __restrict__ info, uint32_t sdim, uint32_t xdiv,
                                                 NearCut cut, uint32_t two_order, uint32_t real_blocks,
                                                 const float4* __restrict__ opos, uint64_t* __restrict__ key_out) {
  constexpr uint32_t G = kSpacingLanes;
  // Each XCD takes a contiguous range of the tile-by-tile source order (as in k_near), so
  // its L2 holds that range's sites.
  const uint32_t vb = xcd_remap(blockIdx.x, gridDim.x);
  if (vb >= real_blocks) return;
  const uint32_t sub = threadIdx.x % G;
  const uint32_t slot = vb * (kBlock / G) + threadIdx.x / G;
  if (slot >= n) return;  // the whole group leaves (groups never straddle a wave)
  const uint32_t src = order[slot];
  const float4 x = opos[slot];
  const float cpu = info->cpu;
  const float fdim = float(sdim), xd = float(xdiv), fdx = float(sdim * xdiv);
  const float lower0 = info->lower[0], lower1 = info->lower[1], lower2 = info->lower[2];
  auto clampc = [&](float v) -> int { return int(fminf(fmaxf(floorf(v), 0.0f), fdim - 1.0f)); };
  auto clampx = [&](float v) -> int { return int(fminf(fmaxf(floorf(v), 0.0f), fdx - 1.0f)); };
  const uint32_t self_rank = src * two_order;
  uint64_t best = ~uint64_t(0);
  // Every site in the cells the chart ball of the cap {y : x.y >= c0} (radius s0 / (c0 + w))
  // crosses, rows trimmed in x.
  auto scan = [&](float c0, float s0) {
    const float denom = c0 + x.x;
    const bool full = !(denom > 1e-3f);
    // Chart position in cell units: y and z in y/z cells, x in x-cells (xdiv per y/z cell).
    float cx = 0.5f * fdx, cy = 0.5f * fdim, cz = 0.5f * fdim, rc = 3e18f;
    if (!full) {
      const float inv = 1.0f / denom;
      cx = (x.y * inv - lower0) * (cpu * xd);
      cy = (x.z * inv - lower1) * cpu;
      cz = (x.w * inv - lower2) * cpu;
      rc = (s0 * inv * 1.001f + 1e-6f) * cpu + 1e-3f;
    }
    const int zlo = clampc(cz - rc), zhi = clampc(cz + rc), ylo = clampc(cy - rc), yhi = clampc(cy + rc);
    const uint32_t ny = uint32_t(yhi - ylo + 1), nrows = ny * uint32_t(zhi - zlo + 1);
    const float rc2 = rc * rc;
    for (uint32_t rb = 0; rb < nrows; rb += G) {
      const uint32_t r = rb + sub;
      uint32_t start = 0, len = 0;
      if (r < nrows) {
        const int yy = ylo + int(r % ny), zz = zlo + int(r / ny);
        const float gy = row_gap(cy, yy, sdim), gz = row_gap(cz, zz, sdim);
        const float rest = rc2 - gy * gy - gz * gz;
        if (rest >= 0.0f) {
          const float span = full ? 3e18f : sqrtf(rest) * xd;  // in x-cells
          const int x0 = clampx(cx - span), x1 = clampx(cx + span);
          const uint32_t row = sdim * xdiv * (uint32_t(yy) + sdim * uint32_t(zz));
          start = offsets[row + x0];
          len = offsets[row + x1 + 1] - start;
        }
      }
      // The group's rows as one run of candidates: candidate t lies in the row k whose
      // inclusive prefix is the first above t (a binary search over the group's lanes), so
      // each lane's candidates load independently of the others'.
      uint32_t incl = len;
      for (uint32_t d = 1; d < G; d <<= 1) {
        const uint32_t below = uint32_t(__shfl_up(int(incl), d, int(G)));
        if (sub >= d) incl += below;
      }
      const uint32_t total = uint32_t(__shfl(int(incl), int(G - 1), int(G)));
      const uint32_t first = start - (incl - len);  // start of the row minus the candidates before it
      // Group-uniform trip count: every lane of the group takes part in the shuffles.  Two
      // candidates per lane per round, both loads issued before either is used.
      auto test = [&](const float4 y, uint32_t rank) {
        const float dm2 = chord2(x, y);
        if (dm2 <= cut.chord2 && rank != self_rank) {
          const uint64_t key = (uint64_t(__float_as_uint(dm2)) << 32) | rank;
          best = key < best ? key : best;
        }
      };
      for (uint32_t base = 0; base < total; base += 2 * G) {
        const uint32_t t0 = base + sub, t1 = t0 + G;
        uint32_t k0 = 0, k1 = 0;
        for (uint32_t step = G / 2; step > 0; step >>= 1) {
          if (uint32_t(__shfl(int(incl), int(k0 + step - 1), int(G))) <= t0) k0 += step;
          if (uint32_t(__shfl(int(incl), int(k1 + step - 1), int(G))) <= t1) k1 += step;
        }
        const uint32_t site0 = uint32_t(__shfl(int(first), int(k0), int(G))) + t0;
        const uint32_t site1 = uint32_t(__shfl(int(first), int(k1), int(G))) + t1;
        float4 y0 = x, y1 = x;
        uint32_t r0 = self_rank, r1 = self_rank;
        if (t0 < total) {
          y0 = sq[site0];
          r0 = srank[site0];
        }
        if (t1 < total) {
          y1 = sq[site1];
          r1 = srank[site1];
        }
        if (t0 < total) test(y0, r0);
        if (t1 < total) test(y1, r1);
      }
    }
    for (uint32_t d = G / 2; d > 0; d >>= 1) {
      const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(best)), int(d), int(G)));
      const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(best >> 32)), int(d), int(G)));
      const uint64_t o = (uint64_t(hi) << 32) | lo;
      best = o < best ? o : best;
    }
  };
  scan(cut.inner_c0, cut.inner_s0);
  // Chords are nonnegative, so their bits order as unsigned integers; best is group-uniform.
  if (uint32_t(best >> 32) > __float_as_uint(cut.inner_chord2)) scan(cut.c0, cut.s0);
  // One 8-byte store per source (the sources' indices are scattered): chord^2 bits and rank.
  if (sub == 0) key_out[src] = best;
}

// Spacing keys (chord^2 bits, rank; all ones for none within the cutoff) to SO(3) degrees
// (as k_degrees computes them) and ranks, and the least degree's bits into *least.
__global__ void k_nearest_unpack(const uint64_t* __restrict__ keys, uint32_t n, float* __restrict__ degrees,
                                 uint32_t* __restrict__ rank_out, uint32_t* __restrict__ least) {
  uint32_t lo = 0xffffffffu;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint64_t key = keys[i];
    const float chord = key == ~uint64_t(0) ? 4.0f : __uint_as_float(uint32_t(key >> 32));
    const float d = 4.0f * asinf(fminf(1.0f, sqrtf(chord) * 0.5f)) * (180.0f / kPiF);
    degrees[i] = d;
    rank_out[i] = key == ~uint64_t(0) ? 0xffffffffu : uint32_t(key);
    lo = min(lo, __float_as_uint(d));
  }
  block_min_commit(least, lo);
}

void launch_near_nearest(Prof& p, hipStream_t s, uint32_t n, uint32_t order, const NearBuffers& b, const NearCut& cut,
                         float* degrees_out, uint32_t* rank_out, uint32_t* least_bits) {
  const uint32_t real = blocks_for(n, kBlock / kSpacingLanes);
  uint64_t* keys = reinterpret_cast<uint64_t*>(b.counts);  // 2n words: one key per source
  p.run("near_nearest", [&] {
    hipLaunchKernelGGL(k_spacing, dim3(xcd_grid(real)), dim3(kBlock), 0, s, n, b.order, b.sq, b.srank, b.offsets, b.info,
                       b.sdim, b.xdiv, cut, 2 * order, real, b.opos, keys);
  });
  launch_fill_u32(p, s, least_bits, 0xffffffffu, 1);
  p.run("degrees", [&] {
    hipLaunchKernelGGL(k_nearest_unpack, dim3(std::min<uint32_t>(2048, blocks_for(n, kBlock))), dim3(kBlock), 0, s, keys,
                       n, degrees_out, rank_out, least_bits);
  });
}

// ---- covering radius (covering.zig:177-280 and :555-577: the largest empty cap) ---------
// The covering radius of the sample is the largest angle from any rotation to its nearest
// orbit point: the farthest vertex of any orbit point's Voronoi cell on S3 (the cloud of all
// signed images).  One wave per source p.  In the gnomonic chart at p, t in the tangent
// space spanned by p.i, p.j, p.k stands for the direction p + t, at angle atan|t| from p;
// the bisector with a neighbour y is the plane n.t <= d, n = (y.(p i), y.(p j), y.(p k)),
// d = 1 - y.p.  The cell's vertices are the points where three planes meet and no plane is
// violated; the kernel enumerates plane triples in FP64 (64 lanes split them), which stays
// exact in the cospherical configurations of lattice grids, where only the largest vertex
// distance matters and not the cell's combinatorics.
//   The cell starts as the chart box |t_k| <= kCoverBox and is clipped by the neighbours in
// distance windows (at most kCoverWindow new planes each): the window's planes join the
// current faces, a first pass over the faces and the 32 nearest new planes bounds the
// cell's radius by theta1 (new planes at angle >= 2 theta1 cannot cut it), a second pass
// gives the cell, and only its faces are kept (the cell only shrinks, so a plane that
// misses it now misses it for good).  The cell is certified when 2 theta is below the
// distance of every unseen neighbour; the windows grow up to c_max.
// Candidates held while a window is chosen (a multiple of 64).  With 192 the wave's LDS
// (26 KB with vertices, 18 KB without) lets 6 (8) waves share a CU instead of 5 with
// 256; 128 narrows the windows so far that some samples' cells are not enclosed.
constexpr uint32_t kCoverGather = 192;
constexpr uint32_t kCoverWindow = 64;   // new planes per window
constexpr uint32_t kCoverPlanes = 320;  // faces kept plus one window
constexpr uint32_t kCoverFirst = 32;
constexpr double kCoverBox = 8.0;       // |t_k| <= 8: cells within ~82 degrees (S3) of p

constexpr uint32_t kCoverVertices = 192;  // vertices of the last cell kept for output

// VERTS: the cell's vertices are kept for output (vertex and tetrahedron lists); a measuring
// pass leaves them out, and the 8 KB it saves let 8 waves share a CU instead of 6.
template <bool VERTS>
struct CoverLds {
  float4 y[kCoverGather];
  float d2[kCoverGather];
  uint32_t code[kCoverGather];  // the candidate's orbit code j*2|K| + 2k + sign
  double pl[kCoverPlanes][4];  // n.x, n.y, n.z, d: current faces, then the window nearest first
  float pd2[kCoverPlanes];     // chord^2 of the plane's neighbour
  uint32_t pcode[kCoverPlanes];  // its orbit code
  uint8_t face[kCoverPlanes];  // the plane bounds the cell (some vertex lies on it)
  float cut;
  double vt[VERTS ? kCoverVertices : 1][4];    // the cell's vertices (t, |t|^2), with `faces`
  uint32_t vcode[VERTS ? kCoverVertices : 1][3];  // the codes of the three planes through each (~0: box)
  uint32_t nv;
};

struct CoverVertex {
  double r2;
  double t[3];
  uint32_t plane[3];  // the three planes through it (indices of the pass; >= m: box)
};

// Plane i of a pass over m neighbour planes: i < m a neighbour, else one of the six faces of
// the box.
template <bool VERTS>
__device__ inline void cover_plane(const CoverLds<VERTS>& L, uint32_t m, uint32_t i, double* n, double& d) {
  if (i < m) {
    n[0] = L.pl[i][0];
    n[1] = L.pl[i][1];
    n[2] = L.pl[i][2];
    d = L.pl[i][3];
    return;
  }
  const uint32_t f = i - m;
  n[0] = n[1] = n[2] = 0.0;
  n[f >> 1] = (f & 1) ? -1.0 : 1.0;
  d = kCoverBox;
}

// The farthest vertex of the cell of m neighbour planes and the box, over the wave; with
// `faces`, every vertex is checked and the planes through one are marked in L.face.
template <bool VERTS>
__device__ CoverVertex cover_cell(CoverLds<VERTS>& L, uint32_t m, bool faces) {
  const uint32_t lane = lane_id();
  const uint32_t total = m + 6;
  CoverVertex best{-1.0, {0.0, 0.0, 0.0}, {~0u, ~0u, ~0u}};
  for (uint32_t a = 0; a + 2 < total; ++a) {
    double na[3], da;
    cover_plane(L, m, a, na, da);
    const uint32_t r = total - a - 1;
    const uint32_t pairs = r * (r - 1) / 2;
    for (uint32_t mm = lane; mm < pairs; mm += 64) {
      // (u, v), u < v < r, the mm-th pair in lexicographic order.
      const double w = double(2 * r - 1);
      uint32_t u = uint32_t((w - sqrt(fmax(w * w - 8.0 * double(mm), 0.0))) * 0.5);
      auto first = [&](uint32_t k) { return k * (2 * r - k - 1) / 2; };
      while (u + 1 < r && first(u + 1) <= mm) ++u;
      while (u > 0 && first(u) > mm) --u;
      const uint32_t v = mm - first(u) + u + 1;
      const uint32_t b = a + 1 + u, c = a + 1 + v;
      double nb[3], db, nc[3], dc;
      cover_plane(L, m, b, nb, db);
      cover_plane(L, m, c, nc, dc);
      const double bc[3] = {nb[1] * nc[2] - nb[2] * nc[1], nb[2] * nc[0] - nb[0] * nc[2], nb[0] * nc[1] - nb[1] * nc[0]};
      const double ca[3] = {nc[1] * na[2] - nc[2] * na[1], nc[2] * na[0] - nc[0] * na[2], nc[0] * na[1] - nc[1] * na[0]};
      const double ab[3] = {na[1] * nb[2] - na[2] * nb[1], na[2] * nb[0] - na[0] * nb[2], na[0] * nb[1] - na[1] * nb[0]};
      const double det = na[0] * bc[0] + na[1] * bc[1] + na[2] * bc[2];
      const double scale = sqrt((na[0] * na[0] + na[1] * na[1] + na[2] * na[2]) *
                                (nb[0] * nb[0] + nb[1] * nb[1] + nb[2] * nb[2]) *
                                (nc[0] * nc[0] + nc[1] * nc[1] + nc[2] * nc[2]));
      if (!(fabs(det) > 1e-13 * scale)) continue;
      const double inv = 1.0 / det;
      const double t[3] = {(da * bc[0] + db * ca[0] + dc * ab[0]) * inv, (da * bc[1] + db * ca[1] + dc * ab[1]) * inv,
                           (da * bc[2] + db * ca[2] + dc * ab[2]) * inv};
      const double r2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
      if (!faces && r2 <= best.r2) continue;  // cannot raise this lane's maximum
      const double tl = sqrt(r2);
      bool inside = true;
      for (uint32_t j = 0; j < m && inside; ++j) {
        if (j == a || j == b || j == c) continue;
        const double sj = L.pl[j][0] * t[0] + L.pl[j][1] * t[1] + L.pl[j][2] * t[2];
        const double nl = sqrt(L.pl[j][0] * L.pl[j][0] + L.pl[j][1] * L.pl[j][1] + L.pl[j][2] * L.pl[j][2]);
        inside = sj <= L.pl[j][3] + 1e-11 * (nl * tl + L.pl[j][3]);
      }
      const double slack = kCoverBox * (1.0 + 1e-12);
      inside = inside && fabs(t[0]) <= slack && fabs(t[1]) <= slack && fabs(t[2]) <= slack;
      if (!inside) continue;
      if (r2 > best.r2) best = {r2, {t[0], t[1], t[2]}, {a, b, c}};
      if (faces) {
        if (a < m) L.face[a] = 1;
        if (b < m) L.face[b] = 1;
        if (c < m) L.face[c] = 1;
        const uint32_t at = VERTS ? atomicAdd(&L.nv, 1u) : kCoverVertices;
        if (VERTS && at < kCoverVertices) {
          L.vt[at][0] = t[0];
          L.vt[at][1] = t[1];
          L.vt[at][2] = t[2];
          L.vt[at][3] = r2;
          L.vcode[at][0] = a < m ? L.pcode[a] : ~0u;
          L.vcode[at][1] = b < m ? L.pcode[b] : ~0u;
          L.vcode[at][2] = c < m ? L.pcode[c] : ~0u;
        }
      }
    }
  }
  // The wave's farthest vertex (lowest lane on a tie).
  double r2 = best.r2;
  uint32_t who = lane;
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(r2, off);
    const uint32_t ow = uint32_t(__shfl_xor(int(who), off));
    if (o > r2 || (o == r2 && ow < who)) {
      r2 = o;
      who = ow;
    }
  }
  CoverVertex out;
  out.r2 = r2;
  for (int k = 0; k < 3; ++k) {
    out.t[k] = __shfl(best.t[k], int(who));
    out.plane[k] = uint32_t(__shfl(int(best.plane[k]), int(who)));
  }
  return out;
}

struct CoverArgs {
  const float4* q;
  uint32_t n;
  const float4* ops;  // all-images mode: every signed image is a candidate
  uint32_t order;
  uint32_t grid;      // grid mode: the near-field sites within the cutoff
  NearBuffers b;
  NearCut cut;        // the near field's cutoff (c_max)
  float c_start, c_max;
  float* theta;
  float4* witness;
  uint32_t* flags;    // 0 certified, 1 needs a cutoff beyond c_max, 2 too many neighbours
  float* faces;       // faces of the cell
  // Optional: every vertex of every certified cell (direction with w >= 0, and its S3
  // radius), appended up to vcap; *vcount counts all of them.
  float4* vq;
  float* vr;
  uint32_t* vcount;
  uint32_t vcap;
  // Optional: the Delaunay tetrahedra of every certified cell (the source and the three
  // neighbours whose planes meet at a vertex), as orbit codes, up to tcap; *tcount counts.
  uint4* tets;
  uint32_t* tcount;
  uint32_t tcap;
  // Optional: per source, the orbit codes of the four points whose cap is its cell's
  // farthest vertex (the source first; ~0 for a box face).
  uint4* wcode;
};

__device__ inline float chord2_of_angle(float a) {
  const float h = sinf(0.5f * fminf(a, 3.1415f));
  return 4.0f * h * h;
}
__device__ inline double angle_of_chord2(float d2) { return 2.0 * asin(fmin(1.0, 0.5 * sqrt(double(d2)))); }

template <bool VERTS>
__global__ __launch_bounds__(64) void k_cover(CoverArgs a) {
  __shared__ CoverLds<VERTS> L;
  const uint32_t slot = blockIdx.x;
  if (slot >= a.n) return;
  const uint32_t lane = lane_id();
  const uint32_t src = a.grid ? a.b.order[slot] : slot;
  const float4 x = a.q[src];
  const double xn = 1.0 / sqrt(double(x.x) * x.x + double(x.y) * x.y + double(x.z) * x.z + double(x.w) * x.w);
  const double p[4] = {x.x * xn, x.y * xn, x.z * xn, x.w * xn};
  // Tangent basis p i, p j, p k.
  const double e[3][4] = {{-p[1], p[0], p[3], -p[2]}, {-p[2], -p[3], p[0], p[1]}, {-p[3], p[2], -p[1], p[0]}};
  // Grid-mode row walk setup: every row within the field's cutoff (as in k_spacing).
  const NearBuffers& b = a.b;
  float cx = 0, cy = 0, cz = 0, rc = 0;
  bool full = true;
  int zlo = 0, ylo = 0;
  uint32_t ny = 0, nrows = 0;
  if (a.grid) {
    const float fdim = float(b.sdim);
    const float cpu = b.info->cpu;
    const float denom = a.cut.c0 + x.x;
    full = !(denom > 1e-3f);
    cx = cy = cz = 0.5f * fdim;
    rc = 3e18f;
    if (!full) {
      const float inv = 1.0f / denom;
      cx = (x.y * inv - b.info->lower[0]) * cpu;
      cy = (x.z * inv - b.info->lower[1]) * cpu;
      cz = (x.w * inv - b.info->lower[2]) * cpu;
      rc = (a.cut.s0 * inv * 1.001f + 1e-6f) * cpu + 1e-3f;
    }
    auto clampc = [&](float v) -> int { return int(fminf(fmaxf(floorf(v), 0.0f), fdim - 1.0f)); };
    zlo = clampc(cz - rc);
    ylo = clampc(cy - rc);
    const int zhi = clampc(cz + rc), yhi = clampc(cy + rc);
    ny = uint32_t(yhi - ylo + 1);
    nrows = ny * uint32_t(zhi - zlo + 1);
  }
  uint32_t faces = 0, flag = 1, nfaces = 0;
  float lo2 = -1.0f, hi = fminf(a.c_start, a.c_max);
  CoverVertex cell{0.0, {0.0, 0.0, 0.0}, {~0u, ~0u, ~0u}};
  uint32_t wc[3] = {~0u, ~0u, ~0u};
  for (int round = 0; round < 256; ++round) {
    float hi2 = chord2_of_angle(hi);
    // Gather the window (lo2, hi2].
    uint32_t count = 0;
    auto take = [&](bool ok, const float4 y, float d2, uint32_t code) {
      const uint64_t mask = __ballot(ok);
      const uint32_t pos = count + mbcnt(mask);
      if (ok && pos < kCoverGather) {
        L.y[pos] = y;
        L.d2[pos] = d2;
        L.code[pos] = code;
      }
      count += uint32_t(__popcll(mask));
    };
    if (a.grid) {
      const uint32_t sdim = b.sdim;
      const float fdim = float(sdim);
      auto clampc = [&](float v) -> int { return int(fminf(fmaxf(floorf(v), 0.0f), fdim - 1.0f)); };
      const float rc2 = rc * rc;
      const uint32_t self_rank = src * 2 * a.order;
      for (uint32_t rb = 0; rb < nrows; rb += 64) {
        const uint32_t r = rb + lane;
        uint32_t start = 0, len = 0;
        if (r < nrows) {
          const int yy = ylo + int(r % ny), zz = zlo + int(r / ny);
          const float gy = row_gap(cy, yy, sdim), gz = row_gap(cz, zz, sdim);
          const float rest = rc2 - gy * gy - gz * gz;
          if (rest >= 0.0f) {
            const float span = full ? 3e18f : sqrtf(rest);
            const int x0 = clampc(cx - span), x1 = clampc(cx + span);
            const uint32_t row = sdim * (uint32_t(yy) + sdim * uint32_t(zz));
            start = b.offsets[row + x0];
            len = b.offsets[row + x1 + 1] - start;
          }
        }
        const uint32_t rows_here = min(64u, nrows - rb);
        for (uint32_t k = 0; k < rows_here; ++k) {
          const uint32_t n_row = uint32_t(__shfl(int(len), int(k)));
          const uint32_t s_row = uint32_t(__shfl(int(start), int(k)));
          for (uint32_t c0 = 0; c0 < n_row; c0 += 64) {
            const uint32_t c = c0 + lane;
            bool ok = false;
            float4 y = make_float4(0.f, 0.f, 0.f, 0.f);
            float d2 = 0.f;
            uint32_t code = 0;
            if (c < n_row) {
              y = b.sq[s_row + c];
              d2 = chord2(x, y);
              code = b.srank[s_row + c];
              ok = d2 > lo2 && d2 <= hi2 && code != self_rank;
            }
            take(ok, y, d2, code);
          }
        }
      }
    } else {
      const uint32_t images = a.n * a.order;
      for (uint32_t base = 0; base < images; base += 64) {
        const uint32_t idx = base + lane;
        bool ok = false;
        float4 y = make_float4(0.f, 0.f, 0.f, 0.f);
        float d2 = 0.f;
        uint32_t code = 0;
        if (idx < images) {
          const uint32_t j = idx / a.order, k = idx % a.order;
          y = image(a.q[j], a.ops[k], k);
          uint32_t sign = 0;
          if (dot4(x, y) < 0.0f) {
            y = neg4(y);
            sign = 1;
          }
          d2 = chord2(x, y);
          code = j * 2 * a.order + 2 * k + sign;
          ok = d2 > lo2 && d2 <= hi2 && !(j == src && k == 0);
        }
        take(ok, y, d2, code);
      }
    }
    if (count > kCoverGather) {
      // Too many in the window: narrow it (a third of the volume) and gather again.
      const float lo = lo2 < 0.0f ? 0.0f : float(angle_of_chord2(lo2));
      const float narrower = lo + (hi - lo) * 0.6f;
      if (!(narrower > lo * (1.0f + 1e-6f) + 1e-9f)) {
        flag = 2;
        break;
      }
      hi = narrower;
      continue;
    }
    __syncthreads();
    // Ranks by chord^2 (ties by gather order); the window keeps the nearest kCoverWindow and
    // everything as near as the last of them.
    uint32_t my_rank[kCoverGather / 64];
    for (uint32_t s = 0; s < kCoverGather / 64; ++s) {
      const uint32_t i = lane + 64 * s;
      my_rank[s] = 0xffffffffu;
      if (i < count) {
        const float di = L.d2[i];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < count; ++j) rank += (L.d2[j] < di || (L.d2[j] == di && j < i)) ? 1u : 0u;
        my_rank[s] = rank;
        if (rank == kCoverWindow - 1) L.cut = di;
      }
    }
    __syncthreads();
    uint32_t take_n = count;
    if (count > kCoverWindow) {
      const float cut = L.cut;
      uint32_t sel = 0;
      for (uint32_t i = lane; i < count; i += 64) sel += L.d2[i] <= cut ? 1u : 0u;
      for (int off = 32; off > 0; off >>= 1) sel += uint32_t(__shfl_xor(int(sel), off));
      take_n = sel;
      hi2 = cut;
    }
    if (faces + take_n > kCoverPlanes) {
      flag = 2;
      break;
    }
    for (uint32_t s = 0; s < kCoverGather / 64; ++s) {
      const uint32_t i = lane + 64 * s;
      if (i < count && my_rank[s] < take_n) {
        const uint32_t at = faces + my_rank[s];
        const float4 yf = L.y[i];
        const double yn = 1.0 / sqrt(double(yf.x) * yf.x + double(yf.y) * yf.y + double(yf.z) * yf.z + double(yf.w) * yf.w);
        const double y[4] = {yf.x * yn, yf.y * yn, yf.z * yn, yf.w * yn};
        for (int k = 0; k < 3; ++k) L.pl[at][k] = y[0] * e[k][0] + y[1] * e[k][1] + y[2] * e[k][2] + y[3] * e[k][3];
        L.pl[at][3] = 1.0 - (y[0] * p[0] + y[1] * p[1] + y[2] * p[2] + y[3] * p[3]);
        L.pd2[at] = L.d2[i];
        L.pcode[at] = L.code[i];
      }
    }
    __syncthreads();
    // Bound, prune the window, then the cell with its faces marked.
    const uint32_t total = faces + take_n;
    cell = cover_cell(L, min(total, faces + kCoverFirst), false);
    const double limit = 2.0 * atan(sqrt(fmax(cell.r2, 0.0))) * (1.0 + 1e-5) + 1e-7;
    uint32_t keep = 0;
    for (uint32_t i = faces + lane; i < total; i += 64) keep += angle_of_chord2(L.pd2[i]) < limit ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) keep += uint32_t(__shfl_xor(int(keep), off));
    keep += faces;
    for (uint32_t i = lane; i < keep; i += 64) L.face[i] = 0;
    if (lane == 0) L.nv = 0;
    __syncthreads();
    cell = cover_cell(L, keep, true);
    __syncthreads();
    for (int k = 0; k < 3; ++k) wc[k] = cell.plane[k] < keep ? L.pcode[cell.plane[k]] : ~0u;
    const double theta = atan(sqrt(fmax(cell.r2, 0.0)));
    const double reach = angle_of_chord2(hi2);
    // Keep only the faces (stable compaction; a chunk is read before it is written).
    uint32_t out = 0;
    for (uint32_t base = 0; base < keep; base += 64) {
      const uint32_t i = base + lane;
      const bool f = i < keep && L.face[i] != 0;
      double v[4] = {0, 0, 0, 0};
      float d = 0.f;
      uint32_t code = 0;
      if (f) {
        for (int k = 0; k < 4; ++k) v[k] = L.pl[i][k];
        d = L.pd2[i];
        code = L.pcode[i];
      }
      const uint64_t mask = __ballot(f);
      const uint32_t pos = out + mbcnt(mask);
      __syncthreads();
      if (f) {
        for (int k = 0; k < 4; ++k) L.pl[pos][k] = v[k];
        L.pd2[pos] = d;
        L.pcode[pos] = code;
      }
      out += uint32_t(__popcll(mask));
      __syncthreads();
    }
    faces = out;
    nfaces = out;
    if (2.0 * theta < reach * (1.0 - 1e-6) - 1e-9) {
      flag = 0;
      break;
    }
    if (reach >= double(a.c_max) * (1.0 - 1e-6)) break;
    lo2 = hi2;
    hi = fminf(a.c_max, fmaxf(1.5f * float(reach), float(2.1 * theta)));
  }
  if (VERTS && a.vq && flag == 0) {
    // The certified cell's vertices (its last pass listed them; a cell with more than
    // kCoverVertices lists the first found).
    const uint32_t nv = min(L.nv, kCoverVertices);
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(a.vcount, nv);
    base = uint32_t(__shfl(int(base), 0));
    for (uint32_t i = lane; i < nv; i += 64) {
      if (base + i >= a.vcap) break;
      double v[4];
      for (int k = 0; k < 4; ++k) v[k] = p[k] + L.vt[i][0] * e[0][k] + L.vt[i][1] * e[1][k] + L.vt[i][2] * e[2][k];
      const double vn = (v[0] < 0.0 ? -1.0 : 1.0) / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
      a.vq[base + i] = make_float4(float(v[0] * vn), float(v[1] * vn), float(v[2] * vn), float(v[3] * vn));
      a.vr[base + i] = float(atan(sqrt(fmax(L.vt[i][3], 0.0))));
    }
  }
  if (VERTS && a.tets && flag == 0) {
    const uint32_t nv = min(L.nv, kCoverVertices);
    const uint32_t self_code = src * 2 * a.order;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(a.tcount, nv);
    base = uint32_t(__shfl(int(base), 0));
    for (uint32_t i = lane; i < nv; i += 64) {
      if (base + i >= a.tcap) break;
      a.tets[base + i] = make_uint4(self_code, L.vcode[i][0], L.vcode[i][1], L.vcode[i][2]);
    }
  }
  if (lane == 0) {
    const double theta = atan(sqrt(fmax(cell.r2, 0.0)));
    a.theta[src] = float(theta);
    double v[4];
    for (int k = 0; k < 4; ++k) v[k] = p[k] + cell.t[0] * e[0][k] + cell.t[1] * e[1][k] + cell.t[2] * e[2][k];
    const double vn = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
    a.witness[src] = make_float4(float(v[0] * vn), float(v[1] * vn), float(v[2] * vn), float(v[3] * vn));
    a.flags[src] = flag;
    a.faces[src] = float(nfaces);
    if (a.wcode) a.wcode[src] = make_uint4(src * 2 * a.order, wc[0], wc[1], wc[2]);
  }
}

void launch_cover(Prof& p, hipStream_t s, const float4* q, uint32_t n, const float4* ops, uint32_t order,
                  const NearBuffers* grid, const NearCut& cut, float c_start, float c_max, float* theta,
                  float4* witness, uint32_t* flags, float* faces, float4* vq, float* vr, uint32_t* vcount,
                  uint32_t vcap, uint4* tets, uint32_t* tcount, uint32_t tcap, uint4* wcode) {
  CoverArgs a{};
  a.q = q;
  a.n = n;
  a.ops = ops;
  a.order = order;
  a.grid = grid ? 1u : 0u;
  if (grid) a.b = *grid;
  a.cut = cut;
  a.c_start = c_start;
  a.c_max = c_max;
  a.theta = theta;
  a.witness = witness;
  a.flags = flags;
  a.faces = faces;
  a.vq = vq;
  a.vr = vr;
  a.vcount = vcount;
  a.vcap = vcap;
  a.tets = tets;
  a.tcount = tcount;
  a.tcap = tcap;
  a.wcode = wcode;
  p.run("cover", [&] {
    if (a.vq || a.tets)
      hipLaunchKernelGGL(k_cover<true>, dim3(n), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL(k_cover<false>, dim3(n), dim3(64), 0, s, a);
  });
}

// Nearest-neighbour angles in SO(3) degrees, 4 asin(min(1, sqrt(c)/2)) 180/pi in FP32
// (nearest_degrees, pairs.wgsl:428-434), and their minimum (positive floats order as their
// bits; 0xffffffff marks an empty set) into *least.
__global__ void k_degrees(const float* __restrict__ chord, uint32_t n, float* __restrict__ out,
                          uint32_t* __restrict__ least) {
  uint32_t lo = 0xffffffffu;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float d = 4.0f * asinf(fminf(1.0f, sqrtf(chord[i]) * 0.5f)) * (180.0f / kPiF);
    out[i] = d;
    lo = min(lo, __float_as_uint(d));
  }
  block_min_commit(least, lo);
}

void launch_degrees(Prof& p, hipStream_t s, const float* chord, uint32_t n, float* out, uint32_t* least) {
  launch_fill_u32(p, s, least, 0xffffffffu, 1);
  p.run("degrees", [&] {
    hipLaunchKernelGGL(k_degrees, dim3(std::min<uint32_t>(2048, blocks_for(n, kBlock))), dim3(kBlock), 0, s, chord, n,
                       out, least);
  });
}

hipError_t launch_spacing_table(hipStream_t s, const float4* q, uint32_t n, uint32_t group, uint64_t* partial,
                                float* chord, uint32_t* rank_out, float* degrees_out, uint32_t* least_bits) {
  const uint32_t order = kGroupOrder[group];
  hipLaunchKernelGGL(k_allpairs_nearest<true>, dim3(blocks_for(n, kBlock), order), dim3(kBlock), 0, s, q, n,
                     nullptr, kGroupOffset[group], order, partial);
  hipLaunchKernelGGL(k_nearest_reduce, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, s, partial, n, order, chord,
                     rank_out);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, s, least_bits, 0xffffffffu, size_t(1));
  hipLaunchKernelGGL(k_degrees, dim3(std::min<uint32_t>(2048, blocks_for(n, kBlock))), dim3(kBlock), 0, s, chord, n,
                     degrees_out, least_bits);
  return hipGetLastError();
}

// ---- BB step (relax.wgsl:29-165) -----------------------------------------------------
// Every kernel takes a JobSet: point-indexed kernels map workgroups to jobs through the
// job's point-block prefix, the one-workgroup stages (sums, median pick) take one
// workgroup per job.  A job's arithmetic and reduction order do not depend on the other
// jobs in the set.
uint32_t bb_partials(uint32_t n) { return blocks_for(n, kBlock); }

__global__ void k_bb_reset(BBState* st) {
  const uint32_t t = threadIdx.x;
  if (t == 0) {
    st->iteration = 0;
    st->flags = 0;
    st->step_bits = 0;
    st->step = 0;
    st->sy = st->yy = st->ss = st->ff = st->energy = 0.0;
    st->rank[0] = st->rank[1] = 0;
    st->prefix[0] = st->prefix[1] = 0;
    st->median = 0.0f;
    st->alpha = 0.0f;
    st->alpha64 = 0.0;
  }
  for (uint32_t k = t; k < kDigitBins; k += blockDim.x) {
    st->hist0[k] = 0;
    st->hist1[0][k] = st->hist1[1][k] = 0;
  }
  for (uint32_t k = t; k < kDigitBins / 2; k += blockDim.x) st->hist2[0][k] = st->hist2[1][k] = 0;
}
void launch_bb_reset(Prof& p, hipStream_t s, BBState* state) {
  p.run("bb_reset", [&] { hipLaunchKernelGGL(k_bb_reset, dim3(1), dim3(kBlock), 0, s, state); });
}

__device__ inline void block_sum5(double v[5], double* out) {
  __shared__ double sh[5][kBlock / 64];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const double s = wave_sum(v[c]);
    if (lane == 0) sh[c][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double s = 0.0;
    for (uint32_t w = 0; w < kBlock / 64; ++w) s += sh[threadIdx.x][w];
    out[threadIdx.x] = s;
  }
}

// The exact median of max(curvature, 0) (relax.wgsl:79-113) by radix selection over both
// middle ranks, in three digits of 11, 11 and 10 bits of the key's FP32 bits (nonnegative
// floats order as their bits).  A digit's histogram is complete only when the launch that
// counts it has ended, so each later launch's workgroups pick the digits they need from
// the finished histograms themselves (the same arithmetic in every workgroup): no
// workgroup waits for another within a launch, and no device-wide fence is needed.
__device__ inline uint32_t curvature_bits(float w) { return __float_as_uint(w > 0.0f ? w : 0.0f); }

// One workgroup finds, in a histogram of `bins` counts, the bin holding rank `want`: each
// thread sums bins / 256 of them, a block scan finds the thread whose run holds the rank,
// and that thread the bin.  Returns (bin, rank within the bin) to every thread.
__device__ uint2 select_bin(const uint32_t* __restrict__ hist, uint32_t bins, uint32_t want) {
  __shared__ uint32_t wave_total[kBlock / 64];
  __shared__ uint2 found;
  const uint32_t per = bins / kBlock;  // 8 or 4
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t counts[8], mine = 0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    counts[k] = k < per ? hist[threadIdx.x * per + k] : 0u;
    mine += counts[k];
  }
  const uint32_t incl = wave_incl_scan(mine);
  if (lane == 63) wave_total[wave] = incl;
  if (threadIdx.x == 0) found = make_uint2(bins - 1, 0u);  // not kept when the counts hold the rank
  __syncthreads();
  uint32_t before = incl - mine;
  for (uint32_t w = 0; w < wave; ++w) before += wave_total[w];
  if (want >= before && want < before + mine) {
    uint32_t left = want - before, bin = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k)
      if (bin == k && k + 1 < per && left >= counts[k]) {
        left -= counts[k];
        bin = k + 1;
      }
    found = make_uint2(threadIdx.x * per + bin, left);
  }
  __syncthreads();
  const uint2 out = found;
  __syncthreads();
  return out;
}

// The first digit of both middle ranks (from hist0), with their remaining ranks.
__device__ inline void first_digits(const BBState* st, uint32_t n, uint32_t prefix[2], uint32_t rank[2]) {
  const uint32_t want[2] = {(n - 1) / 2, n / 2};
  for (uint32_t r = 0; r < 2; ++r) {
    const uint2 b = select_bin(st->hist0, kDigitBins, want[r]);
    prefix[r] = b.x << 21;
    rank[r] = b.y;
  }
}
// The second digit of both middle ranks (from hist1), after the first.
__device__ inline void second_digits(const BBState* st, uint32_t prefix[2], uint32_t rank[2]) {
  for (uint32_t r = 0; r < 2; ++r) {
    const uint2 b = select_bin(st->hist1[r], kDigitBins, rank[r]);
    prefix[r] |= b.x << 10;
    rank[r] = b.y;
  }
}

// BB stage 1, one thread per point: the step sums' partials (s.y, y.y, s.s, f.f, E) per
// workgroup, the flags and the first digit's histogram of the median key.
__global__ __launch_bounds__(256) void k_bb_stats(JobSet s) {
  __shared__ uint32_t local[kDigitBins];
  const uint32_t jb = find_job(s.point_start, s.count, blockIdx.x);
  const JobDesc d = load_job(s, jb);
  const uint32_t block = blockIdx.x - job_start(s.point_start, jb);
  const uint32_t i = block * kBlock + threadIdx.x;
  BBState* __restrict__ st = d.state;
  for (uint32_t k = threadIdx.x; k < kDigitBins; k += kBlock) local[k] = 0;
  if (block == 0)  // the third digit's histogram, used last step, is counted again in stage 3
    for (uint32_t k = threadIdx.x; k < kDigitBins / 2; k += kBlock) st->hist2[0][k] = st->hist2[1][k] = 0;
  __syncthreads();
  const bool first = st->iteration == 0;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  uint32_t fl = 0;
  if (i < d.n) {
    const float4 f = d.rows.force[i];
    const bool finite = (f.x == f.x) && (f.y == f.y) && (f.z == f.z) && fabsf(f.x) <= 3.4e38f &&
                        fabsf(f.y) <= 3.4e38f && fabsf(f.z) <= 3.4e38f;
    if (!finite) fl |= 1u;
    if (!(d.rows.nearest[i] > 0.0f)) fl |= 2u;
    const double fx = f.x, fy = f.y, fz = f.z;
    v[3] = fx * fx + fy * fy + fz * fz;
    v[4] = d.rows.energy[i];
    if (!first && i + 1 != d.fixed) {
      const float4 pf = d.prev[2 * size_t(i)], pd = d.prev[2 * size_t(i) + 1];
      const double sx = pd.x, sy_ = pd.y, sz = pd.z;
      const double yx = fx - double(pf.x), yy_ = fy - double(pf.y), yz = fz - double(pf.z);
      v[0] = sx * yx + sy_ * yy_ + sz * yz;
      v[1] = yx * yx + yy_ * yy_ + yz * yz;
      v[2] = sx * sx + sy_ * sy_ + sz * sz;
    }
    atomicAdd(&local[curvature_bits(f.w) >> 21], 1u);
  }
  if (fl) atomicOr(&st->flags, fl);
  block_sum5(v, d.bb_partials + 5 * size_t(block));
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < kDigitBins; k += kBlock)
    if (local[k]) atomicAdd(&st->hist0[k], local[k]);
}

// BB stages 2 and 3 (median_blocks workgroups per job, reading the curvature word of each
// force row): the keys matching each rank's digits so far are counted by their next digit.
// Stage 2's first workgroup also adds the partials in a fixed order (the step sums) and
// opens the step; stage 3's first workgroup records both ranks' first two digits.
__global__ __launch_bounds__(256) void k_median_round(JobSet s, uint32_t round) {
  __shared__ uint32_t local[2 * kDigitBins];
  const uint32_t jb = find_job(s.median_start, s.count, blockIdx.x);
  const JobDesc d = load_job(s, jb);
  const uint32_t block = blockIdx.x - job_start(s.median_start, jb);
  BBState* __restrict__ st = d.state;
  uint32_t prefix[2], rank[2];
  first_digits(st, d.n, prefix, rank);
  if (round == 2) second_digits(st, prefix, rank);
  const uint32_t bins = round == 1 ? kDigitBins : kDigitBins / 2, shift = round == 1 ? 10u : 0u;
  const uint32_t fixed = round == 1 ? 0xffe00000u : 0xfffffc00u;  // the key bits chosen so far
  for (uint32_t k = threadIdx.x; k < 2 * bins; k += kBlock) local[k] = 0;
  __syncthreads();
  const float* __restrict__ curvature = reinterpret_cast<const float*>(d.rows.force) + 3;
  for (uint32_t i = block * kBlock + threadIdx.x; i < d.n; i += d.median_blocks * kBlock) {
    const uint32_t bits = curvature_bits(curvature[4 * size_t(i)]);
    const uint32_t digit = (bits >> shift) & (bins - 1);
    if ((bits & fixed) == prefix[0]) atomicAdd(&local[digit], 1u);
    if ((bits & fixed) == prefix[1]) atomicAdd(&local[bins + digit], 1u);
  }
  __syncthreads();
  uint32_t* out0 = round == 1 ? st->hist1[0] : st->hist2[0];
  uint32_t* out1 = round == 1 ? st->hist1[1] : st->hist2[1];
  for (uint32_t k = threadIdx.x; k < bins; k += kBlock) {
    if (local[k]) atomicAdd(&out0[k], local[k]);
    if (local[bins + k]) atomicAdd(&out1[k], local[bins + k]);
  }
  if (block != 0) return;
  if (round == 1) {
    double w[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t b = threadIdx.x; b < d.point_blocks; b += kBlock)
      for (int c = 0; c < 5; ++c) w[c] += d.bb_partials[5 * size_t(b) + c];
    __shared__ double sums[5];
    block_sum5(w, sums);
    __syncthreads();
    if (threadIdx.x == 0) {
      st->sy = sums[0];
      st->yy = sums[1];
      st->ss = sums[2];
      st->ff = sums[3];
      st->energy = sums[4];
      st->step_bits = 0;
      st->step = st->iteration;
      st->iteration = st->step + 1;
    }
  } else if (threadIdx.x == 0) {
    st->prefix[0] = prefix[0];
    st->prefix[1] = prefix[1];
    st->rank[0] = rank[0];
    st->rank[1] = rank[1];
  }
}

// BB stage 4, one thread per point: every workgroup takes the third digit from the finished
// histogram, the median = max(1e-12, mean of the two middle keys) and the step length
// alpha (relax.wgsl:115-128: alpha0 = (0.8/3.4) r^3 first, then BB1 on odd and BB2 on even
// steps, clamped to [1e-12, max(1, 10 r)]); then the move (relax.wgsl:140-165): scale
// alpha / max(1e-6, c / median + 1e-2), clipped to half the mean radius, q <- exp(d) q
// (sinc series below 1e-4), sign fixed, the held identity exact; then the fold into the
// zone.  A flagged step moves nothing (the points are folded as they are).  The first
// workgroup clears the first two digits' histograms for the next step.
__global__ __launch_bounds__(256) void k_bb_move(JobSet s) {
  const uint32_t jb = find_job(s.point_start, s.count, blockIdx.x);
  const JobDesc d = load_job(s, jb);
  const uint32_t block = blockIdx.x - job_start(s.point_start, jb);
  const uint32_t i = block * kBlock + threadIdx.x;
  BBState* __restrict__ st = d.state;
  uint32_t key[2];
  for (uint32_t r = 0; r < 2; ++r) key[r] = st->prefix[r] | select_bin(st->hist2[r], kDigitBins / 2, st->rank[r]).x;
  const float median = fmaxf(1e-12f, 0.5f * (__uint_as_float(key[0]) + __uint_as_float(key[1])));
  const uint32_t it = st->step;
  const double radius = d.radius;
  const double alpha0 = (0.8 / 3.4) * radius * radius * radius;
  double alpha64 = alpha0;
  if (it > 0) {
    const double a1 = fabs(st->yy) > 1e-20 ? fabs(st->sy / st->yy) : alpha0;
    const double a2 = fabs(st->sy) > 1e-20 ? fabs(st->ss / st->sy) : alpha0;
    alpha64 = fmin(fmax(it % 2 == 1 ? a1 : a2, 1e-12), fmax(1.0, 10.0 * radius));
  }
  const float alpha = float(alpha64);
  if (block == 0) {
    if (threadIdx.x == 0) {
      st->median = median;
      st->alpha = alpha;
      st->alpha64 = alpha64;
    }
    for (uint32_t k = threadIdx.x; k < kDigitBins; k += kBlock) {
      st->hist0[k] = 0;
      st->hist1[0][k] = st->hist1[1][k] = 0;
    }
  }
  float moved_size = 0.0f;
  if (i < d.n) {
    float4 m = d.q[i];
    if (st->flags != 0) {
      // no move
    } else if (i + 1 == d.fixed) {
      d.prev[2 * size_t(i)] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      d.prev[2 * size_t(i) + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      m = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    } else {
      const float r = float(d.radius);
      const float4 f = d.rows.force[i];
      const float scale = alpha * d.step_scale / fmaxf(1e-6f, f.w / median + 1e-2f);
      const float size = sqrtf(f.x * f.x + f.y * f.y + f.z * f.z) * scale;
      const float clip = fminf(1.0f, 0.5f * r / fmaxf(size, 1e-12f));
      const float3 dv = make_float3(f.x * (scale * clip), f.y * (scale * clip), f.z * (scale * clip));
      d.prev[2 * size_t(i)] = make_float4(f.x, f.y, f.z, 0.0f);
      d.prev[2 * size_t(i) + 1] = make_float4(dv.x, dv.y, dv.z, 0.0f);
      moved_size = size * clip;
      const float angle = sqrtf(dv.x * dv.x + dv.y * dv.y + dv.z * dv.z);
      const float sinc = angle > 1e-4f ? sinf(angle) / angle : 1.0f - angle * angle / 6.0f;
      const float4 rot = make_float4(cosf(angle), dv.x * sinc, dv.y * sinc, dv.z * sinc);
      m = normalize4(qmul(rot, m));
      if (m.x < 0.0f) m = neg4(m);
    }
    // Right-action folding keeps the left tangent frame of the BB history (engine.zig:914-916).
    d.q[i] = fold_point(m, d.group, d.ops, d.order);
  }
  block_max_commit(&st->step_bits, __float_as_uint(moved_size));
}

// One BB step of every job in the set, resident, in four launches: stats with the first
// median digit, two radix rounds (with the step sums), move and fold.
void launch_bb_step(Prof& p, hipStream_t s, const JobSet& jobs) {
  if (!jobs.count) return;
  p.run("bb_stats", [&] { hipLaunchKernelGGL(k_bb_stats, dim3(jobs.point_total), dim3(kBlock), 0, s, jobs); });
  for (uint32_t round = 1; round <= 2; ++round)
    p.run("median_round", [&] {
      hipLaunchKernelGGL(k_median_round, dim3(jobs.median_total), dim3(kBlock), 0, s, jobs, round);
    });
  p.run("bb_move", [&] { hipLaunchKernelGGL(k_bb_move, dim3(jobs.point_total), dim3(kBlock), 0, s, jobs); });
}

// ---- issue-rate probe (zo_issue_probe) --------------------------------------------------
// Long independent chains of one instruction kind per lane, on 8 waves per SIMD over the
// whole chip, timed with events: the rate the VALU issues non-packed FP32 FMAs (8 chains)
// or transcendentals (v_rsq_f32 and v_rcp_f32, 4 chains), or the matrix cores FP32 Gram
// tiles (v_mfma_f32_16x16x4f32, 4 chains), in wave64 instructions per second.
// The chains stay finite (an FMA contracting toward 1, reciprocals of values in [0.5, 2]).
__global__ __launch_bounds__(256) void k_probe_fma(float* __restrict__ out, uint32_t iters) {
  const float m = 0.999f, c = 0.001f;
  float a[8];
  for (int k = 0; k < 8; ++k) a[k] = 1.0f + float(threadIdx.x + k) * 1e-3f;
  for (uint32_t i = 0; i < iters; ++i)
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = fmaf(a[k], m, c);
  out[blockIdx.x * blockDim.x + threadIdx.x] = a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7];
}
__global__ __launch_bounds__(256) void k_probe_trans(float* __restrict__ out, uint32_t iters) {
  float a0 = 1.0f + float(threadIdx.x) * 1e-3f, a1 = a0 + 0.25f, a2 = a0 + 0.5f, a3 = a0 + 0.75f;
  for (uint32_t i = 0; i < iters; ++i) {
    a0 = __builtin_amdgcn_rsqf(a0);
    a1 = __builtin_amdgcn_rcpf(a1);
    a2 = __builtin_amdgcn_rsqf(a2);
    a3 = __builtin_amdgcn_rcpf(a3);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3;
}

// FP32 Gram tiles: v_mfma_f32_16x16x4f32, 16 x 16 dot products of 4-vectors each (the
// shape of a candidate prefilter's tile of 16 sources by 16 sites), 4 independent
// accumulators per wave.
__global__ __launch_bounds__(256) void k_probe_mfma(float* __restrict__ out, uint32_t iters) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const float a = 1.0f + float(threadIdx.x) * 1e-3f, b = 0.5f - float(threadIdx.x) * 1e-4f;
  f32x4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = c0, c2 = c0, c3 = c0;
  for (uint32_t i = 0; i < iters; ++i) {
    c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, a, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f32_16x16x4f32(b, b, c3, 0, 0, 0);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = c0.x + c1.y + c2.z + c3.w;
}

double probe_issue_rate(hipStream_t s, uint32_t kind, uint32_t iters, float* scratch, uint32_t blocks) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.0;
  auto launch = [&](uint32_t it) {
    if (kind == 0)
      hipLaunchKernelGGL(k_probe_fma, dim3(blocks), dim3(kBlock), 0, s, scratch, it);
    else if (kind == 1)
      hipLaunchKernelGGL(k_probe_trans, dim3(blocks), dim3(kBlock), 0, s, scratch, it);
    else
      hipLaunchKernelGGL(k_probe_mfma, dim3(blocks), dim3(kBlock), 0, s, scratch, it);
  };
  launch(64);  // warm-up
  (void)hipEventRecord(a, s);
  launch(iters);
  (void)hipEventRecord(b, s);
  float ms = 0.0f;
  const bool ok = hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (!ok || ms <= 0.0f) return -1.0;
  const double waves = double(blocks) * kBlock / 64.0;
  return waves * double(iters) * (kind == 0 ? 8.0 : 4.0) / (double(ms) * 1e-3);  // 4 per round for 1 and 2
}

}  // namespace zxt
