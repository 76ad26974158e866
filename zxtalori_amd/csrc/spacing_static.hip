// zo_spacing_static: zo_spacing's all-pairs path (every point's nearest orbit point, its
// SO(3) angle and the sample's separation) on caller-owned device buffers, allocating
// nothing.  Its scratch (the per-operator nearest keys and the chords) is one file-scope
// device array of fixed size; each device holds its own copy, whose address is looked up
// once per device and kept.  Every call on a device shares that copy: calls that overlap
// on one device race, and the array keeps the last call's contents (nothing in it needs
// to start at zero: every key and chord a call reads, it has written first).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdint>
#include <mutex>

#include "../../include/zxtalori.h"
#include "groups_table.h"
#include "kernels.h"

namespace {

// Nearest keys (|K| x n u64) then chords (n f32): every group up to 20,000 points.
constexpr size_t kScratchBytes = size_t(10) << 20;
constexpr uint32_t kMaxStaticCount = 20000;  // zo_spacing's all-pairs range
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_lock;
void* g_scratch[kMaxDevices] = {};

int32_t fail(const char* text) {
  zxt::set_create_error(text);
  return -1;
}

// This device's copy of scratch (the device must be current), looked up on first use.
void* scratch_address(int device) {
  std::lock_guard<std::mutex> hold(g_lock);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch[device] = ptr;
  }
  return g_scratch[device];
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" int32_t zo_spacing_static(uint32_t group, const float* q, uint32_t n, uint32_t* rank_out,
                                     float* degrees_out, float* least_out, void* stream) {
  if (group >= zxt::kGroupCount) return fail("Spacing: UnknownGroup");
  if (!q || !rank_out || !degrees_out || !least_out) return fail("Spacing: NullPointer");
  if ((reinterpret_cast<uintptr_t>(q) & 15u) || (reinterpret_cast<uintptr_t>(least_out) & 3u))
    return fail("Spacing: MisalignedBuffer");
  if (n < 2) return fail("Spacing: TooFewOrientations");
  const uint32_t order = zxt::kGroupOrder[group];
  const size_t keys = align256(size_t(order) * n * sizeof(uint64_t));
  if (n > kMaxStaticCount || keys + size_t(n) * sizeof(float) > kScratchBytes) {
    char text[160];
    snprintf(text, sizeof(text), "Spacing: CountExceedsStaticScratch (%u points of a group of order %u; at most %u)", n,
             order, kMaxStaticCount);
    return fail(text);
  }
  int device = -1;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) {
    (void)hipGetLastError();
    return fail("Spacing: GpuError: no current HIP device");
  }
  unsigned char* base = static_cast<unsigned char*>(scratch_address(device));
  if (!base) return fail("Spacing: GpuError: scratch address lookup failed");
  const hipError_t err = zxt::launch_spacing_table(
      static_cast<hipStream_t>(stream), reinterpret_cast<const float4*>(q), n, group,
      reinterpret_cast<uint64_t*>(base), reinterpret_cast<float*>(base + keys), rank_out, degrees_out,
      reinterpret_cast<uint32_t*>(least_out));
  if (err != hipSuccess) {
    char text[200];
    snprintf(text, sizeof(text), "Spacing: GpuError: %s", hipGetErrorString(err));
    return fail(text);
  }
  return 0;
}
