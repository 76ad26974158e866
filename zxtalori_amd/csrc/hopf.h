// Inverse Hopf product grids of SO(3)/K (the role of src/hopf.zig in the reference): the
// grid is built on the host in FP64, rounded once to FP32, and the engine folds it into
// the zone on the GPU.
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace zxt::hopf {

constexpr uint32_t kMaxFrequency = 512;      // hopf.zig:22
constexpr uint64_t kMaxCount = 10000000;     // hopf.zig:23

// Group ids: C2 C3 C4 C6 D2 D3 D4 D6 T O I C1 = 0..11.  The root grid of a group is the
// icosahedral one wherever the group embeds in I (C1 included), else octahedral (C4, D4,
// O) or the six-fold ring grid (C6, D6) (hopf.zig:29-36).
uint32_t root_of(uint32_t group);

// Grid size at frequency f: fibre rotations x base nodes / |K| (hopf.zig:61-76).
uint64_t raw_count(uint32_t group, uint32_t f);

// raw_count, or EngineError InvalidGridResolution (f outside 1..512) / GridTooLarge.
uint32_t count(uint32_t group, uint32_t f);

// The frequency whose count is nearest `target` by ratio, the lowest on a tie (hopf.zig:82-97).
uint32_t nearest_frequency(uint32_t group, uint32_t target);

// The generator text of the stats record (engine.zig:271-275).
const char* description(uint32_t group);

// The group's elements in FP64, scalar first: the closure of its generators (the frames
// of the operator tables), breadth first, scalar part made nonnegative and ordered by
// decreasing scalar part, stably (hopf.zig:161-213).
std::vector<std::array<double, 4>> elements(uint32_t group);

// Writes the frequency-f grid to out as count(group, f) FP32 rows, scalar first, 4 floats
// per point (out holds that many); each point is one image of its orbit (the engine's fold
// picks the zone representative).  Computed on all host threads, the rows independent of
// the split.  Throws EngineError on invalid input.
void points(uint32_t group, uint32_t f, float* out);

}  // namespace zxt::hopf
