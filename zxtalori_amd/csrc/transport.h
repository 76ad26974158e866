// FCC transport grids of T, O and I (src/grids.zig in the reference): the host side.
// The lattice and the per-point map are in transport_math.h; the kernel in kernels.hip.
#pragma once
#include <cstdint>
#include <vector>

#include "transport_math.h"

namespace zxt::transport {

constexpr uint32_t kMaxSemiEdge = 128;  // grids.zig:18

// Lattice points of the group's source at semi-edge h (grids.zig:17-38); EngineError
// InvalidGridResolution (h outside 1..128) or TransportNeedsTOrOOrI.
uint64_t count(uint32_t group, uint32_t h);

// The launch description and the lattice's row offsets (rows are the (i, j) pairs,
// i-major; offsets has side^2 + 1 entries).
struct Plan {
  fcc::Job job;
  std::vector<uint32_t> offsets;
};
Plan plan(uint32_t group, uint32_t h);

// The map tables (zone-side azimuth distributions, Gauss-Legendre rule, icosahedral
// chambers), built once from the geometry.
const fcc::Tables& tables();

// The generator text of the stats record (engine.zig:276-281).
const char* description(uint32_t group);

}  // namespace zxt::transport
