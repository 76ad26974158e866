// All-pairs tile plans per operator count and point count (2^8 .. 2^15), written by
// zxtalori_amd/tune.py on an gfx950 device (2026-12-04).  target: workgroups per
// pass; ops: operators per workgroup (0: automatic).  Regenerate with
// `python -m zxtalori_amd.tune` on the GPU and rebuild.
#pragma once
#include <cstdint>

namespace zxt {

constexpr int kTunedOrderCount = 9;
constexpr int kTunedLogMin = 8;
constexpr int kTunedSizeCount = 8;
constexpr uint32_t kTunedOrders[kTunedOrderCount] = {1, 2, 3, 4, 6, 8, 12, 24, 60};
constexpr uint32_t kTunedTarget[kTunedOrderCount][kTunedSizeCount] = {
    {2048, 2048, 2048, 2048, 16384, 2048, 8192, 8192},
    {2048, 2048, 2048, 2048, 16384, 2048, 8192, 8192},
    {2048, 2048, 2048, 16384, 16384, 4096, 8192, 8192},
    {2048, 2048, 2048, 16384, 2048, 4096, 8192, 8192},
    {2048, 2048, 2048, 16384, 16384, 16384, 16384, 16384},
    {2048, 2048, 2048, 16384, 16384, 16384, 16384, 16384},
    {2048, 2048, 16384, 2048, 16384, 16384, 16384, 16384},
    {2048, 2048, 16384, 4096, 8192, 16384, 16384, 16384},
    {2048, 16384, 2048, 4096, 16384, 16384, 16384, 16384}};
constexpr uint32_t kTunedOps[kTunedOrderCount][kTunedSizeCount] = {
    {0, 0, 0, 0, 8, 0, 2, 8},
    {0, 0, 0, 0, 8, 0, 4, 4},
    {0, 0, 0, 1, 2, 2, 8, 8},
    {0, 0, 0, 2, 0, 4, 8, 8},
    {0, 0, 0, 2, 4, 4, 4, 4},
    {0, 0, 0, 2, 2, 4, 4, 4},
    {0, 0, 0, 0, 2, 8, 8, 8},
    {0, 0, 2, 2, 2, 8, 8, 8},
    {0, 2, 0, 8, 8, 8, 8, 8}};

}  // namespace zxt
