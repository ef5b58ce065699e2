// dtb_devgen.hpp -- seeded fan-out networks generated on the device.
#pragma once

#include "dtb_types.hpp"

namespace dtb {

// Sort width of one row (power of two >= K).
uint32_t fanout_sort_width(uint32_t K);

// Writes rowptr [n+1], col [n K] (channel bit in bit 31) and w [n K] of the
// fan-out network (n, K, seed), and adds to tile_syn [ceil(n / 256)] the
// synapses that target each tile of 256 neurons (zero-filled by the
// caller).  K <= 8192.
hipError_t launch_gen_fanout(uint32_t n, uint32_t K, uint64_t seed, int64_t *rowptr, uint32_t *col, float2 *w,
                             uint32_t *tile_syn, hipStream_t stream);

}  // namespace dtb
