// dtb_gen.hpp -- seeded synthetic networks (host, multi-threaded).
#pragma once

#include <cstdint>

namespace dtb {

// n neurons, 80% excitatory (population sub_base) then 20% inhibitory
// (sub_base + 1); every neuron has exactly K in-edges from uniformly drawn
// presynaptic neurons (no self-connections).  Outputs: props [n][22],
// rowptr [n+1], col [n K], w [n K][2], kind_bits [ceil(n K / 32)].
void gen_uniform(int64_t n, uint32_t K, uint64_t seed, const float g_ui[4], uint32_t sub_base, float *props,
                 int64_t *rowptr, uint32_t *col, float *w, uint32_t *kind_bits);

// n neurons, 80% excitatory (population sub_base) then 20% inhibitory
// (sub_base + 1); every presynaptic neuron r has exactly K synapses: draw k
// picks target t = mulhi(x0, n - 1) (+1 at or above r) and weights
// (x1 >> 8, x2 >> 8) 2^-24 from Philox4x32-10 (counter (r, k, 1, generator
// tag), key seed); each row is ordered by (t, k).  The mean in-degree is K.
// The same network is generated on the device by launch_gen_fanout.
void gen_fanout(int64_t n, uint32_t K, uint64_t seed, const float g_ui[4], uint32_t sub_base, float *props,
                int64_t *rowptr, uint32_t *col, float *w, uint32_t *kind_bits);

// Property matrix of gen_uniform / gen_fanout networks (populations
// sub_base, sub_base + 1 for the first 80% and the rest).
void gen_default_props(int64_t n, const float g_ui[4], uint32_t sub_base, float *props);

// nvox voxels with sizes proportional to lognormal(0, 0.5), each an E/I pair
// (80/20, populations 2v and 2v+1).  K in-edges per neuron: round(K
// intra_frac) from the neuron's own voxel, the rest from a source voxel drawn
// from a sparse lognormal voxel-to-voxel matrix of the given density.
void gen_voxel(int64_t n, uint32_t nvox, uint32_t K, float intra_frac, float density, uint64_t seed,
               const float g_ui[4], float *props, int64_t *rowptr, uint32_t *col, float *w, uint32_t *kind_bits,
               uint32_t *vox_sizes);

// Gamma(alpha, rate beta) draws for neurons first .. first + count - 1 of
// stream `call` (property.cpp:470-560 semantics: alpha < 1 is boosted by
// U^(1/alpha); alpha = 0 gives 0).  Marsaglia and Tsang (2000), "A simple
// method for generating gamma variables", ACM TOMS 26(3), with the normal
// and uniform deviates of Philox4x32-10, counter (neuron, call, attempt,
// gamma tag), key = seed.
float gamma_draw(uint64_t seed, uint32_t call, uint32_t neuron, float alpha, float beta);
void gen_gamma(uint64_t seed, uint32_t call, uint32_t first, uint32_t count, float alpha, float beta, float *out);

}  // namespace dtb
