// dtb_rng.hpp -- counter-based random numbers for the DTB step (device side).
//
// The reference keeps one 64-byte hiprandStatePhilox4_32_10_t per neuron and
// re-seeds it from /dev/urandom on every Run (src/stage/random.cpp:7-19,
// src/block.cpp:129-134, src/common.cpp:19-37); the OU kernel draws
// hiprand_normal(&states[idx]) (src/stage/extern_stimuli.cpp:58).  Here the
// generator is stateless: every normal is a pure function of
// (seed, neuron, substep), so a run is reproducible, resumable and
// independent of launch geometry, and no RNG state crosses HBM.
//
// The normal transform is Box-Muller built only from IEEE add, mul, fma,
// division and square root (all correctly rounded on gfx950 with
// -fhip-fp32-correctly-rounded-divide-sqrt and on the host), so the CPU
// oracle reproduces every bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtb {

enum : uint32_t { kTagOU = 1u, kTagNoise = 2u, kTagGen = 3u, kTagGamma = 4u };

struct Philox4 {
    uint32_t x, y, z, w;
};

// Philox4x32-10 (Salmon et al., SC'11).
__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                 uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Philox4{c0, c1, c2, c3};
}

__host__ __device__ inline uint32_t f32_bits(float x)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.f = x;
    return v.u;
}

__host__ __device__ inline float bits_f32(uint32_t u)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.u = u;
    return v.f;
}

// ln(u) for positive normal u.  u = m 2^e, m in [sqrt(1/2), sqrt(2));
// ln m = 2 atanh(s), s = (m - 1)/(m + 1), odd series to s^9.
__host__ __device__ inline float rng_log(float u)
{
    const uint32_t b = f32_bits(u);
    int32_t e = (int32_t)((b >> 23) & 0xffu) - 127;
    float m = bits_f32((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float z = s * s;
    float p = 0.111111111f;
    p = fmaf(p, z, 0.142857143f);
    p = fmaf(p, z, 0.2f);
    p = fmaf(p, z, 0.333333333f);
    p = fmaf(p, z, 1.0f);
    const float lnm = (s + s) * p;
    return fmaf((float)e, 0.693147180559945309f, lnm);
}

// sin(pi x / 2) and cos(pi x / 2) for x in [0, 1/2] (Taylor, degree 9 / 10).
__host__ __device__ inline float rng_sin_halfpi(float x)
{
    const float y = x * x;
    float p = 0.000160441184757172367f;
    p = fmaf(p, y, -0.00468175413531868810f);
    p = fmaf(p, y, 0.0796926262461670451f);
    p = fmaf(p, y, -0.645964097506246254f);
    p = fmaf(p, y, 1.57079632679489662f);
    return x * p;
}

__host__ __device__ inline float rng_cos_halfpi(float x)
{
    const float y = x * x;
    float p = -0.0000252020423730606054f;
    p = fmaf(p, y, 0.000919260274839426027f);
    p = fmaf(p, y, -0.0208634807633529609f);
    p = fmaf(p, y, 0.253669507901048014f);
    p = fmaf(p, y, -1.23370055013616983f);
    return fmaf(p, y, 1.0f);
}

// z = sqrt(-2 ln u1) cos(2 pi u2); u1 from the top 23 bits of x0 (never 0
// or 1), the angle from the top 24 bits of x1 split into quadrant and
// fraction so that no argument reduction is needed.
__host__ __device__ inline float normal_from_bits(uint32_t x0, uint32_t x1)
{
    const float u1 = ((float)(x0 >> 9) + 0.5f) * 0x1p-23f;
    const float r = sqrtf(-2.0f * rng_log(u1));
    const uint32_t k = x1 >> 8;
    const uint32_t q = k >> 22;
    const float f = ((float)(k & 0x3fffffu) + 0.5f) * 0x1p-22f;
    float a, c;
    if (f <= 0.5f) {
        a = rng_sin_halfpi(f);
        c = rng_cos_halfpi(f);
    } else {
        const float g = 1.0f - f;
        a = rng_cos_halfpi(g);
        c = rng_sin_halfpi(g);
    }
    float t = (q & 1u) ? a : c;
    t = ((q + 1u) & 2u) ? -t : t;   // q = 1, 2 negate
    return r * t;
}

__host__ __device__ inline float uniform_from_bits(uint32_t x0)
{
    return (float)(x0 >> 8) * 0x1p-24f;
}

// Counter layout: (neuron, substep lo, substep hi, stream tag); key = seed.
__host__ __device__ inline float ou_normal(uint64_t seed, uint32_t neuron, uint64_t substep)
{
    const Philox4 x = philox4x32_10(neuron, (uint32_t)substep, (uint32_t)(substep >> 32), kTagOU,
                                    (uint32_t)seed, (uint32_t)(seed >> 32));
    return normal_from_bits(x.x, x.y);
}

__host__ __device__ inline float noise_uniform(uint64_t seed, uint32_t neuron, uint64_t step)
{
    const Philox4 x = philox4x32_10(neuron, (uint32_t)step, (uint32_t)(step >> 32), kTagNoise,
                                    (uint32_t)seed, (uint32_t)(seed >> 32));
    return uniform_from_bits(x.x);
}

}  // namespace dtb
