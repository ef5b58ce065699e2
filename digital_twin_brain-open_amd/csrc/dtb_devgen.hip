// dtb_devgen.hip -- seeded fan-out networks generated directly in HBM.
//
// Serving admits networks faster than a host can build and upload them, so
// this generator writes the presynaptic-major CSR on the device.  Each
// presynaptic neuron r has exactly K synapses; draw k in [0, K) picks the
// target t = mulhi(x0, n - 1) (+1 at or above r: no self-connection) and the
// weight pair (x1 >> 8, x2 >> 8) 2^-24 from Philox4x32-10 with counter
// (r, k, 1, generator tag) and key = seed.  Row r's synapses are ordered by
// (t, k); the channel pair is fixed by the row (inhibitory rows r >= 0.8 n);
// rows of up to 2048 synapses are ordered by a counting sort, longer ones by
// a bitonic network.
// The synapses per target tile (queue capacities) are counted in LDS per
// workgroup and added to the global counts once per workgroup when the
// counts fit beside the sort buffer, else one global atomic per synapse.
// gen_fanout in dtb_gen.cpp is the same generator on the host.
#include "dtb_devgen.hpp"
#include "dtb_rng.hpp"

namespace dtb {

namespace {

__device__ __forceinline__ uint32_t dmulhi32(uint32_t a, uint32_t b)
{
    return __umulhi(a, b);
}

__device__ __forceinline__ unsigned long long fanout_key(uint64_t seed, uint32_t n, uint32_t r, uint32_t k)
{
    const Philox4 x = philox4x32_10(r, k, 1u, kTagGen, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t t = dmulhi32(x.x, n - 1u);
    t += (t >= r) ? 1u : 0u;
    return ((unsigned long long)t << 16) | (unsigned long long)k;
}

__global__ __launch_bounds__(kBlock) void k_fanout_rowptr(uint32_t n, uint32_t K, int64_t *__restrict__ rowptr)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i <= n)
        rowptr[i] = (int64_t)i * (int64_t)K;
}

// One workgroup per row (grid-stride): the row's P keys (P = power of two
// >= K, padding sorts last) are sorted in LDS by a bitonic network, then the
// first K are written out.  lds_tiles: count the synapses per target tile in
// LDS (ntiles counters after the keys) and flush them at the end.
__global__ __launch_bounds__(kBlock) void k_fanout_rows(uint32_t n, uint32_t K, uint32_t P, uint64_t seed,
                                                        uint32_t n_exc, uint32_t *__restrict__ col,
                                                        float2 *__restrict__ w, uint32_t *__restrict__ tile_syn,
                                                        uint32_t ntiles, uint32_t lds_tiles)
{
    extern __shared__ unsigned long long keys[];
    uint32_t *tcount = reinterpret_cast<uint32_t *>(keys + P);
    const uint32_t tid = threadIdx.x;
    if (lds_tiles) {
        for (uint32_t i = tid; i < ntiles; i += kBlock)
            tcount[i] = 0u;
    }
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        for (uint32_t i = tid; i < P; i += kBlock)
            keys[i] = (i < K) ? fanout_key(seed, n, r, i) : ~0ull;
        __syncthreads();
        for (uint32_t size = 2; size <= P; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t i = tid; i < P / 2; i += kBlock) {
                    const uint32_t lo = 2 * i - (i & (stride - 1));
                    const uint32_t hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const unsigned long long a = keys[lo], b = keys[hi];
                    if ((a > b) == up) {
                        keys[lo] = b;
                        keys[hi] = a;
                    }
                }
                __syncthreads();
            }
        }
        const uint32_t kind = (r >= n_exc) ? kKindBit : 0u;
        const int64_t base = (int64_t)r * (int64_t)K;
        for (uint32_t i = tid; i < K; i += kBlock) {
            const unsigned long long key = keys[i];
            const uint32_t t = (uint32_t)(key >> 16);
            const uint32_t k = (uint32_t)(key & 0xffffu);
            const Philox4 x = philox4x32_10(r, k, 1u, kTagGen, (uint32_t)seed, (uint32_t)(seed >> 32));
            col[base + i] = t | kind;
            w[base + i] = make_float2((float)(x.y >> 8) * 0x1p-24f, (float)(x.z >> 8) * 0x1p-24f);
            if (lds_tiles)
                atomicAdd(&tcount[t >> kTileBits], 1u);
            else
                atomicAdd(&tile_syn[t >> kTileBits], 1u);
        }
        __syncthreads();   // keys are reused by the next row
    }
    if (lds_tiles) {
        for (uint32_t i = tid; i < ntiles; i += kBlock)
            if (tcount[i] != 0u)
                atomicAdd(&tile_syn[i], tcount[i]);
    }
}

// k_fanout_rows with a counting sort instead of the bitonic network (rows
// of up to 2048 synapses): the row's K keys are counted into B buckets of
// consecutive targets (target >> shift, B >= K, so about one key per
// bucket), the buckets are laid out by a workgroup prefix sum, the keys
// scattered into them and each bucket sorted by its owner thread (a few
// keys).  Six barriers per row instead of one per bitonic stage; the same
// order (targets ascending, then draw index).  LDS: kin[K], kout[K] (64-bit
// keys), cnt[B], off[B], 4 wave totals, then the tile counts.
__global__ __launch_bounds__(kBlock) void k_fanout_rows_count(uint32_t n, uint32_t K, uint32_t P, uint32_t B,
                                                              uint32_t shift, uint64_t seed, uint32_t n_exc,
                                                              uint32_t *__restrict__ col, float2 *__restrict__ w,
                                                              uint32_t *__restrict__ tile_syn, uint32_t ntiles,
                                                              uint32_t lds_tiles)
{
    extern __shared__ unsigned long long keys[];
    unsigned long long *kin = keys;
    unsigned long long *kout = keys + P;
    uint32_t *cnt = reinterpret_cast<uint32_t *>(keys + 2 * P);
    uint32_t *off = cnt + B;
    uint32_t *wsum = off + B;
    uint32_t *tcount = wsum + 4;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t wv = tid / kWave;
    const uint32_t m = B / kBlock;   // buckets per thread (B is a multiple of kBlock)
    if (lds_tiles) {
        for (uint32_t i = tid; i < ntiles; i += kBlock)
            tcount[i] = 0u;
    }
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {
        for (uint32_t i = tid; i < B; i += kBlock)
            cnt[i] = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < K; i += kBlock) {
            const unsigned long long key = fanout_key(seed, n, r, i);
            kin[i] = key;
            atomicAdd(&cnt[(uint32_t)(key >> 16) >> shift], 1u);
        }
        __syncthreads();
        // bucket starts: thread tid owns buckets [tid m, tid m + m)
        uint32_t loc = 0;
        for (uint32_t j = 0; j < m; ++j)
            loc += cnt[tid * m + j];
        uint32_t incl = loc;
#pragma unroll
        for (uint32_t o = 1; o < (uint32_t)kWave; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, kWave);
            if (lane >= o)
                incl += v;
        }
        if (lane == kWave - 1)
            wsum[wv] = incl;
        __syncthreads();
        uint32_t pre = incl - loc;
        for (uint32_t q = 0; q < wv; ++q)
            pre += wsum[q];
        for (uint32_t j = 0; j < m; ++j) {
            off[tid * m + j] = pre;
            pre += cnt[tid * m + j];
        }
        __syncthreads();
        for (uint32_t i = tid; i < K; i += kBlock) {
            const unsigned long long key = kin[i];
            kout[atomicAdd(&off[(uint32_t)(key >> 16) >> shift], 1u)] = key;
        }
        __syncthreads();
        // off[b] is now the end of bucket b: sort its cnt[b] keys in place
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t b = tid * m + j;
            const uint32_t c = cnt[b];
            if (c > 1u) {
                const uint32_t s0 = off[b] - c;
                for (uint32_t x = s0 + 1u; x < s0 + c; ++x) {
                    const unsigned long long v = kout[x];
                    uint32_t y = x;
                    while (y > s0 && kout[y - 1u] > v) {
                        kout[y] = kout[y - 1u];
                        --y;
                    }
                    kout[y] = v;
                }
            }
        }
        __syncthreads();
        const uint32_t kind = (r >= n_exc) ? kKindBit : 0u;
        const int64_t base = (int64_t)r * (int64_t)K;
        for (uint32_t i = tid; i < K; i += kBlock) {
            const unsigned long long key = kout[i];
            const uint32_t t = (uint32_t)(key >> 16);
            const uint32_t k = (uint32_t)(key & 0xffffu);
            const Philox4 x = philox4x32_10(r, k, 1u, kTagGen, (uint32_t)seed, (uint32_t)(seed >> 32));
            col[base + i] = t | kind;
            w[base + i] = make_float2((float)(x.y >> 8) * 0x1p-24f, (float)(x.z >> 8) * 0x1p-24f);
            if (lds_tiles)
                atomicAdd(&tcount[t >> kTileBits], 1u);
            else
                atomicAdd(&tile_syn[t >> kTileBits], 1u);
        }
        __syncthreads();   // the LDS arrays are reused by the next row
    }
    if (lds_tiles) {
        for (uint32_t i = tid; i < ntiles; i += kBlock)
            if (tcount[i] != 0u)
                atomicAdd(&tile_syn[i], tcount[i]);
    }
}

}  // namespace

uint32_t fanout_sort_width(uint32_t K)
{
    uint32_t P = 1;
    while (P < K)
        P <<= 1;
    return P < 2 ? 2 : P;
}

hipError_t launch_gen_fanout(uint32_t n, uint32_t K, uint64_t seed, int64_t *rowptr, uint32_t *col, float2 *w,
                             uint32_t *tile_syn, hipStream_t stream)
{
    const uint32_t P = fanout_sort_width(K);
    const uint32_t n_exc = (uint32_t)((uint64_t)n * 8u / 10u);
    hipLaunchKernelGGL(k_fanout_rowptr, dim3((n + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, n, K, rowptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    const uint32_t ntiles = (n + kBlock - 1) / kBlock;
    const size_t lds_max = (size_t)64 << 10;
    if (P <= 2048u) {
        // counting sort: B >= P buckets of 2^shift consecutive targets
        const uint32_t B = P < (uint32_t)kBlock ? (uint32_t)kBlock : P;
        uint32_t shift = 0;
        while (((n - 1u) >> shift) >= B)
            ++shift;
        const size_t lds_keys = 2 * sizeof(unsigned long long) * P + 2 * sizeof(uint32_t) * B + 4 * sizeof(uint32_t);
        const size_t lds_all = lds_keys + sizeof(uint32_t) * ntiles;
        const uint32_t lds_tiles = lds_all <= lds_max ? 1u : 0u;
        const uint32_t cap = lds_tiles ? 2048u : 65536u;
        const uint32_t grid = n < cap ? n : cap;
        hipLaunchKernelGGL(k_fanout_rows_count, dim3(grid), dim3(kBlock), lds_tiles ? lds_all : lds_keys, stream, n,
                           K, P, B, shift, seed, n_exc, col, w, tile_syn, ntiles, lds_tiles);
        return hipGetLastError();
    }
    const size_t lds_keys = sizeof(unsigned long long) * P;
    const size_t lds_all = lds_keys + sizeof(uint32_t) * ntiles;
    const uint32_t lds_tiles = lds_all <= lds_max ? 1u : 0u;
    // with LDS counts, fewer workgroups each take many rows, so the flush of
    // ntiles counters per workgroup stays small against the synapses
    const uint32_t cap = lds_tiles ? 2048u : 65536u;
    const uint32_t grid = n < cap ? n : cap;
    hipLaunchKernelGGL(k_fanout_rows, dim3(grid), dim3(kBlock), lds_tiles ? lds_all : lds_keys, stream, n, K, P,
                       seed, n_exc, col, w, tile_syn, ntiles, lds_tiles);
    return hipGetLastError();
}

}  // namespace dtb
