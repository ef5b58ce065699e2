// dtb_kernels.hip -- gfx950 kernels of the DTB single-block step.
//
// Reference stage order for one step with t_steps substeps
// (tools/dist_simulator.cpp:2072-2315):
//   substeps: OU current (extern_stimuli.cpp:39-112) -> LIF membrane update
//             (membrane_voltage.cpp:114-195) -> [t_steps > 1: OR fired flags,
//             J decay, I_syn]
//   propagation of the fired rows into the delta buffers
//             (presynaptic_voltage.cpp:122-225)
//   J = J * decay + delta (presynaptic_voltage.cpp:46-120), I_syn
//             (synaptic_current.cpp:7-105)
//
// One launch per step here (k_step; see the end of this comment for the
// gather schedule's second launch).  Every workgroup of 256 threads owns a
// tile of 256 consecutive neurons of one instance, and per tile:
//   1. reads the tile's queue (the synapses delivered to it: target index,
//      channel pair, weight pair; contiguous, so the reads coalesce) and adds
//      the weights into exact Q23.40 accumulators in LDS (64-bit integer LDS
//      atomics, so the sum does not depend on the order);
//   2. folds the delta into J (J = J * decay + delta), computes I_syn, runs
//      OU + LIF for every substep;
//   3. lists the tile's fired rows in LDS and deals their passes of 128
//      synapses round-robin to its four waves; a wave reads a pass (synapses
//      ordered by target), reserves one contiguous segment per target tile
//      with an atomic add on that tile's count and writes each synapse's
//      entry into the target tile's queue for the next launch (12 bytes per
//      synapse).
// The queues are double-buffered: the launch of parity p drains set p and
// fills set p ^ 1, so step k consumes exactly the synapses queued from the
// spikes of step k - 1.  No global atomics touch the delta.
//   With the gather schedule (instances with group offsets) a k_gather
//   launch precedes each k_step: per group of 2048 target neurons it gathers
//   the last step's fired rows into the same exact LDS sums and writes the
//   delta, which k_step folds instead of draining a queue; k_step lists its
//   fired rows for the next k_gather instead of binning them.
//   k_materialize writes the pending delta of every neuron (without consuming
//   the queues) for sampling and state read-back; k_sample gathers samples.
#include <hip/hip_ext.h>

#include <type_traits>

#include "dtb_kernels.hpp"
#include "dtb_rng.hpp"

namespace dtb {

// Pointers read from the instance descriptors are generic; accessing them
// through the global address space lets the compiler emit global (not flat)
// memory instructions, which do not share the LDS wait counter.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gp(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

typedef float gfloat4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 ld4(const float4 *p, size_t i)
{
    const gfloat4 v = *(const __attribute__((address_space(1))) gfloat4 *)(p + i);
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ void st4(float4 *p, size_t i, const float4 v)
{
    gfloat4 g;
    g.x = v.x;
    g.y = v.y;
    g.z = v.z;
    g.w = v.w;
    *(__attribute__((address_space(1))) gfloat4 *)(p + i) = g;
}

__device__ __forceinline__ float q40_to_f32(long long q)
{
    return (float)((double)q * 0x1p-40);
}

__device__ __forceinline__ unsigned long long f32_to_q40(float w)
{
    return (unsigned long long)(long long)rintf(w * 0x1p40f);
}


// uint8 weight codes: u -> (float)u / 255.f, the correctly rounded quotient
// (presynaptic_voltage.cpp:171-176), folded at compile time (IEEE division).
struct U8Weights {
    float v[256];
};
constexpr U8Weights make_u8_weights()
{
    U8Weights t{};
    for (int u = 0; u < 256; ++u)
        t.v[u] = (float)u / 255.0f;
    return t;
}
__constant__ U8Weights c_u8_weight = make_u8_weights();

// Weight pairs (presynaptic_voltage.cpp:153-176): raw bits of pair j of an
// array in format fmt (at most 64 bits) and the pair they encode.
__device__ __forceinline__ unsigned long long w_raw(const void *w, size_t j, uint32_t fmt)
{
    if (fmt == kWeightF32)
        return gp(static_cast<const unsigned long long *>(w))[j];
    if (fmt == kWeightF16)
        return gp(static_cast<const uint32_t *>(w))[j];
    return gp(static_cast<const uint16_t *>(w))[j];
}

__device__ __forceinline__ float2 w_decode(unsigned long long b, uint32_t fmt)
{
    if (fmt == kWeightF32)
        return make_float2(__uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32)));
    if (fmt == kWeightF16)
        return make_float2((float)__builtin_bit_cast(_Float16, (uint16_t)b),
                           (float)__builtin_bit_cast(_Float16, (uint16_t)(b >> 16)));
    // uchar2 codes over 255
    return make_float2(c_u8_weight.v[b & 0xffu], c_u8_weight.v[(b >> 8) & 0xffu]);
}

// K weight pairs at indices jj[] of array w (one branch on the format for
// all K loads, so they are in flight together).
template <uint32_t K, typename Idx>
__device__ __forceinline__ void w_load(unsigned long long (&out)[K], const void *w, const Idx (&jj)[K],
                                       uint32_t fmt)
{
    if (fmt == kWeightF32) {
        const auto p = gp(static_cast<const unsigned long long *>(w));
#pragma unroll
        for (uint32_t k = 0; k < K; ++k)
            out[k] = p[jj[k]];
    } else if (fmt == kWeightF16) {
        const auto p = gp(static_cast<const uint32_t *>(w));
#pragma unroll
        for (uint32_t k = 0; k < K; ++k)
            out[k] = p[jj[k]];
    } else {
        const auto p = gp(static_cast<const uint16_t *>(w));
#pragma unroll
        for (uint32_t k = 0; k < K; ++k)
            out[k] = p[jj[k]];
    }
}

__device__ __forceinline__ const void *w_row(const void *w, int64_t first, uint32_t fmt)
{
    return static_cast<const char *>(w) + (size_t)first * weight_bytes(fmt);
}

// Row of the parameter records that holds neuron li's parameters.
__device__ __forceinline__ uint32_t param_index(const InstDev &I, uint32_t li)
{
    return (li & I.pmask) + (uint32_t)gp(I.cls)[li];
}

// synaptic_current.cpp:39-43 in source order, no contraction.
__device__ __forceinline__ float syn_current(const float4 g, const float4 vu, const float4 J, const float V)
{
    const float a = g.x * (vu.x - V) * J.x;
    const float b = g.y * (vu.y - V) * J.y;
    const float c = g.z * (vu.z - V) * J.z;
    const float d = g.w * (vu.w - V) * J.w;
    return a + b + c + d;
}

// presynaptic_voltage.cpp:62-73: J = J * decay + delta (fused multiply-add).
__device__ __forceinline__ float4 fold_delta(const float4 J, const float4 dc, const float4 dl)
{
    float4 r;
    r.x = fmaf(J.x, dc.x, dl.x);
    r.y = fmaf(J.y, dc.y, dl.y);
    r.z = fmaf(J.z, dc.z, dl.z);
    r.w = fmaf(J.w, dc.w, dl.w);
    return r;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

// The LDS sums are channel-major: channel c of row i at acc[c * rows + i].
// A wave's 64-bit atomics go to random rows; with the rows 8 bytes apart
// (not 32, as row-major would put them) they spread over twice as many LDS
// banks.
__device__ __forceinline__ uint32_t acc_at(uint32_t row, uint32_t chan, uint32_t rows)
{
    return chan * rows + row;
}

// Adds the queued synapses of one tile into acc[4][256] (Q23.40, LDS): every
// thread takes every 256th entry of the tile's queue (contiguous, so the
// reads coalesce) and adds its weight pair with 64-bit integer LDS atomics,
// so the sum does not depend on the order.  Threads tid < npre already hold
// entry tid in `pre`.
__device__ __forceinline__ void accumulate_tile(unsigned long long *acc, const QSyn *__restrict__ q, uint32_t cnt,
                                                const QSyn &pre, uint32_t npre, uint32_t tid)
{
    for (uint32_t k = tid; k < cnt; k += kBlock) {
        const QSyn e = (k < npre) ? pre : q[k];
        const uint32_t row = e.key & 0xffu, c0 = (e.key & kQSynKind) ? 2u : 0u;
        atomicAdd(&acc[acc_at(row, c0, kBlock)], f32_to_q40(e.w0));
        atomicAdd(&acc[acc_at(row, c0 + 1u, kBlock)], f32_to_q40(e.w1));
    }
}

// Zeroes row tid of a tile's sums (every row by the whole workgroup).
__device__ __forceinline__ void clear_row(unsigned long long *acc, uint32_t tid)
{
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c)
        acc[acc_at(tid, c, kBlock)] = 0ull;
}

// Adds neuron li's spill accumulator of queue set `set` into the neuron's own
// LDS sums (row tid of acc); `consume` clears it for the set's next use.
// Called by the neuron's thread after the queue accumulation.
__device__ __forceinline__ void take_spill(unsigned long long *acc, const InstDev &I, uint32_t set, uint32_t li,
                                           uint32_t tid, bool consume)
{
    const auto s = gp(I.spill) + ((size_t)set * I.n + li) * 4u;
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) {
        acc[acc_at(tid, c, kBlock)] += s[c];
        if (consume)
            s[c] = 0ull;
    }
}

// Pull propagation: adds into acc (the tile's Q23.40 LDS sums) the weight
// pairs of tile t's in-edges whose source row was emitted by the launch that
// filled set `set`: every emitted row (`all`: the queue overflowed and is
// ignored) or only the rows of tiles that deferred them (dbits).  An in-edge
// is one record (key, then the weight pair's bits: kW words), read
// coalesced; the source's bit comes from a per-wave bitmap (n / 8 bytes,
// cache resident; read as 32-bit words), so a batch is two dependent loads
// (records, bit words); each thread keeps kPullItems in-edges in flight.
// Called by the whole workgroup.
constexpr uint32_t kPullItems = 4;
template <uint32_t kFmt>
__device__ __forceinline__ void pull_loop(unsigned long long *acc, const uint32_t *rec,
                                          const unsigned long long *bits, uint32_t len, uint32_t tid)
{
    constexpr uint32_t U = kPullItems;
    constexpr uint32_t kW = kFmt == kWeightF32 ? 3u : 2u;   // grec_words(kFmt)
    // uniform bases and 32-bit byte offsets (a tile's in-edges span less
    // than 4 GB; the bitmap less than 1 MB): one VGPR per address
    using gbyte = const __attribute__((address_space(1))) unsigned char;
    gbyte *rb = reinterpret_cast<gbyte *>(gp(rec));
    gbyte *bb = reinterpret_cast<gbyte *>(gp(reinterpret_cast<const uint32_t *>(bits)));
    for (uint32_t k0 = tid; k0 < len; k0 += kBlock * U) {
        uint32_t kv[U];
        unsigned long long wr[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t j = k0 + u * kBlock;
            const auto r = reinterpret_cast<const __attribute__((address_space(1))) uint32_t *>(
                rb + (j < len ? j : len - 1u) * (kW * 4u));
            kv[u] = r[0];
            wr[u] = kW == 3u ? ((unsigned long long)r[1] | ((unsigned long long)r[2] << 32)) : r[1];
        }
        uint32_t bw[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
            bw[u] = *reinterpret_cast<const __attribute__((address_space(1))) uint32_t *>(
                bb + ((kv[u] & kPullRowMask) >> 5) * 4u);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            if (k0 + u * kBlock < len && ((bw[u] >> (kv[u] & 31u)) & 1u) != 0u) {
                const float2 wp = w_decode(wr[u], kFmt);
                const uint32_t row = (kv[u] >> kPullRowBits) & 0xffu, c0 = (kv[u] & kKindBit) ? 2u : 0u;
                atomicAdd(&acc[acc_at(row, c0, kBlock)], f32_to_q40(wp.x));
                atomicAdd(&acc[acc_at(row, c0 + 1u, kBlock)], f32_to_q40(wp.y));
            }
        }
    }
}

// Not inlined: the pull is the rare path (volleys, overflowing queues), and
// as a call it keeps its registers out of the step kernel's budget.  Its
// registers are the ones the caller must keep free across the call, so a
// kernel specialised for one weight format calls the loop of that format
// only (kFmt; kWeightAny: the instance's format at run time).
template <uint32_t kFmt>
__device__ __attribute__((noinline)) void pull_tile(unsigned long long *acc, const InstDev &I, uint32_t set,
                                                    uint32_t t, bool all, uint32_t tid)
{
    const size_t nw = (I.n + kWave - 1) / kWave;
    const unsigned long long *bits = (all ? I.ebits : I.dbits) + (size_t)set * nw;
    const int64_t b = gp(I.cptr)[t];
    const uint32_t len = (uint32_t)(gp(I.cptr)[t + 1] - b);   // a tile has fewer than 2^32 in-edges
    const uint32_t fmt = kFmt != kWeightAny ? kFmt : I.wfmt;
    const uint32_t *rec = I.crec + (size_t)b * grec_words(fmt);
    if (fmt == kWeightF32)
        pull_loop<kWeightF32>(acc, rec, bits, len, tid);
    else if (fmt == kWeightF16)
        pull_loop<kWeightF16>(acc, rec, bits, len, tid);
    else
        pull_loop<kWeightU8>(acc, rec, bits, len, tid);
}

// A statistic's value as the exact integer rint(x * 2^24).
__device__ __forceinline__ long long stat_q(float x)
{
    return (long long)rintf(x * 0x1p24f);
}

// Receptor currents g_u (V_u - V) J_u (synaptic_current.cpp:39-42 terms).
__device__ __forceinline__ float4 receptor_currents(const float4 g, const float4 vu, const float4 J, const float V)
{
    float4 r;
    r.x = g.x * (vu.x - V) * J.x;
    r.y = g.y * (vu.y - V) * J.y;
    r.z = g.z * (vu.z - V) * J.z;
    r.w = g.w * (vu.w - V) * J.w;
    return r;
}

// Adds this wave's per-population sums of the statistics selected by kmask
// (bit k: field k of kStatFields) of the lanes with `valid` into
// row[pop * kStatFields + k]: one 64-bit atomic per population present in
// the wave and field.  Called by the whole wave.
__device__ __forceinline__ void wave_pop_stats(long long *row, bool valid, uint32_t p, const long long *v,
                                               uint32_t kmask, uint32_t lane)
{
    unsigned long long rest = __ballot(valid);
    while (rest != 0ull) {
        const int src = __ffsll((long long)rest) - 1;
        const uint32_t ps = __shfl(p, src, kWave);
        const bool mine = valid && p == ps;
        rest &= ~__ballot(mine);
#pragma unroll
        for (int k = 0; k < kStatFields; ++k) {
            if (!(kmask & (1u << k)))
                continue;
            const unsigned long long s = wave_sum_u64(mine ? (unsigned long long)v[k] : 0ull);
            if ((int)lane == src)
                atomicAdd(reinterpret_cast<unsigned long long *>(row + (size_t)ps * kStatFields + k), s);
        }
    }
}

// Field mask of wave_pop_stats for the I_syn-based statistics of stat_mask.
__device__ __forceinline__ uint32_t current_fields(uint32_t stat_mask)
{
    return ((stat_mask & kStatI) ? 2u : 0u) | ((stat_mask & kStatReceptor) ? 0x3cu : 0u);
}

__device__ __forceinline__ float4 acc_to_delta(const unsigned long long *acc, uint32_t i, uint32_t rows = kBlock)
{
    float4 r;
    r.x = q40_to_f32((long long)acc[acc_at(i, 0u, rows)]);
    r.y = q40_to_f32((long long)acc[acc_at(i, 1u, rows)]);
    r.z = q40_to_f32((long long)acc[acc_at(i, 2u, rows)]);
    r.w = q40_to_f32((long long)acc[acc_at(i, 3u, rows)]);
    return r;
}

// Lane l gets lane l - 1's value (lane 0: fill), within the wave (DPP
// wave_shr:1, no LDS round trip).
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t fill)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xf, 0xf, false);
}

// Lane l gets lane l + 1's value (lane 63: fill) (DPP wave_shl:1).
__device__ __forceinline__ uint32_t wave_shl1(uint32_t v, uint32_t fill)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x130, 0xf, 0xf, false);
}

// Inclusive max-scan over the 64 lanes with DPP: shifts by 1, 2, 4, 8 within
// each row of 16 lanes (lanes shifted in from outside the row read 0, the
// identity), then the last lane of row 0 / 2 into rows 1 / 3 and the last
// lane of row 1 into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_max_scan(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));   // row_shr:1
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));   // row_shr:2
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));   // row_shr:4
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));   // row_shr:8
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));   // row_bcast:15
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));   // row_bcast:31
    return v;
}

// Adds the wave's synapses past their queue's capacity (`over` per lane) to
// the engine's diagnostic count: one atomic per wave, and none when no lane
// has any (the usual case).
__device__ __forceinline__ void count_over(StepCtl *__restrict__ ctl, uint32_t over, uint32_t lane)
{
    if (__ballot(over != 0u) != 0ull) {
        const unsigned long long s = wave_sum_u64((unsigned long long)over);
        if (lane == 0)
            atomicAdd(&ctl->slot[blockIdx.x & (kCtlSlots - 1u)].spilled, s);
    }
}

// bin_pass with the loads and reservations of a lane's synapses one after the
// other (measured faster for one synapse per lane).
template <uint32_t K, uint32_t kFmt>
__device__ __forceinline__ void bin_pass_serial(const uint32_t *__restrict__ col_all, const void *w_all, uint32_t wfmt,
                                        int64_t row_begin, uint32_t len, uint32_t c0, uint32_t tile_base,
                                        const TileDesc *__restrict__ tiles, uint32_t *__restrict__ qcnt, size_t qset, uint32_t cshift,
                                        QSyn *__restrict__ queue, unsigned long long *__restrict__ spill,
                                        StepCtl *__restrict__ ctl, unsigned long long *seg_dst, uint32_t *seg_room,
                                        uint32_t lane)
{
    if (kFmt != kWeightAny)
        wfmt = kFmt;   // the launch's instances share one format: decoded without branches
    constexpr uint32_t kNone = 0xffffffffu;   // "no start" marker for positions
    const auto col = gp(col_all) + row_begin;
    const void *wt = w_row(w_all, row_begin, wfmt);
    {
        const uint32_t j0 = c0 + lane * K;
        uint32_t cv[K], tl[K];
        float2 wv[K];
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            if (j0 + k < len) {
                cv[k] = col[j0 + k];
                wv[k] = w_decode(w_raw(wt, j0 + k, wfmt), wfmt);
                tl[k] = (cv[k] & kColMask) >> kTileBits;
            } else {
                cv[k] = 0u;
                wv[k] = make_float2(0.0f, 0.0f);
                tl[k] = 0xfffffff0u + k;
            }
        }
        // neighbours across lanes; the pass boundaries end segments
        const uint32_t prev0 = wave_shr1(tl[K - 1], 0xffffffe0u);
        const uint32_t nextK = wave_shl1(tl[0], 0xffffffe1u);
        uint32_t last_start = kNone;
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            const uint32_t pv = (k == 0) ? prev0 : tl[k - 1];
            if (j0 + k < len && tl[k] != pv)
                last_start = j0 + k;
        }
        // exclusive max-scan over lanes of last_start (kNone counts as -1)
        const uint32_t sc = wave_max_scan((last_start == kNone) ? 0u : last_start + 1u);
        const uint32_t before = wave_shr1(sc, 0u);
        uint32_t open = (before != 0u) ? before - 1u : c0;   // segment open at the lane's first synapse
        uint32_t st[K];
        uint32_t over = 0;   // synapses of this lane's segments past their queue's capacity
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            const uint32_t j = j0 + k;
            const uint32_t pv = (k == 0) ? prev0 : tl[k - 1];
            const uint32_t nx = (k == K - 1) ? nextK : tl[k + 1];
            if (j < len && tl[k] != pv)
                open = j;
            st[k] = open;
            if (j < len && tl[k] != nx) {
                const uint32_t tg = tile_base + tl[k];
                const uint32_t n = j + 1u - open;
                const uint32_t base = atomicAdd(&qcnt[count_slot(qset + tg, cshift)], n);
                const TileDesc td = tiles[tg];
                const uint32_t room = td.qcap > base ? min(n, td.qcap - base) : 0u;
                over += n - room;
                seg_dst[open - c0] = td.qoff + base;
                seg_room[open - c0] = room;
            }
        }
        count_over(ctl, over, lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            const uint32_t j = j0 + k;
            if (j < len) {
                const uint32_t at = j - st[k];
                if (at < seg_room[st[k] - c0]) {
                    QSyn e;
                    e.key = (cv[k] & 0xffu) | ((cv[k] & kKindBit) ? kQSynKind : 0u);
                    e.w0 = wv[k].x;
                    e.w1 = wv[k].y;
                    queue[seg_dst[st[k] - c0] + at] = e;
                } else if (spill != nullptr) {
                    unsigned long long *a = spill + (size_t)(cv[k] & kColMask) * 4u + ((cv[k] & kKindBit) ? 2u : 0u);
                    atomicAdd(a, f32_to_q40(wv[k].x));
                    atomicAdd(a + 1, f32_to_q40(wv[k].y));
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}


// Queues synapses c0 .. c0 + 64 K - 1 of one fired row (one pass) for the next
// launch: each synapse's target index within its tile, channel pair and
// weight pair go into the target tile's queue.  The row's synapses are
// ordered by target, so the synapses of one tile form segments; each segment
// reserves its slots with one atomic add on the tile's count and its
// synapses are stored contiguously.  Called by a whole wave with
// wave-uniform arguments.  Lane l holds synapses c0 + K l .. c0 + K l + K - 1
// of the row (segments end at pass boundaries).  A synapse's segment start is
// the last start at or before it (within the lane, else from a wave max-scan
// of the lanes' last starts); the lane holding the segment's last synapse
// reserves and publishes the destination and the number of the segment's
// synapses that fit the queue in seg_dst / seg_room (this wave's 64 K LDS
// entries, indexed by segment start within the pass).  Synapses past the
// queue's capacity add their weights into the target's spill accumulator
// instead (64-bit integer atomics, Q23.40 like the LDS sums, so the total
// does not depend on which path a synapse took).
template <uint32_t K, uint32_t kFmt>
__device__ __forceinline__ void bin_pass(const uint32_t *__restrict__ col_all, const void *w_all, uint32_t wfmt,
                                        int64_t row_begin, uint32_t len, uint32_t c0, uint32_t tile_base,
                                        const TileDesc *__restrict__ tiles, uint32_t *__restrict__ qcnt, size_t qset, uint32_t cshift,
                                        QSyn *__restrict__ queue, unsigned long long *__restrict__ spill,
                                        StepCtl *__restrict__ ctl, unsigned long long *seg_dst, uint32_t *seg_room,
                                        uint32_t lane)
{
    if (kFmt != kWeightAny)
        wfmt = kFmt;   // the launch's instances share one format: decoded without branches
    constexpr uint32_t kNone = 0xffffffffu;   // "no start" marker for positions
    const auto col = gp(col_all) + row_begin;
    const void *wt = w_row(w_all, row_begin, wfmt);
    const uint32_t j0 = c0 + lane * K;
    // Every lane loads K synapses without a branch (past the row's end it
    // re-reads the last one and marks it invalid), so all loads are in
    // flight together.
    uint32_t cv[K], tl[K], jj[K];
    float2 wv[K];
    bool ok[K];
    unsigned long long wr[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        ok[k] = j0 + k < len;
        jj[k] = ok[k] ? j0 + k : len - 1u;
        cv[k] = col[jj[k]];
    }
    w_load<K, uint32_t>(wr, wt, jj, wfmt);
#pragma unroll
    for (uint32_t k = 0; k < K; ++k)
        wv[k] = w_decode(wr[k], wfmt);
#pragma unroll
    for (uint32_t k = 0; k < K; ++k)
        tl[k] = ok[k] ? (cv[k] & kColMask) >> kTileBits : 0xfffffff0u + k;
    // neighbours across lanes; the pass boundaries end segments
    const uint32_t prev0 = wave_shr1(tl[K - 1], 0xffffffe0u);
    const uint32_t nextK = wave_shl1(tl[0], 0xffffffe1u);
    uint32_t last_start = kNone;
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t pv = (k == 0) ? prev0 : tl[k - 1];
        if (ok[k] && tl[k] != pv)
            last_start = j0 + k;
    }
    // exclusive max-scan over lanes of last_start (kNone counts as -1)
    const uint32_t sc = wave_max_scan((last_start == kNone) ? 0u : last_start + 1u);
    const uint32_t before = wave_shr1(sc, 0u);
    uint32_t open = (before != 0u) ? before - 1u : c0;   // segment open at the lane's first synapse
    uint32_t st[K];
    bool end[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t pv = (k == 0) ? prev0 : tl[k - 1];
        const uint32_t nx = (k == K - 1) ? nextK : tl[k + 1];
        if (ok[k] && tl[k] != pv)
            open = j0 + k;
        st[k] = open;
        end[k] = ok[k] && tl[k] != nx;
    }
    // Reservations: every segment-end synapse issues its atomic and its
    // tile's queue range before any result is used, so they overlap.
    uint32_t base[K];
    unsigned long long qoff[K];
    uint32_t qcap[K];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        base[k] = 0u;
        qoff[k] = 0ull;
        qcap[k] = 0u;
        if (end[k]) {
            const uint32_t tg = tile_base + tl[k];
            base[k] = atomicAdd(&qcnt[count_slot(qset + tg, cshift)], j0 + k + 1u - st[k]);
            qoff[k] = tiles[tg].qoff;
            qcap[k] = tiles[tg].qcap;
        }
    }
    uint32_t over = 0;   // synapses of this lane's segments past their queue's capacity
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        if (end[k]) {
            const uint32_t n = j0 + k + 1u - st[k];
            const uint32_t room = qcap[k] > base[k] ? min(n, qcap[k] - base[k]) : 0u;
            over += n - room;
            seg_dst[st[k] - c0] = qoff[k] + base[k];
            seg_room[st[k] - c0] = room;
        }
    }
    count_over(ctl, over, lane);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        if (ok[k]) {
            const uint32_t at = j0 + k - st[k];
            if (at < seg_room[st[k] - c0]) {
                QSyn e;
                e.key = (cv[k] & 0xffu) | ((cv[k] & kKindBit) ? kQSynKind : 0u);
                e.w0 = wv[k].x;
                e.w1 = wv[k].y;
                queue[seg_dst[st[k] - c0] + at] = e;
            } else if (spill != nullptr) {
                unsigned long long *a = spill + (size_t)(cv[k] & kColMask) * 4u + ((cv[k] & kKindBit) ? 2u : 0u);
                atomicAdd(a, f32_to_q40(wv[k].x));
                atomicAdd(a + 1, f32_to_q40(wv[k].y));
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// One tile's step (the body of k_step).  kUniform: every neuron of the tile
// has parameter class d.pcls - 1, so the parameter records are read once per
// tile through the scalar unit instead of per lane.  K: synapses per lane in
// a binning pass (a pass is 64 K synapses).
template <bool kUniform, uint32_t K, uint32_t kFmt>
__device__ __forceinline__ void step_tile(const TileDesc &d, const TileDesc *__restrict__ tiles,
                                          const InstDev *__restrict__ insts, const QueueSet &qs, uint32_t par,
                                          StepCtl *__restrict__ ctl, unsigned long long *acc,
                                          unsigned long long *seg_dst, uint32_t *seg_room, long long *row_rb,
                                          uint32_t *row_len, uint32_t &nrows, uint32_t &s_fit)
{
    const InstDev &I = insts[d.inst];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t w = tid / kWave;
    const bool gat = I.segidx != nullptr;   // gather propagation: k_gather wrote the delta into dlt
    const QSyn *__restrict__ q_in = I.queue + (size_t)par * I.qslots;
    const size_t cslot_in = count_slot((size_t)par * qs.ntiles + blockIdx.x, qs.cshift);
    QSyn *__restrict__ q_out = I.queue + (size_t)(par ^ 1u) * I.qslots;

    // Loads are issued in order of their dependence chains, before anything
    // waits on them: the queue count and step index, the parameter class,
    // the neuron state, the parameter records (need the class), the queue
    // head, the deferral stamp.
    const uint32_t cnt_raw = qs.count[cslot_in];
    const int64_t step = gp(I.step)[par];
    const bool pull_on = I.crec != nullptr;
    const uint32_t li = d.begin + tid;
    const bool active = tid < d.count;
    const uint32_t lc = active ? li : d.begin;   // in-range index for idle lanes
    // A tile whose neurons share one parameter class (the usual case: a
    // population spans many tiles) reads its records without the class load.
    const uint32_t c = kUniform ? d.pcls - 1u : (lc & I.pmask) + (uint32_t)gp(I.cls)[lc];
    float4 J = ld4(I.J, lc);
    float V = gp(I.V)[lc];
    int32_t t = gp(I.tcnt)[lc];
    uint32_t fl = gp(I.flag)[lc];
    float iou = gp(I.iou)[lc];
    float4 pa = ld4(I.pA, c);
    float4 pb = ld4(I.pB, c);
    float4 g = ld4(I.gU, c);
    float4 vu = ld4(I.vU, c);
    float4 dc = ld4(I.dec, c);
    // The ends of the tile's first two populations (a tile rarely spans
    // more) arrive with the state, not after the fold.
    const uint32_t pop_hi = min(d.pop0 + 1u, I.npop - 1u);
    uint32_t end_a = I.pops[d.pop0].end;
    uint32_t end_b = I.pops[pop_hi].end;
    // The queue is complete when the launch starts; entries past the count
    // are ignored (the queue buffer has kQueuePrefetch slots of slack at its
    // end, so the read needs no branch).
    QSyn head{};
    float4 dl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (gat)
        dl = ld4(I.dlt, lc);
    else
        head = q_in[d.qoff + min(tid, kQueuePrefetch - 1u)];
    const int64_t stamp = pull_on ? gp(I.dstamp)[par] : 0;
    // A tile pulls (the pull propagation, never with gather) when its queue
    // overflowed or rows were deferred into this set: rare (volleys).
    const bool over = !gat && cnt_raw > d.qcap;
    const bool pull = !gat && pull_on && (over || stamp == step + 1);
    clear_row(acc, tid);
    if (tid == 0)
        nrows = 0u;
    if (pull) {
        __syncthreads();
        // An overflowed queue is ignored: every emitted row then comes
        // through the in-edges.
        accumulate_tile(acc, q_in + d.qoff, over ? 0u : cnt_raw, QSyn{}, 0u, tid);
        pull_tile<kFmt>(acc, I, par, d.begin >> kTileBits, over, tid);
        if (tid == 0)
            atomicAdd(&ctl->slot[blockIdx.x & (kCtlSlots - 1u)].pulled, 1ull);
        // loaded again, so that nothing loaded above is live across the call
        J = ld4(I.J, lc);
        V = gp(I.V)[lc];
        t = gp(I.tcnt)[lc];
        fl = gp(I.flag)[lc];
        iou = gp(I.iou)[lc];
        pa = ld4(I.pA, c);
        pb = ld4(I.pB, c);
        g = ld4(I.gU, c);
        vu = ld4(I.vU, c);
        dc = ld4(I.dec, c);
        end_a = I.pops[d.pop0].end;
        end_b = I.pops[pop_hi].end;
        dl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // (no gather here)
    } else {
        const uint32_t cnt = gat ? 0u : min(cnt_raw, d.qcap);
        __syncthreads();
        accumulate_tile(acc, q_in + d.qoff, cnt, head, kQueuePrefetch, tid);
    }
    __syncthreads();
    if (!pull_on && over && active)
        take_spill(acc, I, par, li, tid, true);   // synapses past the queue's capacity
    if (tid == 0)
        qs.count[cslot_in] = 0u;   // this launch consumed the queue
    if (d.begin == 0 && tid == 0)
        gp(I.step)[par ^ 1u] = step + 1;   // no tile of this launch reads this slot

    bool emit = false;
    uint32_t p = d.pop0;
    float isyn = 0.0f;
    if (active) {
        if (!(fl & kFlagFolded))
            J = fold_delta(J, dc, gat ? dl : acc_to_delta(acc, tid));
        isyn = syn_current(g, vu, J, V);
        // the population of neuron li (populations in order of their ends)
        if (li >= end_a && p + 1 < I.npop) {
            p = pop_hi;
            if (li >= end_b)
                while (p + 1 < I.npop && li >= I.pops[p].end)
                    ++p;
        }
    }
    // stat_result.cpp:266-391: population sums, blocked neurons excluded.
    // The I_syn-based statistics belong to the previous step (its J is
    // complete only now: J folded, V before this step's update); the V
    // statistic (after the update, below) to this one.
    const int64_t row_v = step - I.rec_step0;
    const bool stat_valid = I.rec_stat != nullptr && active && !gp(I.blocked)[lc];
    if (I.rec_stat != nullptr) {
        const uint32_t cf = current_fields(I.stat_mask);
        if (cf != 0u && row_v - 1 >= 0 && row_v - 1 < (int64_t)I.rec_len) {
            long long stv[kStatFields] = {};
            const float4 rc = receptor_currents(g, vu, J, V);
            stv[1] = stat_q(isyn);
            stv[2] = stat_q(rc.x);
            stv[3] = stat_q(rc.y);
            stv[4] = stat_q(rc.z);
            stv[5] = stat_q(rc.w);
            wave_pop_stats(I.rec_stat + (size_t)(row_v - 1) * I.npop * kStatFields, stat_valid, p, stv, cf, lane);
        }
    }
    if (active) {
        uint32_t f = fl & kFlagSpike;
        const PopDev pop = I.pops[p];
        const bool ou = I.use_ou && pop.ou_on;
        const int32_t T = I.t_steps;
        const int32_t rsteps = __float_as_int(pb.w);

        for (int32_t s = 0; s < T; ++s) {
            float Ieff = isyn;
            if (ou) {
                const float xi = ou_normal(I.seed, li, (uint64_t)step * (uint64_t)T + (uint64_t)s);
                iou = iou + (pop.f1 * (pop.mu - iou) + pop.f2s * xi);
                Ieff = isyn + iou;
            }
            bool spk = false;
            if (t < 0) {
                if (f) {
                    V = pa.x;
                    f = 0;
                }
                t = t + 1;
            } else {
                float cvi = pb.x * (pa.w - V);
                cvi = cvi + Ieff;
                cvi = cvi + pb.y;
                V = V + I.dt_sub * pa.z * cvi;
                spk = V >= pa.y;
                if (spk) {
                    V = pa.y;
                    t = 1 - rsteps;
                }
                f = spk ? 1u : 0u;
            }
            emit = emit || spk;
            if (s < T - 1) {
                J.x = J.x * dc.x;
                J.y = J.y * dc.y;
                J.z = J.z * dc.z;
                J.w = J.w * dc.w;
                isyn = syn_current(g, vu, J, V);
            }
        }
        if (!I.use_ou) {
            // spike.cpp:55-62: f_rec = f | (U < noise_rate)
            const float u = noise_uniform(I.seed, li, (uint64_t)step);
            emit = f || (u < pb.z);
        }
        gp(I.V)[li] = V;
        gp(I.tcnt)[li] = t;
        gp(I.flag)[li] = (uint8_t)(f | (emit ? kFlagEmit : 0u));
        gp(I.iou)[li] = iou;
        st4(I.J, li, J);
    }
    if (I.rec_stat != nullptr && (I.stat_mask & kStatV) && row_v >= 0 && row_v < (int64_t)I.rec_len) {
        long long stv[kStatFields] = {};
        stv[0] = stat_q(V);
        wave_pop_stats(I.rec_stat + (size_t)row_v * I.npop * kStatFields, stat_valid, p, stv, 1u, lane);
    }

    // Spike propagation (presynaptic_voltage.cpp:122-225): the tile's fired
    // rows are listed in LDS, and their passes of 64 K synapses are dealt
    // round-robin to the four waves, so one long row does not keep a single
    // wave busy while the others idle.
    const unsigned long long m = __ballot(emit);
    const bool wave_in = d.begin + w * kWave < I.n;   // this wave's neurons exist
    const size_t nw = (I.n + kWave - 1) / kWave;
    const size_t bw = (size_t)(par ^ 1u) * nw + (d.begin >> 6) + w;   // this wave's bitmap word of set par ^ 1
    if (pull_on && lane == 0 && wave_in)
        gp(I.ebits)[bw] = m;
    if (m != 0ull) {
        const uint32_t cnt_w = (uint32_t)__popcll(m);
        const int leader = __ffsll((long long)m) - 1;
        long long rb = 0;
        uint32_t len = 0;
        if (emit) {
            rb = gp(I.rowptr)[li];
            len = (uint32_t)(gp(I.rowptr)[li + 1] - rb);
        }
        // events = sum of the fired rows' lengths: few per wave, read lane by
        // lane through the scalar unit
        unsigned long long ev = 0;
        for (unsigned long long r = m; r != 0ull; r &= r - 1ull)
            ev += (uint32_t)__builtin_amdgcn_readlane((int)len, __ffsll((long long)r) - 1);
        uint32_t base = 0;
        if ((int)lane == leader) {
            // the wave's own counters, added without waiting for a read
            // (a load-add-store would stall the wave before its binning)
            unsigned long long *c = I.wave_ctr + 2u * ((d.begin >> 6) + w);
            atomicAdd(c, ev);
            atomicAdd(c + 1, (unsigned long long)cnt_w);
            if (!gat)
                base = atomicAdd(&nrows, cnt_w);
        }
        if (gat) {
            // gather propagation: the rows with synapses go to this tile's
            // fired list of set par ^ 1 for the next launch's k_gather (one
            // atomic per wave, on a count shared with 1/64 of the tiles)
            const bool has = emit && len != 0u;
            const unsigned long long mg = __ballot(has);
            if (mg != 0ull) {
                const uint32_t list = (par ^ 1u) * kFiredLists + ((d.begin >> kTileBits) & (kFiredLists - 1u));
                const int lg = __ffsll((long long)mg) - 1;
                uint32_t gb = 0;
                if ((int)lane == lg)
                    gb = atomicAdd(I.fcount + (size_t)list * kFiredCountStride, (uint32_t)__popcll(mg));
                gb = (uint32_t)__builtin_amdgcn_readlane((int)gb, lg);
                if (has)
                    gp(I.fired)[(size_t)list * I.fcap + gb + (uint32_t)__popcll(mg & ((1ull << lane) - 1ull))] = li;
            }
        } else {
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
            if (emit) {
                const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                row_rb[k] = rb;
                row_len[k] = len;
            }
        }
        if (I.rec_counts != nullptr) {
            // stat_result.cpp:161-206: one atomic per population present in
            // the wave (blocked neurons are not counted)
            const int64_t row = step - I.rec_step0;
            if (row >= 0 && row < (int64_t)I.rec_len) {
                const bool counted = emit && !gp(I.blocked)[li];
                unsigned long long rest = __ballot(counted);
                while (rest != 0ull) {
                    const int src = __ffsll((long long)rest) - 1;
                    const uint32_t ps = __shfl(p, src, kWave);
                    const unsigned long long same = __ballot(counted && p == ps);
                    if ((int)lane == src)
                        atomicAdd(&I.rec_counts[(size_t)row * I.npop + ps], (uint32_t)__popcll(same));
                    rest &= ~same;
                }
            }
        }
    }
    __syncthreads();
    const uint32_t nr = nrows;
    // A tile with many emitted rows (a synchronous volley) leaves them to
    // the consumers' pull: binning them would cost one reservation per
    // (row, target tile) and overflow the queues.  A tile that pulls in
    // this launch (the volley is under way) defers only from four times
    // the count: the volley's tail, where a few tiles still emit that many,
    // would otherwise cost the next step a pull of every in-edge.
    const bool defer = pull_on && I.defer_min != 0u && nr >= (pull ? 4u * I.defer_min : I.defer_min);
    if (pull_on) {
        if (lane == 0 && wave_in)
            gp(I.dbits)[bw] = defer ? m : 0ull;
        if (defer && tid == 0) {
            gp(I.dstamp)[par ^ 1u] = step + 2;   // consumer's step + 1 (0: none)
            atomicAdd(&ctl->slot[blockIdx.x & (kCtlSlots - 1u)].deferred, 1ull);
        }
    }
    const uint32_t nbin = defer ? 0u : nr;
    unsigned long long *spill_out = pull_on ? nullptr : I.spill + (size_t)(par ^ 1u) * I.n * 4u;
    // Split binning: the tile's passes (rows in LDS order, passes in row
    // order) go to the engine's pass list for the k_bin launch that follows,
    // where every wave of the GPU takes one; the tile bins here only the
    // passes past the list's capacity.
    uint32_t first_here = 0;   // tile pass index from which this tile bins itself
    if (qs.pass != nullptr && nbin != 0u) {
        if (w == 0) {
            constexpr uint32_t kPass = kWave * K;
            uint32_t total = 0;
            for (uint32_t r0 = 0; r0 < nbin; r0 += kWave) {
                const uint32_t r = r0 + lane;
                const uint32_t np = r < nbin ? (row_len[r] + kPass - 1u) / kPass : 0u;
                total += (uint32_t)wave_sum_u64((unsigned long long)np);
            }
            uint32_t base = 0;
            if (lane == 0)
                base = atomicAdd(qs.pcount + (size_t)par * kPassCountStride, total);
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, 0);
            const uint32_t fit = qs.pcap > base ? min(total, qs.pcap - base) : 0u;
            uint32_t off = 0;   // passes of the rows of earlier chunks
            for (uint32_t r0 = 0; r0 < nbin; r0 += kWave) {
                const uint32_t r = r0 + lane;
                const uint32_t len = r < nbin ? row_len[r] : 0u;
                const uint32_t np = (len + kPass - 1u) / kPass;
                uint32_t incl = np;
#pragma unroll
                for (uint32_t o = 1; o < (uint32_t)kWave; o <<= 1) {
                    const uint32_t v = __shfl_up(incl, o, kWave);
                    if (lane >= o)
                        incl += v;
                }
                const uint32_t ex = off + incl - np;   // this row's first tile pass
                const long long rb = r < nbin ? row_rb[r] : 0;
                for (uint32_t j = 0; j < np && ex + j < fit; ++j) {
                    PassEntry e;
                    e.first = (uint64_t)(rb + (long long)j * kPass);
                    e.len = min(kPass, len - j * kPass);
                    e.inst = d.inst;
                    qs.pass[base + ex + j] = e;
                }
                off += (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
            }
            if (lane == 0)
                s_fit = fit;
        }
        __syncthreads();
        first_here = s_fit;
    }
    uint32_t pass = 0;
    for (uint32_t r = 0; r < nbin; ++r) {
        const long long rb_s = row_rb[r];
        const uint32_t len_s = row_len[r];
        for (uint32_t c0 = 0; c0 < len_s; c0 += kWave * K, ++pass)
            if ((pass & 3u) == w && pass >= first_here)
            {
                if constexpr (K == 1)
                    bin_pass_serial<K, kFmt>(I.col, I.w, I.wfmt, (int64_t)rb_s, len_s, c0, I.tile_base, tiles, qs.count,
                                       (size_t)(par ^ 1u) * qs.ntiles, qs.cshift, q_out,
                                       spill_out, ctl, seg_dst + w * kWave * K,
                                       seg_room + w * kWave * K, lane);
                else
                    bin_pass<K, kFmt>(I.col, I.w, I.wfmt, (int64_t)rb_s, len_s, c0, I.tile_base, tiles, qs.count,
                                (size_t)(par ^ 1u) * qs.ntiles, qs.cshift, q_out,
                                spill_out, ctl, seg_dst + w * kWave * K,
                                seg_room + w * kWave * K, lane);
            }
    }
}

// The step of launch parity `par`: drains queue set par, fills set par ^ 1,
// reads the step index from step[par] and advances it into step[par ^ 1].
// K: synapses per lane in a binning pass (the schedule; every K gives the
// same results).  Eight waves per SIMD for K = 1, 2; K = 4's larger LDS
// segment tables allow six.
template <uint32_t K, uint32_t kFmt>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(K == 4 ? 6 : 8))) void k_step(const TileDesc *__restrict__ tiles,
                                                 const InstDev *__restrict__ insts,
                                                 QueueSet qs,
                                                 uint32_t par,
                                                 StepCtl *__restrict__ ctl)
{
    __shared__ unsigned long long acc[kBlock * 4];
    __shared__ unsigned long long seg_dst[4 * kWave * K];   // per wave: one pass
    __shared__ uint32_t seg_room[4 * kWave * K];
    __shared__ long long row_rb[kBlock];                  // fired rows of the tile: first synapse
    __shared__ uint32_t row_len[kBlock];                  // and length
    __shared__ uint32_t nrows;
    __shared__ uint32_t s_fit;                            // split binning: passes in the list

    const TileDesc d = tiles[blockIdx.x];
    if (d.pcls != 0u)
        step_tile<true, K, kFmt>(d, tiles, insts, qs, par, ctl, acc, seg_dst, seg_room, row_rb, row_len, nrows, s_fit);
    else
        step_tile<false, K, kFmt>(d, tiles, insts, qs, par, ctl, acc, seg_dst, seg_room, row_rb, row_len, nrows,
                                  s_fit);
}

// Split binning: bins the passes that the k_step launch of parity `par`
// listed (into queue set par ^ 1, as k_step would have), one pass per wave
// at a time over the whole grid, so a step's propagation is spread over
// every CU instead of the tiles whose neurons fired.  Also empties the other
// parity's list count (its last reader was the previous k_bin; the next
// k_step appends to it).
template <uint32_t K, uint32_t kFmt>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_bin(const TileDesc *__restrict__ tiles,
                                                const InstDev *__restrict__ insts,
                                                QueueSet qs,
                                                uint32_t par,
                                                StepCtl *__restrict__ ctl)
{
    __shared__ unsigned long long seg_dst[4 * kWave * K];
    __shared__ uint32_t seg_room[4 * kWave * K];
    const uint32_t w = threadIdx.x / kWave;
    const uint32_t lane = __lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0)
        qs.pcount[(size_t)(par ^ 1u) * kPassCountStride] = 0u;
    const uint32_t total = min(qs.pcount[(size_t)par * kPassCountStride], qs.pcap);
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);
    for (uint32_t i = blockIdx.x * (kBlock / kWave) + w; i < total; i += nwaves) {
        const PassEntry e = qs.pass[i];
        const InstDev &I = insts[e.inst];
        QSyn *q_out = I.queue + (size_t)(par ^ 1u) * I.qslots;
        unsigned long long *spill_out = I.crec != nullptr ? nullptr : I.spill + (size_t)(par ^ 1u) * I.n * 4u;
        if constexpr (K == 1)
            bin_pass_serial<K, kFmt>(I.col, I.w, I.wfmt, (int64_t)e.first, e.len, 0u, I.tile_base, tiles, qs.count,
                               (size_t)(par ^ 1u) * qs.ntiles, qs.cshift, q_out, spill_out, ctl,
                               seg_dst + w * kWave * K, seg_room + w * kWave * K, lane);
        else
            bin_pass<K, kFmt>(I.col, I.w, I.wfmt, (int64_t)e.first, e.len, 0u, I.tile_base, tiles, qs.count,
                        (size_t)(par ^ 1u) * qs.ntiles, qs.cshift, q_out, spill_out, ctl,
                        seg_dst + w * kWave * K, seg_room + w * kWave * K, lane);
    }
}

// Gather propagation (instances with InstDev::segidx; their k_step reads
// the delta from dlt): the delta of the rows fired by the last launch (the
// fired lists of set `par`) for one group of kGroupTiles tiles, written into
// dlt.  A wave takes
// 64 fired rows at a time: lane l reads row l's segment bounds for the group
// from the row's group offsets, a wave prefix sum lays the segments end to
// end, and the wave then reads them with consecutive lanes on consecutive
// synapses (a segment is contiguous in the row), adding each weight pair
// into the group's Q23.40 LDS sums (64-bit integer atomics, so the result
// does not depend on the order); each lane has kGatherItems synapses in
// flight.  Groups of 2048 neurons give segments of about 20 synapses at the
// stress shape (in-degree 5000 on 500k neurons), which is what coalesces the
// reads; one 1024-thread workgroup per group.  Also empties the fired lists
// of set par ^ 1 (the ones the k_step launch after this one fills; their
// last reader was the previous k_gather).  Consecutive groups go to one XCD (workgroups
// are dealt to the 8 XCDs round-robin by index), so the segments of
// neighbouring groups, which share cache lines, meet in one L2.
template <uint32_t kFmt>
__global__ __launch_bounds__(kGatherBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_gather(const GroupDesc *__restrict__ groups,
                                                         const InstDev *__restrict__ insts, uint32_t par)
{
    constexpr uint32_t kWaves = kGatherBlock / kWave;
    constexpr uint32_t kNeurons = 1u << kGroupBits;
    __shared__ unsigned long long acc[kNeurons * 4];
    __shared__ uint32_t s_off[kWaves][kWave];             // per wave: first item of each row's segment
    __shared__ unsigned long long s_base[kWaves][kWave];  // and the segment's first synapse
    __shared__ uint32_t s_list[kWaves][kWave];            // per wave: first row of each fired list
    const uint32_t nb = gridDim.x, x = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const GroupDesc d = groups[x * (nb >> 3) + min(x, nb & 7u) + slot];
    const InstDev &I = insts[d.inst];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t wv = tid / kWave;
    for (uint32_t i = tid; i < kNeurons * 4; i += kGatherBlock)
        acc[i] = 0ull;
    // the fired lists of set par laid end to end: lane l holds list l's count
    const uint32_t lcnt = gp(I.fcount)[((size_t)par * kFiredLists + lane) * kFiredCountStride];
    uint32_t lincl = lcnt;
#pragma unroll
    for (uint32_t o = 1; o < (uint32_t)kWave; o <<= 1) {
        const uint32_t v = __shfl_up(lincl, o, kWave);
        if (lane >= o)
            lincl += v;
    }
    const uint32_t nf = (uint32_t)__builtin_amdgcn_readlane((int)lincl, kWave - 1);
    s_list[wv][lane] = lincl - lcnt;
    if (d.g == 0 && tid < kFiredLists)
        gp(I.fcount)[((size_t)(par ^ 1u) * kFiredLists + tid) * kFiredCountStride] = 0u;
    __syncthreads();
    const auto fired = gp(I.fired) + (size_t)par * kFiredLists * I.fcap;
    const uint32_t wfmt = kFmt != kWeightAny ? kFmt : I.wfmt;
    const uint32_t stride = I.ngroups + 1u;
    const uint32_t first = d.g << kGroupBits;
    for (uint32_t c0 = wv * kWave; c0 < nf; c0 += kGatherBlock) {
        const uint32_t k = c0 + lane;
        const bool valid = k < nf;
        const uint32_t kc = valid ? k : nf - 1u;
        // the list holding row kc: the last list starting at or before it
        uint32_t ls = 0;
#pragma unroll
        for (uint32_t st = kWave / 2; st > 0; st >>= 1)
            if (s_list[wv][ls + st] <= kc)
                ls += st;
        const uint32_t r = fired[(size_t)ls * I.fcap + (kc - s_list[wv][ls])];
        const auto sg = gp(I.segidx) + (size_t)r * stride + d.g;
        const uint32_t a = sg[0];
        const uint32_t e = sg[1];
        const int64_t rb = gp(I.rowptr)[r];
        const uint32_t len = (valid && e > a) ? e - a : 0u;
        uint32_t incl = len;
#pragma unroll
        for (uint32_t o = 1; o < (uint32_t)kWave; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o, kWave);
            if (lane >= o)
                incl += v;
        }
        const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
        s_off[wv][lane] = incl - len;
        s_base[wv][lane] = (unsigned long long)(rb + (int64_t)a);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t it = 0; it < total; it += kGatherItems * kWave) {
            uint32_t cv[kGatherItems];
            size_t jw[kGatherItems];
            unsigned long long wb[kGatherItems];
            bool ok[kGatherItems];
#pragma unroll
            for (uint32_t u = 0; u < kGatherItems; ++u) {
                const uint32_t item = it + u * kWave + lane;
                ok[u] = item < total;
                const uint32_t ic = ok[u] ? item : total - 1u;
                // the row whose segment holds item: the last lane whose
                // segment starts at or before it (empty segments before a
                // non-empty one start at the same item and lose to it)
                uint32_t lo = 0;
#pragma unroll
                for (uint32_t st = kWave / 2; st > 0; st >>= 1)
                    if (s_off[wv][lo + st] <= ic)
                        lo += st;
                jw[u] = (size_t)s_base[wv][lo] + (ic - s_off[wv][lo]);
            }
            // one record per synapse: column and weight pair side by side
            // (3 or 2 words at a word stride of grec_words(wfmt))
            const auto rec = gp(I.grec);
            if (wfmt == kWeightF32) {
#pragma unroll
                for (uint32_t u = 0; u < kGatherItems; ++u) {
                    const auto r = rec + jw[u] * 3u;
                    cv[u] = r[0];
                    wb[u] = (unsigned long long)r[1] | ((unsigned long long)r[2] << 32);
                }
            } else {
#pragma unroll
                for (uint32_t u = 0; u < kGatherItems; ++u) {
                    const auto r = rec + jw[u] * 2u;
                    cv[u] = r[0];
                    wb[u] = r[1];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < kGatherItems; ++u) {
                if (ok[u]) {
                    const float2 wp = w_decode(wb[u], wfmt);
                    const uint32_t row = (cv[u] & kColMask) - first, c0 = (cv[u] & kKindBit) ? 2u : 0u;
                    atomicAdd(&acc[acc_at(row, c0, kNeurons)], f32_to_q40(wp.x));
                    atomicAdd(&acc[acc_at(row, c0 + 1u, kNeurons)], f32_to_q40(wp.y));
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    for (uint32_t i = tid; i < d.count; i += kGatherBlock)
        st4(I.dlt, first + i, acc_to_delta(acc, i, kNeurons));
}

// Gather records: synapse j's column and weight-pair bits side by side.
__global__ __launch_bounds__(kBlock) void k_build_grec(const uint32_t *__restrict__ col, const void *__restrict__ w,
                                                       uint32_t wfmt, uint64_t nsyn, uint32_t *__restrict__ rec)
{
    const uint32_t nw = grec_words(wfmt);
    for (uint64_t j = (uint64_t)blockIdx.x * kBlock + threadIdx.x; j < nsyn; j += (uint64_t)gridDim.x * kBlock) {
        const unsigned long long b = w_raw(w, j, wfmt);
        uint32_t *r = rec + j * nw;
        r[0] = col[j];
        r[1] = (uint32_t)b;
        if (nw == 3u)
            r[2] = (uint32_t)(b >> 32);
    }
}

// Group offsets of the gather propagation, one wave per row: segidx[r][g] =
// offset in row r of the first synapse whose target lies in group g or a
// later one, segidx[r][ngroups] = row length.  Rows must be ordered by
// target group (checked on the host).
__global__ __launch_bounds__(kBlock) void k_build_segidx(const int64_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ col, uint32_t n,
                                                         uint32_t ngroups, uint16_t *__restrict__ segidx)
{
    const uint32_t r = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const uint32_t lane = __lane_id();
    if (r >= n)
        return;
    const int64_t rb = rowptr[r];
    const uint32_t len = (uint32_t)(rowptr[r + 1] - rb);
    uint16_t *s = segidx + (size_t)r * (ngroups + 1u);
    for (uint32_t j = lane; j < len; j += kWave) {
        const uint32_t gc = (col[rb + j] & kColMask) >> kGroupBits;
        const uint32_t g0 = j == 0 ? 0u : ((col[rb + j - 1] & kColMask) >> kGroupBits) + 1u;
        for (uint32_t g = g0; g <= gc; ++g)
            s[g] = (uint16_t)j;
    }
    const uint32_t tail = len == 0 ? 0u : ((col[rb + len - 1] & kColMask) >> kGroupBits) + 1u;
    for (uint32_t g = tail + lane; g <= ngroups; g += kWave)
        s[g] = (uint16_t)len;
}

// In-edges of the pull propagation: every synapse of source row r goes to
// its target tile t's range of records (from cptr[t]) as (key = source row |
// target in tile | channel bit, weight pair bits).  A workgroup takes
// kBuildRows consecutive rows: it counts their synapses per target tile in
// LDS, reserves each tile's share with one atomic on the tile's cursor
// (cursor[t]: records placed so far, from 0), then reads the rows again and
// places every record at its tile's reserved range, ranked by an LDS
// atomic.  One device atomic per (workgroup, target tile) instead of one
// per synapse: on the 2M-neuron voxel block a voxel's rows share their
// target tiles, and per-synapse atomics on the same cursors serialised.
// The order within a tile follows the atomics; the pull's sums are exact
// integers, so it does not matter.  Instances of more than kBuildMaxTiles
// tiles are built in windows of that many target tiles (the rows are read
// again for each window).
constexpr uint32_t kBuildBlock = 1024, kBuildRows = 256, kBuildMaxTiles = 16384;
__global__ __launch_bounds__(kBuildBlock) void k_build_pull(const int64_t *__restrict__ rowptr,
                                                            const uint32_t *__restrict__ col,
                                                            const void *__restrict__ w, uint32_t wfmt, uint32_t n,
                                                            uint32_t ntiles, const int64_t *__restrict__ cptr,
                                                            uint32_t *__restrict__ cursor, uint32_t *__restrict__ rec)
{
    extern __shared__ uint32_t hist[];   // min(ntiles, kBuildMaxTiles) words
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t wv = tid / kWave;
    constexpr uint32_t kWaves = kBuildBlock / kWave;
    const uint32_t r0 = blockIdx.x * kBuildRows;
    const uint32_t r1 = min(n, r0 + kBuildRows);
    const uint32_t nw = grec_words(wfmt);
    for (uint32_t t0 = 0; t0 < ntiles; t0 += kBuildMaxTiles) {
        const uint32_t tw = min(kBuildMaxTiles, ntiles - t0);
        for (uint32_t t = tid; t < tw; t += kBuildBlock)
            hist[t] = 0u;
        __syncthreads();
        for (uint32_t r = r0 + wv; r < r1; r += kWaves) {
            const int64_t rb = rowptr[r];
            const uint32_t len = (uint32_t)(rowptr[r + 1] - rb);
            for (uint32_t j = lane; j < len; j += kWave) {
                const uint32_t t = ((col[rb + j] & kColMask) >> kTileBits) - t0;
                if (t < tw)
                    atomicAdd(&hist[t], 1u);
            }
        }
        __syncthreads();
        for (uint32_t t = tid; t < tw; t += kBuildBlock) {
            const uint32_t h = hist[t];
            if (h != 0u)
                hist[t] = atomicAdd(&cursor[t0 + t], h);
        }
        __syncthreads();
        for (uint32_t r = r0 + wv; r < r1; r += kWaves) {
            const int64_t rb = rowptr[r];
            const uint32_t len = (uint32_t)(rowptr[r + 1] - rb);
            for (uint32_t j = lane; j < len; j += kWave) {
                const uint32_t c = col[rb + j];
                const uint32_t t = ((c & kColMask) >> kTileBits) - t0;
                if (t >= tw)
                    continue;
                const uint32_t off = atomicAdd(&hist[t], 1u);
                const unsigned long long b = w_raw(w, (size_t)(rb + j), wfmt);
                uint32_t *e = rec + ((size_t)cptr[t0 + t] + off) * nw;
                e[0] = r | ((c & (kBlock - 1u)) << kPullRowBits) | (c & kKindBit);
                e[1] = (uint32_t)b;
                if (nw == 3u)
                    e[2] = (uint32_t)(b >> 32);
            }
        }
        __syncthreads();   // the window's last ranks are taken before the next window zeroes the counts
    }
}

// Writes the pending delta (the runs queued in queue set `set` for the next
// k_step) of every neuron of every tile without consuming the queues: used
// before sampling I_syn and before reading the state back.
__global__ __launch_bounds__(kBlock) void k_materialize(const TileDesc *__restrict__ tiles,
                                                        const InstDev *__restrict__ insts,
                                                        QueueSet qs,
                                                        uint32_t set)
{
    __shared__ unsigned long long acc[kBlock * 4];
    const TileDesc d = tiles[blockIdx.x];
    const InstDev &I = insts[d.inst];
    if (I.segidx != nullptr)
        return;   // gather propagation: k_gather has written dlt
    const uint32_t tid = threadIdx.x;
    clear_row(acc, tid);
    const uint32_t cnt_raw = qs.count[count_slot((size_t)set * qs.ntiles + blockIdx.x, qs.cshift)];
    const bool over = cnt_raw > d.qcap;
    const bool pull_on = I.crec != nullptr;
    const bool pull = pull_on && (over || gp(I.dstamp)[set] == gp(I.step)[set] + 1);
    const uint32_t cnt = (pull_on && over) ? 0u : min(cnt_raw, d.qcap);
    __syncthreads();
    accumulate_tile(acc, I.queue + (size_t)set * I.qslots + d.qoff, cnt, QSyn{}, 0u, tid);
    if (pull)
        pull_tile<kWeightAny>(acc, I, set, d.begin >> kTileBits, over, tid);
    __syncthreads();
    if (tid < d.count) {
        if (!pull_on && over)
            take_spill(acc, I, set, d.begin + tid, tid, false);
        st4(I.dlt, d.begin + tid, acc_to_delta(acc, tid));
    }
}

// Folds the pending delta of queue set `set` into J now (J = J * decay +
// delta, with the decay of the step that produced it), marks J as folded and
// empties the queues.  Runs before any change of parameters or of the
// instance set, so no pending spike is lost or folded with a later decay.
__global__ __launch_bounds__(kBlock) void k_settle(const TileDesc *__restrict__ tiles,
                                                   const InstDev *__restrict__ insts,
                                                   QueueSet qs,
                                                   uint32_t set)
{
    __shared__ unsigned long long acc[kBlock * 4];
    const TileDesc d = tiles[blockIdx.x];
    const InstDev &I = insts[d.inst];
    const uint32_t tid = threadIdx.x;
    if (I.segidx != nullptr) {
        // gather propagation: k_gather has written the pending delta into
        // dlt; after the fold the fired list of this set is empty
        if (tid < d.count) {
            const uint32_t li = d.begin + tid;
            const uint32_t fl = gp(I.flag)[li];
            if (!(fl & kFlagFolded)) {
                st4(I.J, li, fold_delta(ld4(I.J, li), ld4(I.dec, param_index(I, li)), ld4(I.dlt, li)));
                gp(I.flag)[li] = (uint8_t)(fl | kFlagFolded);
            }
        }
        if (d.begin == 0 && tid < kFiredLists)
            gp(I.fcount)[((size_t)set * kFiredLists + tid) * kFiredCountStride] = 0u;
        return;
    }
    uint32_t *cnt_p = qs.count + count_slot((size_t)set * qs.ntiles + blockIdx.x, qs.cshift);
    clear_row(acc, tid);
    const uint32_t cnt_raw = *cnt_p;
    const bool over = cnt_raw > d.qcap;
    const bool pull_on = I.crec != nullptr;
    const bool pull = pull_on && (over || gp(I.dstamp)[set] == gp(I.step)[set] + 1);
    const uint32_t cnt = (pull_on && over) ? 0u : min(cnt_raw, d.qcap);
    __syncthreads();
    accumulate_tile(acc, I.queue + (size_t)set * I.qslots + d.qoff, cnt, QSyn{}, 0u, tid);
    if (pull)
        pull_tile<kWeightAny>(acc, I, set, d.begin >> kTileBits, over, tid);
    __syncthreads();
    if (tid < d.count) {
        const uint32_t li = d.begin + tid;
        if (!pull_on && over)
            take_spill(acc, I, set, li, tid, true);
        const uint32_t fl = gp(I.flag)[li];
        if (!(fl & kFlagFolded)) {
            st4(I.J, li, fold_delta(ld4(I.J, li), ld4(I.dec, param_index(I, li)), acc_to_delta(acc, tid)));
            gp(I.flag)[li] = (uint8_t)(fl | kFlagFolded);
        }
    }
    if (tid == 0)
        *cnt_p = 0u;
}

// I_syn-based statistics (imean, receptor imeans) of the last step run: the
// k_step launch of the next step would record them, so after the last step
// they are taken here from the materialised pending delta.  `par` is the
// parity of the next launch (its step slot holds the next step index).
__global__ __launch_bounds__(kBlock) void k_stat_tail(const TileDesc *__restrict__ tiles,
                                                      const InstDev *__restrict__ insts,
                                                      uint32_t par)
{
    const TileDesc d = tiles[blockIdx.x];
    const InstDev &I = insts[d.inst];
    const uint32_t cf = current_fields(I.stat_mask);
    if (I.rec_stat == nullptr || cf == 0u)
        return;
    const int64_t row = gp(I.step)[par] - 1 - I.rec_step0;
    if (row < 0 || row >= (int64_t)I.rec_len)
        return;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t li = d.begin + tid;
    const bool active = tid < d.count;
    long long stv[kStatFields];
    uint32_t p = d.pop0;
    bool valid = false;
    if (active) {
        while (p + 1 < I.npop && li >= I.pops[p].end)
            ++p;
        const uint32_t c = param_index(I, li);
        const float V = gp(I.V)[li];
        float4 J = ld4(I.J, li);
        if (!(gp(I.flag)[li] & kFlagFolded))
            J = fold_delta(J, ld4(I.dec, c), ld4(I.dlt, li));
        const float4 g = ld4(I.gU, c), vu = ld4(I.vU, c);
        const float4 rc = receptor_currents(g, vu, J, V);
        stv[1] = stat_q(syn_current(g, vu, J, V));
        stv[2] = stat_q(rc.x);
        stv[3] = stat_q(rc.y);
        stv[4] = stat_q(rc.z);
        stv[5] = stat_q(rc.w);
        valid = !gp(I.blocked)[li];
    }
    wave_pop_stats(I.rec_stat + (size_t)row * I.npop * kStatFields, valid, p, stv, cf, lane);
}

// stat_result.cpp:393-433 (sampled spike flag, V, I_syn, i_ou) of the step
// just run by the k_step launch of parity `par`.  Runs after k_materialize
// (dlt holds that step's delta, folded in exactly as the next k_step will
// fold it).
__global__ __launch_bounds__(kBlock) void k_sample(const InstDev *__restrict__ insts,
                                                   const SampleDev *__restrict__ samples,
                                                   uint32_t nsamples,
                                                   const SampleRecDev *__restrict__ recs,
                                                   uint32_t par)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nsamples)
        return;
    const SampleDev s = samples[i];
    const InstDev &I = insts[s.inst];
    const SampleRecDev R = recs[s.inst];
    const int64_t row = gp(I.step)[par] - I.rec_step0;
    if (row < 0 || row >= (int64_t)I.rec_len)
        return;
    const uint32_t li = s.neuron;
    const size_t o = (size_t)row * R.nsample + s.slot;
    const float V = gp(I.V)[li];
    if (R.spike)
        R.spike[o] = (gp(I.flag)[li] & kFlagEmit) ? 1 : 0;
    if (R.v)
        R.v[o] = V;
    if (R.isyn) {
        const uint32_t c = param_index(I, li);
        const float4 J = fold_delta(ld4(I.J, li), ld4(I.dec, c), ld4(I.dlt, li));   // flag bit 2 is clear here
        R.isyn[o] = syn_current(ld4(I.gU, c), ld4(I.vU, c), J, V);
    }
    if (R.iou)
        R.iou[o] = gp(I.iou)[li];
}

// Verification entry: normals of the OU stream for a range of neurons at one
// substep (checks CPU/GPU bit identity of the generator).
__global__ __launch_bounds__(kBlock) void k_normals(uint64_t seed, uint32_t neuron0, uint64_t substep,
                                                    uint32_t count, float *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count)
        out[i] = ou_normal(seed, neuron0 + i, substep);
}

// ---- launchers ----------------------------------------------------------

// Calls f(kernel) with the kernel instantiated for pass size pass_k and
// weight format wfmt (a format of every launched instance, or kWeightAny).
template <template <uint32_t, uint32_t> class Sel, typename F>
hipError_t with_kernel(uint32_t pass_k, uint32_t wfmt, F f)
{
    auto by_fmt = [&](auto kk) -> hipError_t {
        constexpr uint32_t K = decltype(kk)::value;
        switch (wfmt) {
        case kWeightF32: f(Sel<K, kWeightF32>::get()); break;
        case kWeightF16: f(Sel<K, kWeightF16>::get()); break;
        case kWeightU8: f(Sel<K, kWeightU8>::get()); break;
        case kWeightAny: f(Sel<K, kWeightAny>::get()); break;
        default: return hipErrorInvalidValue;
        }
        return hipSuccess;
    };
    switch (pass_k) {
    case 1: return by_fmt(std::integral_constant<uint32_t, 1>{});
    case 2: return by_fmt(std::integral_constant<uint32_t, 2>{});
    case 4: return by_fmt(std::integral_constant<uint32_t, 4>{});
    default: return hipErrorInvalidValue;
    }
}

template <uint32_t K, uint32_t F>
struct StepSel {
    static auto get() { return k_step<K, F>; }
};
template <uint32_t K, uint32_t F>
struct BinSel {
    static auto get() { return k_bin<K, F>; }
};

hipError_t launch_step_kernel(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                              uint32_t par, StepCtl *ctl, uint32_t pass_k, uint32_t wfmt, hipStream_t stream)
{
    if (ntiles == 0)
        return hipSuccess;
    const hipError_t e = with_kernel<StepSel>(pass_k, wfmt, [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(ntiles), dim3(kBlock), 0, stream, tiles, insts, qs, par, ctl);
    });
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_step_kernel_timed(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                                    uint32_t par, StepCtl *ctl, uint32_t pass_k, uint32_t wfmt, hipEvent_t start,
                                    hipEvent_t stop, hipStream_t stream)
{
    if (ntiles == 0)
        return hipSuccess;
    const hipError_t e = with_kernel<StepSel>(pass_k, wfmt, [&](auto kern) {
        hipExtLaunchKernelGGL(kern, dim3(ntiles), dim3(kBlock), 0, stream, start, stop, 0, tiles, insts, qs, par, ctl);
    });
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_bin_kernel(const TileDesc *tiles, uint32_t grid, const InstDev *insts, const QueueSet &qs,
                             uint32_t par, StepCtl *ctl, uint32_t pass_k, uint32_t wfmt, hipEvent_t start,
                             hipEvent_t stop, hipStream_t stream)
{
    if (grid == 0 || qs.pass == nullptr)
        return hipSuccess;
    const hipError_t e = with_kernel<BinSel>(pass_k, wfmt, [&](auto kern) {
        if (start != nullptr)
            hipExtLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, start, stop, 0, tiles, insts, qs, par,
                                  ctl);
        else
            hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, tiles, insts, qs, par, ctl);
    });
    return e != hipSuccess ? e : hipGetLastError();
}

namespace {
template <typename F>
hipError_t with_gather(uint32_t wfmt, F f)
{
    switch (wfmt) {
    case kWeightF32: f(k_gather<kWeightF32>); break;
    case kWeightF16: f(k_gather<kWeightF16>); break;
    case kWeightU8: f(k_gather<kWeightU8>); break;
    case kWeightAny: f(k_gather<kWeightAny>); break;
    default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}
}  // namespace

hipError_t launch_gather(const GroupDesc *groups, uint32_t ngroups, const InstDev *insts, uint32_t par, uint32_t wfmt,
                         hipStream_t stream)
{
    if (ngroups == 0)
        return hipSuccess;
    const hipError_t e = with_gather(wfmt, [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(ngroups), dim3(kGatherBlock), 0, stream, groups, insts, par);
    });
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_gather_timed(const GroupDesc *groups, uint32_t ngroups, const InstDev *insts, uint32_t par,
                               uint32_t wfmt, hipEvent_t start, hipEvent_t stop, hipStream_t stream)
{
    if (ngroups == 0)
        return hipSuccess;
    const hipError_t e = with_gather(wfmt, [&](auto kern) {
        hipExtLaunchKernelGGL(kern, dim3(ngroups), dim3(kGatherBlock), 0, stream, start, stop, 0, groups, insts, par);
    });
    return e != hipSuccess ? e : hipGetLastError();
}

hipError_t launch_build_grec(const uint32_t *col, const void *w, uint32_t wfmt, uint64_t nsyn, uint32_t *rec,
                             hipStream_t stream)
{
    if (nsyn == 0)
        return hipSuccess;
    const uint64_t nb = std::min<uint64_t>((nsyn + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_build_grec, dim3((uint32_t)nb), dim3(kBlock), 0, stream, col, w, wfmt, nsyn, rec);
    return hipGetLastError();
}

hipError_t launch_build_segidx(const int64_t *rowptr, const uint32_t *col, uint32_t n, uint32_t ngroups,
                               uint16_t *segidx, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t rows_per_block = kBlock / kWave;
    const uint32_t nb = (n + rows_per_block - 1) / rows_per_block;
    hipLaunchKernelGGL(k_build_segidx, dim3(nb), dim3(kBlock), 0, stream, rowptr, col, n, ngroups, segidx);
    return hipGetLastError();
}

hipError_t launch_build_pull(const int64_t *rowptr, const uint32_t *col, const void *w, uint32_t wfmt, uint32_t n,
                             const int64_t *cptr, uint32_t *cursor, uint32_t *rec, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t ntiles = (n + kBlock - 1) / kBlock;
    const uint32_t nb = (n + kBuildRows - 1) / kBuildRows;
    hipLaunchKernelGGL(k_build_pull, dim3(nb), dim3(kBuildBlock), min(ntiles, kBuildMaxTiles) * sizeof(uint32_t),
                       stream, rowptr, col, w, wfmt, n, ntiles, cptr, cursor, rec);
    return hipGetLastError();
}

hipError_t launch_materialize(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                              uint32_t set, hipStream_t stream)
{
    if (ntiles == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_materialize, dim3(ntiles), dim3(kBlock), 0, stream, tiles, insts, qs, set);
    return hipGetLastError();
}

hipError_t launch_settle(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, const QueueSet &qs,
                         uint32_t set, hipStream_t stream)
{
    if (ntiles == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_settle, dim3(ntiles), dim3(kBlock), 0, stream, tiles, insts, qs, set);
    return hipGetLastError();
}

hipError_t launch_stat_tail(const TileDesc *tiles, uint32_t ntiles, const InstDev *insts, uint32_t par,
                            hipStream_t stream)
{
    if (ntiles == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_stat_tail, dim3(ntiles), dim3(kBlock), 0, stream, tiles, insts, par);
    return hipGetLastError();
}

hipError_t launch_sample(const InstDev *insts, const SampleDev *samples, uint32_t nsamples,
                         const SampleRecDev *recs, uint32_t par, hipStream_t stream)
{
    if (nsamples == 0)
        return hipSuccess;
    const uint32_t nb = (nsamples + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_sample, dim3(nb), dim3(kBlock), 0, stream, insts, samples, nsamples, recs, par);
    return hipGetLastError();
}

hipError_t launch_normals(uint64_t seed, uint32_t neuron0, uint64_t substep, uint32_t count, float *out,
                          hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const uint32_t nb = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_normals, dim3(nb), dim3(kBlock), 0, stream, seed, neuron0, substep, count, out);
    return hipGetLastError();
}

}  // namespace dtb
