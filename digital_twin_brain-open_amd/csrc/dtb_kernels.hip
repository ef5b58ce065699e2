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
// Two launches per step here:
//   k_neuron_step  one pass over the neurons that first finishes the previous
//                  step (J fold-in of the pending delta, I_syn), then runs
//                  OU + LIF for every substep and appends the step's fired
//                  rows to a compact list (wave ballot + lane prefix, one
//                  global reservation per wave);
//   k_propagate    walks only the fired rows of the presynaptic-major CSR and
//                  adds their weights into exact Q23.40 fixed-point
//                  accumulators with 64-bit integer atomics, so the result is
//                  independent of scheduling order.
// A sample kernel gathers per-step sample records when samples are set.
#include "dtb_kernels.hpp"
#include "dtb_rng.hpp"

namespace dtb {

__device__ __forceinline__ float q40_to_f32(long long q)
{
    return (float)((double)q * 0x1p-40);
}

__device__ __forceinline__ unsigned long long f32_to_q40(float w)
{
    return (unsigned long long)(long long)rintf(w * 0x1p40f);
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

__device__ __forceinline__ float4 fold_delta(const float4 J, const float4 dc, const long long *dq)
{
    const int4 lo = reinterpret_cast<const int4 *>(dq)[0];
    const int4 hi = reinterpret_cast<const int4 *>(dq)[1];
    const long long q0 = (long long)(((unsigned long long)(uint32_t)lo.y << 32) | (uint32_t)lo.x);
    const long long q1 = (long long)(((unsigned long long)(uint32_t)lo.w << 32) | (uint32_t)lo.z);
    const long long q2 = (long long)(((unsigned long long)(uint32_t)hi.y << 32) | (uint32_t)hi.x);
    const long long q3 = (long long)(((unsigned long long)(uint32_t)hi.w << 32) | (uint32_t)hi.z);
    float4 r;
    r.x = fmaf(J.x, dc.x, q40_to_f32(q0));
    r.y = fmaf(J.y, dc.y, q40_to_f32(q1));
    r.z = fmaf(J.z, dc.z, q40_to_f32(q2));
    r.w = fmaf(J.w, dc.w, q40_to_f32(q3));
    return r;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

__global__ __launch_bounds__(kBlock) void k_neuron_step(const WgDesc *__restrict__ wgs,
                                                        const InstDev *__restrict__ insts,
                                                        FiredDev *__restrict__ fired,
                                                        uint32_t fired_cap,
                                                        StepCtl *__restrict__ ctl)
{
    const WgDesc d = wgs[blockIdx.x];
    const InstDev &I = insts[d.inst];
    const uint32_t parity = ctl->tick & 1u;
    if (blockIdx.x == 0 && threadIdx.x == 0)
        ctl->fired_count[parity ^ 1u] = 0u;   // consumed by the previous launch pair

    const uint32_t li = d.begin + threadIdx.x;
    const bool active = threadIdx.x < d.count;
    const int64_t step = *I.step;
    bool emit = false;
    uint32_t p = d.pop0;

    if (active) {
        const float4 pa = I.pA[li];
        const float4 pb = I.pB[li];
        const float4 g = I.gU[li];
        const float4 vu = I.vU[li];
        const float4 dc = I.dec[li];
        long long *dq = I.dq + 4 * (size_t)li;
        const uint32_t fl = I.flag[li];
        float4 J = I.J[li];
        if (!(fl & kFlagFolded))
            J = fold_delta(J, dc, dq);
        reinterpret_cast<int4 *>(dq)[0] = make_int4(0, 0, 0, 0);
        reinterpret_cast<int4 *>(dq)[1] = make_int4(0, 0, 0, 0);

        float V = I.V[li];
        int32_t t = I.tcnt[li];
        uint32_t f = fl & kFlagSpike;
        float iou = I.iou[li];
        float isyn = syn_current(g, vu, J, V);

        while (p + 1 < I.npop && li >= I.pops[p].end)
            ++p;
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
        I.V[li] = V;
        I.tcnt[li] = t;
        I.flag[li] = (uint8_t)(f | (emit ? kFlagEmit : 0u));
        I.iou[li] = iou;
        I.J[li] = J;
    }

    // Fired-list compaction: one ballot per wave, lane prefix by popcount.
    const unsigned long long m = __ballot(emit);
    if (m != 0ull) {
        const uint32_t lane = __lane_id();
        const uint32_t cnt = (uint32_t)__popcll(m);
        const uint32_t pre = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        const int leader = __ffsll((long long)m) - 1;
        int64_t rb = 0;
        uint32_t len = 0;
        if (emit) {
            rb = I.rowptr[li];
            len = (uint32_t)(I.rowptr[li + 1] - rb);
        }
        const unsigned long long ev = wave_sum_u64((unsigned long long)len);
        uint32_t base = 0;
        if ((int)lane == leader) {
            base = atomicAdd(&ctl->fired_count[parity], cnt);
            atomicAdd(I.events, ev);
            atomicAdd(I.spikes, (unsigned long long)cnt);
        }
        base = __shfl(base, leader, kWave);
        if (emit) {
            FiredDev e;
            e.inst = d.inst;
            e.len = len;
            e.row_begin = rb;
            fired[(size_t)parity * fired_cap + base + pre] = e;
            if (I.rec_counts != nullptr && !I.blocked[li]) {
                const int64_t row = step - I.rec_step0;
                if (row >= 0 && row < (int64_t)I.rec_len)
                    atomicAdd(&I.rec_counts[(size_t)row * I.npop + p], 1u);
            }
        }
    }
}

__device__ __forceinline__ void scatter_one(unsigned long long *__restrict__ dq, uint32_t c, float2 w)
{
    const uint32_t post = c & kColMask;
    const uint32_t ch = (c >> 31) << 1;
    unsigned long long *a = dq + 4 * (size_t)post + ch;
    atomicAdd(a, f32_to_q40(w.x));
    atomicAdd(a + 1, f32_to_q40(w.y));
}

// Wave-per-row over the fired list (grid-stride), 4 synapses per lane in
// flight.  The last workgroup to finish advances the launch tick and every
// instance's step counter.
__global__ __launch_bounds__(kBlock) void k_propagate(const InstDev *__restrict__ insts,
                                                      uint32_t ninst,
                                                      const FiredDev *__restrict__ fired,
                                                      uint32_t fired_cap,
                                                      StepCtl *__restrict__ ctl)
{
    const uint32_t parity = ctl->tick & 1u;
    const uint32_t F = ctl->fired_count[parity];
    const FiredDev *list = fired + (size_t)parity * fired_cap;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const uint32_t nwaves = gridDim.x * (kBlock / kWave);

    for (uint32_t r = wave; r < F; r += nwaves) {
        const FiredDev e = list[r];
        if (e.inst >= ninst)
            continue;   // never written by the neuron step; kept as a bound
        const InstDev &I = insts[e.inst];
        const uint32_t *__restrict__ col = I.col + e.row_begin;
        const float2 *__restrict__ w = I.w + e.row_begin;
        unsigned long long *__restrict__ dq = reinterpret_cast<unsigned long long *>(I.dq);
        const uint32_t len = e.len;
        uint32_t j = lane;
        for (; j + 3 * kWave < len; j += 4 * kWave) {
            const uint32_t c0 = col[j], c1 = col[j + kWave], c2 = col[j + 2 * kWave], c3 = col[j + 3 * kWave];
            const float2 w0 = w[j], w1 = w[j + kWave], w2 = w[j + 2 * kWave], w3 = w[j + 3 * kWave];
            scatter_one(dq, c0, w0);
            scatter_one(dq, c1, w1);
            scatter_one(dq, c2, w2);
            scatter_one(dq, c3, w3);
        }
        for (; j < len; j += kWave)
            scatter_one(dq, col[j], w[j]);
    }

    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t prev = atomicAdd(&ctl->done_blocks, 1u);
        if (prev == gridDim.x - 1) {
            ctl->done_blocks = 0u;
            ctl->tick = ctl->tick + 1u;
            for (uint32_t i = 0; i < ninst; ++i)
                *insts[i].step = *insts[i].step + 1;
            __threadfence();
        }
    }
}

// stat_result.cpp:393-433 (sampled spike flag, V, I_syn, i_ou).  Runs after
// k_propagate, so the instance step counter already counts this step and the
// pending delta of this step is folded in exactly as the next neuron step
// will fold it.
__global__ __launch_bounds__(kBlock) void k_sample(const InstDev *__restrict__ insts,
                                                   const SampleDev *__restrict__ samples,
                                                   uint32_t nsamples,
                                                   const SampleRecDev *__restrict__ recs)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nsamples)
        return;
    const SampleDev s = samples[i];
    const InstDev &I = insts[s.inst];
    const SampleRecDev R = recs[s.inst];
    const int64_t row = *I.step - 1 - I.rec_step0;
    if (row < 0 || row >= (int64_t)I.rec_len)
        return;
    const uint32_t li = s.neuron;
    const size_t o = (size_t)row * R.nsample + s.slot;
    const float V = I.V[li];
    if (R.spike)
        R.spike[o] = (I.flag[li] & kFlagEmit) ? 1 : 0;
    if (R.v)
        R.v[o] = V;
    if (R.isyn) {
        const float4 J = fold_delta(I.J[li], I.dec[li], I.dq + 4 * (size_t)li);   // flag bit 2 is clear here
        R.isyn[o] = syn_current(I.gU[li], I.vU[li], J, V);
    }
    if (R.iou)
        R.iou[o] = I.iou[li];
}

// Debug / verification entry: normals of the OU stream for a range of
// neurons at one substep (used to check CPU/GPU bit identity).
__global__ __launch_bounds__(kBlock) void k_normals(uint64_t seed, uint32_t neuron0, uint64_t substep,
                                                    uint32_t count, float *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count)
        out[i] = ou_normal(seed, neuron0 + i, substep);
}

// ---- launchers ----------------------------------------------------------

hipError_t launch_neuron_step(const WgDesc *wgs, uint32_t nwg, const InstDev *insts, FiredDev *fired,
                              uint32_t fired_cap, StepCtl *ctl, hipStream_t stream)
{
    if (nwg == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_neuron_step, dim3(nwg), dim3(kBlock), 0, stream, wgs, insts, fired, fired_cap, ctl);
    return hipGetLastError();
}

hipError_t launch_propagate(const InstDev *insts, uint32_t ninst, const FiredDev *fired, uint32_t fired_cap,
                            StepCtl *ctl, uint32_t nblocks, hipStream_t stream)
{
    hipLaunchKernelGGL(k_propagate, dim3(nblocks), dim3(kBlock), 0, stream, insts, ninst, fired, fired_cap, ctl);
    return hipGetLastError();
}

hipError_t launch_sample(const InstDev *insts, const SampleDev *samples, uint32_t nsamples,
                         const SampleRecDev *recs, hipStream_t stream)
{
    if (nsamples == 0)
        return hipSuccess;
    const uint32_t nb = (nsamples + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_sample, dim3(nb), dim3(kBlock), 0, stream, insts, samples, nsamples, recs);
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
