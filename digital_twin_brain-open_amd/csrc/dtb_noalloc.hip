// dtb_noalloc.hip -- device side of dtb_run_noalloc: the static array that
// holds one call's per-step records, and the small kernels that clear it and
// turn its statistic sums into means without host buffers.
#include <mutex>

#include "dtb_engine.hpp"

namespace dtb {

namespace {

// One copy per GPU, at its own address on each; host code passes the
// address from hipGetSymbolAddress (scratch_address), never the name.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {};

}  // namespace

void *scratch_address(int device)
{
    if (device < 0 || device >= kMaxDevices)
        throw Error(kNoDevice, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_addr[device] == nullptr) {
        int cur = 0;
        DTB_HIP(hipGetDevice(&cur));
        if (cur != device)
            throw Error(kInternal, "scratch_address: the device is not current");
        void *p = nullptr;
        DTB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_addr[device] = p;
    }
    return g_addr[device];
}

__global__ __launch_bounds__(256) void k_zero_words(uint32_t *p, size_t words)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (size_t)gridDim.x * 256)
        p[i] = 0u;
}

hipError_t launch_zero_words(void *p, size_t bytes, hipStream_t stream)
{
    const size_t words = bytes / 4;
    if (words == 0)
        return hipSuccess;
    const unsigned grid = (unsigned)std::min<size_t>((words + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_words, dim3(grid), dim3(256), 0, stream, static_cast<uint32_t *>(p), words);
    return hipGetLastError();
}

// stat_result.cpp:266-391 divides a population's sum by its neurons minus
// the blocked ones: one workgroup per population counts them.
__global__ __launch_bounds__(256) void k_pop_unblocked(const uint8_t *__restrict__ blocked,
                                                       const PopDev *__restrict__ pops,
                                                       unsigned long long *__restrict__ cnt)
{
    __shared__ unsigned long long part[256 / 64];
    const PopDev p = pops[blockIdx.x];
    unsigned long long c = 0;
    for (uint32_t i = p.begin + threadIdx.x; i < p.end; i += 256)
        c += blocked[i] ? 0ull : 1ull;
    for (int off = 32; off > 0; off >>= 1)
        c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63u) == 0)
        part[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        cnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

struct MeanOut {
    float *f[kStatFields];
};

// mean = (float)((double)sum * 2^-24 / count), as the host conversion of
// dtb_run computes it (IEEE double multiply and divide, one rounding to
// float).
__global__ __launch_bounds__(256) void k_stat_means(const long long *__restrict__ sums,
                                                    const unsigned long long *__restrict__ cnt, uint32_t npop,
                                                    size_t cells, MeanOut out)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < cells; i += (size_t)gridDim.x * 256) {
        const double c = (double)cnt[i % npop];
#pragma unroll
        for (int k = 0; k < kStatFields; ++k)
            if (out.f[k] != nullptr)
                out.f[k][i] = (float)((double)sums[i * kStatFields + k] * 0x1p-24 / c);
    }
}

hipError_t launch_stat_means(const uint8_t *blocked, const PopDev *pops, uint32_t npop, const long long *sums,
                             uint32_t rows, unsigned long long *cnt, float *const *out, hipStream_t stream)
{
    if (npop == 0 || rows == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_pop_unblocked, dim3(npop), dim3(256), 0, stream, blocked, pops, cnt);
    MeanOut m;
    for (int k = 0; k < kStatFields; ++k)
        m.f[k] = out[k];
    const size_t cells = (size_t)rows * npop;
    const unsigned grid = (unsigned)std::min<size_t>((cells + 255) / 256, 4096);
    hipLaunchKernelGGL(k_stat_means, dim3(grid), dim3(256), 0, stream, sums, cnt, npop, cells, m);
    return hipGetLastError();
}

}  // namespace dtb
