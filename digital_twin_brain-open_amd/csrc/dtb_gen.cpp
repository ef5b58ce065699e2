// dtb_gen.cpp -- seeded synthetic networks in the reference layout.
//
// The reference builds blocks offline with python/make_block.py (property
// columns make_block.py:441-465, random in-edges make_block.py:280-330, npz
// members make_block.py:608-614) and loads them as a presynaptic-major CSR
// (src/configuration.cpp:1079-1392).  This generator produces the same
// layout directly from counter-based random numbers, so every synapse is a
// pure function of (seed, postsynaptic neuron, in-edge slot): the network is
// identical for any thread count, and within each row synapses are ordered
// by (postsynaptic id, slot).
#include "dtb_gen.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "dtb_engine.hpp"
#include "dtb_rng.hpp"

namespace dtb {

namespace {

inline uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

inline float weight24(uint32_t x)
{
    return (float)(x >> 8) * 0x1p-24f;   // k 2^-24, exact in Q23.40
}

inline Philox4 gen_bits(uint64_t seed, uint32_t a, uint32_t b, uint32_t c)
{
    return philox4x32_10(a, b, c, kTagGen, (uint32_t)seed, (uint32_t)(seed >> 32));
}

int thread_count(int64_t work)
{
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0)
        hw = 4;
    hw = std::min(hw, 16u);
    const int64_t by_work = std::max<int64_t>(1, work / (1 << 16));
    return (int)std::min<int64_t>(hw, by_work);
}

template <typename F>
void parallel_ranges(int64_t n, int nt, F f)
{
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const int64_t a = n * t / nt, b = n * (t + 1) / nt;
        th.emplace_back([=]() { f(t, a, b); });
    }
    for (auto &x : th)
        x.join();
}

// Edge source: fills (pre, w0, w1) of in-edge k of postsynaptic neuron i.
struct EdgeFn {
    virtual ~EdgeFn() = default;
    virtual uint32_t pre(uint32_t i, uint32_t k, float *w0, float *w1) const = 0;
    virtual bool inhibitory(uint32_t pre) const = 0;
};

// Counting sort of in-edges into a presynaptic-major CSR, deterministic for
// any thread count: thread t owns postsynaptic range [a_t, b_t) and writes
// its synapses of row r after those of threads < t.
void build_csr(int64_t n, uint32_t K, const EdgeFn &fn, int64_t *rowptr, uint32_t *col, float *w,
               uint32_t *kind_bits)
{
    const int nt = thread_count(n * (int64_t)K);
    std::vector<std::vector<uint32_t>> cnt(nt, std::vector<uint32_t>((size_t)n, 0u));
    parallel_ranges(n, nt, [&](int t, int64_t a, int64_t b) {
        uint32_t *c = cnt[t].data();
        float w0, w1;
        for (int64_t i = a; i < b; ++i)
            for (uint32_t k = 0; k < K; ++k)
                ++c[fn.pre((uint32_t)i, k, &w0, &w1)];
    });
    rowptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
        int64_t s = 0;
        for (int t = 0; t < nt; ++t)
            s += cnt[t][(size_t)r];
        rowptr[r + 1] = rowptr[r] + s;
    }
    // Convert counts into per-thread cursors.
    parallel_ranges(n, nt, [&](int, int64_t a, int64_t b) {
        for (int64_t r = a; r < b; ++r) {
            int64_t off = rowptr[r];
            for (int t = 0; t < nt; ++t) {
                const uint32_t c = cnt[t][(size_t)r];
                cnt[t][(size_t)r] = (uint32_t)(off - rowptr[r]);
                off += c;
            }
        }
    });
    const int64_t nsyn = rowptr[n];
    parallel_ranges(n, nt, [&](int t, int64_t a, int64_t b) {
        uint32_t *c = cnt[t].data();
        float w0, w1;
        for (int64_t i = a; i < b; ++i)
            for (uint32_t k = 0; k < K; ++k) {
                const uint32_t p = fn.pre((uint32_t)i, k, &w0, &w1);
                const int64_t j = rowptr[p] + c[p]++;
                col[j] = (uint32_t)i;
                w[2 * j] = w0;
                w[2 * j + 1] = w1;
            }
    });
    // Kind bits, word by word (no two threads share a word).  The channel of
    // a synapse is fixed by its presynaptic row.
    const int64_t words = (nsyn + 31) / 32;
    parallel_ranges(words, thread_count(nsyn), [&](int, int64_t a, int64_t b) {
        if (a >= b)
            return;
        int64_t r = std::upper_bound(rowptr, rowptr + n + 1, a * 32) - rowptr - 1;
        for (int64_t wd = a; wd < b; ++wd) {
            uint32_t bits = 0;
            const int64_t j1 = std::min<int64_t>(nsyn, wd * 32 + 32);
            for (int64_t j = wd * 32; j < j1; ++j) {
                while (rowptr[r + 1] <= j)
                    ++r;
                if (fn.inhibitory((uint32_t)r))
                    bits |= 1u << (j & 31);
            }
            kind_bits[wd] = bits;
        }
    });
}

struct UniformEdges : EdgeFn {
    uint64_t seed;
    uint32_t n, n_exc;
    uint32_t pre(uint32_t i, uint32_t k, float *w0, float *w1) const override
    {
        const Philox4 x = gen_bits(seed, i, k, 0u);
        uint32_t p = mulhi32(x.x, n - 1);
        p += (p >= i) ? 1u : 0u;   // no self-connection
        *w0 = weight24(x.y);
        *w1 = weight24(x.z);
        return p;
    }
    bool inhibitory(uint32_t p) const override { return p >= n_exc; }
};

struct VoxelEdges : EdgeFn {
    uint64_t seed;
    uint32_t K, K_intra;
    std::vector<uint32_t> vox_begin, vox_exc_end;   // per voxel
    std::vector<uint32_t> vox_of;                   // neuron -> voxel
    std::vector<uint32_t> cdf_off;                  // [V+1] offsets into targets
    std::vector<uint32_t> cdf_u;                    // source voxel
    std::vector<uint32_t> cdf_p;                    // cumulative probability, 2^32 scale
    std::vector<uint8_t> inh;                       // neuron -> inhibitory
    uint32_t pre(uint32_t i, uint32_t k, float *w0, float *w1) const override
    {
        const Philox4 x = gen_bits(seed, i, k, 0u);
        const uint32_t v = vox_of[i];
        *w0 = weight24(x.y);
        *w1 = weight24(x.z);
        if (k < K_intra) {
            const uint32_t b = vox_begin[v], m = vox_begin[v + 1] - b;
            uint32_t p = b + mulhi32(x.x, m - 1);
            p += (p >= i) ? 1u : 0u;
            return p;
        }
        uint32_t lo = cdf_off[v], hi = cdf_off[v + 1] - 1;
        while (lo < hi) {   // first entry with cdf > x.w
            const uint32_t mid = (lo + hi) / 2;
            if (cdf_p[mid] > x.w)
                hi = mid;
            else
                lo = mid + 1;
        }
        const uint32_t u = cdf_u[lo];
        const uint32_t b = vox_begin[u], m = vox_begin[u + 1] - b;
        return b + mulhi32(x.x, m);
    }
    bool inhibitory(uint32_t p) const override { return inh[p] != 0; }
};

void fill_props(float *props, uint32_t i, uint32_t sub, const float *g_ui)
{
    // make_block.py:414-465 column defaults; refractory 2 ms.
    float *p = props + (size_t)i * kNumCols;
    std::memset(p, 0, sizeof(float) * kNumCols);
    p[kColNoiseRate] = 0.0f;
    p[kColBlocked] = 0.0f;
    p[kColIExt] = 0.0f;
    p[kColSubIdx] = (float)sub;
    p[kColC] = 1.0f;
    p[kColTRef] = 2.0f;
    p[kColGLeak] = 0.03f;
    p[kColVLeak] = -75.0f;
    p[kColVTh] = -50.0f;
    p[kColVReset] = -65.0f;
    const float vu[4] = {0.0f, 0.0f, -70.0f, -100.0f};
    const float tau[4] = {2.0f, 40.0f, 10.0f, 50.0f};
    for (int c = 0; c < 4; ++c) {
        p[kColGU + c] = g_ui[c];
        p[kColVU + c] = vu[c];
        p[kColTauU + c] = tau[c];
    }
}

}  // namespace

void gen_uniform(int64_t n, uint32_t K, uint64_t seed, const float g_ui[4], uint32_t sub_base, float *props,
                 int64_t *rowptr, uint32_t *col, float *w, uint32_t *kind_bits)
{
    if (n < 2 || n > (int64_t)kColMask || K == 0)
        throw Error(kInvalidParameter, "gen_uniform: need n >= 2 and K >= 1");
    UniformEdges fn;
    fn.seed = seed;
    fn.n = (uint32_t)n;
    fn.n_exc = (uint32_t)(n * 8 / 10);
    for (uint32_t i = 0; i < (uint32_t)n; ++i)
        fill_props(props, i, sub_base + (i >= fn.n_exc ? 1u : 0u), g_ui);
    build_csr(n, K, fn, rowptr, col, w, kind_bits);
}

void gen_default_props(int64_t n, const float g_ui[4], uint32_t sub_base, float *props)
{
    // two rows (excitatory, inhibitory population) copied into place
    const uint32_t n_exc = (uint32_t)(n * 8 / 10);
    float rows[2 * kNumCols];
    fill_props(rows, 0, sub_base, g_ui);
    fill_props(rows, 1, sub_base + 1u, g_ui);
    parallel_ranges(n, thread_count(n * 4), [&](int, int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i)
            std::memcpy(props + (size_t)i * kNumCols, rows + (i >= n_exc ? kNumCols : 0), sizeof(float) * kNumCols);
    });
}

void gen_fanout(int64_t n, uint32_t K, uint64_t seed, const float g_ui[4], uint32_t sub_base, float *props,
                int64_t *rowptr, uint32_t *col, float *w, uint32_t *kind_bits)
{
    if (n < 2 || n > (int64_t)kColMask || K == 0 || K > 8192)
        throw Error(kInvalidParameter, "gen_fanout: need n >= 2 and 1 <= K <= 8192");
    gen_default_props(n, g_ui, sub_base, props);
    const uint32_t nn = (uint32_t)n, n_exc = (uint32_t)(n * 8 / 10);
    for (int64_t r = 0; r <= n; ++r)
        rowptr[r] = r * (int64_t)K;
    parallel_ranges(n, thread_count(n * (int64_t)K), [&](int, int64_t a, int64_t b) {
        std::vector<uint64_t> keys(K);
        for (int64_t r = a; r < b; ++r) {
            for (uint32_t k = 0; k < K; ++k) {
                const Philox4 x = philox4x32_10((uint32_t)r, k, 1u, kTagGen, (uint32_t)seed, (uint32_t)(seed >> 32));
                uint32_t t = mulhi32(x.x, nn - 1);
                t += (t >= (uint32_t)r) ? 1u : 0u;
                keys[k] = ((uint64_t)t << 16) | k;
            }
            std::sort(keys.begin(), keys.end());
            for (uint32_t i = 0; i < K; ++i) {
                const uint32_t t = (uint32_t)(keys[i] >> 16), k = (uint32_t)(keys[i] & 0xffffu);
                const Philox4 x = philox4x32_10((uint32_t)r, k, 1u, kTagGen, (uint32_t)seed, (uint32_t)(seed >> 32));
                const int64_t j = r * (int64_t)K + i;
                col[j] = t;
                w[2 * j] = weight24(x.y);
                w[2 * j + 1] = weight24(x.z);
            }
        }
    });
    const int64_t nsyn = n * (int64_t)K, words = (nsyn + 31) / 32;
    for (int64_t wd = 0; wd < words; ++wd) {
        uint32_t bits = 0;
        for (int64_t j = wd * 32; j < std::min<int64_t>(nsyn, wd * 32 + 32); ++j)
            if (j / K >= n_exc)
                bits |= 1u << (j & 31);
        kind_bits[wd] = bits;
    }
}

void gen_voxel(int64_t n, uint32_t nvox, uint32_t K, float intra_frac, float density, uint64_t seed,
               const float g_ui[4], float *props, int64_t *rowptr, uint32_t *col, float *w, uint32_t *kind_bits,
               uint32_t *vox_sizes)
{
    if (nvox < 2 || n < (int64_t)nvox * 8 || n > (int64_t)kColMask || K == 0)
        throw Error(kInvalidParameter, "gen_voxel: need nvox >= 2, n >= 8 nvox, K >= 1");
    VoxelEdges fn;
    fn.seed = seed;
    fn.K = K;
    fn.K_intra = (uint32_t)std::lround((double)K * (double)intra_frac);
    // Voxel sizes proportional to lognormal(0, 0.5).
    std::vector<double> s(nvox);
    double tot = 0.0;
    for (uint32_t v = 0; v < nvox; ++v) {
        const Philox4 x = gen_bits(seed, v, 0u, 1u);
        s[v] = std::exp(0.5 * (double)normal_from_bits(x.x, x.y));
        tot += s[v];
    }
    std::vector<uint32_t> sz(nvox);
    int64_t used = 0;
    for (uint32_t v = 0; v < nvox; ++v) {
        sz[v] = std::max<uint32_t>(8u, (uint32_t)std::floor((double)(n - 8 * (int64_t)nvox) * s[v] / tot) + 8u);
        used += sz[v];
    }
    for (uint32_t v = 0; used < n; v = (v + 1) % nvox, ++used)
        ++sz[v];
    for (uint32_t v = 0; used > n; v = (v + 1) % nvox)
        if (sz[v] > 8u) {
            --sz[v];
            --used;
        }
    fn.vox_begin.assign(nvox + 1, 0u);
    fn.vox_exc_end.assign(nvox, 0u);
    for (uint32_t v = 0; v < nvox; ++v) {
        fn.vox_begin[v + 1] = fn.vox_begin[v] + sz[v];
        fn.vox_exc_end[v] = fn.vox_begin[v] + sz[v] * 8 / 10;
        if (vox_sizes)
            vox_sizes[v] = sz[v];
    }
    fn.vox_of.resize((size_t)n);
    fn.inh.resize((size_t)n);
    for (uint32_t v = 0; v < nvox; ++v)
        for (uint32_t i = fn.vox_begin[v]; i < fn.vox_begin[v + 1]; ++i) {
            fn.vox_of[i] = v;
            fn.inh[i] = i >= fn.vox_exc_end[v];
            fill_props(props, i, 2 * v + (fn.inh[i] ? 1u : 0u), g_ui);
        }
    // Sparse voxel-to-voxel probabilities: each off-diagonal entry present
    // with probability `density`, value lognormal(0, 1), rows normalised.
    fn.cdf_off.assign(nvox + 1, 0u);
    for (uint32_t v = 0; v < nvox; ++v) {
        std::vector<std::pair<uint32_t, double>> row;
        for (uint32_t u = 0; u < nvox; ++u) {
            if (u == v)
                continue;
            const Philox4 x = gen_bits(seed, v, u, 2u);
            if ((double)x.x * 0x1p-32 < (double)density)
                row.emplace_back(u, std::exp((double)normal_from_bits(x.y, x.z)));
        }
        if (row.empty()) {
            const Philox4 x = gen_bits(seed, v, nvox, 2u);
            const uint32_t u = (v + 1 + x.x % (nvox - 1)) % nvox;
            row.emplace_back(u, 1.0);
        }
        double rs = 0.0;
        for (auto &e : row)
            rs += e.second;
        double acc = 0.0;
        for (size_t k = 0; k < row.size(); ++k) {
            acc += row[k].second / rs;
            const double c = (k + 1 == row.size()) ? 4294967296.0 : std::min(4294967295.0, acc * 4294967296.0);
            fn.cdf_u.push_back(row[k].first);
            fn.cdf_p.push_back(k + 1 == row.size() ? 0xffffffffu : (uint32_t)c);
        }
        fn.cdf_off[v + 1] = (uint32_t)fn.cdf_u.size();
    }
    build_csr(n, K, fn, rowptr, col, w, kind_bits);
}

float gamma_draw(uint64_t seed, uint32_t call, uint32_t neuron, float alpha, float beta)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (alpha == 0.0f)
        return 0.0f;
    float scale = 1.0f;
    if (alpha < 1.0f) {
        // boost: Gamma(alpha) = Gamma(alpha + 1) * U^(1/alpha)
        const Philox4 b = philox4x32_10(neuron, call, 0xffffffffu, kTagGamma, k0, k1);
        const float u = 1.0f - uniform_from_bits(b.x);   // (0, 1]
        scale = std::pow(u, 1.0f / alpha);
        alpha += 1.0f;
    }
    const float d = alpha - 1.0f / 3.0f;
    const float c = 1.0f / std::sqrt(9.0f * d);
    for (uint32_t attempt = 0;; ++attempt) {
        const Philox4 x4 = philox4x32_10(neuron, call, attempt, kTagGamma, k0, k1);
        const float x = normal_from_bits(x4.x, x4.y);
        const float y = 1.0f + c * x;
        if (y <= 0.0f)
            continue;
        const float v = y * y * y;
        const float u = 1.0f - uniform_from_bits(x4.z);   // (0, 1]
        const float xx = x * x;
        if (u < 1.0f - 0.0331f * xx * xx || std::log(u) < 0.5f * xx + d * (1.0f - v + std::log(v)))
            return scale * d * v / beta;
    }
}

void gen_gamma(uint64_t seed, uint32_t call, uint32_t first, uint32_t count, float alpha, float beta, float *out)
{
    if (!(alpha >= 0.0f) || !(beta > 0.0f) || !std::isfinite(alpha) || !std::isfinite(beta))
        throw Error(kInvalidParameter, "gamma needs alpha >= 0 and beta > 0");
    for (uint32_t i = 0; i < count; ++i)
        out[i] = gamma_draw(seed, call, first + i, alpha, beta);
}

}  // namespace dtb
