// dtb_engine.cpp -- host runtime of the DTB step (see dtb_engine.hpp).
//
// Host arithmetic that produces kernel inputs (1/C, refractory steps,
// exp(-dt/tau), OU factors, V0) is done here once per run configuration, in
// the order the CPU oracle uses, with contraction disabled, so that the GPU
// and the oracle start from bit-identical parameters.
#include "dtb_engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "dtb_devgen.hpp"
#include "dtb_gen.hpp"
#include "dtb_kernels.hpp"

namespace dtb {

void hip_check(hipError_t e, const char *what, const char *file, int line)
{
    if (e == hipSuccess)
        return;
    std::string msg = std::string(hipGetErrorString(e)) + " in " + what + " (" + file + ":" + std::to_string(line) + ")";
    if (e == hipErrorOutOfMemory)
        throw Error(kOutOfMemory, msg);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
        throw Error(kNoDevice, msg);
    throw Error(kHipError, msg);
}

int prop_type_column(int prop)
{
    // dist_blockwrapper.py:32-33 maps columns to PropType with
    // [19, -1, 0, -1, 1, 2, ..., 18]; this is the inverse.
    static const int kCol[kNumPropTypes] = {kColIExt,  kColC,      kColTRef,   kColGLeak, kColVLeak,
                                            kColVTh,   kColVReset, kColGU + 0, kColGU + 1, kColGU + 2,
                                            kColGU + 3, kColVU + 0, kColVU + 1, kColVU + 2, kColVU + 3,
                                            kColTauU + 0, kColTauU + 1, kColTauU + 2, kColTauU + 3, kColNoiseRate};
    if (prop < 0 || prop >= kNumPropTypes)
        return -1;
    return kCol[prop];
}

size_t data_size(DataType t)
{
    switch (t) {
    case DOUBLE:
        return 8;
    case FLOAT:
    case INT32:
        return 4;
    case FLOAT16:
        return 2;
    case INT8:
        return 1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Host-side parameter derivation.  Mirrors the oracle exactly.

static inline float host_expf(float x)
{
    return (float)std::exp((double)x);
}

static inline int32_t refractory_substeps(float t_ref, float dt, int32_t t_steps)
{
    // property.cpp:45-53 (t_ref / dt at init, block.cpp:150) and
    // membrane_voltage.cpp:177 ((int)t_ref * t_steps).
    const float steps = t_ref / dt;
    return (int32_t)steps * t_steps;
}

static inline void ou_factors(float dt_sub, float tau, float sigma, float *f1, float *f2s)
{
    // extern_stimuli.cpp:87-91
    const float factor = (-dt_sub) / tau;
    const float a = 1.0f - host_expf(factor);
    const float b = std::sqrt(1.0f - host_expf(2.0f * factor));
    *f1 = a;
    *f2s = b * sigma;
}

static inline float bits_as_float(int32_t v)
{
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// ---------------------------------------------------------------------------

Instance::Instance(const NetworkHost &net, float dt, uint64_t seed, hipStream_t s)
{
    if (net.n <= 0 || net.n > (int64_t)kColMask)
        throw Error(kInvalidNetwork, "neuron count out of range");
    if (net.props == nullptr || net.rowptr == nullptr)
        throw Error(kInvalidParameter, "props and rowptr are required");
    if (net.nsyn < 0 || (net.nsyn > 0 && (net.col == nullptr || net.w == nullptr || net.kind_bits == nullptr)))
        throw Error(kInvalidParameter, "synapse arrays are required");
    if (net.wfmt != kWeightF32 && net.wfmt != kWeightF16 && net.wfmt != kWeightU8)
        throw Error(kInvalidParameter, "unknown weight format");
    n_ = (uint32_t)net.n;
    nsyn_ = net.nsyn;
    wfmt_ = net.wfmt;

    // CSR validation (rows ascending, columns in range).
    if (net.rowptr[0] != 0 || net.rowptr[net.n] != net.nsyn)
        throw Error(kInvalidNetwork, "rowptr must start at 0 and end at nsyn");
    if (net.nsyn >= kMaxSynapses)
        throw Error(kInvalidNetwork, "too many synapses for one block");
    for (int64_t i = 0; i < net.n; ++i)
        if (net.rowptr[i + 1] < net.rowptr[i] || net.rowptr[i + 1] - net.rowptr[i] >= kMaxRowLength)
            throw Error(kInvalidNetwork, "rowptr must be non-decreasing with rows below 2^31 synapses");
    for (int64_t j = 0; j < net.nsyn; ++j)
        if (net.col[j] >= n_)
            throw Error(kInvalidNetwork, "column index out of range");

    init_common(net.props, dt, seed, s);

    rowptr_.upload(net.rowptr, (size_t)n_ + 1, s);
    // Pack the channel bit into bit 31 of the column index so that one
    // 4-byte load per synapse carries both.
    const int64_t chunk = (int64_t)1 << 24;
    std::vector<uint32_t> stage((size_t)std::min<int64_t>(std::max<int64_t>(nsyn_, 1), chunk));
    for (int64_t base = 0; base < nsyn_; base += chunk) {
        const int64_t m = std::min<int64_t>(chunk, nsyn_ - base);
        for (int64_t j = 0; j < m; ++j) {
            const int64_t g = base + j;
            const uint32_t kind = (net.kind_bits[g >> 5] >> (g & 31)) & 1u;
            stage[(size_t)j] = net.col[g] | (kind ? kKindBit : 0u);
        }
        DTB_HIP(hipMemcpyAsync(col_.mutable_gpu_data() + base, stage.data(), (size_t)m * sizeof(uint32_t),
                               hipMemcpyHostToDevice, s));
        DTB_HIP(hipStreamSynchronize(s));
    }
    if (nsyn_ > 0)
        w_.upload(static_cast<const uint8_t *>(net.w), (size_t)nsyn_ * weight_bytes(wfmt_), s);
    count_tile_synapses(net);
    DTB_HIP(hipStreamSynchronize(s));
}

void Instance::init_common(const float *props, float dt, uint64_t seed, hipStream_t s)
{
    init_common(std::vector<float>(props, props + (size_t)n_ * kNumCols), dt, seed, s);
}

void Instance::init_common(std::vector<float> &&props, float dt, uint64_t seed, hipStream_t s)
{
    if (!(dt > 0.0f))
        throw Error(kInvalidParameter, "dt must be positive");
    dt_ = dt;
    seed_ = seed;
    props_ = std::move(props);

    // Populations: contiguous runs of the sub-block column
    // (configuration.cpp:1902-1930 requires each sub-block to be contiguous).
    std::unordered_set<uint32_t> seen;
    for (uint32_t i = 0; i < n_; ++i) {
        const float v = props_[(size_t)i * kNumCols + kColSubIdx];
        if (!(v >= 0.0f) || v > 4294967040.0f)
            throw Error(kInvalidNetwork, "sub-block id must be a non-negative integer");
        const uint32_t id = (uint32_t)v;
        if (i == 0 || id != pop_ids_.back()) {
            if (!seen.insert(id).second)
                throw Error(kInvalidNetwork, "sub-block ids must form contiguous ranges");
            pop_ids_.push_back(id);
            pop_begin_.push_back(i);
        }
    }
    pop_begin_.push_back(n_);
    const size_t np = pop_ids_.size();
    ou_on_.assign(np, 0);
    ou_mean_.assign(np, 0.0f);
    ou_sigma_.assign(np, 0.0f);
    ou_tau_.assign(np, 1.0f);

    J_.resize(n_);
    blocked_.resize(n_);
    flag_.resize(n_);
    V_.resize(n_);
    iou_.resize(n_);
    tcnt_.resize(n_);
    dlt_.resize(n_);
    spill_.resize((size_t)n_ * 8);
    rowptr_.resize((size_t)n_ + 1);
    col_.resize((size_t)std::max<int64_t>(nsyn_, 1));
    w_.resize((size_t)std::max<int64_t>(nsyn_, 1) * weight_bytes(wfmt_));
    pops_.resize(np);
    step_.resize(2);
    counters_.resize(2 * (size_t)((n_ + kWave - 1) / kWave));

    derive_params(s);
    init_state(s);

}

Instance::Instance(const FanoutSpec &f, float dt, uint64_t seed, hipStream_t s)
{
    if (f.n < 2 || f.n > (int64_t)kColMask || f.K == 0 || f.K > 8192)
        throw Error(kInvalidParameter, "fan-out network: need n >= 2 and 1 <= K <= 8192");
    if (f.n * (int64_t)f.K >= kMaxSynapses)
        throw Error(kInvalidNetwork, "too many synapses for one block");
    n_ = (uint32_t)f.n;
    nsyn_ = f.n * (int64_t)f.K;
    // The synapses are generated on the device while the host builds the
    // property matrix and parameter records (init_common's uploads follow
    // the generation on the same stream).
    rowptr_.resize((size_t)n_ + 1);
    col_.resize((size_t)nsyn_);
    wfmt_ = kWeightF32;
    w_.resize((size_t)nsyn_ * weight_bytes(wfmt_));
    const uint32_t ntile = (n_ + (1u << kTileBits) - 1) >> kTileBits;
    DataAllocator<uint32_t> tsyn(ntile);
    DTB_HIP(launch_gen_fanout(n_, f.K, f.net_seed, rowptr_.mutable_gpu_data(), col_.mutable_gpu_data(),
                              reinterpret_cast<float2 *>(w_.mutable_gpu_data()), tsyn.mutable_gpu_data(), s));
    std::vector<float> props((size_t)n_ * kNumCols);
    gen_default_props(f.n, f.g_ui, f.sub_base, props.data());
    init_common(std::move(props), dt, seed, s);
    tile_syn_.assign(ntile, 0u);
    tsyn.download(tile_syn_.data(), ntile, s);
    DTB_HIP(hipStreamSynchronize(s));
    // expected distinct tiles (groups) among K uniform targets, as a fraction
    row_coverage_ = 1.0 - std::pow(1.0 - 1.0 / ntile, (double)f.K);
    group_coverage_ = 1.0 - std::pow(1.0 - 1.0 / ngroups(), (double)f.K);
    // rows are ordered by target, K <= 8192
    gather_ok_ = (double)n_ * (ngroups() + 1.0) * 2.0 <= 12.0 * (double)nsyn_;
}

Instance::~Instance() = default;

void Instance::count_tile_synapses(const NetworkHost &net)
{
    // Synapses per target tile: a tile's queue holds at most one entry per
    // synapse that targets it (every row fires at most once per step), so
    // these counts are its exact capacity.
    const uint32_t nfine = (n_ + (1u << kTileBits) - 1) >> kTileBits;
    tile_syn_.assign(nfine, 0u);
    unsigned hw = std::thread::hardware_concurrency();
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(std::min(hw ? hw : 4u, 16u), nsyn_ / (1 << 20)));
    std::vector<std::vector<uint32_t>> part(nt, std::vector<uint32_t>(nfine, 0u));
    std::vector<uint64_t> segs(nt, 0), gsegs(nt, 0);
    std::vector<char> sorted(nt, 1);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            const int64_t a = (int64_t)n_ * t / nt, b = (int64_t)n_ * (t + 1) / nt;
            uint32_t *c = part[t].data();
            uint64_t sg = 0, gs = 0;
            bool ordered = true;
            for (int64_t r = a; r < b; ++r) {
                uint32_t prev = 0xffffffffu, gprev = 0xffffffffu;
                for (int64_t j = net.rowptr[r]; j < net.rowptr[r + 1]; ++j) {
                    const uint32_t tl = net.col[j] >> kTileBits;
                    const uint32_t gr = net.col[j] >> kGroupBits;
                    ++c[tl];
                    sg += tl != prev;
                    gs += gr != gprev;
                    ordered = ordered && (gprev == 0xffffffffu || gr >= gprev);
                    prev = tl;
                    gprev = gr;
                }
                ordered = ordered && net.rowptr[r + 1] - net.rowptr[r] <= kGatherMaxRow;
            }
            segs[t] = sg;
            gsegs[t] = gs;
            sorted[t] = ordered ? 1 : 0;
        });
    }
    for (auto &x : th)
        x.join();
    uint64_t sg = 0, gs = 0;
    bool ordered = true;
    for (int t = 0; t < nt; ++t) {
        sg += segs[t];
        gs += gsegs[t];
        ordered = ordered && sorted[t];
        for (uint32_t k = 0; k < nfine; ++k)
            tile_syn_[k] += part[t][k];
    }
    row_coverage_ = (double)sg / ((double)n_ * (double)nfine);
    group_coverage_ = (double)gs / ((double)n_ * (double)ngroups());
    gather_ok_ = ordered && nsyn_ > 0 && (double)n_ * (ngroups() + 1.0) * 2.0 <= 12.0 * (double)nsyn_;
}

uint32_t Instance::fired_capacity() const
{
    const uint32_t ntile = (n_ + kBlock - 1) / kBlock;
    return (ntile + kFiredLists - 1) / kFiredLists * kBlock;
}

void Instance::ensure_gather(hipStream_t s)
{
    if (!gather_ok_)
        throw Error(kInvalidParameter, "this network does not allow gather propagation");
    if (gather_built_)
        return;
    try {
        fcount_.resize(2 * (size_t)kFiredLists * kFiredCountStride);
        fcount_.mutable_gpu_data();
        fired_.resize(2 * (size_t)kFiredLists * fired_capacity());
        fired_.mutable_gpu_data();
        segidx_.resize((size_t)n_ * (ngroups() + 1u));
        DTB_HIP(launch_build_segidx(rowptr_.gpu_data(), col_.gpu_data(), n_, ngroups(), segidx_.mutable_gpu_data(),
                                    s));
        grec_.resize((size_t)nsyn_ * grec_words(wfmt_));
        DTB_HIP(launch_build_grec(col_.gpu_data(), w_.gpu_data(), wfmt_, (uint64_t)nsyn_, grec_.mutable_gpu_data(),
                                  s));
        DTB_HIP(hipStreamSynchronize(s));
    } catch (...) {
        // out of device memory (or a failed build): nothing half-built stays
        segidx_.resize(0);
        grec_.resize(0);
        fired_.resize(0);
        fcount_.resize(0);
        throw;
    }
    gather_built_ = true;
}

void Instance::release_gather()
{
    segidx_.resize(0);
    grec_.resize(0);
    fired_.resize(0);
    fcount_.resize(0);
    gather_built_ = false;
}

void Instance::clear_fired(hipStream_t s)
{
    if (gather_built_)
        DTB_HIP(hipMemsetAsync(fcount_.mutable_gpu_data(), 0, fcount_.size(), s));
}

uint64_t Instance::pull_bytes() const
{
    const uint64_t ntile = (n_ + kBlock - 1) / kBlock;
    const uint64_t nw = (n_ + kWave - 1) / kWave;
    return (uint64_t)nsyn_ * sizeof(uint32_t) * grec_words(wfmt_) + (ntile + 1) * sizeof(int64_t) +
           4 * nw * sizeof(unsigned long long) + 2 * sizeof(int64_t);
}

void Instance::ensure_pull(hipStream_t s)
{
    if (!pull_allowed())
        throw Error(kInvalidParameter, "this network does not allow pull propagation");
    if (pull_built_)
        return;
    const uint32_t ntile = (n_ + kBlock - 1) / kBlock;
    const size_t nw = (n_ + kWave - 1) / kWave;
    try {
        // first in-edge of every target tile: prefix sums of the synapses
        // targeting the tiles
        std::vector<int64_t> ptr((size_t)ntile + 1, 0);
        for (uint32_t t = 0; t < ntile; ++t)
            ptr[t + 1] = ptr[t] + tile_syn_.at(t);
        if (ptr[ntile] != nsyn_)
            throw Error(kInternal, "tile synapse counts do not add up to the synapses");
        cptr_.resize((size_t)ntile + 1);
        cptr_.upload(ptr.data(), ptr.size(), s);
        crec_.resize((size_t)nsyn_ * grec_words(wfmt_));
        ebits_.resize(2 * nw);
        dbits_.resize(2 * nw);
        dstamp_.resize(2);
        ebits_.mutable_gpu_data();
        dbits_.mutable_gpu_data();
        dstamp_.mutable_gpu_data();
        DataAllocator<uint32_t> cursor(ntile);
        DTB_HIP(hipMemsetAsync(cursor.mutable_gpu_data(), 0, (size_t)ntile * sizeof(uint32_t), s));
        DTB_HIP(launch_build_pull(rowptr_.gpu_data(), col_.gpu_data(), w_.gpu_data(), wfmt_, n_, cptr_.gpu_data(),
                                  cursor.mutable_gpu_data(), crec_.mutable_gpu_data(), s));
        DTB_HIP(hipStreamSynchronize(s));
    } catch (...) {
        release_pull();
        throw;
    }
    pull_built_ = true;
}

void Instance::release_pull()
{
    crec_.resize(0);
    cptr_.resize(0);
    ebits_.resize(0);
    dbits_.resize(0);
    dstamp_.resize(0);
    pull_built_ = false;
}

void Instance::clear_deferred(hipStream_t s)
{
    if (pull_built_)
        DTB_HIP(hipMemsetAsync(dstamp_.mutable_gpu_data(), 0, dstamp_.size(), s));
}

uint32_t Instance::tile_synapses(uint32_t tile) const
{
    return tile_syn_.at(tile);
}

namespace {

// The kernel-side parameter record of one neuron (five float4 columns).
struct ParamRec {
    float4 a, b, g, v, d;
};

bool same_bits(const ParamRec &x, const ParamRec &y)
{
    return std::memcmp(&x, &y, sizeof(ParamRec)) == 0;
}

struct ParamRecHash {
    size_t operator()(const ParamRec &r) const
    {
        uint32_t w[sizeof(ParamRec) / 4];
        std::memcpy(w, &r, sizeof(ParamRec));
        uint64_t h = 0x9e3779b97f4a7c15ull;
        for (uint32_t x : w)
            h = (h ^ x) * 0x100000001b3ull;
        return (size_t)(h ^ (h >> 29));
    }
};

struct ParamRecEq {
    bool operator()(const ParamRec &x, const ParamRec &y) const { return same_bits(x, y); }
};

}  // namespace

void Instance::derive_params(hipStream_t s)
{
    const float dt_sub = dt_ / (float)t_steps_;
    const float *prev = nullptr;
    ParamRec cur{};
    auto derive = [&](uint32_t i) -> const ParamRec & {
        const float *p = &props_[(size_t)i * kNumCols];
        if (prev != nullptr && std::memcmp(p, prev, sizeof(float) * kNumCols) == 0)
            return cur;   // same property row as the neuron before
        prev = p;
        cur.a = make_float4(p[kColVReset], p[kColVTh], 1.0f / p[kColC], p[kColVLeak]);
        cur.b = make_float4(p[kColGLeak], p[kColIExt], p[kColNoiseRate],
                            bits_as_float(refractory_substeps(p[kColTRef], dt_, t_steps_)));
        cur.g = make_float4(p[kColGU + 0], p[kColGU + 1], p[kColGU + 2], p[kColGU + 3]);
        cur.v = make_float4(p[kColVU + 0], p[kColVU + 1], p[kColVU + 2], p[kColVU + 3]);
        // property.cpp:19-29: exp(-dt / tau)
        cur.d = make_float4(host_expf((-dt_sub) / p[kColTauU + 0]), host_expf((-dt_sub) / p[kColTauU + 1]),
                            host_expf((-dt_sub) / p[kColTauU + 2]), host_expf((-dt_sub) / p[kColTauU + 3]));
        return cur;
    };

    // Deduplicate the records: class per neuron + one record per class.
    std::vector<ParamRec> recs;
    std::vector<uint16_t> cls(n_);
    std::unordered_map<ParamRec, uint32_t, ParamRecHash, ParamRecEq> index;
    bool dense = false;
    for (uint32_t i = 0; i < n_ && !dense; ++i) {
        const ParamRec &r = derive(i);
        if (!recs.empty() && same_bits(r, recs[cls[i - 1]])) {
            cls[i] = cls[i - 1];
            continue;
        }
        auto it = index.find(r);
        if (it == index.end()) {
            if (recs.size() == kMaxParamClasses) {
                dense = true;
                break;
            }
            it = index.emplace(r, (uint32_t)recs.size()).first;
            recs.push_back(r);
        }
        cls[i] = (uint16_t)it->second;
    }
    dense_params_ = dense;
    const size_t nrec = dense ? n_ : recs.size();
    pA_.resize(nrec);
    pB_.resize(nrec);
    gU_.resize(nrec);
    vU_.resize(nrec);
    dec_.resize(nrec);
    cls_.resize(n_);
    float4 *pa = pA_.mutable_cpu_data();
    float4 *pb = pB_.mutable_cpu_data();
    float4 *gu = gU_.mutable_cpu_data();
    float4 *vu = vU_.mutable_cpu_data();
    float4 *dc = dec_.mutable_cpu_data();
    auto put = [&](size_t k, const ParamRec &r) {
        pa[k] = r.a;
        pb[k] = r.b;
        gu[k] = r.g;
        vu[k] = r.v;
        dc[k] = r.d;
    };
    if (dense) {
        prev = nullptr;
        uint16_t *cl = cls_.mutable_cpu_data();
        for (uint32_t i = 0; i < n_; ++i) {
            put(i, derive(i));
            cl[i] = (uint16_t)(i & (kBlock - 1));
        }
    } else {
        for (size_t k = 0; k < recs.size(); ++k)
            put(k, recs[k]);
        std::memcpy(cls_.mutable_cpu_data(), cls.data(), sizeof(uint16_t) * n_);
    }
    const uint32_t ntile = (n_ + kBlock - 1) / kBlock;
    tile_cls_.assign(ntile, 0u);
    if (!dense) {
        for (uint32_t t = 0; t < ntile; ++t) {
            const uint32_t b = t * kBlock, e = std::min<uint32_t>(n_, b + kBlock);
            bool same = true;
            for (uint32_t i = b + 1; i < e && same; ++i)
                same = cls[i] == cls[b];
            tile_cls_[t] = same ? (uint32_t)cls[b] + 1u : 0u;
        }
    }
    uint8_t *bl = blocked_.mutable_cpu_data();
    for (uint32_t i = 0; i < n_; ++i)
        bl[i] = props_[(size_t)i * kNumCols + kColBlocked] != 0.0f;
    pA_.to_gpu(s);
    pB_.to_gpu(s);
    gU_.to_gpu(s);
    vU_.to_gpu(s);
    dec_.to_gpu(s);
    cls_.to_gpu(s);
    blocked_.to_gpu(s);
    derive_pops(s);
}

void Instance::derive_pops(hipStream_t s)
{
    // extern_stimuli.cpp:84-110: per-population OU factors at the substep dt.
    const float dt_sub = dt_ / (float)t_steps_;
    PopDev *pd = pops_.mutable_cpu_data();
    for (size_t q = 0; q < pop_ids_.size(); ++q) {
        PopDev r{};
        r.begin = pop_begin_[q];
        r.end = pop_begin_[q + 1];
        r.ou_on = ou_on_[q];
        if (ou_on_[q]) {
            ou_factors(dt_sub, ou_tau_[q], ou_sigma_[q], &r.f1, &r.f2s);
            r.mu = ou_mean_[q];
        }
        pd[q] = r;
    }
    pops_.to_gpu(s);
    DTB_HIP(hipStreamSynchronize(s));
}

void Instance::init_state(hipStream_t s)
{
    // block.cpp:136-179 and membrane_voltage.cpp:18.
    float *V = V_.mutable_cpu_data();
    for (uint32_t i = 0; i < n_; ++i) {
        const float *p = &props_[(size_t)i * kNumCols];
        V[i] = (p[kColVTh] + p[kColVReset]) / 2.0f;
    }
    V_.to_gpu(s);
    float *io = iou_.mutable_cpu_data();
    std::fill(io, io + n_, 0.0f);
    iou_.to_gpu(s);
    DTB_HIP(hipMemsetAsync(tcnt_.mutable_gpu_data(), 0, tcnt_.size(), s));
    DTB_HIP(hipMemsetAsync(flag_.mutable_gpu_data(), 0, flag_.size(), s));
    DTB_HIP(hipMemsetAsync(J_.mutable_gpu_data(), 0, J_.size(), s));
    DTB_HIP(hipMemsetAsync(dlt_.mutable_gpu_data(), 0, dlt_.size(), s));
    DTB_HIP(hipMemsetAsync(spill_.mutable_gpu_data(), 0, spill_.size(), s));
    DTB_HIP(hipMemsetAsync(step_.mutable_gpu_data(), 0, step_.size(), s));
    DTB_HIP(hipMemsetAsync(counters_.mutable_gpu_data(), 0, counters_.size(), s));
    DTB_HIP(hipStreamSynchronize(s));
    V_.release_cpu();
    iou_.release_cpu();
}

uint32_t Instance::find_pop(uint32_t brain_id) const
{
    for (size_t q = 0; q < pop_ids_.size(); ++q)
        if (pop_ids_[q] == brain_id)
            return (uint32_t)q;
    return UINT32_MAX;
}

void Instance::set_ou(const uint32_t *brain_ids, const float *mean, const float *sigma, const float *tau,
                      size_t n, hipStream_t s)
{
    // block.cpp:222-281.  A population seen for the first time has its OU
    // current reset to the mean (reset_partial_gpu); a known one only has its
    // parameters replaced.  Ids that are not in this block are skipped.
    for (size_t k = 0; k < n; ++k) {
        if (!(tau[k] > 0.0f))
            throw Error(kInvalidParameter, "OU correlation time must be positive");
    }
    for (size_t k = 0; k < n; ++k) {
        const uint32_t q = find_pop(brain_ids[k]);
        if (q == UINT32_MAX)
            continue;
        const bool first = !ou_on_[q];
        ou_on_[q] = 1;
        ou_mean_[q] = mean[k];
        ou_sigma_[q] = sigma[k];
        ou_tau_[q] = tau[k];
        if (first) {
            const uint32_t b = pop_begin_[q], e = pop_begin_[q + 1];
            std::vector<float> fill(e - b, mean[k]);
            DTB_HIP(hipMemcpyAsync(iou_.mutable_gpu_data() + b, fill.data(), fill.size() * sizeof(float),
                                   hipMemcpyHostToDevice, s));
            DTB_HIP(hipStreamSynchronize(s));
        }
    }
    derive_pops(s);   // the neurons' parameter records do not involve the OU current
}

void Instance::update_props(const uint32_t *neuron_ids, const uint32_t *prop_ids, const float *vals, size_t n,
                            hipStream_t s)
{
    // property.cpp:57-183 sets one property of one neuron; the columns of the
    // host property matrix are updated and the kernel inputs re-derived.
    for (size_t k = 0; k < n; ++k) {
        if (prop_type_column((int)prop_ids[k]) < 0)
            throw Error(kInvalidParameter, "unknown property id");
    }
    for (size_t k = 0; k < n; ++k) {
        if (neuron_ids[k] >= n_)
            continue;   // property.cpp:73 ignores out-of-range neurons
        props_[(size_t)neuron_ids[k] * kNumCols + prop_type_column((int)prop_ids[k])] = vals[k];
    }
    derive_params(s);
}

void Instance::update_prop_cols(const uint32_t *prop_ids, const uint32_t *brain_ids, const float *vals, size_t n,
                                bool assign, hipStream_t s)
{
    // property.cpp:185-466: per-population assign (=) or multiply (*=).
    for (size_t k = 0; k < n; ++k) {
        if (prop_type_column((int)prop_ids[k]) < 0)
            throw Error(kInvalidParameter, "unknown property id");
    }
    for (size_t k = 0; k < n; ++k) {
        const uint32_t q = find_pop(brain_ids[k]);
        if (q == UINT32_MAX)
            continue;
        const int c = prop_type_column((int)prop_ids[k]);
        for (uint32_t i = pop_begin_[q]; i < pop_begin_[q + 1]; ++i) {
            float &x = props_[(size_t)i * kNumCols + c];
            x = assign ? vals[k] : x * vals[k];
        }
    }
    derive_params(s);
}

void Instance::gamma_prop_cols(const uint32_t *prop_ids, const uint32_t *brain_ids, const float *alpha,
                               const float *beta, size_t n, hipStream_t s)
{
    // block.cpp:1472-1536, property.cpp:470-640: every neuron of the
    // population gets its own Gamma(alpha, rate beta) draw as the property.
    for (size_t k = 0; k < n; ++k) {
        if (prop_type_column((int)prop_ids[k]) < 0)
            throw Error(kInvalidParameter, "unknown property id");
        if (!(alpha[k] >= 0.0f) || !(beta[k] > 0.0f) || !std::isfinite(alpha[k]) || !std::isfinite(beta[k]))
            throw Error(kInvalidParameter, "gamma needs concentration >= 0 and rate > 0");
    }
    std::vector<float> draws;
    for (size_t k = 0; k < n; ++k) {
        const uint32_t call = gamma_calls_++;
        const uint32_t q = find_pop(brain_ids[k]);
        if (q == UINT32_MAX)
            continue;
        const int c = prop_type_column((int)prop_ids[k]);
        const uint32_t b = pop_begin_[q], e = pop_begin_[q + 1];
        draws.resize(e - b);
        gen_gamma(seed_, call, b, e - b, alpha[k], beta[k], draws.data());
        for (uint32_t i = b; i < e; ++i)
            props_[(size_t)i * kNumCols + c] = draws[i - b];
    }
    derive_params(s);
}

void Instance::read_props(float *out) const
{
    std::memcpy(out, props_.data(), sizeof(float) * props_.size());
}

void Instance::set_samples(const uint32_t *ids, size_t n)
{
    for (size_t k = 0; k < n; ++k)
        if (ids[k] >= n_)
            throw Error(kInvalidParameter, "sample id out of range");
    samples_.assign(ids, ids + n);
}

bool Instance::configure_run(int32_t t_steps, bool use_ou, hipStream_t s)
{
    if (t_steps < 1)
        throw Error(kInvalidParameter, "t_steps must be >= 1");
    if (t_steps > 1 && !use_ou)
        throw Error(kInvalidParameter, "t_steps > 1 requires the OU background");   // dist_simulator.cpp:3225-3234
    const bool changed = t_steps != t_steps_ || use_ou != use_ou_;
    use_ou_ = use_ou;
    if (t_steps != t_steps_) {
        t_steps_ = t_steps;
        derive_params(s);
    }
    return changed;
}

void Instance::begin_record(int64_t steps, bool freq, bool samples, uint32_t stat_mask, hipStream_t s)
{
    if (samples && samples_.empty())
        throw Error(kUnsetSampleFailure, "samples requested but not set");
    rec_step0_ = step(s);
    rec_len_ = steps;
    rec_freq_ = freq;
    rec_samples_ = samples;
    rec_stat_mask_ = stat_mask;
    if (stat_mask != 0) {
        rec_stat_.resize((size_t)steps * npop() * kStatFields);
        DTB_HIP(hipMemsetAsync(rec_stat_.mutable_gpu_data(), 0, rec_stat_.size(), s));
    }
    if (freq) {
        rec_counts_.resize((size_t)steps * npop());
        DTB_HIP(hipMemsetAsync(rec_counts_.mutable_gpu_data(), 0, rec_counts_.size(), s));
    }
    if (samples) {
        const size_t m = (size_t)steps * samples_.size();
        rec_spike_.resize(m);
        rec_v_.resize(m);
        rec_isyn_.resize(m);
        rec_iou_.resize(m);
        rec_spike_.mutable_gpu_data();
        rec_v_.mutable_gpu_data();
        rec_isyn_.mutable_gpu_data();
        rec_iou_.mutable_gpu_data();
    }
}

void Instance::begin_record_at(int64_t steps, bool freq, bool samples, uint32_t stat_mask, const RecordBuffers &b,
                               hipStream_t s)
{
    if (samples && samples_.empty())
        throw Error(kUnsetSampleFailure, "samples requested but not set");
    if ((freq && !b.counts) || (stat_mask != 0 && !b.stat) || (samples && !(b.spike && b.v && b.isyn && b.iou)))
        throw Error(kInvalidParameter, "record buffer missing");
    rec_step0_ = step(s);
    rec_len_ = steps;
    rec_freq_ = freq;
    rec_samples_ = samples;
    rec_stat_mask_ = stat_mask;
    rec_ext_ = b;
    rec_ext_on_ = true;
    if (stat_mask != 0)
        DTB_HIP(launch_zero_words(b.stat, (size_t)steps * npop() * kStatFields * sizeof(long long), s));
    if (freq)
        DTB_HIP(launch_zero_words(b.counts, (size_t)steps * npop() * sizeof(uint32_t), s));
}

void Instance::end_record()
{
    rec_ext_on_ = false;
    rec_ext_ = RecordBuffers{};
    rec_freq_ = rec_samples_ = false;
    rec_stat_mask_ = 0;
    rec_len_ = 0;
}

void Instance::fetch_record(uint32_t *freq, uint8_t *spike, float *v, float *isyn, float *iou, float *const *means,
                            hipStream_t s)
{
    if (rec_ext_on_)
        throw Error(kInternal, "the recording is in caller buffers");
    if (means != nullptr && rec_stat_mask_ != 0) {
        // stat_result.cpp:266-391 divides by the population size minus its
        // blocked neurons: mean = (double) sum * 2^-24 / count, as float.
        std::vector<long long> sums(rec_stat_.count());
        rec_stat_.download(sums.data(), sums.size(), s);
        DTB_HIP(hipStreamSynchronize(s));
        const uint32_t np = npop();
        std::vector<double> cnt(np, 0.0);
        for (uint32_t q = 0; q < np; ++q)
            for (uint32_t i = pop_begin_[q]; i < pop_begin_[q + 1]; ++i)
                cnt[q] += props_[(size_t)i * kNumCols + kColBlocked] != 0.0f ? 0.0 : 1.0;
        for (int k = 0; k < kStatFields; ++k) {
            if (means[k] == nullptr)
                continue;
            for (int64_t r = 0; r < rec_len_; ++r)
                for (uint32_t q = 0; q < np; ++q)
                    means[k][(size_t)r * np + q] =
                        (float)((double)sums[((size_t)r * np + q) * kStatFields + k] * 0x1p-24 / cnt[q]);
        }
    }
    if (freq && rec_freq_)
        rec_counts_.download(freq, rec_counts_.count(), s);
    if (rec_samples_) {
        if (spike)
            rec_spike_.download(spike, rec_spike_.count(), s);
        if (v)
            rec_v_.download(v, rec_v_.count(), s);
        if (isyn)
            rec_isyn_.download(isyn, rec_isyn_.count(), s);
        if (iou)
            rec_iou_.download(iou, rec_iou_.count(), s);
    }
    DTB_HIP(hipStreamSynchronize(s));
}

int64_t Instance::step(hipStream_t s)
{
    int64_t v[2] = {0, 0};
    step_.download(v, 2, s);
    DTB_HIP(hipStreamSynchronize(s));
    return std::max(v[0], v[1]);
}

void Instance::set_step(int64_t step, hipStream_t s)
{
    const int64_t v[2] = {step, step};
    step_.upload(v, 2, s);
    DTB_HIP(hipStreamSynchronize(s));
}

void Instance::counters(int64_t *step, uint64_t *events, uint64_t *spikes, hipStream_t s)
{
    std::vector<unsigned long long> w(counters_.count());
    int64_t st[2] = {0, 0};
    counters_.download(w.data(), w.size(), s);
    step_.download(st, 2, s);
    DTB_HIP(hipStreamSynchronize(s));
    if (step)
        *step = std::max(st[0], st[1]);
    unsigned long long c[2] = {0, 0};
    for (size_t k = 0; k < w.size(); k += 2) {
        c[0] += w[k];
        c[1] += w[k + 1];
    }
    if (events)
        *events = c[0];
    if (spikes)
        *spikes = c[1];
}

void Instance::read_state(float *V, int32_t *t, uint8_t *f, uint8_t *emitted, float *iou, float *J, float *isyn,
                          hipStream_t s)
{
    std::vector<float> v(n_);
    std::vector<uint8_t> fl(n_);
    V_.download(v.data(), n_, s);
    flag_.download(fl.data(), n_, s);
    if (t)
        tcnt_.download(t, n_, s);
    if (iou)
        iou_.download(iou, n_, s);
    std::vector<float4> j, dl;
    if (J || isyn) {
        j.resize(n_);
        dl.resize(n_);
        J_.download(j.data(), n_, s);
        dlt_.download(dl.data(), n_, s);
    }
    DTB_HIP(hipStreamSynchronize(s));
    if (V)
        std::memcpy(V, v.data(), sizeof(float) * n_);
    for (uint32_t i = 0; i < n_; ++i) {
        if (f)
            f[i] = fl[i] & kFlagSpike;
        if (emitted)
            emitted[i] = (fl[i] & kFlagEmit) ? 1 : 0;
    }
    if (J || isyn) {
        // Fold the pending delta exactly as the next neuron step will.
        const float4 *dcs = dec_.cpu_data();
        const float4 *gs = gU_.cpu_data();
        const float4 *vus = vU_.cpu_data();
        const uint16_t *cl = cls_.cpu_data();
        const uint32_t pmask = dense_params_ ? ~(kBlock - 1) : 0u;
        for (uint32_t i = 0; i < n_; ++i) {
            const size_t c = (i & pmask) + cl[i];
            const float4 *dc = dcs + c, *g = gs + c, *vu = vus + c;
            float4 r = j[i];
            if (!(fl[i] & kFlagFolded)) {
                r.x = std::fma(j[i].x, dc->x, dl[i].x);
                r.y = std::fma(j[i].y, dc->y, dl[i].y);
                r.z = std::fma(j[i].z, dc->z, dl[i].z);
                r.w = std::fma(j[i].w, dc->w, dl[i].w);
            }
            if (J) {
                J[(size_t)i * 4 + 0] = r.x;
                J[(size_t)i * 4 + 1] = r.y;
                J[(size_t)i * 4 + 2] = r.z;
                J[(size_t)i * 4 + 3] = r.w;
            }
            if (isyn) {
                const float a = g->x * (vu->x - v[i]) * r.x;
                const float b = g->y * (vu->y - v[i]) * r.y;
                const float c2 = g->z * (vu->z - v[i]) * r.z;
                const float d = g->w * (vu->w - v[i]) * r.w;
                isyn[i] = a + b + c2 + d;
            }
        }
    }
}

void Instance::write_state(const float *V, const int32_t *t, const uint8_t *f, const float *iou, const float *J,
                           hipStream_t s)
{
    // Checkpoint restore.  J is the reference-visible J of the last step, so
    // the pending delta is cleared and the flag byte marks J as folded: the
    // next neuron step then uses it as is.
    if (V)
        V_.upload(V, n_, s);
    if (t)
        tcnt_.upload(t, n_, s);
    if (iou)
        iou_.upload(iou, n_, s);
    if (f || J) {
        std::vector<uint8_t> fl(n_);
        flag_.download(fl.data(), n_, s);
        DTB_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < n_; ++i) {
            uint8_t x = fl[i];
            if (f)
                x = (uint8_t)((x & ~kFlagSpike) | (f[i] ? kFlagSpike : 0));
            if (J)
                x = (uint8_t)(x | kFlagFolded);
            fl[i] = x;
        }
        flag_.upload(fl.data(), n_, s);
        DTB_HIP(hipStreamSynchronize(s));
    }
    if (J) {
        J_.upload(reinterpret_cast<const float4 *>(J), n_, s);
        DTB_HIP(hipMemsetAsync(dlt_.mutable_gpu_data(), 0, dlt_.size(), s));
    }
    DTB_HIP(hipStreamSynchronize(s));
}

void Instance::ensure_queue(uint64_t slots)
{
    slots = std::max<uint64_t>(slots, 1);
    if (slots > qslots_ || queue_.count() == 0) {
        queue_.resize((size_t)(2 * slots + kQueuePrefetch));
        queue_.mutable_gpu_data();
        qslots_ = slots;
    }
}

void Instance::clear_spill(hipStream_t s)
{
    DTB_HIP(hipMemsetAsync(spill_.mutable_gpu_data(), 0, spill_.size(), s));
}

InstDev Instance::dev(uint32_t tile_base, bool gather, bool pull, uint32_t defer_min)
{
    InstDev d{};
    if (pull) {
        if (!pull_built_ || gather)
            throw Error(kInternal, "pull propagation needs the in-edges and no gather");
        d.crec = crec_.gpu_data();
        d.cptr = cptr_.gpu_data();
        d.ebits = ebits_.mutable_gpu_data();
        d.dbits = dbits_.mutable_gpu_data();
        d.dstamp = dstamp_.mutable_gpu_data();
        d.defer_min = defer_min;
    }
    if (gather) {
        if (!gather_built_)
            throw Error(kInvalidParameter, "gather buffers not built");
        d.segidx = segidx_.gpu_data();
        d.grec = grec_.gpu_data();
        d.fired = fired_.mutable_gpu_data();
        d.fcount = fcount_.mutable_gpu_data();
        d.ngroups = ngroups();
        d.fcap = fired_capacity();
    }
    d.pA = pA_.gpu_data();
    d.pB = pB_.gpu_data();
    d.gU = gU_.gpu_data();
    d.vU = vU_.gpu_data();
    d.dec = dec_.gpu_data();
    d.cls = cls_.gpu_data();
    d.pmask = dense_params_ ? ~(kBlock - 1) : 0u;
    d.blocked = blocked_.gpu_data();
    d.V = V_.mutable_gpu_data();
    d.tcnt = tcnt_.mutable_gpu_data();
    d.flag = flag_.mutable_gpu_data();
    d.iou = iou_.mutable_gpu_data();
    d.J = J_.mutable_gpu_data();
    d.dlt = dlt_.mutable_gpu_data();
    d.spill = spill_.mutable_gpu_data();
    d.queue = queue_.mutable_gpu_data();
    d.qslots = qslots_;
    d.rowptr = rowptr_.gpu_data();
    d.col = col_.gpu_data();
    d.w = w_.gpu_data();
    d.wfmt = wfmt_;
    d.pops = pops_.gpu_data();
    d.rec_counts = (rec_freq_ && rec_len_ > 0) ? (rec_ext_on_ ? rec_ext_.counts : rec_counts_.mutable_gpu_data())
                                                : nullptr;
    d.rec_stat = (rec_stat_mask_ != 0 && rec_len_ > 0) ? (rec_ext_on_ ? rec_ext_.stat : rec_stat_.mutable_gpu_data())
                                                       : nullptr;
    d.stat_mask = rec_stat_mask_;
    d.wave_ctr = counters_.mutable_gpu_data();
    d.step = step_.mutable_gpu_data();
    d.seed = seed_;
    d.rec_step0 = rec_step0_;
    d.n = n_;
    d.npop = npop();
    d.rec_len = (uint32_t)rec_len_;
    d.t_steps = t_steps_;
    d.use_ou = use_ou_ ? 1 : 0;
    d.dt_sub = dt_ / (float)t_steps_;
    d.tile_base = tile_base;
    return d;
}

SampleRecDev Instance::sample_rec()
{
    SampleRecDev r{};
    if (rec_samples_ && rec_len_ > 0 && rec_ext_on_) {
        r.spike = rec_ext_.spike;
        r.v = rec_ext_.v;
        r.isyn = rec_ext_.isyn;
        r.iou = rec_ext_.iou;
        r.nsample = (uint32_t)samples_.size();
    } else if (rec_samples_ && rec_len_ > 0) {
        r.spike = rec_spike_.mutable_gpu_data();
        r.v = rec_v_.mutable_gpu_data();
        r.isyn = rec_isyn_.mutable_gpu_data();
        r.iou = rec_iou_.mutable_gpu_data();
        r.nsample = (uint32_t)samples_.size();
    }
    return r;
}

// ---------------------------------------------------------------------------

}  // namespace dtb
