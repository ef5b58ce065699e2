// dtb_engine.hpp -- C++ host runtime: device buffers, instances and the
// step engine that drives the gfx950 kernels.
//
// Vocabulary follows the reference host library:
//   DataAllocator<T>  include/data_allocator.hpp:8-38 (device buffer + host
//                     mirror, zero-filled, explicit copies)
//   Weights/DataType  include/weights.hpp:14-94 (two values per synapse)
//   Instance          one BrainBlock<float,float2> (include/block.hpp:21-889)
//   Engine            the per-GPU step loop of snn_run
//                     (tools/dist_simulator.cpp:2072-2315), here able to step
//                     many instances in one launch pair.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "dtb_types.hpp"

namespace dtb {

// Status codes: 0-5 are protos/snn.proto:44-51 SnnStatus.
enum StatusCode : int {
    kOk = 0,
    kInvalidParameter = 1,
    kUnsetSampleFailure = 2,
    kUnsupported = 3,
    kNotEnableMetric = 4,
    kInvalidNetwork = 5,
    kHipError = 100,
    kOutOfMemory = 101,
    kNoDevice = 102,
    kInternal = 103,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what, const char *file, int line);
#define DTB_HIP(x) ::dtb::hip_check((x), #x, __FILE__, __LINE__)

// Property-matrix columns (src/configuration.cpp:1778-2156).
enum Col : int {
    kColNoiseRate = 0,
    kColBlocked = 1,
    kColIExt = 2,
    kColSubIdx = 3,
    kColC = 4,
    kColTRef = 5,
    kColGLeak = 6,
    kColVLeak = 7,
    kColVTh = 8,
    kColVReset = 9,
    kColGU = 10,
    kColVU = 14,
    kColTauU = 18,
    kNumCols = 22,
};

// include/stage.hpp:320-341 == protos/snn.proto:94-115.
enum PropType : int {
    EXT_STIMULI_I = 0,
    MEMBRANE_C = 1,
    REF_T = 2,
    LEAKAGE_G = 3,
    LEAKAGE_V = 4,
    THRESHOLD_V = 5,
    RESET_V = 6,
    CONDUCT_G_AMPA = 7,
    CONDUCT_G_NMDA = 8,
    CONDUCT_G_GABAa = 9,
    CONDUCT_G_GABAb = 10,
    MEMBRANE_V_AMPA = 11,
    MEMBRANE_V_NMDA = 12,
    MEMBRANE_V_GABAa = 13,
    MEMBRANE_V_GABAb = 14,
    TAO_AMPA = 15,
    TAO_NMDA = 16,
    TAO_GABAa = 17,
    TAO_GABAb = 18,
    NOISE_RATE = 19,
    kNumPropTypes = 20,
};

int prop_type_column(int prop);   // PropType -> property column, -1 if unknown

// include/weights.hpp:14-21
enum DataType : int { DOUBLE = 0, FLOAT = 1, FLOAT16 = 2, INT32 = 3, INT8 = 4 };
size_t data_size(DataType t);

// Device memory that stays with the library once freed (dtb_devmem.cpp).
namespace devmem {
void *alloc(size_t bytes);           // on the current device
void release(void *p);               // block from alloc; may still be in use by queued work
void reserve(size_t bytes);          // keeps at least this much for blocks of 1 MB and more
void trim();                         // returns all unused memory to the runtime
size_t cached_bytes(int device);     // bytes held in the cache
// While one is alive on a thread, alloc on that thread throws instead of
// taking memory (the entry points that promise not to allocate).
struct NoAlloc {
    NoAlloc();
    ~NoAlloc();
    NoAlloc(const NoAlloc &) = delete;
    NoAlloc &operator=(const NoAlloc &) = delete;
};
}  // namespace devmem

// The static device array of dtb_run_noalloc (kScratchBytes) on the current
// device: its address, looked up once per device (dtb_noalloc.hip).
void *scratch_address(int device);
// Zeroes `bytes` (a multiple of 4) at p with a fill kernel.
hipError_t launch_zero_words(void *p, size_t bytes, hipStream_t stream);
// Per-population means of a recording: cnt[q] = non-blocked neurons of
// population q (counted on the device from the instance's blocked flags and
// population table), then out[k][r][q] = (float)((double)sums[r][q][k] *
// 2^-24 / cnt[q]) for every field k with out[k] != null.
hipError_t launch_stat_means(const uint8_t *blocked, const PopDev *pops, uint32_t npop, const long long *sums,
                             uint32_t rows, unsigned long long *cnt, float *const *out, hipStream_t stream);

// Device buffer with a lazily created, zero-filled host mirror.
template <typename T>
class DataAllocator {
public:
    explicit DataAllocator(size_t count = 0) { resize(count); }
    ~DataAllocator() { release(); }
    DataAllocator(const DataAllocator &) = delete;
    DataAllocator &operator=(const DataAllocator &) = delete;

    void resize(size_t count)
    {
        if (count == count_)
            return;
        release();
        count_ = count;
    }
    size_t count() const { return count_; }
    size_t size() const { return count_ * sizeof(T); }

    const T *gpu_data() { return mutable_gpu_data(); }
    T *mutable_gpu_data()
    {
        if (gpu_ == nullptr && count_ > 0) {
            gpu_ = static_cast<T *>(devmem::alloc(size()));
            // The zero fill runs on the null stream, which does not order
            // with the engine's non-blocking stream: wait for it here so a
            // later upload on the engine stream cannot be overwritten.
            DTB_HIP(hipMemsetAsync(gpu_, 0, size(), nullptr));
            DTB_HIP(hipStreamSynchronize(nullptr));
        }
        return gpu_;
    }
    const T *cpu_data() { return mutable_cpu_data(); }
    T *mutable_cpu_data()
    {
        if (cpu_.size() != count_)
            cpu_.assign(count_, T{});
        return cpu_.data();
    }
    void to_gpu(hipStream_t s)
    {
        if (count_ == 0)
            return;
        DTB_HIP(hipMemcpyAsync(mutable_gpu_data(), mutable_cpu_data(), size(), hipMemcpyHostToDevice, s));
    }
    void to_cpu(hipStream_t s)
    {
        if (count_ == 0)
            return;
        DTB_HIP(hipMemcpyAsync(mutable_cpu_data(), mutable_gpu_data(), size(), hipMemcpyDeviceToHost, s));
    }
    void upload(const T *src, size_t n, hipStream_t s)
    {
        if (n > count_)
            throw Error(kInvalidParameter, "upload larger than buffer");
        if (n == 0)
            return;
        DTB_HIP(hipMemcpyAsync(mutable_gpu_data(), src, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *dst, size_t n, hipStream_t s)
    {
        if (n > count_)
            throw Error(kInvalidParameter, "download larger than buffer");
        if (n == 0)
            return;
        DTB_HIP(hipMemcpyAsync(dst, mutable_gpu_data(), n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void release_cpu() { std::vector<T>().swap(cpu_); }

private:
    void release()
    {
        devmem::release(gpu_);
        gpu_ = nullptr;
        std::vector<T>().swap(cpu_);
        count_ = 0;
    }
    size_t count_ = 0;
    T *gpu_ = nullptr;
    std::vector<T> cpu_;
};

// Host description of a network in the reference layout: property matrix
// [n][22] and a presynaptic-major CSR whose synapses carry a weight pair and a
// channel bit (0: AMPA/NMDA pair, 1: GABA_A/GABA_B pair), bit-packed in
// 32-bit words as in presynaptic_voltage.cpp:152.
struct NetworkHost {
    int64_t n = 0;
    const float *props = nullptr;
    int64_t nsyn = 0;
    const int64_t *rowptr = nullptr;
    const uint32_t *col = nullptr;
    const void *w = nullptr;            // [nsyn][2] in format wfmt
    const uint32_t *kind_bits = nullptr;  // [ceil(nsyn/32)]
    uint32_t wfmt = kWeightF32;         // kWeightF32 / kWeightF16 / kWeightU8
};

// Per-run options (protos/snn.proto:63-77 RunRequest).
struct RunOptions {
    int32_t t_steps = 1;
    bool use_ou = true;
    bool record_freq = false;
    bool record_samples = false;
};

class Engine;

// One simulated block.
// A fan-out network generated on the device (gen_fanout / launch_gen_fanout).
struct FanoutSpec {
    int64_t n = 0;
    uint32_t K = 0;
    uint64_t net_seed = 0;
    float g_ui[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t sub_base = 0;
};

class Instance {
public:
    Instance(const NetworkHost &net, float dt, uint64_t seed, hipStream_t stream);
    // The network is generated in device memory (no host copy of the
    // synapses, nothing crosses PCIe but the property matrix).
    Instance(const FanoutSpec &spec, float dt, uint64_t seed, hipStream_t stream);
    ~Instance();

    uint32_t n() const { return n_; }
    int64_t nsyn() const { return nsyn_; }
    uint32_t weight_format() const { return wfmt_; }
    uint32_t npop() const { return (uint32_t)pop_ids_.size(); }
    const std::vector<uint32_t> &pop_ids() const { return pop_ids_; }
    const std::vector<uint32_t> &pop_begin() const { return pop_begin_; }
    float dt() const { return dt_; }
    uint64_t seed() const { return seed_; }
    int32_t t_steps() const { return t_steps_; }
    bool use_ou() const { return use_ou_; }

    // extern_stimuli.cpp / block.cpp:222-281 (set_I_ou_current_param).
    void set_ou(const uint32_t *brain_ids, const float *mean, const float *sigma, const float *tau, size_t n,
                hipStream_t s);
    // property.cpp:57-466 (update_props / assign_prop_cols / update_prop_cols).
    void update_props(const uint32_t *neuron_ids, const uint32_t *prop_ids, const float *vals, size_t n,
                      hipStream_t s);
    void update_prop_cols(const uint32_t *prop_ids, const uint32_t *brain_ids, const float *vals, size_t n,
                          bool assign, hipStream_t s);
    // Per-population Gamma(concentration, rate) resampling of properties:
    // draw stream = (seed, call count of this instance, neuron).
    void gamma_prop_cols(const uint32_t *prop_ids, const uint32_t *brain_ids, const float *alpha, const float *beta,
                         size_t n, hipStream_t s);
    // The [n][22] property matrix as updated so far.
    void read_props(float *out) const;
    void set_samples(const uint32_t *ids, size_t n);
    const std::vector<uint32_t> &samples() const { return samples_; }

    // Configure t_steps / OU mode for the next run (recomputes substep decays).
    bool configure_run(int32_t t_steps, bool use_ou, hipStream_t s);
    // Recording window of `steps` rows starting at the current step;
    // stat_mask selects per-population means (kStatV | kStatI | kStatReceptor).
    void begin_record(int64_t steps, bool freq, bool samples, uint32_t stat_mask, hipStream_t s);
    // The same window recorded into caller-provided device buffers (no
    // allocation): counts [steps][npop] and sums [steps][npop][kStatFields]
    // are zeroed here; the sample rows [steps][nsample] need no clearing.
    struct RecordBuffers {
        uint32_t *counts;
        long long *stat;
        uint8_t *spike;
        float *v, *isyn, *iou;
    };
    void begin_record_at(int64_t steps, bool freq, bool samples, uint32_t stat_mask, const RecordBuffers &b,
                         hipStream_t s);
    void end_record();
    // Device descriptors the statistics need (the instance's blocked flags
    // and population table).
    const uint8_t *blocked_dev() { return blocked_.gpu_data(); }
    const PopDev *pops_dev() { return pops_.gpu_data(); }
    uint32_t stat_mask() const { return rec_stat_mask_; }
    // means: [6] arrays [steps][npop] (vmean, imean, AMPA, NMDA, GABA_A,
    // GABA_B), any may be null.
    void fetch_record(uint32_t *freq, uint8_t *spike, float *v, float *isyn, float *iou, float *const *means,
                      hipStream_t s);

    // Reference-visible state after the last completed step: V, refractory
    // counter, spike flag, emitted flag, i_ou, J[4] (pending delta folded
    // in), I_syn (without i_ou).
    void read_state(float *V, int32_t *t, uint8_t *f, uint8_t *emitted, float *iou, float *J, float *isyn,
                    hipStream_t s);
    void write_state(const float *V, const int32_t *t, const uint8_t *f, const float *iou, const float *J,
                     hipStream_t s);
    // The step index is kept in two device slots (read and written by
    // alternating launch parities); the current step is the larger one.
    int64_t step(hipStream_t s);
    void set_step(int64_t step, hipStream_t s);
    void counters(int64_t *step, uint64_t *events, uint64_t *spikes, hipStream_t s);

    // Kernel descriptor; `gather`: propagate by gather (ensure_gather first);
    // defer_min > 0 with the in-edges built: pull propagation (a tile with
    // at least defer_min emitted rows defers them to the pull).
    InstDev dev(uint32_t tile_base, bool gather, bool pull, uint32_t defer_min);
    // Distinct parameter records (classes); n() when stored per neuron.
    uint32_t param_classes() const { return (uint32_t)pA_.count(); }
    SampleRecDev sample_rec();
    // Synapses whose target lies in tile `tile` (2^kTileBits neurons): the
    // exact queue bound of the binning.
    uint32_t tile_synapses(uint32_t tile) const;
    // 1 + the parameter class of every neuron of tile `tile` if they share
    // one, else 0.
    // Zeroes both spill accumulators (the queues of this instance were
    // cleared).
    void clear_spill(hipStream_t s);
    // Propagation queues of `slots` entries per parity (grow-only; the
    // contents need no initialisation: only the first `count` entries of a
    // tile's queue are ever read).
    void ensure_queue(uint64_t slots);
    uint32_t tile_class(uint32_t tile) const { return tile_cls_.at(tile); }
    // Mean over rows of the fraction of this instance's tiles that the row's
    // synapses reach (tile segments of the row / tiles).
    double row_coverage() const { return row_coverage_; }

    // Gather propagation (k_gather): possible when every row is ordered by
    // target group, no row is longer than kGatherMaxRow and the group offsets
    // ([n][ngroups + 1] x 2 bytes) are no larger than the synapse arrays.
    bool gather_allowed() const { return gather_ok_; }
    // Mean over rows of the fraction of the groups that the row reaches.
    double group_coverage() const { return group_coverage_; }
    uint32_t ngroups() const { return (n_ + (1u << kGroupBits) - 1) >> kGroupBits; }
    // Rows per fired list (the tiles of one list hold at most this many).
    uint32_t fired_capacity() const;
    // Builds the group offsets and fired lists if not built yet.
    void ensure_gather(hipStream_t s);
    // Returns them (no table may still reference them).
    void release_gather();
    bool gather_built() const { return gather_built_; }
    // Empties both fired lists (checkpoint restore).
    void clear_fired(hipStream_t s);

    // Pull propagation: a copy of the synapses grouped by target tile (the
    // in-edges, 12 bytes per synapse), through which a tile whose queue
    // overflowed, or whose sources deferred their rows, takes its input.
    // Possible for instances of at most kPullMaxNeurons neurons.
    bool pull_allowed() const { return n_ <= kPullMaxNeurons && nsyn_ > 0; }
    // Device bytes ensure_pull allocates.
    uint64_t pull_bytes() const;
    void ensure_pull(hipStream_t s);
    void release_pull();
    bool pull_built() const { return pull_built_; }
    // Forgets deferred rows (after the pending input was folded or dropped).
    void clear_deferred(hipStream_t s);

private:
    void init_common(const float *props, float dt, uint64_t seed, hipStream_t s);
    void init_common(std::vector<float> &&props, float dt, uint64_t seed, hipStream_t s);
    void derive_params(hipStream_t s);
    void derive_pops(hipStream_t s);     // population table (OU factors) only
    void init_state(hipStream_t s);
    void count_tile_synapses(const NetworkHost &net);
    uint32_t find_pop(uint32_t brain_id) const;

    uint32_t n_ = 0;
    int64_t nsyn_ = 0;
    float dt_ = 0.1f;
    uint64_t seed_ = 0;
    int32_t t_steps_ = 1;
    bool use_ou_ = true;

    std::vector<float> props_;           // host property matrix [n][22]
    std::vector<uint32_t> pop_ids_;      // brain id per population
    std::vector<uint32_t> pop_begin_;    // [npop+1]
    std::vector<uint8_t> ou_on_;
    std::vector<float> ou_mean_, ou_sigma_, ou_tau_;
    std::vector<uint32_t> samples_;

    DataAllocator<float4> pA_, pB_, gU_, vU_, dec_;    // parameter records per class
    DataAllocator<uint16_t> cls_;                      // parameter class per neuron
    bool dense_params_ = false;                        // one class per neuron
    DataAllocator<float4> J_;
    DataAllocator<uint8_t> blocked_, flag_;
    DataAllocator<float> V_, iou_;
    DataAllocator<int32_t> tcnt_;
    DataAllocator<float4> dlt_;
    DataAllocator<unsigned long long> spill_;      // [2][n][4] Q23.40
    DataAllocator<QSyn> queue_;                    // [2][qslots_] + kQueuePrefetch
    uint64_t qslots_ = 0;
    DataAllocator<int64_t> rowptr_;
    DataAllocator<uint32_t> col_;
    DataAllocator<uint8_t> w_;                     // [nsyn] weight pairs in format wfmt_
    uint32_t wfmt_ = kWeightF32;
    DataAllocator<PopDev> pops_;
    DataAllocator<int64_t> step_;                   // [2]
    DataAllocator<unsigned long long> counters_;   // per wave: events, spikes
    DataAllocator<uint32_t> rec_counts_;
    DataAllocator<long long> rec_stat_;            // [steps][npop][kStatFields]
    uint32_t rec_stat_mask_ = 0;
    DataAllocator<uint8_t> rec_spike_;
    DataAllocator<float> rec_v_, rec_isyn_, rec_iou_;
    int64_t rec_step0_ = 0;
    int64_t rec_len_ = 0;
    bool rec_freq_ = false, rec_samples_ = false;
    bool rec_ext_on_ = false;            // recording into rec_ext_ (begin_record_at)
    RecordBuffers rec_ext_{};
    std::vector<uint32_t> tile_syn_;      // synapses per target tile
    double row_coverage_ = 0.0;           // mean fraction of the tiles one row reaches
    double group_coverage_ = 0.0;         // mean fraction of the gather groups one row reaches
    bool gather_ok_ = false;              // gather_allowed()
    bool gather_built_ = false;           // segidx_, fired_, fcount_ built
    DataAllocator<uint16_t> segidx_;      // [n][ngroups + 1] group offsets (built on demand)
    DataAllocator<uint32_t> grec_;        // [nsyn][grec_words] gather records (built with segidx_)
    DataAllocator<uint32_t> fired_;       // [2][kFiredLists][fired_capacity()] fired rows
    DataAllocator<uint32_t> fcount_;      // [2][kFiredLists][kFiredCountStride]
    bool pull_built_ = false;             // crec_, cptr_, bitmaps built
    DataAllocator<uint32_t> crec_;        // [nsyn][grec_words] in-edges: key (source row | target in tile |
                                          // channel bit), weight pair bits
    DataAllocator<int64_t> cptr_;         // [ntiles + 1] first in-edge of each target tile
    DataAllocator<unsigned long long> ebits_, dbits_;   // [2][ceil(n / 64)] emitted / deferred rows
    DataAllocator<int64_t> dstamp_;       // [2] consuming step of deferred rows
    uint32_t gamma_calls_ = 0;            // gamma draw streams used
    std::vector<uint32_t> tile_cls_;      // per tile: 1 + shared parameter class, or 0
};

// Steps a set of instances together.  One launch per step covers every
// neuron of every active instance (one workgroup per tile of 256 neurons),
// preceded by one k_gather launch when some instance gathers.  The step
// loop can be captured into HIP graphs (one per starting parity).
class Engine {
public:
    explicit Engine(int device);
    ~Engine();

    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }

    int add(std::unique_ptr<Instance> inst);
    std::unique_ptr<Instance> remove(int id);
    Instance *get(int id);
    size_t active() const;

    // Runs `steps` steps of every active instance.  graph_chunk > 1 captures
    // that many steps into one HIP graph and replays it.
    void run(int64_t steps, int graph_chunk);
    // Since the engine was made: synapses past a queue's capacity (spilled,
    // or dropped and taken by the pull; not an error), tiles that deferred
    // their rows, tiles that pulled.
    CtlTotals propagation_stats();
    uint64_t spilled() { return propagation_stats().spilled; }
    // Step-kernel schedule: synapses per lane in a binning pass (1, 2, 4) and
    // the propagation layout: queues with dense or spread counts
    // (kLayoutAuto: by row coverage), or gather (kLayoutGather: every
    // instance that allows it; the others use dense queues).  Every schedule
    // gives the same results.
    void set_schedule(uint32_t pass_k, int layout);
    // Whether instance p propagates by gather under the current layout.
    bool use_gather(const Instance &p) const;
    uint32_t pass_k() const { return pass_k_; }
    int count_layout() const { return count_layout_; }
    uint32_t count_shift() const { return cshift_; }
    // Times every schedule on about `trial_steps` real steps of all
    // instances, in turns of kAutotuneWindow steps (each preceded by 2 steps
    // that apply it, and one untimed turn first), keeps the fastest; us[7] =
    // mean microseconds per step, [dense, spread] x [K = 1, 2, 4], then
    // gather (K = 2; a candidate only when some instance allows it, else its
    // turns run the default schedule and us[6] = 0).  Steps run:
    // autotune_steps(trial_steps).
    static constexpr int64_t kAutotuneWindow = 4;
    static int64_t autotune_steps(int64_t trial_steps)
    {
        const int64_t rounds = (trial_steps + kAutotuneWindow - 1) / kAutotuneWindow;
        return 7 * (2 + rounds * (2 + kAutotuneWindow));
    }
    // Some instance allows gather propagation.
    bool gather_possible() const;
    // Some instance propagates by gather under the current layout.
    bool any_gather() const;
    void autotune(int64_t trial_steps, uint32_t *best_k, int *best_layout, float *us);
    // Pull propagation: on = every instance that allows it gets its
    // in-edges (when added, or at the next table build) and propagates
    // through them where queues overflow or tiles defer; defer_min = emitted
    // rows from which a tile defers (0: tiles never defer, overflowing
    // queues still pull).  Off: queues only, past-capacity synapses through
    // the spill accumulators.  Every setting gives the same results.
    void set_pull(bool on, uint32_t defer_min);
    // Split binning: the fired rows' binning passes run in a second launch
    // per step (k_bin) spread over every CU, instead of in the tiles whose
    // neurons fired.  Results never depend on it.
    // cap: entries of the pass list (0: sized by the instances); passes past
    // it are binned by their own tile.
    void set_split(bool on, uint32_t cap)
    {
        if (on != split_ || cap != split_cap_) {
            split_ = on;
            split_cap_ = cap;
            dirty_ = true;
        }
    }
    bool split() const { return split_; }
    uint32_t split_cap() const { return split_cap_; }
    bool pull_on() const { return pull_on_; }
    uint32_t defer_min() const { return defer_min_; }
    bool uses_pull(const Instance &p) const { return pull_on_ && p.pull_built() && !use_gather(p); }
    // Queue capacity per tile: min(synapses targeting it, floor +
    // (synapses >> shift)); shift 0 keeps the exact worst case.
    void set_queue_capacity(uint32_t floor, uint32_t shift);
    uint32_t queue_capacity(uint32_t synapses) const;
    // Queue slots per parity of instance p at the current capacity.
    uint64_t queue_slots(const Instance &p) const;
    void sync();
    void mark_dirty() { dirty_ = true; }
    // Stream for building instances: their set-up runs beside the steps of
    // the instances already in the tables.
    hipStream_t admit_stream() const { return admit_stream_; }
    // Instances added since the last table build do not step yet: changing
    // them needs no settle, and their uploads go on the admission stream.
    bool staged(int id) const { return staged_.count(id) != 0; }
    void settle_for(int id)
    {
        if (!staged(id))
            settle();
    }
    hipStream_t stream_for(int id) const { return staged(id) ? admit_stream_ : stream_; }
    // Writes the pending delta of every instance into its dlt buffer (the
    // queues are left for the next step), then waits.
    void materialize();
    // Drops the pending delta of one instance (checkpoint restore).
    void clear_queues(int id);
    // Records the I_syn-based statistics of the last step run (imean,
    // receptor imeans), which the next step's launch would otherwise take.
    void stat_tail();
    // Folds every pending delta into J with the current tables and
    // parameters.  Must run before parameters, samples, recording or the
    // instance set change (a table rebuild empties the queues).
    void settle();

    // Kernel-time measurement: events recorded around every step kernel
    // (total in ms over `steps` steps).
    void run_timed(int64_t steps, float *ms_step);
    // run(steps, graph_chunk) between two events on the engine stream; *ms =
    // the GPU time between them (the step kernels and the gaps between).
    void run_span(int64_t steps, int graph_chunk, float *ms);
    // run(steps, graph_chunk) with an event after every unit (a step run
    // eagerly or a replayed graph chunk); *ms = the GPU time from before the
    // first unit to after the last, ms_steps[steps] (optional) = each step's
    // share of its unit's time (a chunk's time divided by its steps).
    void run_span_units(int64_t steps, int graph_chunk, float *ms, float *ms_steps);

private:
    void rebuild();
    void launch_step();        // one step at parity par_, then par_ ^= 1
    void launch_step_k();
    void launch_aux();         // materialize + sample after a step, if samples are set
    void drop_graphs();
    QueueSet queues();

    int device_ = 0;
    hipStream_t stream_ = nullptr;
    std::vector<std::unique_ptr<Instance>> insts_;
    std::vector<uint32_t> tile_base_;      // per instance slot, first tile in the engine table
    bool dirty_ = true;
    hipStream_t admit_stream_ = nullptr;
    std::unordered_set<int> staged_;
    bool pending_ = false;                 // queues may hold an unfolded delta
    uint32_t par_ = 0;                     // parity of the next step launch (= pending queue set)
    uint32_t ninst_ = 0, nsamples_ = 0, ntiles_ = 0;
    uint32_t cshift_ = kCountShiftDense;   // queue count layout (count_slot)
    int count_layout_ = kLayoutAuto;       // requested layout
    uint32_t pass_k_ = kPassKDefault;      // synapses per lane in a binning pass
    uint32_t wfmt_ = kWeightF32;           // weight format of every instance in the tables, or kWeightAny
    uint32_t queue_floor_ = 4096, queue_shift_ = 3;   // queue_capacity
    bool tuning_ = false;                  // autotune running: gather buffers are kept
    bool pull_on_ = true;                  // pull propagation where in-edges fit
    bool split_ = false;                   // split binning (k_bin)
    uint32_t split_cap_ = 0;               // requested pass-list entries (0: automatic)
    DataAllocator<PassEntry> d_pass_;      // pass list of split binning
    DataAllocator<uint32_t> d_pcount_;     // its two counts, kPassCountStride words apart
    uint32_t pass_cap_ = 0;                // entries of d_pass_ in use (0: split binning off)
    uint32_t bin_grid_ = 0;                // k_bin workgroups
    uint32_t defer_min_ = kDeferMinDefault;
    void try_pull(Instance &p, hipStream_t s);   // builds p's in-edges if on and they fit
    DataAllocator<InstDev> d_insts_;
    DataAllocator<StepCtl> d_ctl_;
    DataAllocator<SampleDev> d_samples_;
    DataAllocator<SampleRecDev> d_recs_;
    DataAllocator<TileDesc> d_tiles_;
    DataAllocator<uint32_t> d_qcnt_;       // count_words(2 * ntiles)
    DataAllocator<GroupDesc> d_groups_;    // gather groups of the gather instances
    uint32_t ngroups_ = 0;
    void launch_gather_k(uint32_t par);    // k_gather of parity par, if any instance gathers
    hipGraphExec_t graph_[2] = {nullptr, nullptr};   // by starting parity
    int graph_steps_[2] = {0, 0};
};

}  // namespace dtb
