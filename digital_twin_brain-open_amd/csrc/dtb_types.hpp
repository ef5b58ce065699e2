// dtb_types.hpp -- device-visible data layout of a simulation instance.
//
// One instance is one reference "block" (BrainBlock<float,float2>,
// include/block.hpp:782-885).  Its per-neuron state and parameters are
// structure-of-arrays in HBM; parameters that the step reads together are
// packed into float4 records so that one 64-lane wave issues 16-byte loads
// over a contiguous 1 KiB.  Many instances can be stepped by one launch: the
// kernels find their instance through small descriptor tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtb {

constexpr int kBlock = 256;            // threads per workgroup (4 wave64)
constexpr int kWave = 64;
constexpr uint32_t kKindBit = 0x80000000u;   // channel-pair bit packed into the column index
constexpr uint32_t kColMask = 0x7fffffffu;
constexpr uint32_t kTileBits = 8;            // target tile = one step workgroup = 256 neurons
static_assert((1 << kTileBits) == kBlock, "a tile is one workgroup of neurons");

// Flag byte per neuron: bit 0 = membrane spike flag f (reference
// f_inner_actives_), bit 1 = spike emitted to the synapses in the last step
// (f_inner_actives_ or f_recorded_inner_actives_, dist_simulator.cpp:2118-2131).
constexpr uint8_t kFlagSpike = 1u;
constexpr uint8_t kFlagEmit = 2u;
// bit 2: J already holds the reference-visible J of the last step (set when a
// checkpoint is restored or pending input is settled); the next step uses it
// without the decay-and-delta fold.
constexpr uint8_t kFlagFolded = 4u;

// Per-population record (reference sub-block, sub_bcounts_ ranges and the
// OU_Background_Current_Param map, include/stage.hpp:14-19).
struct alignas(16) PopDev {
    uint32_t begin;     // first neuron (instance-local)
    uint32_t end;       // one past the last neuron
    float mu;           // OU mean
    float f1;           // 1 - exp(-dt/tau)
    float f2s;          // sqrt(1 - exp(-2 dt/tau)) * sigma
    uint32_t ou_on;     // OU configured for this population
    uint32_t pad0, pad1;
};

// One delivered synapse queued for its target tile: the target's index in
// the tile (bits 0..7), the channel pair (bit 8: 0 AMPA/NMDA, 1
// GABA_A/GABA_B) and the weight pair.
struct QSyn {
    uint32_t key;
    float w0, w1;
};
constexpr uint32_t kQSynKind = 1u << 8;
constexpr int64_t kMaxSynapses = (int64_t)1 << 40;
constexpr int64_t kMaxRowLength = (int64_t)1 << 31;

// Per-population statistics (stat_result.cpp:266-391): which sums a run
// records.  Each neuron's value enters as the exact integer rint(x * 2^24),
// so the sums do not depend on the order of the additions.
constexpr uint32_t kStatV = 1u;          // membrane potential (vmean)
constexpr uint32_t kStatI = 2u;          // synaptic current (imean)
constexpr uint32_t kStatReceptor = 4u;   // per-receptor currents (receptor imeans)
constexpr int kStatFields = 6;           // V, I_syn, AMPA, NMDA, GABA_A, GABA_B

// Largest number of distinct parameter records kept as classes (2-byte class
// index per neuron); beyond it the records are stored per neuron.
constexpr size_t kMaxParamClasses = 65536;

// Everything a kernel needs about one instance.
struct alignas(16) InstDev {
    // Parameter records, one float4 each, indexed by parameter class:
    //   pA = (v_reset, v_th, 1/C, v_leak)
    //   pB = (g_leak, i_ext, noise_rate, bits(refractory substeps))
    //   gU, vU, dec = (AMPA, NMDA, GABA_A, GABA_B)
    // Neurons with bit-identical records share a class (in practice one per
    // population), so the step reads 2 bytes of class per neuron instead of
    // 80 bytes of parameters.  The record of neuron i is
    // (i & pmask) + cls[i]: pmask = 0 with shared classes; when the records
    // have more than 65536 values every neuron has its own (pmask = ~0xff,
    // cls[i] = i & 0xff, record = i).
    const float4 *pA, *pB, *gU, *vU, *dec;
    const uint16_t *cls;        // [n] parameter class
    const uint8_t *blocked;     // excluded from population statistics
    // state
    float *V;
    int32_t *tcnt;              // refractory counter (reference t_actives_)
    uint8_t *flag;              // kFlagSpike | kFlagEmit | kFlagFolded
    float *iou;                 // OU current (reference i_ou_background_stimuli_)
    float4 *J;                  // J_u after the previous step's decay (delta below pending)
    float4 *dlt;                // materialised pending delta (sampling / read-back only)
    unsigned long long *spill;  // [2][n][4] Q23.40 sums of the synapses past a queue's capacity
    QSyn *queue;                // [2][qslots] propagation queues (+ read-ahead slack)
    uint64_t qslots;            // queue slots per parity (tile t's start: TileDesc::qoff)
    // connectivity: presynaptic-major CSR, 64-bit row offsets
    const int64_t *rowptr;      // [n+1]
    const uint32_t *col;        // [S] postsynaptic id | kKindBit
    const void *w;              // [S] weight pair in format wfmt (kWeightF32 / F16 / U8)
    // Gather propagation (null: queue propagation).  segidx[r][g] = offset in
    // row r of its first synapse whose target lies in group g or later
    // (segidx[r][ngroups] = row length).  The rows with synapses fired by the
    // launch of parity p ^ 1 are in the kFiredLists lists of set p: list s
    // takes the rows of tiles t with t % kFiredLists == s, so it never holds
    // more than fcap rows; its count is word (p * kFiredLists + s) *
    // kFiredCountStride of fcount (one cache line per count).
    const uint16_t *segidx;     // [n][ngroups + 1]
    // The synapses again as one record each (column, then the weight pair's
    // raw bits: grec_words(wfmt) 32-bit words), so a gathered segment is one
    // contiguous stream instead of two.
    const uint32_t *grec;       // [S][grec_words(wfmt)]
    uint32_t *fired;            // [2][kFiredLists][fcap]
    uint32_t *fcount;           // [2][kFiredLists][kFiredCountStride]
    // Pull propagation (null crec: off; never together with gather).  The
    // in-edges of target tile t are records [cptr[t], cptr[t + 1]) of crec,
    // grec_words(wfmt) words each: key = source row (bits 0..22) | target's
    // index in the tile (bits 23..30) | channel pair (bit 31), then the
    // weight pair's raw bits, so one load brings both.  Per wave of 64
    // neurons, ebits[s] holds the rows emitted by the launch that filled
    // queue set s and dbits[s] those of them whose tile deferred its
    // propagation to the pull; dstamp[s] = 1 + the step index of the launch
    // that consumes set s when some tile deferred into it (0: none; a stale
    // stamp never equals a later consumer's).  A consuming tile pulls
    // when its queue overflowed (then it ignores the queue and takes every
    // emitted row through its in-edges) or when rows were deferred (then it
    // adds the deferred rows to its queue).
    const uint32_t *crec;       // [S][grec_words(wfmt)]
    const int64_t *cptr;        // [ntiles + 1]
    unsigned long long *ebits;  // [2][ceil(n / 64)]
    unsigned long long *dbits;  // [2][ceil(n / 64)]
    int64_t *dstamp;            // [2]
    const PopDev *pops;         // [npop]
    // recording (device buffers, row = step - rec_step0)
    uint32_t *rec_counts;       // [rec_len][npop] spikes per population
    long long *rec_stat;        // [rec_len][npop][6] sums of V, I_syn, 4 receptor currents (x 2^24)
    // counters
    // per wave of 64 neurons: (synaptic events delivered = sum of fired
    // out-degrees, spikes emitted); each slot is updated only by the wave
    // that owns those neurons, so no atomics are needed
    unsigned long long *wave_ctr;   // [ceil(n/64)][2]
    int64_t *step;              // [2] step index by launch parity: the launch of
                                // parity p reads step[p] and writes step[p^1] = step[p] + 1
    // scalars
    uint64_t seed;
    int64_t rec_step0;
    uint32_t n;
    uint32_t npop;
    uint32_t rec_len;
    int32_t t_steps;
    int32_t use_ou;
    float dt_sub;
    uint32_t tile_base;         // engine-wide index of this instance's first tile
    uint32_t stat_mask;         // kStatV | kStatI | kStatReceptor: statistics recorded in rec_stat
    uint32_t pmask;             // parameter record of neuron i: (i & pmask) + cls[i]
    uint32_t ngroups;           // gather groups of the instance (kGroupTiles tiles each)
    uint32_t fcap;              // rows per fired list
    uint32_t defer_min;         // pull: a tile with this many emitted rows defers them (0: never)
    uint32_t wfmt;              // weight pair format of w and cw
    uint32_t pad2[3];
};

// Weight pair formats kept on the device (include/weights.hpp:14-21 and the
// decode of presynaptic_voltage.cpp:153-176): two fp32 (8 bytes), two fp16
// (4 bytes, converted exactly) or two uint8 codes u (2 bytes, weight
// (float)u / 255.f, correctly rounded).  Double weights are stored as fp32,
// the type the reference's float block converts them to.
enum : uint32_t { kWeightF32 = 0, kWeightF16 = 1, kWeightU8 = 2 };
// Kernel specialisation for launches whose instances hold different formats
// (each weight pair is then decoded by its instance's format at run time).
constexpr uint32_t kWeightAny = 3;
__host__ __device__ inline uint32_t weight_bytes(uint32_t fmt)
{
    return fmt == kWeightF32 ? 8u : fmt == kWeightF16 ? 4u : 2u;
}

// 32-bit words of a gather record: the column and the weight pair (fp32
// pairs 2 words; fp16 pairs and uint8 codes 1 word).
__host__ __device__ inline uint32_t grec_words(uint32_t fmt)
{
    return fmt == kWeightF32 ? 3u : 2u;
}

// Pull propagation limits: source rows are 23-bit fields of the in-edge key.
constexpr uint32_t kPullRowBits = 23;
constexpr uint32_t kPullRowMask = (1u << kPullRowBits) - 1u;
constexpr uint32_t kPullMaxNeurons = 1u << kPullRowBits;
// Default deferral threshold: a tile whose step emitted at least this many
// rows leaves them to the consumers' pull instead of binning them.
constexpr uint32_t kDeferMinDefault = 16;

// One tile = one workgroup of the step kernel: up to 256 consecutive neurons
// of one instance, and the queue of the synapses delivered to them.
struct alignas(16) TileDesc {
    uint32_t inst;
    uint32_t begin;     // first instance-local neuron
    uint32_t count;     // neurons in the tile (<= 256)
    uint32_t pop0;      // population of the first neuron
    uint64_t qoff;      // first slot in the instance's queue of one parity
    uint32_t qcap;      // queue capacity: synapses targeting this tile (all rows)
    uint32_t pcls;      // 1 + the parameter class shared by every neuron of the tile, or 0
};

// Gather propagation: one workgroup of kGatherBlock threads gathers the
// synapses of every fired row that target one group of kGroupTiles
// consecutive tiles (2^kGroupBits neurons) into LDS, then writes their delta.
constexpr uint32_t kGroupTiles = 8;
constexpr uint32_t kGroupBits = kTileBits + 3;
static_assert((1u << (kGroupBits - kTileBits)) == kGroupTiles, "a group is kGroupTiles tiles");
constexpr int kGatherBlock = 1024;
// Synapses each lane of k_gather loads before it adds them (in flight together).
constexpr uint32_t kGatherItems = 8;
// Longest row the 16-bit group offsets can index.
constexpr int64_t kGatherMaxRow = 65535;
// Fired lists per set (appends spread over this many counters, one wave of
// k_gather reads one count per lane) and the words between their counts.
constexpr uint32_t kFiredLists = 64;
constexpr uint32_t kFiredCountStride = 32;
static_assert(kFiredLists == 64, "k_gather reads one list count per lane");

struct alignas(16) GroupDesc {
    uint32_t inst;
    uint32_t g;         // group index within the instance (first neuron g << kGroupBits)
    uint32_t count;     // neurons in the group
    uint32_t pad;
};

// Queues of both parities: the launch of parity p drains set p and fills set
// p ^ 1.
// Queue entries per tile that the step kernel reads ahead together with the
// count; the queue buffer has this many slots of slack after both sets.
constexpr uint32_t kQueuePrefetch = 128;

// Word of queue count i (i = set * ntiles + tile): groups of 2^shift
// consecutive counts share one 128-byte line (32 words).  shift = 5 packs the
// counts densely: the reservations of one wave, which mostly hit neighbouring
// tiles, touch few lines.  shift = 0 gives every count a line of its own:
// when every fired row reaches most tiles, many waves reserve in the same
// counts at once, and counts sharing a line would serialise on it.
constexpr uint32_t kCountLineWords = 32;
// Synapses per lane in a binning pass of the step kernel: 1, 2 or 4 (passes
// of 64, 128 or 256 synapses).  Every value gives the same results.
constexpr uint32_t kPassKDefault = 2;

// Row coverage (tile segments per row / tiles) above which the automatic
// layout spreads the counts.
constexpr double kSpreadCoverage = 0.95;

// Requested layout of an engine's queue counts.
// kLayoutGather: instances that allow it propagate by gather (the others
// keep dense queues); it is chosen explicitly or by the autotuner.
enum : int { kLayoutAuto = 0, kLayoutDense = 1, kLayoutSpread = 2, kLayoutGather = 3 };
constexpr uint32_t kCountShiftDense = 5;
constexpr uint32_t kCountShiftSpread = 0;

__host__ __device__ inline size_t count_slot(size_t i, uint32_t shift)
{
    return ((i >> shift) * kCountLineWords) + (i & ((1u << shift) - 1u));
}

__host__ __device__ inline size_t count_words(size_t n, uint32_t shift)
{
    return ((n + (1u << shift) - 1u) >> shift) * kCountLineWords;
}

// One binning pass of a fired row handed from k_step to k_bin (split
// binning): synapses [first, first + len) of instance `inst`'s CSR, len at
// most one pass.
struct PassEntry {
    uint64_t first;
    uint32_t len;
    uint32_t inst;
};
// Words between the two pass-list counts (each on its own 128-byte line).
constexpr uint32_t kPassCountStride = 32;
// Most workgroups of a k_bin launch (8 per CU on 256 CUs: all resident).
constexpr uint32_t kBinGridMax = 2048;

struct QueueSet {
    uint32_t *count;    // [2][ntiles], tile t of set p at word count_slot(p * ntiles + t, cshift)
    uint32_t ntiles;
    uint32_t cshift;    // kCountShiftDense or kCountShiftSpread
    // Split binning (pass == null: off): k_step of parity p appends its
    // tiles' passes to `pass` (count at pcount[p * kPassCountStride]) and the
    // k_bin launch after it bins them; passes past pcap are binned by their
    // own tile.
    PassEntry *pass;
    uint32_t *pcount;
    uint32_t pcap;
    uint32_t pad;
};

// Sample request (reference stat_samples_kernel, stat_result.cpp:393-433).
struct alignas(16) SampleDev {
    uint32_t inst;
    uint32_t neuron;
    uint32_t slot;      // column in the instance's sample record
    uint32_t pad;
};

// Per-instance sample records, row = step - rec_step0.
struct alignas(16) SampleRecDev {
    uint8_t *spike;
    float *v;
    float *isyn;
    float *iou;
    uint32_t nsample;
    uint32_t pad[3];
};

// Size of the static device array that holds the per-step records of one
// dtb_run_noalloc call (one copy per GPU, kept until the process ends).
constexpr size_t kScratchBytes = (size_t)64 << 20;

// Shared per-engine diagnostics in device memory.
// Diagnostic counters of the step launches.  Each is spread over kCtlSlots
// cache lines, a workgroup adding into slot blockIdx.x % kCtlSlots: atomics
// from every tile of a launch on one address would be serialised by the
// memory and hold the launch's end (a volley step defers or pulls in all
// 7,813 tiles of the 2M-neuron block).
constexpr uint32_t kCtlSlots = 64;
struct alignas(128) CtlSlot {
    unsigned long long spilled;   // synapses past a queue's capacity (spilled, or taken by the pull)
    unsigned long long deferred;  // tiles that left their emitted rows to the pull
    unsigned long long pulled;    // tiles that took input through their in-edges (step launches)
    unsigned long long pad[13];
};
struct StepCtl {
    CtlSlot slot[kCtlSlots];
};
// The counters summed over their slots.
struct CtlTotals {
    unsigned long long spilled, deferred, pulled;
};

}  // namespace dtb
