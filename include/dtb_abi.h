/*
 * dtb_abi.h -- C ABI of the MI355X (gfx950) DTB single-block step library
 * (libdtb.so).
 *
 * Plain C types only: opaque handles, pointers and sizes.  Every function
 * returns an int status: 0..5 are the reference's SnnStatus values
 * (protos/snn.proto:44-51), DTB_HIP_ERROR and above are runtime failures.
 * No exception crosses the ABI; dtb_last_error() returns the message of the
 * calling thread's last failure.  One driver thread per engine handle.
 *
 * Citations name the reference interface each entry point replaces, relative
 * to the reference repository (DTB-consortium/Digital_twin_brain-open).
 */
#ifndef DTB_ABI_H
#define DTB_ABI_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTB_ABI_VERSION 1

/* protos/snn.proto:44-51 SnnStatus, then runtime failures. */
enum dtb_status {
    DTB_OK = 0,
    DTB_INVALID_PARAMETER = 1,
    DTB_UNSET_SAMPLE_FAILURE = 2,
    DTB_UNSUPPORTED = 3,
    DTB_NOT_ENABLE_METRIC = 4,
    DTB_INVALID_NETWORK = 5,
    DTB_HIP_ERROR = 100,
    DTB_OUT_OF_MEMORY = 101,
    DTB_NO_DEVICE = 102,
    DTB_INTERNAL = 103
};

/* Run flags (protos/snn.proto:63-77 RunRequest fields). */
enum dtb_run_flags {
    DTB_RUN_USE_OU = 1u << 0,          /* use_ou_background */
    DTB_RUN_FREQ = 1u << 1,            /* output_freq (u32 counts per population) */
    DTB_RUN_SAMPLE_SPIKE = 1u << 2,    /* output_sample_spike */
    DTB_RUN_SAMPLE_VMEMB = 1u << 3,    /* output_sample_vmemb */
    DTB_RUN_SAMPLE_ISYN = 1u << 4,     /* output_sample_isynaptic */
    DTB_RUN_SAMPLE_IOU = 1u << 5,      /* output_sample_iou */
    DTB_RUN_VMEAN = 1u << 6,           /* output_vmean */
    DTB_RUN_IMEAN = 1u << 7,           /* output_imean */
    DTB_RUN_GRAPH = 1u << 8,           /* replay the step loop from a HIP graph */
    DTB_RUN_RECEPTOR_IMEAN = 1u << 9   /* output_receptor_imeas */
};

/* PropType, include/stage.hpp:320-341 == protos/snn.proto:94-115. */
enum dtb_prop_type {
    DTB_EXT_STIMULI_I = 0, DTB_MEMBRANE_C = 1, DTB_REF_T = 2, DTB_LEAKAGE_G = 3,
    DTB_LEAKAGE_V = 4, DTB_THRESHOLD_V = 5, DTB_RESET_V = 6,
    DTB_CONDUCT_G_AMPA = 7, DTB_CONDUCT_G_NMDA = 8, DTB_CONDUCT_G_GABAa = 9, DTB_CONDUCT_G_GABAb = 10,
    DTB_MEMBRANE_V_AMPA = 11, DTB_MEMBRANE_V_NMDA = 12, DTB_MEMBRANE_V_GABAa = 13, DTB_MEMBRANE_V_GABAb = 14,
    DTB_TAO_AMPA = 15, DTB_TAO_NMDA = 16, DTB_TAO_GABAa = 17, DTB_TAO_GABAb = 18,
    DTB_NOISE_RATE = 19
};

typedef struct dtb_engine_s *dtb_engine_t;

/* Weight pair types kept on the device (include/weights.hpp:14-21 DataType
 * FLOAT, FLOAT16, INT8; decoded per synapse as
 * src/stage/presynaptic_voltage.cpp:153-176 does: fp16 converted exactly,
 * uint8 codes u as (float)u / 255.f).  DOUBLE blocks are given as float32,
 * the type the reference's float block converts them to. */
enum dtb_weight_type { DTB_WEIGHTS_F32 = 0, DTB_WEIGHTS_F16 = 1, DTB_WEIGHTS_U8 = 2 };

/* A network in the reference block layout (src/configuration.cpp:1079-1667,
 * 1778-2156): property matrix [n][22] float32 (columns as python/
 * make_block.py:441-465), presynaptic-major CSR with 64-bit row offsets,
 * postsynaptic ids, a weight pair per synapse (include/weights.hpp:30-37) of
 * type weight_type and the channel-pair bit per synapse packed in 32-bit
 * words (bit j of word j/32; src/stage/presynaptic_voltage.cpp:152). */
typedef struct dtb_network {
    int64_t n;
    const float *props;         /* [n][22] */
    int64_t nsyn;
    const int64_t *rowptr;      /* [n+1] */
    const uint32_t *colinds;    /* [nsyn] */
    const void *weights;        /* [nsyn][2] float32 / float16 / uint8 */
    const uint32_t *connkinds;  /* [ceil(nsyn/32)] */
    int32_t weight_type;        /* dtb_weight_type (0: float32) */
} dtb_network;

/* Per-step outputs of dtb_run, caller-allocated, row-major [iters][...];
 * NULL disables an output (RunResponse fields, protos/snn.proto:79-92). */
typedef struct dtb_run_outputs {
    uint32_t *freq;             /* [iters][npop] spikes per population */
    uint8_t *sample_spike;      /* [iters][nsample] */
    float *sample_vmemb;        /* [iters][nsample] */
    float *sample_isynaptic;    /* [iters][nsample] */
    float *sample_iou;          /* [iters][nsample] */
    /* Per-population means over the non-blocked neurons at the end of each
     * step (stat_result.cpp:266-391): V, I_syn, and the receptor currents
     * g_u (V_u - V) J_u; [iters][npop] each.  Each neuron's value enters the
     * sum as a multiple of 2^-24 (exact, order-free sums). */
    float *vmean;
    float *imean;
    float *ampa_imean;
    float *nmda_imean;
    float *gabaa_imean;
    float *gabab_imean;
} dtb_run_outputs;

/* ---- library ------------------------------------------------------------ */

/* include/version.hpp:1-6; tools/dist_simulator.cpp:3916-3922 (GetVersion). */
int dtb_version(void);
/* Message of the calling thread's last failure (empty if none). */
const char *dtb_last_error(void);
/* Number of visible HIP devices (0 without a GPU). */
int dtb_device_count(int *count);
/* No counterpart in the reference (its blocks live for the whole run).
 * Device memory of removed instances and resized buffers stays with the
 * library and is reused (every block is carved best-fit from larger
 * allocations); a server reserves its budget once with
 * dtb_device_memory_reserve so that admitting and retiring instances makes
 * no HIP allocation.  dtb_device_memory_cached reports the unused bytes held
 * on `device`; dtb_device_memory_trim returns them to the HIP runtime. */
int dtb_device_memory_reserve(int device, uint64_t bytes);
int dtb_device_memory_cached(int device, uint64_t *bytes);
int dtb_device_memory_trim(void);

/* ---- engine: one GPU, one stream, any number of instances ---------------- */

/* Replaces the per-GPU worker state of tools/dist_simulator.cpp:513-606
 * (NodeInfo) and the BrainBlock constructor, src/block.cpp:19-34. */
int dtb_engine_create(int device, dtb_engine_t *out);
int dtb_engine_destroy(dtb_engine_t h);

/* Adds a block: BrainBlock::init_connection_table_gpu + init_config_params_gpu
 * + init_all_stages_gpu (src/block.cpp:36-179).  dt in ms; seed keys the
 * Philox streams (the reference seeds from /dev/urandom, src/block.cpp:129-134). */
int dtb_instance_add(dtb_engine_t h, const dtb_network *net, float dt, uint64_t seed, int *inst_id);
/* Adds a block whose network is generated in device memory (the fan-out
 * network of dtb_gen_fanout with the same arguments, bit for bit): no host
 * copy of the synapses.  Used where networks are admitted faster than a host
 * can build and upload them (serving).  k <= 8192. */
int dtb_instance_add_fanout(dtb_engine_t h, int64_t n, uint32_t k, uint64_t net_seed, const float *g_ui,
                            uint32_t sub_base, float dt, uint64_t seed, int *inst_id);
int dtb_instance_remove(dtb_engine_t h, int inst);
/* Neuron and population counts; pop_ids [npop] (brain ids, column 3) and
 * pop_begin [npop+1] may be NULL (sub_bids_/sub_bcounts_, include/block.hpp). */
int dtb_instance_info(dtb_engine_t h, int inst, int64_t *n, int64_t *nsyn, int32_t *npop, uint32_t *pop_ids,
                      uint32_t *pop_begin);

/* BrainBlock::set_I_ou_current_param, src/block.cpp:222-281
 * (Snn.Updateoubackgroundparam, protos/snn.proto:161-167). */
int dtb_set_ou(dtb_engine_t h, int inst, const uint32_t *brain_ids, const float *mean, const float *deviation,
               const float *correlation_time, int32_t n);

/* Snn.Updatesample, tools/dist_simulator.cpp:2703-2739. */
int dtb_set_samples(dtb_engine_t h, int inst, const uint32_t *ids, int32_t n);

/* BrainBlock::update_Props_gpu, src/block.cpp:1281-1331 ->
 * src/stage/property.cpp:57-183 (Snn.Updateprop). */
int dtb_update_props(dtb_engine_t h, int inst, const uint32_t *neuron_ids, const uint32_t *prop_ids,
                     const float *vals, int32_t n);

/* BrainBlock::update_Prop_Cols_gpu (assign = 0, multiply) and
 * assign_Prop_Cols_gpu (assign = 1), src/block.cpp:1333-1470 ->
 * src/stage/property.cpp:185-466 (Snn.Updatehyperpara). */
int dtb_update_prop_cols(dtb_engine_t h, int inst, const uint32_t *prop_ids, const uint32_t *brain_ids,
                         const float *vals, int32_t n, int32_t assign);

/* BrainBlock::update_gamma_Prop_Cols_gpu, src/block.cpp:1472-1536 ->
 * src/stage/property.cpp:470-640 (Snn.Updategamma): every neuron of
 * population brain_ids[k] gets its own Gamma(concentration[k], rate[k])
 * draw as property prop_ids[k] (dtb_gen_gamma's stream, keyed by the
 * instance seed, one stream per (call, k)). */
int dtb_gamma_prop_cols(dtb_engine_t h, int inst, const uint32_t *prop_ids, const uint32_t *brain_ids,
                        const float *concentration, const float *rate, int32_t n);
/* No counterpart in the reference (its properties live on the device): the
 * instance's property matrix [n][22] with every update applied. */
int dtb_read_props(dtb_engine_t h, int inst, float *props);

/* The Run loop of tools/dist_simulator.cpp:1820-2332 (snn_run) for one
 * instance: iters steps of t_steps substeps, per-step outputs into `out`.
 * Every other instance of the engine advances too. */
int dtb_run(dtb_engine_t h, int inst, int32_t iters, int32_t t_steps, uint32_t flags, const dtb_run_outputs *out);

/* Streamed Run (tools/dist_simulator.cpp:3343-3455 writes one RunResponse
 * per step while the loop runs, fed by the report thread :4049-4086): the
 * run of dtb_run in chunks of `chunk` steps; after each chunk `fn` receives
 * the chunk's rows (first_iter = index of its first step within the run,
 * rows = its steps; the outputs selected by `flags`, row-major
 * [rows][...], in library-owned host memory valid during the call).  Device
 * and host memory hold one chunk, whatever iters is.  fn returns 0 to go on,
 * anything else to stop after that chunk (the run then ends with DTB_OK).
 * Results are bit-identical to one dtb_run of the same steps. */
typedef int (*dtb_run_chunk_fn)(void *user, int32_t first_iter, int32_t rows, const dtb_run_outputs *out);
int dtb_run_stream(dtb_engine_t h, int inst, int32_t iters, int32_t t_steps, uint32_t flags, int32_t chunk,
                   dtb_run_chunk_fn fn, void *user);

/* dtb_run without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The per-step
 * records live in a static device array of DTB_NOALLOC_SCRATCH_BYTES per GPU
 * (held until the process ends) instead of buffers allocated per call; the
 * outputs are copied into the caller's host arrays as dtb_run does, with
 * bit-identical values.  The steps are launched one by one (DTB_RUN_GRAPH is
 * ignored: instantiating a graph allocates), with the instance's configured
 * t_steps (another value is rejected).  Fails with DTB_INVALID_PARAMETER,
 * allocating nothing, when the records (counts 4 B x npop, statistic sums
 * 48 B x npop plus 4 B per requested mean x npop, samples 13 B x nsample,
 * per step, each array rounded up to 256 B) exceed the array, or when the
 * engine's tables would need memory first (after instances, samples or the
 * schedule change: one dtb_run or dtb_engine_run builds them).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (from
 * several threads or engines) race. */
#define DTB_NOALLOC_SCRATCH_BYTES (64u << 20)
int dtb_run_noalloc(dtb_engine_t h, int inst, int32_t iters, int32_t t_steps, uint32_t flags,
                    const dtb_run_outputs *out);

/* Batched stepping of every instance of the engine with its last run
 * configuration (t_steps, OU), no per-step outputs; graph_chunk > 1 replays
 * chunks of that many steps from a captured HIP graph. */
int dtb_engine_run(dtb_engine_t h, int64_t steps, int32_t graph_chunk);
/* Per-instance run configuration for dtb_engine_run. */
int dtb_configure(dtb_engine_t h, int inst, int32_t t_steps, int32_t use_ou);
/* Waits for all work of the engine. */
int dtb_engine_sync(dtb_engine_t h);
/* No counterpart in the reference: dtb_engine_run between two HIP events on
 * the engine stream; *ms = the GPU time between them (step kernels and the
 * gaps between them), the engine synchronised on return. */
int dtb_engine_run_span(dtb_engine_t h, int64_t steps, int32_t graph_chunk, float *ms);
/* No counterpart in the reference: dtb_engine_run_span with an event after
 * every unit of work (a step launched eagerly, or a replayed graph chunk);
 * ms_steps[steps] (may be NULL) = each step's share of its unit's GPU time
 * (per step when eager, the chunk's mean when replayed). */
int dtb_engine_run_span_steps(dtb_engine_t h, int64_t steps, int32_t graph_chunk, float *ms, float *ms_steps);
/* Times `steps` steps with HIP events around each step kernel (delta
 * accumulation + neuron update + spike binning); total in ms. */
int dtb_engine_time_kernels(dtb_engine_t h, int64_t steps, float *ms_step);
/* Waits for the engine and reports any runtime failure of its work. */
int dtb_engine_check(dtb_engine_t h);
/* No counterpart in the reference (it scatters into per-neuron buffers).
 * Spikes travel through per-tile queues; a tile's queue holds
 * min(S_t, floor + (S_t >> shift)) entries, S_t = synapses targeting the tile
 * (default floor 4096, shift 3; shift 0 = always room for every synapse).
 * Synapses past the capacity are added into a per-neuron spill accumulator
 * with the same exact integer arithmetic, so results do not depend on the
 * capacity; dtb_engine_spilled counts them.  Takes effect at the next run. */
int dtb_engine_queue_capacity(dtb_engine_t h, uint32_t floor, uint32_t shift);
/* No counterpart in the reference (its scatter costs the same per synapse at
 * any firing rate, presynaptic_voltage.cpp:137-188).  Pull propagation: with
 * on != 0 every instance of at most 2^23 neurons gets a copy of its synapses
 * grouped by target tile (its in-edges, 12 bytes per synapse; built on the
 * device when the instance is added or at the next run; skipped when device
 * memory is short).  A tile whose queue overflowed then takes every emitted
 * row through its in-edges instead of spill atomics, and a tile with at
 * least defer_min emitted rows in a step (a synchronous volley) leaves them
 * to its targets' pull instead of binning them (defer_min 0: never; a tile
 * that itself pulls in that step defers from 4 * defer_min rows, so the
 * volley's tail is binned).  Default on, defer_min 16.  Every setting gives
 * bit-identical results. */
int dtb_engine_pull(dtb_engine_t h, int32_t on, uint32_t defer_min);
/* Split binning (the binning stage of src/stage/presynaptic_voltage.cpp:
 * 122-225, scheduled differently): on = the fired rows' binning passes are
 * listed by the step kernel and binned by a second launch per step over the
 * whole GPU, instead of by the tiles whose neurons fired.  capacity = entries
 * of the pass list (0: sized by the instances); passes past it are binned by
 * their own tile.  Results never depend on either. */
int dtb_engine_split(dtb_engine_t h, int32_t on, uint32_t capacity);
int dtb_engine_get_split(dtb_engine_t h, int32_t *on, uint32_t *capacity);
int dtb_engine_get_pull(dtb_engine_t h, int32_t *on, uint32_t *defer_min);
/* Pull propagation of one instance: allowed (at most 2^23 neurons, has
 * synapses), in-edges built, used under the current settings (not while the
 * instance gathers), device bytes of the in-edges (any pointer may be NULL). */
int dtb_instance_pull_info(dtb_engine_t h, int inst, int32_t *allowed, int32_t *built, int32_t *in_use,
                           uint64_t *bytes);

/* No counterpart in the reference (it has one schedule).  The step kernel's
 * schedule: pass_k synapses per lane in a binning pass (1, 2 or 4: passes of
 * 64, 128 or 256 synapses dealt to a tile's waves) and the propagation
 * layout: queues whose counts are DENSE (32 per cache line) or SPREAD (one
 * per line; DTB_LAYOUT_AUTO picks by row coverage), or GATHER (every
 * instance that allows it gathers, per group of 2048 target neurons, the
 * synapses of the last step's fired rows through per-row group offsets; the
 * others use dense queues).  Every schedule gives bit-identical results.
 * layout_in_use: GATHER when some instance gathers, else DENSE or SPREAD. */
enum { DTB_LAYOUT_AUTO = 0, DTB_LAYOUT_DENSE = 1, DTB_LAYOUT_SPREAD = 2, DTB_LAYOUT_GATHER = 3 };
int dtb_engine_schedule(dtb_engine_t h, int32_t pass_k, int32_t layout);
int dtb_engine_get_schedule(dtb_engine_t h, int32_t *pass_k, int32_t *layout, int32_t *layout_in_use);
/* Autotuner: every schedule (2 dense/spread x 3 pass sizes, then gather with
 * pass size 2) runs about trial_steps timed real steps of all instances, the
 * schedules taking turns in windows of 4 steps (each turn preceded by 2 steps
 * that apply it, and one untimed turn of each first), timed with HIP events;
 * the fastest total is kept.  Gather is a candidate only if some instance
 * allows it; otherwise its turns run the default schedule, untimed for the
 * choice.  Steps run in all: 7 (2 + r (2 + 4)), r = ceil(trial_steps / 4).
 * us_per_step[7] (optional) = mean per schedule, [dense K=1, 2, 4, spread
 * K=1, 2, 4, gather] (gather 0 when not a candidate).  The steps advance the
 * simulation exactly as any other steps would; tune in the regime that will
 * be run (after the initial transient). */
int dtb_engine_autotune(dtb_engine_t h, int64_t trial_steps, int32_t *pass_k, int32_t *layout, float *us_per_step);
/* No counterpart in the reference.  Gather propagation of one instance:
 * whether it allows it (rows ordered by target group, rows of at most 65535
 * synapses, group offsets no larger than the synapse arrays), whether the
 * current layout makes it gather, and the mean fraction of the 2048-neuron
 * groups that one row reaches (any pointer may be null). */
int dtb_instance_propagation(dtb_engine_t h, int inst, int32_t *gather_allowed, int32_t *gather_in_use,
                             float *group_coverage);
int dtb_engine_spilled(dtb_engine_t h, uint64_t *synapses);
/* Since the engine was created: synapses past a queue's capacity (spilled,
 * or taken by the pull), tile steps that deferred their rows to the pull,
 * tile steps that pulled (any pointer may be NULL). */
int dtb_engine_propagation_stats(dtb_engine_t h, uint64_t *spilled, uint64_t *deferred_tiles, uint64_t *pulled_tiles);
/* Synapses targeting each tile of 256 neurons of an instance (the queues'
 * worst case), out[ceil(n / 256)]. */
int dtb_instance_tile_synapses(dtb_engine_t h, int inst, uint32_t *out);

/* Checkpoint: reference-visible state after the last step (v_membranes_,
 * t_actives_, f_inner_actives_, emitted flags, i_ou_background_stimuli_,
 * J (AMPA, NMDA, GABA_A, GABA_B) [n][4], i_synaptics_ without i_ou).
 * Any pointer may be NULL. */
int dtb_read_state(dtb_engine_t h, int inst, float *v, int32_t *t_ref_count, uint8_t *f, uint8_t *emitted,
                   float *i_ou, float *j, float *i_syn);
int dtb_write_state(dtb_engine_t h, int inst, const float *v, const int32_t *t_ref_count, const uint8_t *f,
                    const float *i_ou, const float *j);
/* Sets the step index (the Philox counter of the next step); with
 * dtb_write_state this restores a checkpoint bit-exactly. */
int dtb_set_step(dtb_engine_t h, int inst, int64_t step);
/* Steps completed, synaptic events delivered (sum of fired out-degrees) and
 * spikes emitted since the instance was added. */
int dtb_counters(dtb_engine_t h, int inst, int64_t *step, uint64_t *events, uint64_t *spikes);

/* ---- verification ------------------------------------------------------- */

/* OU-stream normals of neurons [neuron0, neuron0+count) at one substep,
 * computed on the device. */
int dtb_ou_normals(int device, uint64_t seed, uint32_t neuron0, uint64_t substep, uint32_t count, float *out);

/* ---- synthetic networks (host) ----------------------------------------- */

/* Outputs sized by the caller: props [n][22], rowptr [n+1], colinds [n K],
 * weights [n K][2], connkinds [ceil(n K/32)].  g_ui = 4 conductances
 * (columns 10-13). */
int dtb_gen_uniform(int64_t n, uint32_t k, uint64_t seed, const float *g_ui, uint32_t sub_base, float *props,
                    int64_t *rowptr, uint32_t *colinds, float *weights, uint32_t *connkinds);
/* Fan-out network: every presynaptic neuron has exactly k synapses to
 * uniformly drawn targets (no self-connections; mean in-degree k), 80/20 E/I
 * by presynaptic index, rows ordered by target, weights k 2^-24 from
 * Philox4x32-10 keyed by seed.  Same output layout as dtb_gen_uniform. */
int dtb_gen_fanout(int64_t n, uint32_t k, uint64_t seed, const float *g_ui, uint32_t sub_base, float *props,
                   int64_t *rowptr, uint32_t *colinds, float *weights, uint32_t *connkinds);
int dtb_gen_voxel(int64_t n, uint32_t nvox, uint32_t k, float intra_frac, float density, uint64_t seed,
                  const float *g_ui, float *props, int64_t *rowptr, uint32_t *colinds, float *weights,
                  uint32_t *connkinds, uint32_t *vox_sizes);
/* Gamma(alpha, rate beta) draws of neurons first .. first + count - 1 in
 * stream `call` (Marsaglia-Tsang with Philox4x32-10 deviates; the sampler of
 * dtb_gamma_prop_cols, on the host). */
int dtb_gen_gamma(uint64_t seed, uint32_t call, uint32_t first, uint32_t count, float alpha, float beta, float *out);

/* ---- reference block files (host) --------------------------------------- */

/* The reference's block files (python/make_block.py:608-614, np.savez with
 * property / output_neuron_idx / input_block_idx / input_neuron_idx /
 * input_channel_offset / weight; read by src/configuration.cpp:1079-1667,
 * 1778-2156 through BrainBlock::init_connection_table_gpu and
 * init_config_params_gpu, src/block.cpp:36-127).  `path` is one block_<i>.npz
 * or a directory of block_0.npz .. block_<B-1>.npz, whose blocks are
 * concatenated in block order into one network (connections between the
 * blocks become ordinary synapses).  Stored and deflated archives, zip64.
 * First call: totals.  Second call: fills caller-allocated arrays in the
 * dtb_network layout (rows sorted by postsynaptic id, weights converted to
 * fp32 as the reference's propagation does: float64/float16 cast, uint8 / 255);
 * block_begin [nblocks+1] may be NULL. */
int dtb_npz_size(const char *path, int64_t *n, int64_t *nsyn, int32_t *nblocks);
int dtb_npz_read(const char *path, float *props, int64_t *rowptr, uint32_t *colinds, float *weights,
                 uint32_t *connkinds, int64_t *block_begin);
/* The weight type a block path keeps on the device: uint8 for uint8/int8
 * weight files, float16 for half, float32 for single and double
 * (src/configuration.cpp:1411-1435 picks the type by the block path; the
 * reader takes it from the stored dtype, which all files must share for
 * uint8 or float16). */
int dtb_npz_weight_type(const char *path, int32_t *weight_type);
/* dtb_npz_read with the weights written as weight_type ([nsyn][2] of
 * float32, float16 or uint8): float16 and uint8 are the stored codes,
 * unconverted (blocks stored in another type are rejected). */
int dtb_npz_read_typed(const char *path, float *props, int64_t *rowptr, uint32_t *colinds, void *weights,
                       int32_t weight_type, uint32_t *connkinds, int64_t *block_begin);

#ifdef __cplusplus
}
#endif
#endif
