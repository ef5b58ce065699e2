/*
 * dtb_oracle.h -- CPU oracle for the single-block DTB simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker or as the timed CPU
 * baseline, never as the thing measured or shipped.
 *
 * The oracle restates the reference algorithm in plain C (gcc, -O2,
 * -ffp-contract=off) in two precisions:
 *   - fp32 "canonical order": the exact operation sequence the HIP kernels
 *     use, so spike rasters, V, J and i_ou are bit-identical to the GPU;
 *   - fp64: the same model in double, for the 1e-5 relative check on V.
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/stage/extern_stimuli.cpp:39-112   OU background current
 *   src/stage/synaptic_current.cpp:7-71   I_syn = sum_u g_u (V_u - V) J_u
 *   src/stage/membrane_voltage.cpp:114-195 LIF, threshold, refractory
 *   src/stage/presynaptic_voltage.cpp:46-191 J decay + CSR spike scatter
 *   src/stage/property.cpp:19-53          exp(-dt/tau), t_ref / dt
 *   src/stage/spike.cpp:42-96              noise-spike mode
 *   src/block.cpp:136-179                  stage initialisation
 *   tools/dist_simulator.cpp:2072-2315     per-step stage order
 *   src/configuration.cpp:1778-2156        22-column property matrix
 *   models/block_new.py:71-158             CPU model (same equations)
 */
#ifndef DTB_ORACLE_H
#define DTB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Property-matrix columns (src/configuration.cpp:1778-2156;
 * python/make_block.py:441-465; models/block_new.py:164-165). */
enum {
    DTBO_COL_NOISE_RATE = 0,
    DTBO_COL_BLOCKED = 1,
    DTBO_COL_I_EXT = 2,
    DTBO_COL_SUB_IDX = 3,
    DTBO_COL_C = 4,
    DTBO_COL_T_REF = 5,
    DTBO_COL_G_LEAK = 6,
    DTBO_COL_V_LEAK = 7,
    DTBO_COL_V_TH = 8,
    DTBO_COL_V_RESET = 9,
    DTBO_COL_G_U = 10,   /* 10..13: AMPA, NMDA, GABA_A, GABA_B */
    DTBO_COL_V_U = 14,   /* 14..17 */
    DTBO_COL_TAU_U = 18, /* 18..21 */
    DTBO_NCOL = 22
};

/* Philox stream tags (word 3 of the counter). */
enum { DTBO_TAG_OU = 1, DTBO_TAG_NOISE = 2 };

typedef struct dtbo_net {
    int64_t n;                 /* neurons */
    int32_t npop;              /* populations (reference "sub-blocks") */
    const int64_t *pop_begin;  /* [npop+1] contiguous ranges, ascending */
    const float *props;        /* [n][22] property matrix */
    const int64_t *rowptr;     /* [n+1] presynaptic-major CSR */
    const uint32_t *col;       /* [S] postsynaptic neuron ids */
    const uint8_t *kind;       /* [S] 0: (AMPA,NMDA) pair, 1: (GABA_A,GABA_B) */
    const float *w;            /* [S][2] weight pair */
    float dt;                  /* ms, whole step */
    int32_t t_steps;           /* substeps per step */
    uint64_t seed;             /* Philox key */
    const uint8_t *ou_on;      /* [npop] OU configured for the population */
    const float *ou_mean;      /* [npop] */
    const float *ou_sigma;     /* [npop] */
    const float *ou_tau;       /* [npop] correlation time, ms */
} dtbo_net;

typedef struct dtbo_state32 {
    float *V;       /* [n] */
    int32_t *t;     /* [n] refractory counter */
    uint8_t *f;     /* [n] membrane spike flag */
    float *i_ou;    /* [n] */
    float *J;       /* [n][4] */
    float *I_syn;   /* [n] synaptic current without i_ou */
    int64_t step;   /* steps completed */
} dtbo_state32;

typedef struct dtbo_state64 {
    double *V;
    int32_t *t;
    uint8_t *f;
    double *i_ou;
    double *J;
    double *I_syn;
    int64_t step;
} dtbo_state64;

/* ---- random numbers ---------------------------------------------------- */
void dtbo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float dtbo_log(float u);                        /* ln u, u in (0, inf) normal range */
float dtbo_sin_halfpi(float x);                 /* sin(pi x / 2), x in [0, 0.5] */
float dtbo_cos_halfpi(float x);                 /* cos(pi x / 2), x in [0, 0.5] */
float dtbo_normal_from_bits(uint32_t x0, uint32_t x1);
float dtbo_uniform_from_bits(uint32_t x0);
float dtbo_ou_normal(uint64_t seed, uint32_t neuron, uint64_t substep);
float dtbo_noise_uniform(uint64_t seed, uint32_t neuron, uint64_t step);
void dtbo_normals(uint64_t seed, uint32_t neuron0, uint64_t substep, int64_t count, float *out);

/* ---- derived per-neuron / per-population parameters -------------------- */
float dtbo_expf_host(float x);
void dtbo_ou_factors(float dt_sub, float tau, float sigma, float *f1, float *f2s);

/* ---- simulation ---------------------------------------------------------- */
void dtbo_init32(const dtbo_net *net, dtbo_state32 *st);
void dtbo_init64(const dtbo_net *net, dtbo_state64 *st);

/* Runs nsteps steps.  raster: [nsteps][n] emitted spike flags (or NULL);
 * vtrace: [nsteps][n] V after each step (or NULL); pop_counts:
 * [nsteps][npop] emitted spikes per population excluding blocked neurons (or
 * NULL); stat_sums: [nsteps][npop][6] per-population sums over non-blocked
 * neurons of V, I_syn and the four receptor currents g_u (V_u - V) J_u at the
 * end of each step, each value quantized to a multiple of 2^-24 (see
 * dtbo_stat_q) so the sums are exact (or NULL; stat_result.cpp:266-391
 * averages these).  use_ou = 0 selects the noise-spike mode.  Returns 0 on
 * success. */
int dtbo_run32(const dtbo_net *net, dtbo_state32 *st, int32_t nsteps, int32_t use_ou,
               uint8_t *raster, float *vtrace, uint32_t *pop_counts, int64_t *stat_sums, int32_t nthreads);

/* A statistic's value as the exact integer rint(x * 2^24). */
int64_t dtbo_stat_q(float x);

/* fp64 model.  force_raster: [nsteps][n] spike decisions to follow instead of
 * the fp64 threshold test (or NULL for a free run).  mismatch_margin (may be
 * NULL) receives, per step, the largest |V - V_th| over neurons whose fp64
 * threshold test disagreed with force_raster. */
int dtbo_run64(const dtbo_net *net, dtbo_state64 *st, int32_t nsteps, int32_t use_ou,
               const uint8_t *force_raster, uint8_t *raster, double *vtrace,
               double *mismatch_margin, int32_t nthreads);

/* Dense brute-force propagation for tiny nets: delta[n][4] (double) of the
 * fired rows, computed as a dense matrix-vector product W[c][post][pre] @ f. */
void dtbo_dense_delta(const dtbo_net *net, const uint8_t *fired, double *delta);

/* CSR scatter with exact fixed-point accumulation, delta_q[n][4] Q23.40. */
void dtbo_scatter_q40(const dtbo_net *net, const uint8_t *fired, int64_t *delta_q);

#ifdef __cplusplus
}
#endif
#endif
