/*
 * dtb_oracle.c -- CPU oracle for the single-block DTB simulation step.
 *
 * TEST INFRASTRUCTURE ONLY (see dtb_oracle.h).  Build: oracle/Makefile,
 * gcc -O2 -ffp-contract=off -fopenmp.  fmaf() is used only where the
 * canonical order names it; every other expression is evaluated exactly as
 * written (no contraction, SSE arithmetic, round-to-nearest-even).
 */
#include "dtb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Round: (hi,lo)_0 = M0*c0, (hi,lo)_1 = M1*c2;   */
/* c' = (hi_1^c1^k0, lo_1, hi_0^c3^k1, lo_0); key bumped by the Weyl         */
/* constants between rounds.  The reference draws its OU normals from this   */
/* generator through hipRAND (src/stage/extern_stimuli.cpp:58,              */
/* src/stage/random.cpp:7-19).                                               */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void dtbo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* Box-Muller normal with our own log / sin / cos built only from IEEE      */
/* add, mul, fma, div and sqrt, so that the CPU and the GPU produce the same */
/* bits.  (The reference uses hiprand_normal; its stream is re-seeded from   */
/* /dev/urandom on every Run, src/common.cpp:19-37, so no bit-level match    */
/* with it is possible or intended.)                                         */
/* ------------------------------------------------------------------------ */
static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

#define LN2_F 0.693147180559945309f

/* ln(u) for positive normal u: u = m * 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m-1)/(m+1), series to s^9. */
float dtbo_log(float u)
{
    uint32_t b = f2u(u);
    int32_t e = (int32_t)((b >> 23) & 0xffu) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float z = s * s;
    float p = 0.111111111f;              /* 1/9 */
    p = fmaf(p, z, 0.142857143f);        /* 1/7 */
    p = fmaf(p, z, 0.2f);                /* 1/5 */
    p = fmaf(p, z, 0.333333333f);        /* 1/3 */
    p = fmaf(p, z, 1.0f);
    float lnm = (s + s) * p;
    return fmaf((float)e, LN2_F, lnm);
}

/* Taylor coefficients of sin(pi x/2) and cos(pi x/2) in x (|x| <= 1/2). */
#define SP1 1.57079632679489662f
#define SP3 -0.645964097506246254f
#define SP5 0.0796926262461670451f
#define SP7 -0.00468175413531868810f
#define SP9 0.000160441184757172367f
#define CP0 1.0f
#define CP2 -1.23370055013616983f
#define CP4 0.253669507901048014f
#define CP6 -0.0208634807633529609f
#define CP8 0.000919260274839426027f
#define CP10 -0.0000252020423730606054f

float dtbo_sin_halfpi(float x)
{
    float y = x * x;
    float p = SP9;
    p = fmaf(p, y, SP7);
    p = fmaf(p, y, SP5);
    p = fmaf(p, y, SP3);
    p = fmaf(p, y, SP1);
    return x * p;
}

float dtbo_cos_halfpi(float x)
{
    float y = x * x;
    float p = CP10;
    p = fmaf(p, y, CP8);
    p = fmaf(p, y, CP6);
    p = fmaf(p, y, CP4);
    p = fmaf(p, y, CP2);
    return fmaf(p, y, CP0);
}

/* z = sqrt(-2 ln u1) * cos(2 pi u2).
 * u1 = ((x0 >> 9) + 1/2) 2^-23 in (0,1); the angle uses the top 24 bits of
 * x1: quadrant q = bits 23..22, fraction f = ((x1 >> 8 & 0x3fffff) + 1/2)
 * 2^-22, angle = (pi/2)(q + f). */
float dtbo_normal_from_bits(uint32_t x0, uint32_t x1)
{
    float u1 = ((float)(x0 >> 9) + 0.5f) * 0x1p-23f;
    float r = sqrtf(-2.0f * dtbo_log(u1));
    uint32_t k = x1 >> 8;
    uint32_t q = k >> 22;
    float f = ((float)(k & 0x3fffffu) + 0.5f) * 0x1p-22f;
    float a, c;  /* sin(pi f/2), cos(pi f/2) */
    if (f <= 0.5f) {
        a = dtbo_sin_halfpi(f);
        c = dtbo_cos_halfpi(f);
    } else {
        float g = 1.0f - f;
        a = dtbo_cos_halfpi(g);
        c = dtbo_sin_halfpi(g);
    }
    float t;
    switch (q) {
    case 0: t = c; break;
    case 1: t = -a; break;
    case 2: t = -c; break;
    default: t = a; break;
    }
    return r * t;
}

float dtbo_uniform_from_bits(uint32_t x0)
{
    return (float)(x0 >> 8) * 0x1p-24f;
}

float dtbo_ou_normal(uint64_t seed, uint32_t neuron, uint64_t substep)
{
    uint32_t ctr[4] = {neuron, (uint32_t)substep, (uint32_t)(substep >> 32), DTBO_TAG_OU};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    dtbo_philox4x32_10(ctr, key, x);
    return dtbo_normal_from_bits(x[0], x[1]);
}

float dtbo_noise_uniform(uint64_t seed, uint32_t neuron, uint64_t step)
{
    uint32_t ctr[4] = {neuron, (uint32_t)step, (uint32_t)(step >> 32), DTBO_TAG_NOISE};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    dtbo_philox4x32_10(ctr, key, x);
    return dtbo_uniform_from_bits(x[0]);
}

void dtbo_normals(uint64_t seed, uint32_t neuron0, uint64_t substep, int64_t count, float *out)
{
    for (int64_t i = 0; i < count; ++i)
        out[i] = dtbo_ou_normal(seed, neuron0 + (uint32_t)i, substep);
}

/* ------------------------------------------------------------------------ */
/* Derived parameters.  exp() is evaluated in double and rounded once, on   */
/* the host, by the oracle and by the product alike (the reference          */
/* evaluates texp<float> on the device: src/stage/property.cpp:19-43,        */
/* src/stage/extern_stimuli.cpp:87-91).                                      */
/* ------------------------------------------------------------------------ */
float dtbo_expf_host(float x)
{
    return (float)exp((double)x);
}

/* extern_stimuli.cpp:87-91: factor = -dt/tau; f1 = 1 - e^factor;
 * f2 = sqrt(1 - e^(2 factor)); the kernel multiplies f2 * sigma * y. */
void dtbo_ou_factors(float dt_sub, float tau, float sigma, float *f1, float *f2s)
{
    float factor = (-dt_sub) / tau;
    float a = 1.0f - dtbo_expf_host(factor);
    float b = sqrtf(1.0f - dtbo_expf_host(2.0f * factor));
    *f1 = a;
    *f2s = b * sigma;
}

typedef struct {
    int64_t n;
    int32_t T;
    float dt_sub;
    float *c_inv, *v_rst, *v_th, *v_leak, *g_leak, *i_ext, *noise;
    int32_t *r_steps;   /* (int)(t_ref / dt) * t_steps */
    float *g, *vu, *decay;  /* [n][4] */
    uint8_t *blocked;
    int32_t *pop;       /* [n] population of each neuron */
    float *f1, *f2s, *mu;   /* [npop] */
} derived32;

typedef struct {
    int64_t n;
    int32_t T;
    double dt_sub;
    double *c_inv, *v_rst, *v_th, *v_leak, *g_leak, *i_ext, *noise;
    int32_t *r_steps;
    double *g, *vu, *decay;
    uint8_t *blocked;
    int32_t *pop;
    double *f1, *f2s, *mu;
} derived64;

static int32_t *make_pop_index(const dtbo_net *net)
{
    int32_t *pop = (int32_t *)malloc(sizeof(int32_t) * (size_t)(net->n > 0 ? net->n : 1));
    for (int32_t p = 0; p < net->npop; ++p)
        for (int64_t i = net->pop_begin[p]; i < net->pop_begin[p + 1]; ++i)
            pop[i] = p;
    return pop;
}

static int32_t ref_steps(const dtbo_net *net, float t_ref)
{
    /* property.cpp:45-53 divides t_ref by the whole-step dt at init
     * (block.cpp:150); membrane_voltage.cpp:177 uses (int)t_ref * t_steps. */
    float steps = t_ref / net->dt;
    return (int32_t)steps * net->t_steps;
}

static void derive32(const dtbo_net *net, derived32 *d)
{
    int64_t n = net->n;
    size_t nn = (size_t)(n > 0 ? n : 1);
    d->n = n;
    d->T = net->t_steps;
    d->dt_sub = net->dt / (float)net->t_steps;
    d->c_inv = malloc(sizeof(float) * nn);
    d->v_rst = malloc(sizeof(float) * nn);
    d->v_th = malloc(sizeof(float) * nn);
    d->v_leak = malloc(sizeof(float) * nn);
    d->g_leak = malloc(sizeof(float) * nn);
    d->i_ext = malloc(sizeof(float) * nn);
    d->noise = malloc(sizeof(float) * nn);
    d->r_steps = malloc(sizeof(int32_t) * nn);
    d->g = malloc(sizeof(float) * 4 * nn);
    d->vu = malloc(sizeof(float) * 4 * nn);
    d->decay = malloc(sizeof(float) * 4 * nn);
    d->blocked = malloc(nn);
    for (int64_t i = 0; i < n; ++i) {
        const float *p = net->props + (size_t)i * DTBO_NCOL;
        d->noise[i] = p[DTBO_COL_NOISE_RATE];
        d->blocked[i] = p[DTBO_COL_BLOCKED] != 0.0f;
        d->i_ext[i] = p[DTBO_COL_I_EXT];
        d->c_inv[i] = 1.0f / p[DTBO_COL_C];            /* configuration.cpp:1670-1677 */
        d->r_steps[i] = ref_steps(net, p[DTBO_COL_T_REF]);
        d->g_leak[i] = p[DTBO_COL_G_LEAK];
        d->v_leak[i] = p[DTBO_COL_V_LEAK];
        d->v_th[i] = p[DTBO_COL_V_TH];
        d->v_rst[i] = p[DTBO_COL_V_RESET];
        for (int c = 0; c < 4; ++c) {
            d->g[i * 4 + c] = p[DTBO_COL_G_U + c];
            d->vu[i * 4 + c] = p[DTBO_COL_V_U + c];
            /* property.cpp:19-29: exp(-dt / tau) */
            d->decay[i * 4 + c] = dtbo_expf_host((-d->dt_sub) / p[DTBO_COL_TAU_U + c]);
        }
    }
    d->pop = make_pop_index(net);
    size_t np = (size_t)(net->npop > 0 ? net->npop : 1);
    d->f1 = malloc(sizeof(float) * np);
    d->f2s = malloc(sizeof(float) * np);
    d->mu = malloc(sizeof(float) * np);
    for (int32_t q = 0; q < net->npop; ++q) {
        d->f1[q] = d->f2s[q] = d->mu[q] = 0.0f;
        if (net->ou_on && net->ou_on[q]) {
            dtbo_ou_factors(d->dt_sub, net->ou_tau[q], net->ou_sigma[q], &d->f1[q], &d->f2s[q]);
            d->mu[q] = net->ou_mean[q];
        }
    }
}

static void derive64(const dtbo_net *net, derived64 *d)
{
    int64_t n = net->n;
    size_t nn = (size_t)(n > 0 ? n : 1);
    d->n = n;
    d->T = net->t_steps;
    d->dt_sub = (double)net->dt / (double)net->t_steps;
    d->c_inv = malloc(sizeof(double) * nn);
    d->v_rst = malloc(sizeof(double) * nn);
    d->v_th = malloc(sizeof(double) * nn);
    d->v_leak = malloc(sizeof(double) * nn);
    d->g_leak = malloc(sizeof(double) * nn);
    d->i_ext = malloc(sizeof(double) * nn);
    d->noise = malloc(sizeof(double) * nn);
    d->r_steps = malloc(sizeof(int32_t) * nn);
    d->g = malloc(sizeof(double) * 4 * nn);
    d->vu = malloc(sizeof(double) * 4 * nn);
    d->decay = malloc(sizeof(double) * 4 * nn);
    d->blocked = malloc(nn);
    for (int64_t i = 0; i < n; ++i) {
        const float *p = net->props + (size_t)i * DTBO_NCOL;
        d->noise[i] = p[DTBO_COL_NOISE_RATE];
        d->blocked[i] = p[DTBO_COL_BLOCKED] != 0.0f;
        d->i_ext[i] = p[DTBO_COL_I_EXT];
        d->c_inv[i] = 1.0 / (double)p[DTBO_COL_C];
        d->r_steps[i] = ref_steps(net, p[DTBO_COL_T_REF]);
        d->g_leak[i] = p[DTBO_COL_G_LEAK];
        d->v_leak[i] = p[DTBO_COL_V_LEAK];
        d->v_th[i] = p[DTBO_COL_V_TH];
        d->v_rst[i] = p[DTBO_COL_V_RESET];
        for (int c = 0; c < 4; ++c) {
            d->g[i * 4 + c] = p[DTBO_COL_G_U + c];
            d->vu[i * 4 + c] = p[DTBO_COL_V_U + c];
            d->decay[i * 4 + c] = exp(-d->dt_sub / (double)p[DTBO_COL_TAU_U + c]);
        }
    }
    d->pop = make_pop_index(net);
    size_t np = (size_t)(net->npop > 0 ? net->npop : 1);
    d->f1 = malloc(sizeof(double) * np);
    d->f2s = malloc(sizeof(double) * np);
    d->mu = malloc(sizeof(double) * np);
    for (int32_t q = 0; q < net->npop; ++q) {
        d->f1[q] = d->f2s[q] = d->mu[q] = 0.0;
        if (net->ou_on && net->ou_on[q]) {
            double factor = -d->dt_sub / (double)net->ou_tau[q];
            d->f1[q] = 1.0 - exp(factor);
            d->f2s[q] = sqrt(1.0 - exp(2.0 * factor)) * (double)net->ou_sigma[q];
            d->mu[q] = net->ou_mean[q];
        }
    }
}

#define FREE_DERIVED(d)                                                         \
    do {                                                                        \
        free((d)->c_inv); free((d)->v_rst); free((d)->v_th); free((d)->v_leak); \
        free((d)->g_leak); free((d)->i_ext); free((d)->noise);                  \
        free((d)->r_steps); free((d)->g); free((d)->vu); free((d)->decay);      \
        free((d)->blocked); free((d)->pop); free((d)->f1); free((d)->f2s);      \
        free((d)->mu);                                                          \
    } while (0)

/* synaptic_current.cpp:39-43, evaluated left to right, no contraction. */
static inline float isyn32(const float *g, const float *vu, const float *J, float V)
{
    float a = g[0] * (vu[0] - V) * J[0];
    float b = g[1] * (vu[1] - V) * J[1];
    float c = g[2] * (vu[2] - V) * J[2];
    float d = g[3] * (vu[3] - V) * J[3];
    return a + b + c + d;
}

static inline double isyn64(const double *g, const double *vu, const double *J, double V)
{
    double a = g[0] * (vu[0] - V) * J[0];
    double b = g[1] * (vu[1] - V) * J[1];
    double c = g[2] * (vu[2] - V) * J[2];
    double d = g[3] * (vu[3] - V) * J[3];
    return a + b + c + d;
}

/* Q23.40 fixed point: exact for every fp32 weight >= 2^-17 and for every
 * multiple of 2^-40 (presynaptic_voltage.cpp:178-187 accumulates with float
 * atomics instead; integer addition makes the sum order-free). */
static inline int64_t w_to_q40(float w)
{
    return (int64_t)rintf(w * 0x1p40f);
}

static inline float q40_to_f(int64_t q)
{
    return (float)((double)q * 0x1p-40);
}

void dtbo_scatter_q40(const dtbo_net *net, const uint8_t *fired, int64_t *dq)
{
    /* presynaptic_voltage.cpp:137-188: for each fired presynaptic row, each
     * synapse adds w.x / w.y to the (AMPA,NMDA) or (GABA_A,GABA_B) pair of
     * its postsynaptic neuron. */
    int64_t n = net->n;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t r = 0; r < n; ++r) {
        if (!fired[r])
            continue;
        for (int64_t j = net->rowptr[r]; j < net->rowptr[r + 1]; ++j) {
            int64_t post = net->col[j];
            int ch = net->kind[j] ? 2 : 0;
            int64_t qx = w_to_q40(net->w[2 * j]);
            int64_t qy = w_to_q40(net->w[2 * j + 1]);
            __atomic_fetch_add(&dq[post * 4 + ch], qx, __ATOMIC_RELAXED);
            __atomic_fetch_add(&dq[post * 4 + ch + 1], qy, __ATOMIC_RELAXED);
        }
    }
}

void dtbo_dense_delta(const dtbo_net *net, const uint8_t *fired, double *delta)
{
    /* Brute force: build W[c][post][pre] densely, then delta = W @ fired
     * (models/block_new.py:71-84, the dense bmm path). */
    int64_t n = net->n;
    double *W = calloc((size_t)(4 * n * n), sizeof(double));
    for (int64_t r = 0; r < n; ++r)
        for (int64_t j = net->rowptr[r]; j < net->rowptr[r + 1]; ++j) {
            int ch = net->kind[j] ? 2 : 0;
            int64_t post = net->col[j];
            W[((int64_t)ch * n + post) * n + r] += net->w[2 * j];
            W[((int64_t)(ch + 1) * n + post) * n + r] += net->w[2 * j + 1];
        }
    for (int c = 0; c < 4; ++c)
        for (int64_t post = 0; post < n; ++post) {
            double s = 0.0;
            for (int64_t r = 0; r < n; ++r)
                s += W[((int64_t)c * n + post) * n + r] * (double)(fired[r] != 0);
            delta[post * 4 + c] = s;
        }
    free(W);
}

static void scatter64(const dtbo_net *net, const uint8_t *fired, double *delta)
{
    /* Exact in double for the 24-bit weights used by the workloads (sums of
     * multiples of 2^-24 below 2^29); serial so the order is fixed. */
    int64_t n = net->n;
    for (int64_t r = 0; r < n; ++r) {
        if (!fired[r])
            continue;
        for (int64_t j = net->rowptr[r]; j < net->rowptr[r + 1]; ++j) {
            int64_t post = net->col[j];
            int ch = net->kind[j] ? 2 : 0;
            delta[post * 4 + ch] += (double)net->w[2 * j];
            delta[post * 4 + ch + 1] += (double)net->w[2 * j + 1];
        }
    }
}

void dtbo_init32(const dtbo_net *net, dtbo_state32 *st)
{
    /* block.cpp:136-179: V0 = (V_th + V_reset)/2 (membrane_voltage.cpp:18),
     * t = 0, J = 0, I_syn from J = 0, f = 0; OU current starts at its mean
     * (block.cpp:255-259 via reset_partial_gpu). */
    int64_t n = net->n;
    for (int64_t i = 0; i < n; ++i) {
        const float *p = net->props + (size_t)i * DTBO_NCOL;
        float vth = p[DTBO_COL_V_TH], vr = p[DTBO_COL_V_RESET];
        st->V[i] = (vth + vr) / 2.0f;
        st->t[i] = 0;
        st->f[i] = 0;
        st->i_ou[i] = 0.0f;
        for (int c = 0; c < 4; ++c)
            st->J[i * 4 + c] = 0.0f;
        st->I_syn[i] = isyn32(p + DTBO_COL_G_U, p + DTBO_COL_V_U, st->J + i * 4, st->V[i]);
    }
    for (int32_t q = 0; q < net->npop; ++q)
        if (net->ou_on && net->ou_on[q])
            for (int64_t i = net->pop_begin[q]; i < net->pop_begin[q + 1]; ++i)
                st->i_ou[i] = net->ou_mean[q];
    st->step = 0;
}

void dtbo_init64(const dtbo_net *net, dtbo_state64 *st)
{
    int64_t n = net->n;
    for (int64_t i = 0; i < n; ++i) {
        const float *p = net->props + (size_t)i * DTBO_NCOL;
        double vth = p[DTBO_COL_V_TH], vr = p[DTBO_COL_V_RESET];
        st->V[i] = (vth + vr) / 2.0;
        st->t[i] = 0;
        st->f[i] = 0;
        st->i_ou[i] = 0.0;
        for (int c = 0; c < 4; ++c)
            st->J[i * 4 + c] = 0.0;
        st->I_syn[i] = 0.0;
    }
    for (int32_t q = 0; q < net->npop; ++q)
        if (net->ou_on && net->ou_on[q])
            for (int64_t i = net->pop_begin[q]; i < net->pop_begin[q + 1]; ++i)
                st->i_ou[i] = net->ou_mean[q];
    st->step = 0;
}

/* One membrane substep for neuron i (membrane_voltage.cpp:130-190).
 * Returns 1 when the neuron crosses threshold in this substep. */
static inline int lif32(const derived32 *d, dtbo_state32 *st, int64_t i, float I)
{
    int32_t t = st->t[i];
    if (t < 0) {
        if (st->f[i]) {
            st->V[i] = d->v_rst[i];
            st->f[i] = 0;
        }
        st->t[i] = t + 1;
        return 0;
    }
    float V = st->V[i];
    float cvi = d->g_leak[i] * (d->v_leak[i] - V);
    cvi = cvi + I;
    cvi = cvi + d->i_ext[i];
    V = V + d->dt_sub * d->c_inv[i] * cvi;
    int fired = V >= d->v_th[i];
    if (fired) {
        V = d->v_th[i];
        st->t[i] = 1 - d->r_steps[i];
    }
    st->f[i] = (uint8_t)fired;
    st->V[i] = V;
    return fired;
}

int64_t dtbo_stat_q(float x)
{
    return (int64_t)rintf(x * 0x1p24f);
}

int dtbo_run32(const dtbo_net *net, dtbo_state32 *st, int32_t nsteps, int32_t use_ou,
               uint8_t *raster, float *vtrace, uint32_t *pop_counts, int64_t *stat_sums, int32_t nthreads)
{
    if (net->t_steps < 1 || (net->t_steps > 1 && !use_ou))
        return 1;   /* dist_simulator.cpp:3225-3234: t_steps > 1 needs OU */
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    derived32 d;
    derive32(net, &d);
    int64_t n = net->n;
    int32_t T = net->t_steps;
    size_t nn = (size_t)(n > 0 ? n : 1);
    uint8_t *emit = calloc(nn, 1);
    int64_t *dq = calloc(nn * 4, sizeof(int64_t));

    for (int32_t k = 0; k < nsteps; ++k) {
        int64_t step = st->step;
        memset(emit, 0, nn);
        for (int32_t s = 0; s < T; ++s) {
            uint64_t sub = (uint64_t)step * (uint64_t)T + (uint64_t)s;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                float I = st->I_syn[i];
                int32_t q = d.pop[i];
                if (use_ou && net->ou_on && net->ou_on[q]) {
                    /* extern_stimuli.cpp:58-65 */
                    float xi = dtbo_ou_normal(net->seed, (uint32_t)i, sub);
                    float iou = st->i_ou[i];
                    iou = iou + (d.f1[q] * (d.mu[q] - iou) + d.f2s[q] * xi);
                    st->i_ou[i] = iou;
                    I = I + iou;
                }
                int fired = lif32(&d, st, i, I);
                if (T > 1 || !use_ou)
                    emit[i] |= (uint8_t)fired;
                else
                    emit[i] = (uint8_t)fired;
                if (s < T - 1) {
                    /* dist_simulator.cpp:2092-2097: J decay without delta,
                     * then I_syn, between substeps (presynaptic_voltage.cpp:81-90). */
                    float *J = st->J + i * 4;
                    for (int c = 0; c < 4; ++c)
                        J[c] = J[c] * d.decay[i * 4 + c];
                    st->I_syn[i] = isyn32(d.g + i * 4, d.vu + i * 4, J, st->V[i]);
                }
            }
        }
        if (!use_ou) {
            /* spike.cpp:55-62: f_rec = f | (U < noise_rate) */
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                float u = dtbo_noise_uniform(net->seed, (uint32_t)i, (uint64_t)step);
                emit[i] = (uint8_t)(st->f[i] | (uint8_t)(u < d.noise[i]));
            }
        }
        /* Propagation (presynaptic_voltage.cpp:122-191) into exact Q23.40. */
        memset(dq, 0, nn * 4 * sizeof(int64_t));
        dtbo_scatter_q40(net, emit, dq);
        /* J = J * decay + delta (presynaptic_voltage.cpp:62-73, fused
         * multiply-add), then I_syn (synaptic_current.cpp). */
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            float *J = st->J + i * 4;
            for (int c = 0; c < 4; ++c)
                J[c] = fmaf(J[c], d.decay[i * 4 + c], q40_to_f(dq[i * 4 + c]));
            st->I_syn[i] = isyn32(d.g + i * 4, d.vu + i * 4, J, st->V[i]);
        }
        if (raster)
            memcpy(raster + (size_t)k * nn, emit, (size_t)n);
        if (vtrace)
            memcpy(vtrace + (size_t)k * nn, st->V, sizeof(float) * (size_t)n);
        if (pop_counts) {
            /* stat_result.cpp:161-206: per-population spike count without
             * blocked neurons. */
            for (int32_t q = 0; q < net->npop; ++q) {
                uint32_t cnt = 0;
                for (int64_t i = net->pop_begin[q]; i < net->pop_begin[q + 1]; ++i)
                    cnt += (uint32_t)(emit[i] && !d.blocked[i]);
                pop_counts[(size_t)k * (size_t)net->npop + (size_t)q] = cnt;
            }
        }
        if (stat_sums) {
            /* stat_result.cpp:266-391: per-population V, I_syn and receptor
             * currents (synaptic_current.cpp:39-43 terms) at the end of the
             * step, blocked neurons excluded. */
            for (int32_t q = 0; q < net->npop; ++q) {
                int64_t acc[6] = {0, 0, 0, 0, 0, 0};
                for (int64_t i = net->pop_begin[q]; i < net->pop_begin[q + 1]; ++i) {
                    if (d.blocked[i])
                        continue;
                    const float V = st->V[i];
                    const float *J = st->J + i * 4, *g = d.g + i * 4, *vu = d.vu + i * 4;
                    acc[0] += dtbo_stat_q(V);
                    acc[1] += dtbo_stat_q(st->I_syn[i]);
                    for (int c = 0; c < 4; ++c)
                        acc[2 + c] += dtbo_stat_q(g[c] * (vu[c] - V) * J[c]);
                }
                memcpy(stat_sums + ((size_t)k * (size_t)net->npop + (size_t)q) * 6, acc, sizeof(acc));
            }
        }
        st->step = step + 1;
    }
    free(emit);
    free(dq);
    FREE_DERIVED(&d);
    return 0;
}

int dtbo_run64(const dtbo_net *net, dtbo_state64 *st, int32_t nsteps, int32_t use_ou,
               const uint8_t *force_raster, uint8_t *raster, double *vtrace,
               double *mismatch_margin, int32_t nthreads)
{
    if (net->t_steps < 1 || (net->t_steps > 1 && !use_ou))
        return 1;
    if (force_raster && net->t_steps != 1)
        return 1;
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    derived64 d;
    derive64(net, &d);
    int64_t n = net->n;
    int32_t T = net->t_steps;
    size_t nn = (size_t)(n > 0 ? n : 1);
    uint8_t *emit = calloc(nn, 1);
    double *delta = calloc(nn * 4, sizeof(double));

    for (int32_t k = 0; k < nsteps; ++k) {
        int64_t step = st->step;
        double margin = 0.0;
        memset(emit, 0, nn);
        for (int32_t s = 0; s < T; ++s) {
            uint64_t sub = (uint64_t)step * (uint64_t)T + (uint64_t)s;
#pragma omp parallel for schedule(static) reduction(max : margin)
            for (int64_t i = 0; i < n; ++i) {
                double I = st->I_syn[i];
                int32_t q = d.pop[i];
                if (use_ou && net->ou_on && net->ou_on[q]) {
                    double xi = (double)dtbo_ou_normal(net->seed, (uint32_t)i, sub);
                    double iou = st->i_ou[i];
                    iou = iou + (d.f1[q] * (d.mu[q] - iou) + d.f2s[q] * xi);
                    st->i_ou[i] = iou;
                    I = I + iou;
                }
                int fired = 0;
                int32_t t = st->t[i];
                if (t < 0) {
                    if (st->f[i]) {
                        st->V[i] = d.v_rst[i];
                        st->f[i] = 0;
                    }
                    st->t[i] = t + 1;
                } else {
                    double V = st->V[i];
                    double cvi = d.g_leak[i] * (d.v_leak[i] - V);
                    cvi = cvi + I;
                    cvi = cvi + d.i_ext[i];
                    V = V + d.dt_sub * d.c_inv[i] * cvi;
                    fired = V >= d.v_th[i];
                    if (force_raster) {
                        int want = force_raster[(size_t)k * nn + (size_t)i] != 0;
                        if (want != fired) {
                            double m = fabs(V - d.v_th[i]);
                            if (m > margin)
                                margin = m;
                        }
                        fired = want;
                    }
                    if (fired) {
                        V = d.v_th[i];
                        st->t[i] = 1 - d.r_steps[i];
                    }
                    st->f[i] = (uint8_t)fired;
                    st->V[i] = V;
                }
                if (T > 1 || !use_ou)
                    emit[i] |= (uint8_t)fired;
                else
                    emit[i] = (uint8_t)fired;
                if (s < T - 1) {
                    double *J = st->J + i * 4;
                    for (int c = 0; c < 4; ++c)
                        J[c] = J[c] * d.decay[i * 4 + c];
                    st->I_syn[i] = isyn64(d.g + i * 4, d.vu + i * 4, J, st->V[i]);
                }
            }
        }
        if (!use_ou) {
            for (int64_t i = 0; i < n; ++i) {
                double u = (double)dtbo_noise_uniform(net->seed, (uint32_t)i, (uint64_t)step);
                emit[i] = (uint8_t)(st->f[i] | (uint8_t)(u < d.noise[i]));
            }
        }
        memset(delta, 0, nn * 4 * sizeof(double));
        scatter64(net, emit, delta);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double *J = st->J + i * 4;
            for (int c = 0; c < 4; ++c)
                J[c] = J[c] * d.decay[i * 4 + c] + delta[i * 4 + c];
            st->I_syn[i] = isyn64(d.g + i * 4, d.vu + i * 4, J, st->V[i]);
        }
        if (raster)
            memcpy(raster + (size_t)k * nn, emit, (size_t)n);
        if (vtrace)
            memcpy(vtrace + (size_t)k * nn, st->V, sizeof(double) * (size_t)n);
        if (mismatch_margin)
            mismatch_margin[k] = margin;
        st->step = step + 1;
    }
    free(emit);
    free(delta);
    FREE_DERIVED(&d);
    return 0;
}
