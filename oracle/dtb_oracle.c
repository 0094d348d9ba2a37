/*
 * dtb_oracle.c -- CPU oracle for the Digital Twin Brain network step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (digital_twin_brain_amd/) never calls into it.
 *
 * It restates, in plain C, the reference algorithm of:
 *   - models/block.py:30-150          LIF + conductance block step
 *   - models/bold_model_pytorch.py:43-78   Balloon-Windkessel Euler step
 *   - data_assimilation/DataAssimilation.py:40-60  diffusion EnKF,
 *     with the observation noise xi passed in instead of drawn
 *   - Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based
 *     RNG shared by the product and this oracle for noise and network draws
 *   - the seeded fixed-in-degree network generator whose output layout is
 *     the presynaptic-major CSR of cuda/src/configuration.cpp:747-828
 *
 * Arithmetic type: REAL (double by default; -DDTB_REAL_FLOAT builds the FP32
 * variant used only to time the reference algorithm on the CPU).
 *
 * Parity pinning (see tests/test_oracle.py): Philox against the Random123
 * known-answer vectors; the block step against the golden vectors written by
 * the reference models/block.py (tests/golden/) and against closed forms
 * (leak relaxation, refractory hold, exponential synaptic decay, single-spike
 * delivery, dyadic-weight exact sums); BOLD against golden vectors and its
 * fixed point; the EnKF against the textbook scalar Kalman update, the
 * identical-member and sigma->inf invariants and linearity in rho.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef DTB_REAL_FLOAT
typedef float REAL;
#define RSUFFIX(name) name##_f32
#else
typedef double REAL;
#define RSUFFIX(name) name##_f64
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                        */
/* ------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 24-bit uniform in [0,1): exact in float32 */
static inline float u24(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

/* Stream identifiers (word 3 of the counter). */
#define DTB_STREAM_NOISE 0x4E4F4953u /* noise uniforms: ctr = (i_lo, i_hi, step, NOISE) */
#define DTB_STREAM_CONN 0x434F4E4Eu  /* network draws:  ctr = (k, post_lo, post_hi, CONN) */

/* Noise uniform for neuron i at step n (n = 1 for the first step). */
float orc_noise_uniform(uint64_t seed, uint64_t step, uint64_t i)
{
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)step, DTB_STREAM_NOISE};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return u24(o[0]);
}

void orc_noise_uniforms(uint64_t seed, uint64_t step, int64_t n, float* out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_noise_uniform(seed, step, (uint64_t)i);
}

/* ------------------------------------------------------------------ */
/* Fixed in-degree generator (canonical presynaptic-major CSR)          */
/* ------------------------------------------------------------------ */
/*
 * For post i in [0,n) and slot k in [0,K): one Philox call with
 * ctr = (k, i_lo, i_hi, CONN), key = seed gives words r0..r3:
 *   pre   = floor(r0 * (n-1) / 2^32), shifted past i (no self-connection)
 *   w0    = fmaf(w_hi - w_lo, u24(r1), w_lo)
 *   delay = 1 + floor(r2 * D / 2^32)
 *   w1    = fmaf(w_hi - w_lo, u24(r3), w_lo)   (pair weights only)
 *   kind  = pre >= n_exc  (0: AMPA/NMDA pair, 1: GABAa/GABAb pair; the
 *           receptor pair follows the presynaptic type, generation/make_block.py:290-291)
 * Row pre lists its synapses by ascending (post, k).
 * rowptr must hold n+1 entries; the other outputs nnz = n*K entries.
 */
/* Members packed in one network: member m owns neurons [off_m, off_m+n_m),
 * the first nexc_m excitatory, each receiving K_m synapses from inside the
 * member; draws as above with pre = off_m + floor(r0 (n_m-1) / 2^32) shifted
 * past i, and kind = (pre - off_m >= nexc_m). */
/* Regions (or members) with a source-region table: region m owns the next
 * ns[m] neurons, the first nexcs[m] excitatory; each of its neurons receives
 * Ks[m] synapses from inside the region (draws as orc_gen_members) and then
 * Kx[m] synapses from other regions.  Inter slot k: the source region is the
 * smallest s with r < cdf[m*M + s], r from a second Philox call
 * (ctr = (k, post_lo, post_hi, REGION)), and the presynaptic neuron is
 * off_s + floor(r0 n_s / 2^32).  cdf holds cumulative probabilities scaled to
 * 2^32 (the last entry of a row is 2^32). */
#define DTB_STREAM_REGION 0x52454749u /* ctr = (k, post_lo, post_hi, REGION) */

void orc_gen_regions(int32_t M, const int64_t* ns, const int64_t* nexcs, const int32_t* Ks, const int32_t* Kx,
                     const uint64_t* cdf, int32_t D, float w_lo, float w_hi, int32_t pair, uint64_t seed,
                     int64_t* rowptr, int32_t* post, uint8_t* kind, uint8_t* delay, float* w0, float* w1)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    float span = w_hi - w_lo;
    int64_t n = 0;
    int64_t* offs = (int64_t*)calloc((size_t)M + 1, sizeof(int64_t));
    for (int32_t m = 0; m < M; ++m) {
        offs[m] = n;
        n += ns[m];
    }
    offs[M] = n;
    int64_t* cursor = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    memset(rowptr, 0, sizeof(int64_t) * (size_t)(n + 1));
    for (int pass = 0; pass < 2; ++pass) {
        for (int32_t m = 0; m < M; ++m) {
            const int64_t off = offs[m];
            const int32_t kx = Kx ? Kx[m] : 0;
            for (int64_t i = off; i < off + ns[m]; ++i) {
                for (int32_t k = 0; k < Ks[m] + kx; ++k) {
                    uint32_t ctr[4] = {(uint32_t)k, (uint32_t)i, (uint32_t)((uint64_t)i >> 32), DTB_STREAM_CONN};
                    uint32_t r[4];
                    orc_philox4x32_10(ctr, key, r);
                    int32_t src = m;
                    int64_t pl;
                    if (k < Ks[m]) {
                        pl = (int64_t)(((uint64_t)r[0] * (uint64_t)(ns[m] - 1)) >> 32);
                        if (pl >= i - off) pl += 1;
                    } else {
                        uint32_t c2[4] = {(uint32_t)k, (uint32_t)i, (uint32_t)((uint64_t)i >> 32), DTB_STREAM_REGION};
                        uint32_t q[4];
                        orc_philox4x32_10(c2, key, q);
                        src = 0;
                        while (src < M - 1 && (uint64_t)q[0] >= cdf[(int64_t)m * M + src]) ++src;
                        pl = (int64_t)(((uint64_t)r[0] * (uint64_t)ns[src]) >> 32);
                    }
                    int64_t pre = offs[src] + pl;
                    if (pass == 0) {
                        rowptr[pre + 1] += 1;
                        continue;
                    }
                    int64_t e = cursor[pre]++;
                    post[e] = (int32_t)i;
                    kind[e] = (uint8_t)(pl >= nexcs[src]);
                    delay[e] = (uint8_t)(1 + (uint32_t)(((uint64_t)r[2] * (uint64_t)D) >> 32));
                    w0[e] = fmaf(span, u24(r[1]), w_lo);
                    if (w1) w1[e] = pair ? fmaf(span, u24(r[3]), w_lo) : w0[e];
                }
            }
        }
        if (pass == 0) {
            for (int64_t j = 0; j < n; ++j) rowptr[j + 1] += rowptr[j];
            for (int64_t j = 0; j < n; ++j) cursor[j] = rowptr[j];
        }
    }
    free(cursor);
    free(offs);
}

void orc_gen_members(int32_t M, const int64_t* ns, const int64_t* nexcs, const int32_t* Ks, int32_t D, float w_lo,
                     float w_hi, int32_t pair, uint64_t seed, int64_t* rowptr, int32_t* post, uint8_t* kind,
                     uint8_t* delay, float* w0, float* w1)
{
    orc_gen_regions(M, ns, nexcs, Ks, NULL, NULL, D, w_lo, w_hi, pair, seed, rowptr, post, kind, delay, w0, w1);
}

void orc_gen_fixed_indegree(int64_t n, int64_t n_exc, int32_t K, int32_t D, float w_lo, float w_hi,
                            int32_t pair, uint64_t seed, int64_t* rowptr, int32_t* post, uint8_t* kind,
                            uint8_t* delay, float* w0, float* w1)
{
    orc_gen_members(1, &n, &n_exc, &K, D, w_lo, w_hi, pair, seed, rowptr, post, kind, delay, w0, w1);
}

/* ------------------------------------------------------------------ */
/* Block step (models/block.py)                                         */
/* ------------------------------------------------------------------ */
/*
 * Property columns (generation/make_block.py:394-434, models/block.py:156):
 *  0 noise rate (E_I in block.py, unused by block.run), 1 blocked_in_stat,
 *  2 I_ext, 3 sub_idx, 4 C, 5 T_ref, 6 g_L, 7 V_L, 8 V_th, 9 V_reset,
 *  10-13 g_ui, 14-17 V_ui, 18-21 tau_ui.
 */
typedef struct {
    int64_t n, nnz;
    int32_t S, D;
    REAL dt;
    /* properties */
    float* props; /* [n*22] */
    /* presynaptic-major CSR (push) */
    int64_t* rowptr;
    int32_t* post;
    uint8_t* kind;
    uint8_t* delay;
    float *w0, *w1;
    /* postsynaptic-major CSR (pull, the torch.sparse.mm formulation) */
    int64_t* prow;
    int32_t* ppre;
    uint8_t* pkind;
    uint8_t* pdelay;
    float *pw0, *pw1;
    /* state */
    REAL t;
    int64_t step;
    REAL *V, *J, *I_ui, *I_syn, *t_last; /* J, I_ui: [4][n] */
    REAL* ring;                          /* [D][4][n] pending arrivals (push) */
    uint8_t* shist;                      /* [D][n] spike history (pull) */
    uint8_t *active, *spike;
    int32_t* sub;
    /* external Poisson drive: count = #{j : r >= ext_thr[j]}, times ext_w on channels 0 and 1 */
    int32_t ext_n;
    float ext_w;
    uint32_t ext_thr[16];
    /* forced-spike sources (models/block.py:38-50,104-109): srcflag bit0 =
     * source, bit1 = frozen (an input block's neuron: no membrane update and
     * no noise, cuda/src/block.cpp:195-221,277-283); events (ev_step,
     * ev_neuron) sorted by step, fire[] marks the current step's */
    uint8_t* srcflag;
    uint8_t* fire;
    int64_t nev;
    uint32_t* ev_step;
    uint32_t* ev_neuron;
} orc_block;

#define P(b, i, c) ((REAL)(b)->props[(int64_t)(i) * 22 + (c)])

static void* xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

void* RSUFFIX(orc_block_create)(int64_t n, const float* props, int64_t nnz, const int64_t* rowptr,
                                const int32_t* post, const uint8_t* kind, const uint8_t* delay,
                                const float* w0, const float* w1, double dt, int32_t build_pull)
{
    orc_block* b = (orc_block*)xcalloc(1, sizeof(orc_block));
    b->n = n;
    b->nnz = nnz;
    b->dt = (REAL)dt;
    b->props = (float*)malloc(sizeof(float) * (size_t)n * 22);
    memcpy(b->props, props, sizeof(float) * (size_t)n * 22);
    b->rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    memcpy(b->rowptr, rowptr, sizeof(int64_t) * (size_t)(n + 1));
    b->post = (int32_t*)xcalloc((size_t)nnz, sizeof(int32_t));
    b->kind = (uint8_t*)xcalloc((size_t)nnz, 1);
    b->delay = (uint8_t*)xcalloc((size_t)nnz, 1);
    b->w0 = (float*)xcalloc((size_t)nnz, sizeof(float));
    b->w1 = (float*)xcalloc((size_t)nnz, sizeof(float));
    memcpy(b->post, post, sizeof(int32_t) * (size_t)nnz);
    memcpy(b->kind, kind, (size_t)nnz);
    int32_t D = 1;
    for (int64_t e = 0; e < nnz; ++e) {
        b->delay[e] = delay ? delay[e] : 1;
        if (b->delay[e] > D) D = b->delay[e];
    }
    b->D = D;
    memcpy(b->w0, w0, sizeof(float) * (size_t)nnz);
    memcpy(b->w1, w1 ? w1 : w0, sizeof(float) * (size_t)nnz);
    int32_t S = 0;
    b->sub = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        b->sub[i] = (int32_t)props[i * 22 + 3];
        if (b->sub[i] + 1 > S) S = b->sub[i] + 1;
    }
    b->S = S;
    if (build_pull) {
        b->prow = (int64_t*)xcalloc((size_t)n + 1, sizeof(int64_t));
        b->ppre = (int32_t*)xcalloc((size_t)nnz, sizeof(int32_t));
        b->pkind = (uint8_t*)xcalloc((size_t)nnz, 1);
        b->pdelay = (uint8_t*)xcalloc((size_t)nnz, 1);
        b->pw0 = (float*)xcalloc((size_t)nnz, sizeof(float));
        b->pw1 = (float*)xcalloc((size_t)nnz, sizeof(float));
        for (int64_t e = 0; e < nnz; ++e) b->prow[b->post[e] + 1]++;
        for (int64_t i = 0; i < n; ++i) b->prow[i + 1] += b->prow[i];
        int64_t* cur = (int64_t*)xcalloc((size_t)n, sizeof(int64_t));
        for (int64_t i = 0; i < n; ++i) cur[i] = b->prow[i];
        for (int64_t j = 0; j < n; ++j)
            for (int64_t e = rowptr[j]; e < rowptr[j + 1]; ++e) {
                int64_t q = cur[b->post[e]]++;
                b->ppre[q] = (int32_t)j;
                b->pkind[q] = b->kind[e];
                b->pdelay[q] = b->delay[e];
                b->pw0[q] = b->w0[e];
                b->pw1[q] = b->w1[e];
            }
        free(cur);
        b->shist = (uint8_t*)xcalloc((size_t)D * (size_t)n, 1);
    }
    b->V = (REAL*)xcalloc((size_t)n, sizeof(REAL));
    b->J = (REAL*)xcalloc((size_t)n * 4, sizeof(REAL));
    b->I_ui = (REAL*)xcalloc((size_t)n * 4, sizeof(REAL));
    b->I_syn = (REAL*)xcalloc((size_t)n, sizeof(REAL));
    b->t_last = (REAL*)xcalloc((size_t)n, sizeof(REAL));
    b->ring = (REAL*)xcalloc((size_t)D * 4 * (size_t)n, sizeof(REAL));
    b->active = (uint8_t*)xcalloc((size_t)n, 1);
    b->spike = (uint8_t*)xcalloc((size_t)n, 1);
    /* block.__init__ (models/block.py:55-60): t_last=0, V=(V_th+V_reset)/2, J=0, t=0, I_syn from that */
    for (int64_t i = 0; i < n; ++i) {
        b->V[i] = (P(b, i, 8) + P(b, i, 9)) / 2;
    }
    b->t = 0;
    b->step = 0;
    return b;
}

void RSUFFIX(orc_block_destroy)(void* h)
{
    orc_block* b = (orc_block*)h;
    if (!b) return;
    free(b->props); free(b->rowptr); free(b->post); free(b->kind); free(b->delay); free(b->w0); free(b->w1);
    free(b->prow); free(b->ppre); free(b->pkind); free(b->pdelay); free(b->pw0); free(b->pw1); free(b->shist);
    free(b->V); free(b->J); free(b->I_ui); free(b->I_syn); free(b->t_last); free(b->ring);
    free(b->active); free(b->spike); free(b->sub);
    free(b->srcflag); free(b->fire); free(b->ev_step); free(b->ev_neuron);
    free(b);
}

/*
 * One block.run() step (models/block.py:116-150).
 *   noise_mode 0: scalar rate p (block.run's noise_rate argument)
 *   noise_mode 1: per-neuron rate from property column 0 (the HIP engine,
 *                 cuda/src/stage/spike.cpp:156-157)
 *   uniforms: if non-NULL, the n uniforms of this step; else Philox(seed).
 *   pull != 0: accumulate W.s as a full pass over every synapse (what
 *   torch.sparse.mm does, models/block.py:76-81); else event-driven push.
 * Outputs (may be NULL): sum_act[S] (threshold spikes per sub_idx),
 * mean_v[S] (mean V per sub_idx, NaN for an empty sub_idx).
 */
#define DTB_STREAM_EXT 0x45585420u /* external drive: ctr = (i_lo, i_hi, step, EXT) */

uint32_t RSUFFIX(orc_ext_count)(uint64_t seed, uint64_t step, uint64_t i, int32_t n, const uint32_t* thr)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)step, DTB_STREAM_EXT};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    uint32_t c = 0;
    for (int32_t j = 0; j < n; ++j) c += r[0] >= thr[j];
    return c;
}

void RSUFFIX(orc_block_set_external)(void* h, int32_t n, const uint32_t* thr, float w)
{
    orc_block* b = (orc_block*)h;
    b->ext_n = n;
    b->ext_w = w;
    for (int32_t j = 0; j < n && j < 16; ++j) b->ext_thr[j] = thr[j];
}

void RSUFFIX(orc_block_step)(void* h, int32_t noise_mode, float noise_rate, const float* uniforms,
                             uint64_t seed, int32_t isolated, int32_t pull, REAL* sum_act, REAL* mean_v)
{
    orc_block* b = (orc_block*)h;
    const int64_t n = b->n;
    const REAL dt = b->dt;
    b->step += 1;
    b->t += dt; /* models/block.py:133 */
    const REAL t = b->t;
    const int64_t step = b->step;
    /* forced spikes of this step: src[:, iter] with iter = step - 1 (models/block.py:105) */
    if (b->srcflag) {
        memset(b->fire, 0, (size_t)n);
        for (int64_t k = 0; k < b->nev; ++k)
            if ((int64_t)b->ev_step[k] == step - 1) b->fire[b->ev_neuron[k]] = 1;
    }
    /* update_Vi, models/block.py:89-111 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        REAL V = b->V[i];
        REAL main_part = -P(b, i, 6) * (V - P(b, i, 7));
        REAL c_diff = main_part + b->I_syn[i] + P(b, i, 2);
        REAL dV = dt / P(b, i, 4) * c_diff;
        REAL Vn = V + dV;
        int not_sat = t >= b->t_last[i] + P(b, i, 5);
        REAL Vi = not_sat ? Vn : P(b, i, 9);
        uint8_t act = Vi >= P(b, i, 8);
        REAL vth = P(b, i, 8);
        REAL Vnew = Vi < vth ? Vi : vth;
        uint8_t sf = b->srcflag ? b->srcflag[i] : 0;
        if (sf & 1) {
            act = b->fire[i]; /* active[src_neuron] = src[:, iter] */
            if (sf & 2) Vnew = V;
        }
        b->V[i] = Vnew;
        b->active[i] = act;
        float u = uniforms ? uniforms[i] : orc_noise_uniform(seed, (uint64_t)step, (uint64_t)i);
        float p = noise_mode == 1 ? b->props[i * 22 + 0] : noise_rate;
        uint8_t noise = (u < p) && !(sf & 2);
        b->spike[i] = isolated ? noise : (uint8_t)(noise | act);
    }
    /* update_J_ui, models/block.py:68-74: J = J*exp(-dt/tau) + W.s */
    const int32_t D = b->D;
    if (!pull) {
        const int32_t slot_now = (int32_t)(step % D);
        for (int64_t j = 0; j < n; ++j) {
            if (!b->spike[j]) continue;
            for (int64_t e = b->rowptr[j]; e < b->rowptr[j + 1]; ++e) {
                int32_t slot = (int32_t)((step + b->delay[e] - 1) % D);
                REAL* r = b->ring + ((int64_t)slot * 4) * n;
                int c0 = b->kind[e] ? 2 : 0;
                int32_t q = b->post[e];
                r[(int64_t)c0 * n + q] += (REAL)b->w0[e];
                r[(int64_t)(c0 + 1) * n + q] += (REAL)b->w1[e];
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            REAL* r = b->ring + ((int64_t)slot_now * 4) * n;
            REAL ext = b->ext_n ? (REAL)RSUFFIX(orc_ext_count)(seed, (uint64_t)step, (uint64_t)i, b->ext_n,
                                                              b->ext_thr) * (REAL)b->ext_w : 0;
            for (int k = 0; k < 4; ++k) {
                REAL tau = P(b, i, 18 + k);
                REAL J = b->J[(int64_t)k * n + i] * (REAL)exp(-(double)dt / (double)tau);
                J += r[(int64_t)k * n + i];
                if (k < 2) J += ext;
                r[(int64_t)k * n + i] = 0;
                b->J[(int64_t)k * n + i] = J;
            }
        }
    } else {
        /* spike history ring: shist[(step) % D] holds this step's spikes */
        memcpy(b->shist + (int64_t)(step % D) * n, b->spike, (size_t)n);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            REAL acc[4] = {0, 0, 0, 0};
            for (int64_t e = b->prow[i]; e < b->prow[i + 1]; ++e) {
                int64_t src_step = step - b->pdelay[e] + 1;
                if (src_step < 1) continue;
                REAL s = (REAL)b->shist[(src_step % D) * n + b->ppre[e]];
                int c0 = b->pkind[e] ? 2 : 0;
                acc[c0] += (REAL)b->pw0[e] * s;
                acc[c0 + 1] += (REAL)b->pw1[e] * s;
            }
            REAL ext = b->ext_n ? (REAL)RSUFFIX(orc_ext_count)(seed, (uint64_t)step, (uint64_t)i, b->ext_n,
                                                              b->ext_thr) * (REAL)b->ext_w : 0;
            for (int k = 0; k < 4; ++k) {
                REAL tau = P(b, i, 18 + k);
                b->J[(int64_t)k * n + i] = b->J[(int64_t)k * n + i] * (REAL)exp(-(double)dt / (double)tau) + acc[k] +
                                           (k < 2 ? ext : 0);
            }
        }
    }
    /* update_I_syn, models/block.py:83-87 (with the new V); update_t_ik_last :113-114 */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        REAL V = b->V[i];
        REAL sum = 0;
        for (int k = 0; k < 4; ++k) {
            REAL I = P(b, i, 10 + k) * (P(b, i, 14 + k) - V) * b->J[(int64_t)k * n + i];
            b->I_ui[(int64_t)k * n + i] = I;
            sum += I;
        }
        b->I_syn[i] = sum;
        if (b->active[i]) b->t_last[i] = t;
    }
    /* per sub_idx statistics, models/block.py:144-150 */
    if (sum_act || mean_v) {
        int32_t S = b->S;
        REAL* cnt = (REAL*)xcalloc((size_t)S, sizeof(REAL));
        REAL* vs = (REAL*)xcalloc((size_t)S, sizeof(REAL));
        REAL* sz = (REAL*)xcalloc((size_t)S, sizeof(REAL));
        for (int64_t i = 0; i < n; ++i) {
            int32_t s = b->sub[i];
            if (s < 0) continue;
            cnt[s] += b->active[i];
            vs[s] += b->V[i];
            sz[s] += 1;
        }
        for (int32_t s = 0; s < S; ++s) {
            if (sum_act) sum_act[s] = cnt[s];
            if (mean_v) mean_v[s] = vs[s] / sz[s];
        }
        free(cnt); free(vs); free(sz);
    }
}

int32_t RSUFFIX(orc_block_num_sub)(void* h) { return ((orc_block*)h)->S; }
int32_t RSUFFIX(orc_block_max_delay)(void* h) { return ((orc_block*)h)->D; }

/* field: 0 V[n], 1 J[4n], 2 I_ui[4n], 3 I_syn[n], 4 t_last[n] (REAL); 5 active[n], 6 spike[n] (u8); 7 t (1 REAL) */
void RSUFFIX(orc_block_get)(void* h, int32_t field, void* dst)
{
    orc_block* b = (orc_block*)h;
    int64_t n = b->n;
    switch (field) {
    case 0: memcpy(dst, b->V, sizeof(REAL) * (size_t)n); break;
    case 1: memcpy(dst, b->J, sizeof(REAL) * (size_t)n * 4); break;
    case 2: memcpy(dst, b->I_ui, sizeof(REAL) * (size_t)n * 4); break;
    case 3: memcpy(dst, b->I_syn, sizeof(REAL) * (size_t)n); break;
    case 4: memcpy(dst, b->t_last, sizeof(REAL) * (size_t)n); break;
    case 5: memcpy(dst, b->active, (size_t)n); break;
    case 6: memcpy(dst, b->spike, (size_t)n); break;
    case 7: memcpy(dst, &b->t, sizeof(REAL)); break;
    default: break;
    }
}

/* New connectivity, state kept (push form only): w_uij replaced between
 * steps, as block.update_conn_weight does (models/block.py:172-177). */
void RSUFFIX(orc_block_replace_csr)(void* h, int64_t nnz, const int64_t* rowptr, const int32_t* post,
                                    const uint8_t* kind, const float* w0, const float* w1)
{
    orc_block* b = (orc_block*)h;
    free(b->post); free(b->kind); free(b->delay); free(b->w0); free(b->w1);
    b->nnz = nnz;
    memcpy(b->rowptr, rowptr, sizeof(int64_t) * (size_t)(b->n + 1));
    b->post = (int32_t*)xcalloc((size_t)nnz, sizeof(int32_t));
    b->kind = (uint8_t*)xcalloc((size_t)nnz, 1);
    b->delay = (uint8_t*)xcalloc((size_t)nnz, 1);
    b->w0 = (float*)xcalloc((size_t)nnz, sizeof(float));
    b->w1 = (float*)xcalloc((size_t)nnz, sizeof(float));
    memcpy(b->post, post, sizeof(int32_t) * (size_t)nnz);
    memcpy(b->kind, kind, (size_t)nnz);
    memset(b->delay, 1, (size_t)nnz);
    memcpy(b->w0, w0, sizeof(float) * (size_t)nnz);
    memcpy(b->w1, w1 ? w1 : w0, sizeof(float) * (size_t)nnz);
}

/* Forced-spike sources and their schedule: event k fires neuron[k] at the
 * step whose 0-based index is step[k]; frozen[k] != 0 marks an input-block
 * neuron (no membrane update, no noise).  n_src = 0 removes them. */
void RSUFFIX(orc_block_set_sources)(void* h, int64_t n_src, const uint32_t* sources, const uint8_t* frozen,
                                    int64_t n_events, const uint32_t* step, const uint32_t* neuron)
{
    orc_block* b = (orc_block*)h;
    free(b->srcflag); free(b->fire); free(b->ev_step); free(b->ev_neuron);
    b->srcflag = NULL; b->fire = NULL; b->ev_step = NULL; b->ev_neuron = NULL; b->nev = 0;
    if (n_src <= 0) return;
    b->srcflag = (uint8_t*)xcalloc((size_t)b->n, 1);
    b->fire = (uint8_t*)xcalloc((size_t)b->n, 1);
    for (int64_t k = 0; k < n_src; ++k) b->srcflag[sources[k]] = (uint8_t)(1 | (frozen && frozen[k] ? 2 : 0));
    b->nev = n_events;
    b->ev_step = (uint32_t*)xcalloc((size_t)n_events, sizeof(uint32_t));
    b->ev_neuron = (uint32_t*)xcalloc((size_t)n_events, sizeof(uint32_t));
    memcpy(b->ev_step, step, sizeof(uint32_t) * (size_t)n_events);
    memcpy(b->ev_neuron, neuron, sizeof(uint32_t) * (size_t)n_events);
}

/* Replace part of the state between steps (test infrastructure: re-aligns
 * the oracle with another implementation's state after a threshold tie so
 * the following steps are compared from the same starting point).
 * field: 0 V[n], 1 J[4n], 4 t_last[n] (REAL).  I_ui and I_syn are then
 * recomputed from V and J as update_I_syn does (models/block.py:83-87). */
void RSUFFIX(orc_block_set)(void* h, int32_t field, const void* src)
{
    orc_block* b = (orc_block*)h;
    int64_t n = b->n;
    switch (field) {
    case 0: memcpy(b->V, src, sizeof(REAL) * (size_t)n); break;
    case 1: memcpy(b->J, src, sizeof(REAL) * (size_t)n * 4); break;
    case 4: memcpy(b->t_last, src, sizeof(REAL) * (size_t)n); break;
    default: return;
    }
    for (int64_t i = 0; i < n; ++i) {
        REAL V = b->V[i];
        REAL sum = 0;
        for (int k = 0; k < 4; ++k) {
            REAL I = P(b, i, 10 + k) * (P(b, i, 14 + k) - V) * b->J[(int64_t)k * n + i];
            b->I_ui[(int64_t)k * n + i] = I;
            sum += I;
        }
        b->I_syn[i] = sum;
    }
}

/* Spike compaction restated from the CPU twin (cuda/src/block.cpp:1033-1055):
 * the ascending list of positions whose flag is set.  Returns the count. */
int64_t orc_compact(const uint8_t* flags, int64_t n, int32_t* out)
{
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i)
        if (flags[i]) out[k++] = (int32_t)i;
    return k;
}

/* ------------------------------------------------------------------ */
/* Balloon-Windkessel (models/bold_model_pytorch.py:43-78)             */
/* ------------------------------------------------------------------ */
/* par = {epsilon, tao_s, tao_f, tao_0, alpha, E_0, V_0, delta_t} */
void orc_bold_step(int64_t n, const double* par, double* s, double* q, double* v, double* f,
                   const double* u, double* out)
{
    const double eps = par[0], ts = par[1], tf = par[2], t0 = par[3], E0 = par[5], V0 = par[6], dt = par[7];
    const double div_alpha = 1.0 / par[4];
    for (int64_t i = 0; i < n; ++i) {
        double d_s = eps * u[i] - s[i] / ts - (f[i] - 1) / tf;
        double q_part = f[i] > 0 ? 1 - pow(1 - E0, 1 / f[i]) : 1.0;
        double dq = (f[i] * q_part / E0 - q[i] * pow(v[i], div_alpha - 1)) / t0;
        double dv = (f[i] - pow(v[i], div_alpha)) / t0;
        q[i] += dq * dt;
        v[i] += dv * dt;
        f[i] += s[i] * dt;
        f[i] = f[i] > 0 ? f[i] : 0.0;
        s[i] += d_s * dt;
        out[i] = V0 * (7 * E0 * (1 - q[i]) + 2 * (1 - q[i] / v[i]) + (2 * E0 - 0.2) * (1 - v[i]));
    }
}

/* ------------------------------------------------------------------ */
/* Diffusion EnKF (data_assimilation/DataAssimilation.py:40-60)         */
/* ------------------------------------------------------------------ */
/* w: [E][V][S] in/out; y: [V]; xi: [E][V] standard normal draws. */
void orc_diffusion_enkf(int32_t E, int32_t V, int32_t S, double* w, double sigma, const double* y,
                        double rho, const double* xi)
{
    size_t tot = (size_t)E * V * S;
    double* wh = (double*)malloc(sizeof(double) * tot);
    double* wd = (double*)malloc(sizeof(double) * tot);
    double* mean = (double*)calloc((size_t)V * S, sizeof(double));
    double* temp = (double*)malloc(sizeof(double) * (size_t)E * V);
    double* innov = (double*)malloc(sizeof(double) * (size_t)E * V);
    double* cxx = (double*)calloc((size_t)V, sizeof(double));
    double* A = (double*)calloc((size_t)E * E, sizeof(double));
    memcpy(wh, w, sizeof(double) * tot);
#define W3(a, e, v, s) a[((size_t)(e) * V + (v)) * S + (s)]
    for (int e = 0; e < E; ++e)
        for (int v = 0; v < V; ++v)
            for (int s = 0; s < S; ++s) mean[(size_t)v * S + s] += W3(wh, e, v, s);
    for (size_t k = 0; k < (size_t)V * S; ++k) mean[k] /= E;
    for (int e = 0; e < E; ++e)
        for (int v = 0; v < V; ++v)
            for (int s = 0; s < S; ++s) W3(wd, e, v, s) = W3(wh, e, v, s) - mean[(size_t)v * S + s];
    for (int v = 0; v < V; ++v) {
        double acc = 0;
        for (int e = 0; e < E; ++e) acc += W3(wd, e, v, S - 1) * W3(wd, e, v, S - 1);
        cxx[v] = acc / (E - 1) + sigma;
    }
    for (int e = 0; e < E; ++e)
        for (int v = 0; v < V; ++v) {
            temp[(size_t)e * V + v] = W3(wd, e, v, S - 1) / cxx[v] / (E - 1);
            innov[(size_t)e * V + v] = y[v] + sqrt(sigma) * xi[(size_t)e * V + v] - W3(wh, e, v, S - 1);
        }
    /* local (solo) term */
    for (int v = 0; v < V; ++v)
        for (int s = 0; s < S; ++s) {
            double g = 0;
            for (int e = 0; e < E; ++e) g += temp[(size_t)e * V + v] * W3(wd, e, v, s);
            for (int e = 0; e < E; ++e) W3(w, e, v, s) += rho * innov[(size_t)e * V + v] * g;
        }
    /* global term: ((innov @ temp^T) / V) @ wd */
    for (int e = 0; e < E; ++e)
        for (int f2 = 0; f2 < E; ++f2) {
            double acc = 0;
            for (int v = 0; v < V; ++v) acc += innov[(size_t)e * V + v] * temp[(size_t)f2 * V + v];
            A[(size_t)e * E + f2] = acc / V;
        }
    for (int e = 0; e < E; ++e)
        for (int v = 0; v < V; ++v)
            for (int s = 0; s < S; ++s) {
                double acc = 0;
                for (int f2 = 0; f2 < E; ++f2) acc += A[(size_t)e * E + f2] * W3(wd, f2, v, s);
                W3(w, e, v, s) += (1 - rho) * acc;
            }
#undef W3
    free(wh); free(wd); free(mean); free(temp); free(innov); free(cxx); free(A);
}

int32_t orc_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
