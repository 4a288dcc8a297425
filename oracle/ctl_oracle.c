/*
 * ctl_oracle.c -- CPU oracle for the tau-leaping hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  See ctl_oracle.h for the two precisions it runs in.
 *
 * Every function below cites the lines of the reference
 * (2phase_n-tau-leaping.cu = T.cu, kernel.cu = K.cu) whose behaviour it
 * restates.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off -mfma).
 * FP contraction must stay off: the REPLAY precision is compared bit for
 * bit with the gfx950 kernel, which is built the same way, and every fused
 * multiply-add on either side is an explicit fmaf().
 */
#include "ctl_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ------------------------------------------------------------------ */
/* Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: */
/* as easy as 1, 2, 3", SC'11.  Round: (hi,lo) = M*c; key bump Weyl.     */
/* ------------------------------------------------------------------ */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One event stream: counter = {block, event, sim, 0}; the 4 words of a block
 * are consumed in order x, y, z, w. */
typedef struct {
    uint32_t key[2];
    uint32_t sim, ev, blk;
    uint32_t buf[4];
    int pos;
    uint32_t tag;        /* counter word 3: 0, or 1 + lane for the group mode's lane streams */
} orc_rng;

static inline void rng_begin(orc_rng* g, uint32_t ev) { g->ev = ev; g->blk = 0; g->pos = 4; }

static inline uint32_t rng_next(orc_rng* g) {
    if (g->pos == 4) {
        uint32_t ctr[4] = {g->blk, g->ev, g->sim, g->tag};
        orc_philox(ctr, g->key, g->buf);
        g->blk++;
        g->pos = 0;
    }
    return g->buf[g->pos++];
}

/* U(0,1] and U[0,1) on the 2^-24 grid (exact in float). */
static inline float u01_oc(orc_rng* g) { return (float)((rng_next(g) >> 8) + 1u) * 0x1p-24f; }
static inline float u01_co(orc_rng* g) { return (float)(rng_next(g) >> 8) * 0x1p-24f; }

/* ------------------------------------------------------------------ */
/* FP32 log / exp written from the textbook method: Cody-Waite argument  */
/* reduction by ln 2 (split in a 16-bit head and a tail) and the atanh   */
/* (log) / Taylor (exp) series with exact rational coefficients.         */
/* ------------------------------------------------------------------ */
#define ORC_LN2_HI 0x1.62e4p-1f          /* 45426 / 65536 */
#define ORC_LN2_LO 0x1.7f7d1cp-20f       /* ln2 - LN2_HI */
#define ORC_LOG2E  0x1.715476p+0f

/* log(x) for positive normal x: x = 2^e * (1+f), f in (sqrt(1/2)-1, sqrt(2)-1];
 * s = f/(2+f); log(1+f) = f - (f^2/2 - s*(f^2/2 + R(s^2))), R(z) = 2z/3 + 2z^2/5
 * + 2z^3/7 + 2z^4/9 (the atanh series). */
float orc_logf(float x) {
    uint32_t ix = f2u(x);
    int e = (int)(ix >> 23) - 127;
    uint32_t mb = (ix & 0x007fffffu) | 0x3f800000u;
    if (mb > 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
    float f = u2f(mb) - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f);
    R = fmaf(z, R, 0x1.99999ap-2f);
    R = fmaf(z, R, 0x1.555556p-1f);
    R = R * z;
    float hfsq = 0.5f * f * f;
    float ef = (float)e;
    return ef * ORC_LN2_HI + (f - (hfsq - (s * (hfsq + R) + ef * ORC_LN2_LO)));
}

/* exp(x), x clamped to [-87, 88]: k = rint(x log2 e), r = x - k ln2 (two steps),
 * exp(r) by the degree-8 Taylor polynomial, scaled by 2^k. */
float orc_expf(float x) {
    if (x < -87.0f) x = -87.0f;
    if (x > 88.0f) x = 88.0f;
    float kf = rintf(x * ORC_LOG2E);
    float r = (x - kf * ORC_LN2_HI) - kf * ORC_LN2_LO;
    float p = fmaf(r, 0x1.a01a02p-16f, 0x1.a01a02p-13f);
    p = fmaf(r, p, 0x1.6c16c2p-10f);
    p = fmaf(r, p, 0x1.111112p-7f);
    p = fmaf(r, p, 0x1.555556p-5f);
    p = fmaf(r, p, 0x1.555556p-3f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    int k = (int)kf;
    return p * u2f((uint32_t)(k + 127) << 23);
}

/* log(1 - v) for v in [0, 1): the log series above with f = -v exactly while
 * 1 - v > sqrt(1/2), so small arguments keep full relative precision. */
float orc_log1m(float v) {
    if (v < 0x1.2bec34p-2f) {
        float f = -v;
        float s = f / (2.0f + f);
        float z = s * s;
        float R = fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f);
        R = fmaf(z, R, 0x1.99999ap-2f);
        R = fmaf(z, R, 0x1.555556p-1f);
        R = R * z;
        float hfsq = 0.5f * f * f;
        return f - (hfsq - s * (hfsq + R));
    }
    return orc_logf(1.0f - v);
}

/* Exponential(1) variate -log(1 - V): V = r * 2^-32 below 2^-8 (exact) and
 * (r >> 8) * 2^-24 above, so short gaps keep 32-bit resolution and V < 1. */
static inline float exp_rv(orc_rng* g) {
    uint32_t r = rng_next(g);
    float v = r < 0x01000000u ? (float)r * 0x1p-32f : (float)(r >> 8) * 0x1p-24f;
    return -orc_log1m(v);
}

/* ------------------------------------------------------------------ */
/* Poisson sampler (replaces curand_poisson, T.cu:2018, 2040).           */
/*  lam < 10        : inversion by sequential search from 0, one U[0,1);  */
/*                    a search that outruns 63 terms (float cumulative    */
/*                    stalled below u) restarts with a fresh uniform.     */
/*  10 <= lam < 2^23: PTRS, W. Hormann, "The transformed rejection method */
/*                    for generating Poisson random variables", Insurance:*/
/*                    Mathematics and Economics 12 (1993) 39-45; log pmf  */
/*                    via Loader's bd0 deviance + Stirling tail (float-   */
/*                    stable for large lam).                              */
/*  lam >= 2^23     : normal approximation (Marsaglia polar), in double.  */
/* ------------------------------------------------------------------ */
static const float ORC_RECIP[64] = {
    0.0f,        1.0f,        1.0f / 2,  1.0f / 3,  1.0f / 4,  1.0f / 5,  1.0f / 6,  1.0f / 7,
    1.0f / 8,    1.0f / 9,    1.0f / 10, 1.0f / 11, 1.0f / 12, 1.0f / 13, 1.0f / 14, 1.0f / 15,
    1.0f / 16,   1.0f / 17,   1.0f / 18, 1.0f / 19, 1.0f / 20, 1.0f / 21, 1.0f / 22, 1.0f / 23,
    1.0f / 24,   1.0f / 25,   1.0f / 26, 1.0f / 27, 1.0f / 28, 1.0f / 29, 1.0f / 30, 1.0f / 31,
    1.0f / 32,   1.0f / 33,   1.0f / 34, 1.0f / 35, 1.0f / 36, 1.0f / 37, 1.0f / 38, 1.0f / 39,
    1.0f / 40,   1.0f / 41,   1.0f / 42, 1.0f / 43, 1.0f / 44, 1.0f / 45, 1.0f / 46, 1.0f / 47,
    1.0f / 48,   1.0f / 49,   1.0f / 50, 1.0f / 51, 1.0f / 52, 1.0f / 53, 1.0f / 54, 1.0f / 55,
    1.0f / 56,   1.0f / 57,   1.0f / 58, 1.0f / 59, 1.0f / 60, 1.0f / 61, 1.0f / 62, 1.0f / 63};

static const float ORC_LOGFACT[10] = {
    0.0f, 0.0f, 0x1.62e43p-1f, 0x1.cab0cp+0f, 0x1.96ca78p+1f,
    0x1.326644p+2f, 0x1.a51274p+2f, 0x1.10ce2p+3f, 0x1.5358e8p+3f, 0x1.99a892p+3f};

/* bd0(k, lam) = k log(k/lam) + lam - k, diff = k - lam computed exactly. */
static float bd0f(float kf, float lam, float diff) {
    float sum = kf + lam;
    if (fabsf(diff) < 0.1f * sum) {
        float v = diff / sum;
        float s = diff * v;
        float ej = 2.0f * kf * v;
        float v2 = v * v;
        ej = ej * v2; s = s + ej * 0x1.555556p-2f;
        ej = ej * v2; s = s + ej * 0x1.99999ap-3f;
        ej = ej * v2; s = s + ej * 0x1.24924ap-3f;
        ej = ej * v2; s = s + ej * 0x1.c71c72p-4f;
        ej = ej * v2; s = s + ej * 0x1.745d18p-4f;
        return s;
    }
    return (kf * orc_logf(kf / lam) + lam) - kf;
}

static int32_t poisson_ptrs(orc_rng* g, float lam) {
    float slam = sqrtf(lam);
    float loglam = orc_logf(lam);
    float b = 0.931f + 2.53f * slam;
    float a = -0.059f + 0.02483f * b;
    float invalpha = 1.1239f + 1.1328f / (b - 3.4f);
    float vr = 0.9277f - 3.6224f / (b - 2.0f);
    float li = floorf(lam);
    float lf = lam - li;
    for (;;) {
        float U = u01_oc(g) - 0.5f;
        float V = u01_oc(g);
        float us = 0.5f - fabsf(U);
        if (us < 0.013f && V > us) continue;
        float xk = (2.0f * a / us + b) * U + 0.43f;
        float d = floorf(xk + lf);
        if (us >= 0.07f && V <= vr) return (int32_t)li + (int32_t)d;
        if (d < -li) continue;
        if (d > 1073741824.0f) continue;
        int32_t k = (int32_t)li + (int32_t)d;
        float lhs = orc_logf(V * invalpha / (a / (us * us) + b));
        float rhs;
        if (k < 10) {
            rhs = (-lam + (float)k * loglam) - ORC_LOGFACT[k];
        } else {
            float kf = (float)k;
            float kinv = 1.0f / kf;
            float k2 = kinv * kinv;
            float stirl = (0x1.555556p-4f - (0x1.6c16c2p-9f - 0x1.a01a02p-11f * k2) * k2) * kinv;
            rhs = (-bd0f(kf, lam, d - lf) - 0.5f * orc_logf(0x1.921fb6p+2f * kf)) - stirl;
        }
        if (lhs <= rhs) return k;
    }
}

static int32_t poisson_normal(orc_rng* g, float lam) {
    if (!(lam < 2147483648.0f)) return 2147483647;   /* also +inf */
    float u, v, s;
    do {
        u = 2.0f * u01_oc(g) - 1.0f;
        v = 2.0f * u01_oc(g) - 1.0f;
        s = u * u + v * v;
    } while (s >= 1.0f || s == 0.0f);
    float z = u * sqrtf(-2.0f * orc_logf(s) / s);
    double kd = floor((double)lam + sqrt((double)lam) * (double)z + 0.5);
    if (kd < 0.0) return 0;
    if (kd > 2147483647.0) return 2147483647;
    return (int32_t)kd;
}

static int32_t poisson(orc_rng* g, float lam) {
    if (!(lam > 0.0f)) return 0;
    if (lam < 10.0f) {
        float p0 = orc_expf(-lam);
        for (;;) {
            float u = u01_co(g);
            float p = p0, F = p0;
            int k = 0;
            while (u >= F) {
                ++k;
                if (k > 63) break;
                p = (p * lam) * ORC_RECIP[k];
                F = F + p;
            }
            if (k <= 63) return k;
        }
    }
    if (lam < 8388608.0f) return poisson_ptrs(g, lam);
    return poisson_normal(g, lam);
}

void orc_poisson_stream(uint32_t seed_lo, uint32_t seed_hi, uint32_t sim, uint32_t ev,
                        float lam, int count, int32_t* out) {
    orc_rng g = {{seed_lo, seed_hi}, sim, 0, 0, {0, 0, 0, 0}, 4, 0u};
    rng_begin(&g, ev);
    for (int i = 0; i < count; ++i) out[i] = poisson(&g, lam);
}

void orc_uniform_stream(uint32_t seed_lo, uint32_t seed_hi, uint32_t sim, uint32_t ev,
                        int open_closed, int count, float* out) {
    orc_rng g = {{seed_lo, seed_hi}, sim, 0, 0, {0, 0, 0, 0}, 4, 0u};
    rng_begin(&g, ev);
    for (int i = 0; i < count; ++i) out[i] = open_closed ? u01_oc(&g) : u01_co(&g);
}

/* ------------------------------------------------------------------ */
/* Model compile: restates read_stoichiometry, T.cu:122-457.            */
/* ------------------------------------------------------------------ */
void orc_model_free(orc_model* m) {
    free(m->Lptr); free(m->Ls); free(m->Lv);
    free(m->Vptr); free(m->Vr); free(m->Vv);
    free(m->Fptr); free(m->Fs); free(m->Fv);
    free(m->FSptr); free(m->FSr); free(m->FSv);
    free(m->hor); free(m->hort);
    memset(m, 0, sizeof(*m));
}

int orc_compile(const orc_net* net, orc_model* m) {
    const int S = net->S, R = net->R;
    memset(m, 0, sizeof(*m));
    if (S <= 0 || R <= 0) return -1;
    m->S = S; m->R = R;
    int* var = (int*)calloc((size_t)R * S, sizeof(int));
    int* fear = (int*)calloc((size_t)R * S, sizeof(int));
    for (int r = 0; r < R; ++r)
        for (int s = 0; s < S; ++s) {
            int l = net->left[r * S + s];
            if (l < 0 || net->right[r * S + s] < 0) { free(var); free(fear); return -1; }
            /* T.cu:270-286: var = right - left; fear zeroes feed-locked columns */
            var[r * S + s] = net->right[r * S + s] - l;
            fear[r * S + s] = (net->locked && net->locked[s]) ? 0 : var[r * S + s];
            if (l) m->nL++;
            if (var[r * S + s]) m->nV++;
            if (fear[r * S + s]) m->nF++;
        }
    m->Lptr = (int*)malloc(sizeof(int) * (R + 1));
    m->Ls = (int*)malloc(sizeof(int) * (m->nL + 1));
    m->Lv = (int*)malloc(sizeof(int) * (m->nL + 1));
    m->Fptr = (int*)malloc(sizeof(int) * (R + 1));
    m->Fs = (int*)malloc(sizeof(int) * (m->nF + 1));
    m->Fv = (int*)malloc(sizeof(int) * (m->nF + 1));
    m->Vptr = (int*)malloc(sizeof(int) * (S + 1));
    m->Vr = (int*)malloc(sizeof(int) * (m->nV + 1));
    m->Vv = (int*)malloc(sizeof(int) * (m->nV + 1));
    int nl = 0, nf = 0, nv = 0;
    for (int r = 0; r < R; ++r) {                 /* T.cu:219-232, 323-335 */
        m->Lptr[r] = nl; m->Fptr[r] = nf;
        for (int s = 0; s < S; ++s) {
            if (net->left[r * S + s]) { m->Ls[nl] = s; m->Lv[nl] = net->left[r * S + s]; nl++; }
            if (fear[r * S + s]) { m->Fs[nf] = s; m->Fv[nf] = fear[r * S + s]; nf++; }
        }
    }
    m->Lptr[R] = nl; m->Fptr[R] = nf;
    for (int s = 0; s < S; ++s) {                 /* T.cu:308-320 */
        m->Vptr[s] = nv;
        for (int r = 0; r < R; ++r)
            if (var[r * S + s]) { m->Vr[nv] = r; m->Vv[nv] = var[r * S + s]; nv++; }
    }
    m->Vptr[S] = nv;
    /* fear entries species-major, ascending reaction (group algorithm update) */
    m->FSptr = (int*)malloc(sizeof(int) * (S + 1));
    m->FSr = (int*)malloc(sizeof(int) * (m->nF + 1));
    m->FSv = (int*)malloc(sizeof(int) * (m->nF + 1));
    {
        int nfs = 0;
        for (int sp = 0; sp < S; ++sp) {
            m->FSptr[sp] = nfs;
            for (int r = 0; r < R; ++r)
                if (fear[r * S + sp]) { m->FSr[nfs] = r; m->FSv[nfs] = fear[r * S + sp]; nfs++; }
        }
        m->FSptr[S] = nfs;
    }
    /* HOR / HORtype, T.cu:402-432: reaction order = sum of left coefficients
     * (unsigned char); the type is overwritten when a higher order is found and
     * then raised to any larger left coefficient (no braces in the reference). */
    m->hor = (unsigned char*)calloc(S, 1);
    m->hort = (unsigned char*)calloc(S, 1);
    unsigned char* order = (unsigned char*)calloc(R, 1);
    for (int r = 0; r < R; ++r) {
        unsigned char tot = 0;
        for (int s = 0; s < S; ++s) tot = (unsigned char)(tot + net->left[r * S + s]);
        order[r] = tot;
    }
    for (int s = 0; s < S; ++s)
        for (int r = 0; r < R; ++r) {
            int l = net->left[r * S + s];
            if (l > 0 && order[r] > m->hor[s]) { m->hor[s] = order[r]; m->hort[s] = (unsigned char)l; }
            if ((int)m->hort[s] < l) m->hort[s] = (unsigned char)l;
        }
    free(order); free(var); free(fear);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Per-simulation state.                                                */
/* ------------------------------------------------------------------ */
enum { MODE_LEAP = 0, MODE_SSA = 1, MODE_DONE = -1 };

typedef struct {
    const orc_model* m;
    const orc_job* J;
    const double* c;      /* rates row */
    float* cf;            /* rates as float (REPLAY) */
    int32_t *X, *X1;
    float* Af;
    double* Ad;
    unsigned char* C;
    uint32_t* out;        /* K*SS samples of this simulation */
    const uint32_t* fv;   /* feed values of this simulation (nfs) */
    double t, tmax;
    uint32_t cursor, evc;
    int mode;
    orc_rng g;
    orc_stat st;
} simctx;

/* ABSORB, T.cu:1665-1697 (leap) and T.cu:2165-2196 (SSA): the remaining
 * samples copy the last RECORDED sample, not the current state. */
static void absorb(simctx* s) {
    const orc_job* J = s->J;
    const int SS = J->SS, K = J->K;
    if (s->cursor == 0) {
        for (int sk = 0; sk < SS; ++sk) s->out[sk] = (uint32_t)s->X[J->cs[sk]];
        s->cursor = 1;
    }
    const uint32_t* last = s->out + (size_t)(s->cursor - 1) * SS;
    for (int i = (int)s->cursor; i < K; ++i)
        for (int sk = 0; sk < SS; ++sk) s->out[(size_t)i * SS + sk] = last[sk];
    s->t = s->tmax + (double)0.1f;
    s->cursor = (uint32_t)K;
    s->mode = MODE_DONE;
    s->st.absorbed = 1;
}

/* Linear interpolation of the samples crossed by a leap, check_time_instant
 * T.cu:1507-1562; the reference's float->unsigned conversion of a negative
 * rint() saturates to 0, so a decreasing species records its start value. */
static void leap_samples(simctx* s, double tau) {
    const orc_job* J = s->J;
    const int SS = J->SS;
    uint32_t i = s->cursor;
    while (i < (uint32_t)J->K && s->t + tau >= J->ts[i]) {
        for (int sk = 0; sk < SS; ++sk) {
            int sp = J->cs[sk];
            int32_t d = s->X1[sp] - s->X[sp];
            uint32_t val = (uint32_t)s->X[sp];
            if (d > 0) {
                double delta = ((double)d * (J->ts[i] - s->t)) / tau;
                val += (uint32_t)rint(delta);
            }
            s->out[(size_t)i * SS + sk] = val;
        }
        ++i;
    }
    s->cursor = i;
}

/* Piecewise-constant sampling inside SSA, T.cu:2213-2236 (strict '>'). */
static void ssa_samples(simctx* s, double tn) {
    const orc_job* J = s->J;
    const int SS = J->SS;
    uint32_t i = s->cursor;
    while (i < (uint32_t)J->K && tn > J->ts[i]) {
        for (int sk = 0; sk < SS; ++sk) s->out[(size_t)i * SS + sk] = (uint32_t)s->X[J->cs[sk]];
        ++i;
    }
    s->cursor = i;
}

/* Incremental update + negativity test over the fear flat in reaction-major
 * order, test_update T.cu:1300-1351 (active branch T.cu:1320-1331).  Counts
 * are drawn reaction by reaction from the attempt's event stream; the first
 * negative partial sum rejects (the remaining draws are not needed: a retry
 * opens a new event).  Overflow past INT32_MAX also rejects.
 *
 * REFSEM draws an independent Poisson(tau a_j) for every non-critical
 * reaction, like the reference (curand_poisson, T.cu:2011-2023).  REPLAY
 * (the GPU algorithm) draws independent Poisson(lambda_j) only for
 * lambda_j >= 1; the reactions with lambda_j < 1 share one rate-1 Poisson
 * process whose events are walked lazily in reaction order: reaction j owns
 * the interval [L_{j-1}, L_j) of the running sum L of their lambdas and fires
 * once per event in it.  Counts in disjoint intervals are independent
 * Poisson(lambda_j), so the law is the same with far fewer draws.  The
 * running sum and the event positions are kept in double. */
static int apply_leap(simctx* s, int first, int pick, float tauf, double taud, int use_double,
                      uint64_t* fired) {
    const orc_model* m = s->m;
    memcpy(s->X1, s->X, sizeof(int32_t) * m->S);
    *fired = 0;
    double cum = 0.0, pos = -1.0;
    for (int j = 0; j < m->R; ++j) {
        int32_t k;
        if (s->C[j]) {
            k = (first && j == pick) ? 1 : 0;
        } else if (use_double) {
            k = poisson(&s->g, (float)(taud * s->Ad[j]));
            s->st.poisson++;
        } else {
            float lam = tauf * s->Af[j];
            if (lam < 1.0f) {
                k = 0;
                if (lam > 0.0f) {
                    if (pos < 0.0) { pos = (double)exp_rv(&s->g); s->st.gaps++; }
                    cum = cum + (double)lam;
                    while (pos < cum) { ++k; pos = pos + (double)exp_rv(&s->g); s->st.gaps++; }
                }
            } else {
                k = poisson(&s->g, lam);
                s->st.poisson++;
            }
        }
        if (k == 0) continue;
        *fired += (uint64_t)k;
        for (int h = m->Fptr[j]; h < m->Fptr[j + 1]; ++h) {
            int64_t nx = (int64_t)s->X1[m->Fs[h]] + (int64_t)m->Fv[h] * (int64_t)k;
            if (nx < 0 || nx > 2147483647LL) return 0;
            s->X1[m->Fs[h]] = (int32_t)nx;
        }
    }
    return 1;
}

static void swap_state(simctx* s) { int32_t* tmp = s->X; s->X = s->X1; s->X1 = tmp; }

/* Per-simulation feeds (read_feeds T.cu:459-567): after every accepted leap
 * each feed species is reset to the simulation's feed value (T.cu:2072-2079).
 * SSA steps apply no feed: the reference's SSA kernels update with the fear
 * matrix under USE_FEAR_SSA (T.cu:38, 2240-2244, 2420-2424), whose branch has
 * no feed block (the blocks at T.cu:2253-2258 and 2434-2440 sit in its #else).
 * The feed row is the simulation's own row (the reference reads row tid,
 * T.cu:2077, a per-block index; the loader stores rows per simulation,
 * T.cu:529). */
static void apply_feeds(simctx* s) {
    for (int u = 0; u < s->J->nfs; ++u) s->X[s->J->fs[u]] = (int32_t)s->fv[u];
}

/* ======================= REPLAY precision (FP32) ====================== */

/* CalculatePropensity<DEV_FLATTEN_MATRIX_CONST>, T.cu:1443-1470, in FP32;
 * criticality of T.cu:1725-1735 fused in: 0 < floor(X/v) < nc. */
static float prop_one_f32(simctx* s, int j, unsigned char* crit_out) {
    const orc_model* m = s->m;
    const uint32_t nc = (uint32_t)s->J->nc;
    float a = s->cf[j];
    int crit = 0;
    for (int h = m->Lptr[j]; h < m->Lptr[j + 1]; ++h) {
        int32_t x = s->X[m->Ls[h]];
        int v = m->Lv[h];
        if (v == 1) {
            a = a * (float)x;
        } else if (v == 2) {
            a = a * ((x < 2) ? 0.0f : (0.5f * (float)x) * (float)(x - 1));
        } else if (x < v) {
            a = a * 0.0f;
        } else {
            for (int e = 1; e <= v; ++e) a = a * ((float)(x - e + 1) / (float)e);
        }
        uint32_t q = (uint32_t)x / (uint32_t)v;
        if (q > 0u && q < nc) crit = 1;
    }
    *crit_out = (unsigned char)crit;
    return a;
}

static float prop_f32(simctx* s, float* a0c) {
    const orc_model* m = s->m;
    float a0 = 0.0f, ac = 0.0f;
    for (int j = 0; j < m->R; ++j) {
        s->Af[j] = prop_one_f32(s, j, &s->C[j]);
        a0 = a0 + s->Af[j];
    }
    for (int j = 0; j < m->R; ++j)            /* T.cu:1945-1952 */
        if (s->C[j]) ac = ac + s->Af[j];
    if (a0c) *a0c = ac;
    return a0;
}

/* g_i of Cao et al. from HOR/HORtype, T.cu:1741-1789, in FP32.  For X < 34
 * every branch gives eps*X/g < 1 so the max(., 1) below is 1 and the
 * reference's unsigned wrap-around at X <= 2 cannot matter. */
static float g_f32(unsigned char hor, unsigned char hort, int32_t x) {
    float xf = (float)x;
    switch (hor) {
        case 1: return 1.0f;
        case 2: return hort == 2 ? 2.0f + 1.0f / (xf - 1.0f) : 2.0f;
        case 3:
            if (hort == 3) return 3.0f + 1.0f / (xf - 1.0f) + 2.0f / (xf - 2.0f);
            if (hort == 2) return 3.0f + 1.5f / (xf - 1.0f);
            return 3.0f;
        default: return 1.0f;
    }
}

enum { LEAP_DONE = 0, LEAP_OK = 1, LEAP_SWITCH = 2, LEAP_ABSORB = 3 };

/* One adaptive leap attempt, PreCalculations T.cu:1568-2093 (Appendix A). */
static int leap_f32(simctx* s) {
    const orc_model* m = s->m;
    const orc_job* J = s->J;
    if (s->t >= s->tmax) { s->mode = MODE_DONE; return LEAP_DONE; }   /* T.cu:1624-1628 */
    s->mode = MODE_LEAP;
    float a0c;
    float a0 = prop_f32(s, &a0c);
    if (a0 == 0.0f) { absorb(s); return LEAP_ABSORB; }                /* T.cu:1665-1697 */
    /* tau' over the species-major var flat, T.cu:1860-1922 (quirk: every
     * species with a non-zero var column, feed-locked included). */
    float tau1 = FLT_MAX;
    for (int sp = 0; sp < m->S; ++sp) {
        float mu = 0.0f, sg = 0.0f;
        for (int h = m->Vptr[sp]; h < m->Vptr[sp + 1]; ++h) {
            int j = m->Vr[h];
            if (s->C[j]) continue;
            int v = m->Vv[h];
            mu = mu + (float)v * s->Af[j];
            sg = sg + (float)(v * v) * s->Af[j];
        }
        if (mu == 0.0f || sg == 0.0f) continue;
        float gg = g_f32(m->hor[sp], m->hort[sp], s->X[sp]);
        float tmp2 = (J->eps * (float)s->X[sp]) / gg;
        float macs = tmp2 > 1.0f ? tmp2 : 1.0f;
        float t3 = macs / fabsf(mu);
        float t4 = (macs * macs) / sg;
        float tmp = t3 < t4 ? t3 : t4;
        if (tmp < tau1) tau1 = tmp;
    }
    if (tau1 < J->ssa_factor / a0) { s->mode = MODE_SSA; s->st.switches++; return LEAP_SWITCH; }  /* T.cu:1935 */
    rng_begin(&s->g, s->evc++);
    float tau2 = FLT_MAX;
    if (a0c > 0.0f) tau2 = exp_rv(&s->g) / a0c;                         /* T.cu:1962-1965 */
    float tau = tau1 < tau2 ? tau1 : tau2;
    int pick = -1;
    if (tau1 > tau2) {                                                 /* T.cu:1993-2006 */
        float w = u01_oc(&s->g) * a0c;
        float acc = s->Af[0];
        int r = 0;
        while (w > acc && r < m->R - 1) { ++r; if (s->C[r]) acc = acc + s->Af[r]; }
        pick = r;
    }
    uint64_t fired;
    for (int attempt = 0;; ++attempt) {                                /* T.cu:2011-2050 */
        if (attempt > 0) { rng_begin(&s->g, s->evc++); s->st.rejections++; }
        if (apply_leap(s, attempt == 0, pick, tau, 0.0, 0, &fired)) break;
        tau = tau * 0.5f;
    }
    leap_samples(s, (double)tau);                                      /* T.cu:2065 */
    swap_state(s);                                                     /* T.cu:2068-2070 */
    apply_feeds(s);
    s->t = s->t + (double)tau;                                         /* T.cu:2084 */
    s->st.firings += fired;
    s->st.leaps++;
    return LEAP_OK;
}

/* SSA burst, EventualSSASteps T.cu:2099-2273 / LightweightSSA T.cu:2280-2454. */
static void ssa_f32(simctx* s) {
    const orc_model* m = s->m;
    const orc_job* J = s->J;
    rng_begin(&s->g, s->evc++);
    double tacc = 0.0;
    for (int step = 0; step < J->ssa_steps; ++step) {
        float a0 = prop_f32(s, NULL);
        if (a0 == 0.0f) { absorb(s); return; }
        float r1 = u01_oc(&s->g) * a0;
        int j = 0;
        float acc = s->Af[0];
        while (r1 > acc && j < m->R - 1) { ++j; acc = acc + s->Af[j]; }
        float dt = exp_rv(&s->g) / a0;
        tacc = tacc + (double)dt;
        double tn = s->t + tacc;
        ssa_samples(s, tn);
        for (int h = m->Fptr[j]; h < m->Fptr[j + 1]; ++h)
            s->X[m->Fs[h]] = (int32_t)((uint32_t)s->X[m->Fs[h]] + (uint32_t)m->Fv[h]);
        s->st.ssa_steps++;
        if (tn <= s->tmax) s->st.firings++;
    }
    s->t = s->t + tacc;
}

/* ===================== REFSEM precision (reference types) ================ */

/* CalculatePropensity, T.cu:1443-1470: double a_j, unsigned X, the v=2
 * binomial in float (T.cu:1458), v>=3 as successive double factors. */
static double prop_f64(simctx* s, double* a0c) {
    const orc_model* m = s->m;
    const uint32_t nc = (uint32_t)s->J->nc;
    double a0 = 0.0, ac = 0.0;
    for (int j = 0; j < m->R; ++j) {
        double a = s->c[j];
        int crit = 0;
        for (int h = m->Lptr[j]; h < m->Lptr[j + 1]; ++h) {
            uint32_t x = (uint32_t)s->X[m->Ls[h]];
            int v = m->Lv[h];
            if (v == 1) a *= (double)x;
            else if (v == 2) a *= (double)((0.5f * (float)x) * (float)(x - 1u));
            else
                for (int e = 1; e <= v; ++e) a *= ((double)(uint32_t)(x - (uint32_t)e + 1u)) / (double)e;
            uint32_t q = x / (uint32_t)v;
            if (q > 0u && q < nc) crit = 1;
        }
        s->Ad[j] = a;
        s->C[j] = (unsigned char)crit;
        a0 += a;
    }
    for (int j = 0; j < m->R; ++j)
        if (s->C[j]) ac += s->Ad[j];
    if (a0c) *a0c = ac;
    return a0;
}

/* T.cu:1741-1789 with the reference's mixed double / unsigned arithmetic. */
static float g_ref(unsigned char hor, unsigned char hort, uint32_t x) {
    switch (hor) {
        case 1: return 1.0f;
        case 2: return hort == 2 ? (float)(2.0f + 1.0f / ((double)x - 1)) : 2.0f;
        case 3:
            if (hort == 3) return 3.0f + 1.0f / (float)(x - 1u) + 2.0f / (float)(x - 2u);
            if (hort == 2) return 3.0f + 1.5f / (float)(x - 1u);
            return 3.0f;
        default: return 1.0f;
    }
}

static int leap_f64(simctx* s) {
    const orc_model* m = s->m;
    const orc_job* J = s->J;
    if (s->t >= s->tmax) { s->mode = MODE_DONE; return LEAP_DONE; }
    s->mode = MODE_LEAP;
    double a0c;
    double a0 = prop_f64(s, &a0c);
    if (a0 == 0.0) { absorb(s); return LEAP_ABSORB; }
    double tau1 = FLT_MAX;
    float tmp = FLT_MAX;
    for (int sp = 0; sp < m->S; ++sp) {
        if (m->Vptr[sp] == m->Vptr[sp + 1]) continue;
        float mu = 0.0f, sg = 0.0f;
        for (int h = m->Vptr[sp]; h < m->Vptr[sp + 1]; ++h) {
            int j = m->Vr[h];
            int v = m->Vv[h];
            mu = (float)((double)mu + (double)(1 - s->C[j]) * ((double)v * s->Ad[j]));
            sg = (float)((double)sg + (double)(1 - s->C[j]) * ((double)(v * v) * s->Ad[j]));
        }
        float gg = g_ref(m->hor[sp], m->hort[sp], (uint32_t)s->X[sp]);
        float tmp2 = (J->eps * (float)(uint32_t)s->X[sp]) / gg;
        float macs = tmp2 > 1.0f ? tmp2 : 1.0f;
        if (mu != 0.0f && sg != 0.0f) {
            float t3 = macs / fabsf(mu);
            float t4 = (macs * macs) / sg;
            tmp = t3 < t4 ? t3 : t4;
        }
        if ((double)tmp < tau1) tau1 = tmp;
    }
    if (tau1 < (double)J->ssa_factor / a0) { s->mode = MODE_SSA; s->st.switches++; return LEAP_SWITCH; }
    rng_begin(&s->g, s->evc++);
    double tau2 = FLT_MAX;
    if (a0c > 0.0) tau2 = (1.0 / a0c) * (double)logf(1.0f / u01_oc(&s->g));
    double tau = tau1 < tau2 ? tau1 : tau2;
    int pick = -1;
    if (tau1 > tau2) {
        float w = (float)((double)u01_oc(&s->g) * a0c);
        double acc = s->Ad[0];
        int r = 0;
        while ((double)w > acc && r < m->R - 1) { ++r; acc += s->Ad[r] * (double)s->C[r]; }
        pick = r;
    }
    uint64_t fired;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 0) { rng_begin(&s->g, s->evc++); s->st.rejections++; }
        if (apply_leap(s, attempt == 0, pick, 0.0f, tau, 1, &fired)) break;
        tau *= 0.5f;
    }
    leap_samples(s, tau);
    swap_state(s);
    apply_feeds(s);
    s->t += tau;
    s->st.firings += fired;
    s->st.leaps++;
    return LEAP_OK;
}

static void ssa_f64(simctx* s) {
    const orc_model* m = s->m;
    const orc_job* J = s->J;
    rng_begin(&s->g, s->evc++);
    double tacc = 0.0;
    for (int step = 0; step < J->ssa_steps; ++step) {
        double a0 = prop_f64(s, NULL);
        if (a0 == 0.0) { absorb(s); return; }
        double r1 = (double)u01_oc(&s->g) * a0;
        double r2 = (double)u01_oc(&s->g);
        int j = 0;
        double acc = s->Ad[0];
        while (r1 > acc && j < m->R - 1) { ++j; acc += s->Ad[j]; }
        tacc += (1.0f / a0) * log(1.0f / r2);
        double tn = s->t + tacc;
        ssa_samples(s, tn);
        for (int h = m->Fptr[j]; h < m->Fptr[j + 1]; ++h)
            s->X[m->Fs[h]] = (int32_t)((uint32_t)s->X[m->Fs[h]] + (uint32_t)m->Fv[h]);
        s->st.ssa_steps++;
        if (tn <= s->tmax) s->st.firings++;
    }
    s->t += tacc;
}

/* Per-simulation driver: the host do-while of K.cu:242-345 restricted to one
 * simulation (simulations are independent): {leap attempt; SSA burst if it
 * switched} until t >= t_max; force_ssa runs bursts only (K.cu:278-288) and
 * retires the simulation once t >= t_max (the reference keeps stepping it
 * until every simulation has finished, which changes no output). */
static void run_one(simctx* s) {
    const orc_job* J = s->J;
    const int dbl = J->precision == ORC_REFSEM;
    s->mode = J->force_ssa ? MODE_SSA : MODE_LEAP;
    for (;;) {
        if (J->force_ssa) {
            if (s->t >= s->tmax) break;
            if (dbl) ssa_f64(s); else ssa_f32(s);
            if (s->mode == MODE_DONE) break;
            continue;
        }
        int r = dbl ? leap_f64(s) : leap_f32(s);
        if (r == LEAP_DONE || r == LEAP_ABSORB) break;
        if (r == LEAP_SWITCH) {
            if (dbl) ssa_f64(s); else ssa_f32(s);
            if (s->mode == MODE_DONE) break;
        }
    }
    s->st.t_end = s->t;
    s->st.events = s->evc;
}

/* ====================== GROUP algorithm (G lanes) ===================== */
/* The gfx950 group kernel's algorithm, lane by lane (see the header).     */

typedef struct {
    simctx* s;
    int G;
    orc_rng* lane;       /* G lane streams, tag 1 + l */
    int32_t* K;          /* R drawn counts */
    float *P, *Q, *T;    /* G lane partials and scan scratch */
    float* O;            /* G own partials (before the scan) */
} grpctx;

/* Inclusive scan over the G lanes of a group, in the kernels' order: a
 * Hillis-Steele scan (v[l] += v[l - d], d = 1, 2, 4, 8) inside rows of 16
 * lanes, then the row totals carried across rows (G = 32, 64). */
static void g_scan(float* v, float* tmp, int G) {
    /* Hillis-Steele inside rows of 16 lanes (the whole group when G <= 16) */
    const int W = G < 16 ? G : 16;
    for (int d = 1; d < W; d <<= 1) {
        memcpy(tmp, v, sizeof(float) * (size_t)G);
        for (int l = 0; l < G; ++l)
            if (l % W >= d) v[l] = tmp[l - d] + tmp[l];
    }
    if (G >= 32) {            /* rows 1 and 3 add the previous row's last lane */
        memcpy(tmp, v, sizeof(float) * (size_t)G);
        for (int l = 16; l < 32; ++l) v[l] = tmp[15] + tmp[l];
        if (G == 64)
            for (int l = 48; l < 64; ++l) v[l] = tmp[47] + tmp[l];
    }
    if (G == 64) {            /* rows 2 and 3 add lane 31 (the first two rows' total) */
        memcpy(tmp, v, sizeof(float) * (size_t)G);
        for (int l = 32; l < 64; ++l) v[l] = tmp[31] + tmp[l];
    }
}

/* Lane partials of the propensities (all, or the critical-pick weights). */
static void g_partials(grpctx* q, float* P, int which) {
    const simctx* s = q->s;
    const int G = q->G, R = s->m->R;
    for (int l = 0; l < G; ++l) {
        float p = 0.0f;
        for (int j = l; j < R; j += G) {
            if (which == 1 && !s->C[j]) continue;                 /* critical */
            if (which == 2 && !(s->C[j] || j == 0)) continue;     /* pick walk: reaction 0 first */
            p = p + s->Af[j];
        }
        P[l] = p;
    }
}

/* Selection: among the lanes with a positive own partial (Pown), the first
 * whose inclusive sum (Pincl) reaches w, else the last such lane; then the
 * walk through that lane's reactions from the previous lane's inclusive sum.
 * (Lanes with a zero partial are skipped: the scan associates differently
 * per lane, so such a lane's inclusive sum can round above its predecessor's.) */
static int g_select(grpctx* q, const float* Pown, const float* Pincl, float w, int which) {
    const simctx* s = q->s;
    const int G = q->G, R = s->m->R;
    int L = -1, lastc = -1;
    for (int l = 0; l < G; ++l) {
        if (!(Pown[l] > 0.0f)) continue;
        lastc = l;
        if (Pincl[l] >= w) { L = l; break; }
    }
    if (L < 0) L = lastc;
    if (L < 0) return -1;
    float acc = L > 0 ? Pincl[L - 1] : 0.0f;
    int last = -1;
    for (int j = L; j < R; j += G) {
        if (which == 2 && !(s->C[j] || j == 0)) continue;
        const float a = s->Af[j];
        acc = acc + a;
        if (a > 0.0f) last = j;
        if (w <= acc && a > 0.0f) return j;
    }
    return last;
}

static int g_leap(grpctx* q) {
    simctx* s = q->s;
    const orc_model* m = s->m;
    const orc_job* J = s->J;
    const int G = q->G, R = m->R, S = m->S;
    if (s->t >= s->tmax) { s->mode = MODE_DONE; return LEAP_DONE; }
    s->mode = MODE_LEAP;
    for (int j = 0; j < R; ++j) s->Af[j] = prop_one_f32(s, j, &s->C[j]);
    g_partials(q, q->P, 0);
    g_scan(q->P, q->T, G);
    g_partials(q, q->Q, 1);
    g_scan(q->Q, q->T, G);
    const float a0 = q->P[G - 1], a0c = q->Q[G - 1];
    if (a0 == 0.0f) { absorb(s); return LEAP_ABSORB; }
    /* tau': the REPLAY per-species values; their minimum is order-free */
    float tau1 = FLT_MAX;
    for (int sp = 0; sp < S; ++sp) {
        float mu = 0.0f, sg = 0.0f;
        for (int h = m->Vptr[sp]; h < m->Vptr[sp + 1]; ++h) {
            int j = m->Vr[h];
            if (s->C[j]) continue;
            int v = m->Vv[h];
            mu = mu + (float)v * s->Af[j];
            sg = sg + (float)(v * v) * s->Af[j];
        }
        if (mu == 0.0f || sg == 0.0f) continue;
        float gg = g_f32(m->hor[sp], m->hort[sp], s->X[sp]);
        float tmp2 = (J->eps * (float)s->X[sp]) / gg;
        float macs = tmp2 > 1.0f ? tmp2 : 1.0f;
        float t3 = macs / fabsf(mu);
        float t4 = (macs * macs) / sg;
        float tmp = t3 < t4 ? t3 : t4;
        if (tmp < tau1) tau1 = tmp;
    }
    if (tau1 < J->ssa_factor / a0) { s->mode = MODE_SSA; s->st.switches++; return LEAP_SWITCH; }
    uint32_t e = s->evc++;
    rng_begin(&s->g, e);
    float tau2 = FLT_MAX;
    if (a0c > 0.0f) tau2 = exp_rv(&s->g) / a0c;
    float tau = tau1 < tau2 ? tau1 : tau2;
    int pick = -1;
    if (tau1 > tau2) {
        const float w = u01_oc(&s->g) * a0c;
        g_partials(q, q->Q, 2);
        memcpy(q->O, q->Q, sizeof(float) * (size_t)G);
        g_scan(q->Q, q->T, G);
        pick = g_select(q, q->O, q->Q, w, 2);
    }
    uint64_t fired;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 0) { e = s->evc++; s->st.rejections++; }
        fired = 0;
        /* per lane: critical reactions (the pick) and the lane's small-lambda
         * process, whose first gap is drawn up front; reactions with
         * lambda >= 1 draw from their own stream (word 3 = 2^31 | j) */
        for (int l = 0; l < G; ++l) {
            orc_rng* g = &q->lane[l];
            rng_begin(g, e);
            double cum = 0.0, pos = (double)exp_rv(g);
            s->st.gaps++;
            for (int j = l; j < R; j += G) {
                int32_t k;
                if (s->C[j]) {
                    k = (attempt == 0 && j == pick) ? 1 : 0;
                } else {
                    const float lam = tau * s->Af[j];
                    if (lam < 1.0f) {
                        k = 0;
                        if (lam > 0.0f) {
                            cum = cum + (double)lam;
                            while (pos < cum) { ++k; pos = pos + (double)exp_rv(g); s->st.gaps++; }
                        }
                    } else {
                        orc_rng rj = s->g;
                        rj.tag = 0x80000000u | (uint32_t)j;
                        rng_begin(&rj, e);
                        k = poisson(&rj, lam);
                        s->st.poisson++;
                    }
                }
                q->K[j] = k;
                fired += (uint64_t)k;
            }
        }
        int ok = 1;
        for (int sp = 0; sp < S; ++sp) {
            int64_t x = s->X[sp];
            for (int h = m->FSptr[sp]; h < m->FSptr[sp + 1]; ++h) {
                x += (int64_t)m->FSv[h] * (int64_t)q->K[m->FSr[h]];
                if (x < 0 || x > 2147483647LL) { ok = 0; break; }
            }
            s->X1[sp] = (int32_t)x;
        }
        if (ok) break;
        tau = tau * 0.5f;
    }
    leap_samples(s, (double)tau);
    swap_state(s);
    apply_feeds(s);
    s->t = s->t + (double)tau;
    s->st.firings += fired;
    s->st.leaps++;
    return LEAP_OK;
}

static void g_ssa(grpctx* q) {
    simctx* s = q->s;
    const orc_model* m = s->m;
    const orc_job* J = s->J;
    const int G = q->G, R = m->R;
    rng_begin(&s->g, s->evc++);
    double tacc = 0.0;
    for (int step = 0; step < J->ssa_steps; ++step) {
        for (int j = 0; j < R; ++j) s->Af[j] = prop_one_f32(s, j, &s->C[j]);
        g_partials(q, q->P, 0);
        memcpy(q->O, q->P, sizeof(float) * (size_t)G);
        g_scan(q->P, q->T, G);
        const float a0 = q->P[G - 1];
        if (a0 == 0.0f) { absorb(s); return; }
        const float r1 = u01_oc(&s->g) * a0;
        int j = g_select(q, q->O, q->P, r1, 0);
        if (j < 0) { absorb(s); return; }         /* unreachable for a0 > 0 */
        const float dt = exp_rv(&s->g) / a0;
        tacc = tacc + (double)dt;
        const double tn = s->t + tacc;
        ssa_samples(s, tn);
        for (int h = m->Fptr[j]; h < m->Fptr[j + 1]; ++h)
            s->X[m->Fs[h]] = (int32_t)((uint32_t)s->X[m->Fs[h]] + (uint32_t)m->Fv[h]);
        s->st.ssa_steps++;
        if (tn <= s->tmax) s->st.firings++;
    }
    s->t = s->t + tacc;
}

static void run_group(grpctx* q) {
    simctx* s = q->s;
    const orc_job* J = s->J;
    s->mode = J->force_ssa ? MODE_SSA : MODE_LEAP;
    for (;;) {
        if (J->force_ssa) {
            if (s->t >= s->tmax) break;
            g_ssa(q);
            if (s->mode == MODE_DONE) break;
            continue;
        }
        int r = g_leap(q);
        if (r == LEAP_DONE || r == LEAP_ABSORB) break;
        if (r == LEAP_SWITCH) {
            g_ssa(q);
            if (s->mode == MODE_DONE) break;
        }
    }
    s->st.t_end = s->t;
    s->st.events = s->evc;
}

/* Input row of simulation i: one shared row, one row per simulation
 * (rows >= n, the reference's MX_0 / c_matrix), or n / rows consecutive
 * simulations per row (replica groups). */
static int row_of(int rows, int n, int i) {
    if (rows <= 1) return 0;
    if (rows >= n) return i;
    return i / (n / rows);
}

int orc_run(const orc_net* net, const orc_job* J, uint32_t* storage, orc_stat* stats, int32_t* xfinal) {
    orc_model m;
    if (J->K <= 0 || J->SS < 0 || J->n < 0) return -1;
    if (J->group < 0 || J->group > 64 || (J->group & (J->group - 1))) return -1;
    if (J->group > 0 && J->precision != ORC_REPLAY) return -1;
    if (J->nfs < 0 || (J->nfs > 0 && (!J->fs || !J->fv || J->fv_rows < 1)) ||
        (J->nfs > 0 && J->fv_rows > 1 && J->fv_rows < J->n && J->n % J->fv_rows))
        return -1;
    for (int u = 0; u < J->nfs; ++u)
        if (J->fs[u] < 0 || J->fs[u] >= net->S) return -1;
    if ((J->x0_rows > 1 && J->x0_rows < J->n && J->n % J->x0_rows) ||
        (J->c_rows > 1 && J->c_rows < J->n && J->n % J->c_rows))
        return -1;
    for (int sk = 0; sk < J->SS; ++sk)
        if (J->cs[sk] < 0 || J->cs[sk] >= net->S) return -1;
    if (orc_compile(net, &m) != 0) return -1;
    const int S = net->S, R = net->R;
    const double tmax = J->ts[J->K - 1];
#ifdef _OPENMP
    if (J->threads > 0) omp_set_num_threads(J->threads);
#endif
#pragma omp parallel
    {
        int32_t* X = (int32_t*)malloc(sizeof(int32_t) * S);
        int32_t* X1 = (int32_t*)malloc(sizeof(int32_t) * S);
        float* Af = (float*)malloc(sizeof(float) * R);
        double* Ad = (double*)malloc(sizeof(double) * R);
        float* cf = (float*)malloc(sizeof(float) * R);
        unsigned char* C = (unsigned char*)malloc(R);
        const int G = J->group > 0 ? J->group : 1;
        orc_rng* lanes = (orc_rng*)calloc((size_t)G, sizeof(orc_rng));
        int32_t* Kc = (int32_t*)malloc(sizeof(int32_t) * R);
        float* gP = (float*)malloc(sizeof(float) * 4 * (size_t)G);
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < J->n; ++i) {
            simctx s;
            memset(&s, 0, sizeof(s));
            s.m = &m; s.J = J;
            s.c = J->c + (size_t)row_of(J->c_rows, J->n, i) * R;
            for (int r = 0; r < R; ++r) cf[r] = (float)s.c[r];
            s.cf = cf; s.Af = Af; s.Ad = Ad; s.C = C;
            s.X = X; s.X1 = X1;
            const uint32_t* x0 = J->x0 + (size_t)row_of(J->x0_rows, J->n, i) * S;
            for (int sp = 0; sp < S; ++sp) X[sp] = (int32_t)x0[sp];
            s.out = storage + (size_t)i * J->K * J->SS;
            s.fv = J->nfs > 0 ? J->fv + (size_t)row_of(J->fv_rows, J->n, i) * J->nfs : NULL;
            memset(s.out, 0, sizeof(uint32_t) * (size_t)J->K * J->SS);   /* PrepareVariables T.cu:1269-1297 */
            s.t = 0.0; s.tmax = tmax;
            s.g.key[0] = J->seed_lo; s.g.key[1] = J->seed_hi;
            s.g.sim = J->sim_base + (uint32_t)i;
            s.g.pos = 4;
            if (J->group > 0) {
                grpctx q = {&s, G, lanes, Kc, gP, gP + G, gP + 2 * G, gP + 3 * G};
                for (int l = 0; l < G; ++l) {
                    lanes[l] = s.g;
                    lanes[l].tag = 1u + (uint32_t)l;
                }
                run_group(&q);
            } else {
                run_one(&s);
            }
            if (stats) stats[i] = s.st;
            if (xfinal) memcpy(xfinal + (size_t)i * S, s.X, sizeof(int32_t) * S);
            /* run_one may leave X/X1 swapped; keep the buffers this thread owns */
            if (s.X != X) { int32_t* tmp = X; X = X1; X1 = tmp; }
        }
        free(X); free(X1); free(Af); free(Ad); free(cf); free(C);
        free(lanes); free(Kc); free(gP);
    }
    orc_model_free(&m);
    return 0;
}
