// ctl_sim.hip -- gfx950 simulation kernels and the ctl_sim runtime of libctl.
//
// One GPU lane runs one stochastic simulation (Cao's adaptive tau-leaping
// with SSA fallback).  The kernel is persistent: a grid of resident "slots"
// pulls simulation indices from a device queue (work stealing), runs each to
// t_max and writes its samples; a launch runs a bounded number of iterations
// per slot and the host relaunches until every simulation is done (the
// replacement of the reference's per-step host loop and ParallelVoting,
// kernel.cu:242-345 and 2phase_n-tau-leaping.cu:2990-3002).
//
// Layout in HBM (ns = number of slots):
//   per slot, structure of arrays, coalesced across a wavefront:
//     X  int32 [2][S][ns]  state and trial state (parity bit swaps them)
//     A  uint32[R][ns]     propensity (FP32 bits) | critical flag << 31
//     Cs float [R][ns]     rate constants when rows differ between sims
//   per simulation: samples uint32 [sim][K][SS], stats, final X [sim][S]
// The compiled model (flats, HOR) is read through scalar loads: every lane
// of a wavefront walks the same reaction / species loops, so each table
// entry is a wave-uniform constant-cache broadcast.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "ctl_device.h"
#include "ctl_internal.h"

#pragma clang fp contract(off)

namespace ctl {

enum : int { MODE_LEAP = 0, MODE_SSA = 1 };
enum : int { LEAP_DONE = 0, LEAP_OK = 1, LEAP_SWITCH = 2, LEAP_ABSORB = 3 };

struct SlotState {          // SoA, one entry per slot
    int32_t* sim;           // local simulation index, -1 none, -2 queue exhausted
    double* t;
    double* tacc;           // clock advance inside the current SSA burst
    uint32_t* cursor;
    int32_t* mode;
    uint32_t* evc;          // next event number
    uint32_t* ev;           // event of the SSA burst in progress
    uint32_t* blk;
    uint32_t* pos;
    int32_t* ssa_left;      // steps left in the current SSA burst (0 = none)
    uint32_t* xb;           // parity of the current state buffer (global storage)
    uint64_t* firings;
    uint64_t* ssa_steps;
    uint64_t* poisson;
    uint64_t* gaps;
    uint32_t* leaps;
    uint32_t* rejections;
    uint32_t* switches;
};

struct KParams {
    // model
    int S, R, K, SS;
    const uint32_t* Lptr;
    const uint32_t* Lent;
    const uint32_t* Fptr;
    const uint32_t* Fent;
    const uint32_t* Vptr;
    const uint32_t* Vent;
    const uint32_t* Dptr;     // SSA dependency graph: reactions whose propensity
    const uint32_t* Dent;     //   reads a species changed by reaction j
    const uint32_t* horw;     // hor | hort << 8
    const int32_t* cs;
    const double* ts;
    double tmax;
    const float* c_bcast;     // R (rates shared by every simulation) or null
    const float* c_sims;      // [sim][R] or null
    const uint32_t* x0_bcast; // S or null
    const uint32_t* x0_sims;  // [sim][S] or null
    // run
    int n;
    uint32_t sim_base, key0, key1;
    int nc;
    float eps;
    int ssa_steps;
    float ssa_factor;
    int force_ssa;
    int iters;
    int ssa_chunk;
    // slots (global per-lane storage, [.][ns])
    int ns;
    int32_t* X;               // [2][S][ns]
    uint32_t* A;              // [R][ns]
    float* Cs;                // [R][ns] when rates differ between simulations
    SlotState st;
    // outputs
    uint32_t* storage;        // [sim][K][SS]
    ctl_sim_stat* simstat;
    int32_t* xfinal;          // [sim][S]
    double* tfinal;
    // queue
    uint32_t* next_sim;
    uint32_t* done;
};

// Model tables are read through constant-address-space pointers: the loads
// are then invariant, and with the wave-uniform loop indices of the reaction
// and species loops they become scalar loads (s_load, constant cache) instead
// of per-lane vector loads.
typedef const __attribute__((address_space(4))) uint32_t* ctab_u32;
typedef const __attribute__((address_space(4))) int32_t* ctab_i32;
typedef const __attribute__((address_space(4))) float* ctab_f32;
typedef const __attribute__((address_space(4))) double* ctab_f64;
#define CT_U32(ptr) ((ctab_u32)(ptr))
#define CT_I32(ptr) ((ctab_i32)(ptr))
#define CT_F32(ptr) ((ctab_f32)(ptr))
#define CT_F64(ptr) ((ctab_f64)(ptr))

__constant__ float c_recip[64] = {
    0.0f,      1.0f,      1.0f / 2,  1.0f / 3,  1.0f / 4,  1.0f / 5,  1.0f / 6,  1.0f / 7,
    1.0f / 8,  1.0f / 9,  1.0f / 10, 1.0f / 11, 1.0f / 12, 1.0f / 13, 1.0f / 14, 1.0f / 15,
    1.0f / 16, 1.0f / 17, 1.0f / 18, 1.0f / 19, 1.0f / 20, 1.0f / 21, 1.0f / 22, 1.0f / 23,
    1.0f / 24, 1.0f / 25, 1.0f / 26, 1.0f / 27, 1.0f / 28, 1.0f / 29, 1.0f / 30, 1.0f / 31,
    1.0f / 32, 1.0f / 33, 1.0f / 34, 1.0f / 35, 1.0f / 36, 1.0f / 37, 1.0f / 38, 1.0f / 39,
    1.0f / 40, 1.0f / 41, 1.0f / 42, 1.0f / 43, 1.0f / 44, 1.0f / 45, 1.0f / 46, 1.0f / 47,
    1.0f / 48, 1.0f / 49, 1.0f / 50, 1.0f / 51, 1.0f / 52, 1.0f / 53, 1.0f / 54, 1.0f / 55,
    1.0f / 56, 1.0f / 57, 1.0f / 58, 1.0f / 59, 1.0f / 60, 1.0f / 61, 1.0f / 62, 1.0f / 63};

// Per-lane working view of one slot.
struct Lane {
    int slot;
    int sim;                 // local index, -1 none
    double t;
    double tacc;
    uint32_t cursor;
    int mode;
    uint32_t evc;
    int ssa_left;
    uint32_t xb;
    uint64_t firings, ssa_steps, poisson, gaps;
    uint32_t leaps, rejections, switches;
    uint32_t absorbed;
    Rng g;
};

// Where a lane's working arrays live: in LDS (small models, layout
// [index][lane of the workgroup]) or in the global slot arrays
// ([index][slot]).  Either way consecutive lanes of a wavefront touch
// consecutive words.
struct View {
    int32_t* X;              // buffer 0 of this lane; buffer 1 is X + xbuf
    uint32_t* A;
    float* Cs;
    size_t st;               // stride between consecutive indices
    size_t xbuf;             // distance between the two state buffers
};

__device__ __forceinline__ int32_t* xcur(const View& V, const Lane& L) { return V.X + (L.xb ? V.xbuf : 0); }
__device__ __forceinline__ int32_t* xnext(const View& V, const Lane& L) { return V.X + (L.xb ? 0 : V.xbuf); }
__device__ __forceinline__ float aval(uint32_t u) { return __uint_as_float(u & 0x7fffffffu); }

__device__ __forceinline__ float rate(const KParams& p, const View& V, int j) {
    return p.c_bcast ? CT_F32(p.c_bcast)[j] : V.Cs[(size_t)j * V.st];
}

// Propensity of reaction j, CalculatePropensity<DEV_FLATTEN_MATRIX_CONST>
// (T.cu:1443-1470) in FP32, and its criticality (T.cu:1725-1735: some
// reactant with 0 < floor(X/v) < nc).  Returns the FP32 bits | crit << 31.
__device__ __forceinline__ uint32_t prop_one(const KParams& p, const View& V, const int32_t* __restrict__ X, int j,
                                             float* aout) {
    float a = rate(p, V, j);
    bool crit = false;
    const uint32_t nc = (uint32_t)p.nc;
    const uint32_t h1 = CT_U32(p.Lptr)[j + 1];
    for (uint32_t h = CT_U32(p.Lptr)[j]; h < h1; ++h) {
        const uint32_t e = CT_U32(p.Lent)[h];
        const int s = (int)(e & 0xffffu);
        const int v = (int)(e >> 16);
        const int32_t x = X[(size_t)s * V.st];
        if (v == 1) {
            a = a * (float)x;
        } else if (v == 2) {
            a = a * ((x < 2) ? 0.0f : (0.5f * (float)x) * (float)(x - 1));
        } else if (x < v) {
            a = a * 0.0f;
        } else {
            for (int q = 1; q <= v; ++q) a = a * ((float)(x - q + 1) / (float)q);
        }
        const uint32_t fl = (uint32_t)x / (uint32_t)v;
        crit = crit || (fl > 0u && fl < nc);
    }
    *aout = a;
    return __float_as_uint(a) | (crit ? 0x80000000u : 0u);
}

// All propensities; a0 and a0c summed in ascending reaction order (the
// reference's loops T.cu:1467-1468 and T.cu:1945-1952).
__device__ __forceinline__ float propensities(const KParams& p, const View& V, const int32_t* __restrict__ X,
                                              float* a0c) {
    float a0 = 0.0f, ac = 0.0f;
    for (int j = 0; j < p.R; ++j) {
        float a;
        const uint32_t u = prop_one(p, V, X, j, &a);
        V.A[(size_t)j * V.st] = u;
        a0 = a0 + a;
        if (u & 0x80000000u) ac = ac + a;
    }
    *a0c = ac;
    return a0;
}

// Cao's g_i from HOR / HORtype (T.cu:1741-1789) in FP32.
__device__ __forceinline__ float gfun(uint32_t hw, int32_t x) {
    const uint32_t hor = hw & 0xffu, hort = (hw >> 8) & 0xffu;
    const float xf = (float)x;
    if (hor == 2u) return hort == 2u ? 2.0f + 1.0f / (xf - 1.0f) : 2.0f;
    if (hor == 3u) {
        if (hort == 3u) return 3.0f + 1.0f / (xf - 1.0f) + 2.0f / (xf - 2.0f);
        if (hort == 2u) return 3.0f + 1.5f / (xf - 1.0f);
        return 3.0f;
    }
    return 1.0f;
}

// ABSORB (T.cu:1665-1697, 2165-2196): the remaining samples repeat the last
// recorded sample.
__device__ __forceinline__ void absorb(const KParams& p, Lane& L, const View& V, const int32_t* __restrict__ X) {
    uint32_t* out = p.storage + (size_t)L.sim * p.K * p.SS;
    if (L.cursor == 0) {
        for (int sk = 0; sk < p.SS; ++sk) out[sk] = (uint32_t)X[(size_t)CT_I32(p.cs)[sk] * V.st];
        L.cursor = 1;
    }
    const uint32_t* last = out + (size_t)(L.cursor - 1) * p.SS;
    for (int i = (int)L.cursor; i < p.K; ++i)
        for (int sk = 0; sk < p.SS; ++sk) out[(size_t)i * p.SS + sk] = last[sk];
    L.t = p.tmax + (double)0.1f;
    L.cursor = (uint32_t)p.K;
    L.absorbed = 1;
}

// One adaptive leap attempt, PreCalculations (T.cu:1568-2093).
__device__ __forceinline__ int leap(const KParams& p, Lane& L, const View& V, const float* __restrict__ recip) {
    if (L.t >= p.tmax) return LEAP_DONE;
    int32_t* X = xcur(V, L);
    float a0c;
    const float a0 = propensities(p, V, X, &a0c);
    if (a0 == 0.0f) { absorb(p, L, V, X); return LEAP_ABSORB; }
    // tau' over the species-major var flat (T.cu:1860-1922)
    float tau1 = 3.40282347e+38f;
    for (int sp = 0; sp < p.S; ++sp) {
        const uint32_t h0 = CT_U32(p.Vptr)[sp], h1 = CT_U32(p.Vptr)[sp + 1];
        if (h0 == h1) continue;
        float mu = 0.0f, sg = 0.0f;
        for (uint32_t h = h0; h < h1; ++h) {
            const uint32_t e = CT_U32(p.Vent)[h];
            const int j = (int)(e & 0xffffu);
            const int v = (int)(int16_t)(e >> 16);
            const uint32_t u = V.A[(size_t)j * V.st];
            if (u & 0x80000000u) continue;
            const float a = __uint_as_float(u);
            mu = mu + (float)v * a;
            sg = sg + (float)(v * v) * a;
        }
        if (mu == 0.0f || sg == 0.0f) continue;
        const int32_t x = X[(size_t)sp * V.st];
        const float gg = gfun(CT_U32(p.horw)[sp], x);
        const float tmp2 = (p.eps * (float)x) / gg;
        const float macs = tmp2 > 1.0f ? tmp2 : 1.0f;
        const float t3 = macs / fabsf(mu);
        const float t4 = (macs * macs) / sg;
        const float tmp = t3 < t4 ? t3 : t4;
        if (tmp < tau1) tau1 = tmp;
    }
    if (tau1 < p.ssa_factor / a0) { L.switches++; return LEAP_SWITCH; }   // T.cu:1935-1941
    rng_event(L.g, L.evc++);
    float tau2 = 3.40282347e+38f;
    if (a0c > 0.0f) tau2 = exp_rv(L.g) / a0c;                             // T.cu:1962-1965
    float tau = tau1 < tau2 ? tau1 : tau2;
    int pick = -1;
    if (tau1 > tau2) {                                                     // T.cu:1993-2006
        const float w = rng_oc(L.g) * a0c;
        float acc = aval(V.A[0]);
        int r = 0;
        while (w > acc && r < p.R - 1) {
            ++r;
            const uint32_t u = V.A[(size_t)r * V.st];
            if (u & 0x80000000u) acc = acc + aval(u);
        }
        pick = r;
    }
    int32_t* X1 = xnext(V, L);
    uint64_t fired = 0;
    for (int attempt = 0;; ++attempt) {                                    // T.cu:2011-2050
        if (attempt > 0) { rng_event(L.g, L.evc++); L.rejections++; }
        for (int s = 0; s < p.S; ++s) X1[(size_t)s * V.st] = X[(size_t)s * V.st];
        bool ok = true;
        fired = 0;
        // Counts: independent Poisson draws for lambda_j >= 1; one shared
        // rate-1 Poisson process over the running sum of the lambda_j < 1,
        // walked in reaction order (double precision positions).  Applied
        // in test_update's order (T.cu:1300-1351): reaction-major fear flat,
        // reject on the first negative partial sum or an int32 overflow.
        double cum = 0.0, pos = -1.0;
        for (int j = 0; j < p.R && ok; ++j) {
            const uint32_t u = V.A[(size_t)j * V.st];
            int32_t k;
            if (u & 0x80000000u) {
                k = (attempt == 0 && j == pick) ? 1 : 0;
            } else {
                const float lam = tau * __uint_as_float(u);
                if (lam < 1.0f) {
                    k = 0;
                    if (lam > 0.0f) {
                        if (pos < 0.0) { pos = (double)exp_rv(L.g); L.gaps++; }
                        cum = cum + (double)lam;
                        while (pos < cum) { ++k; pos = pos + (double)exp_rv(L.g); L.gaps++; }
                    }
                } else {
                    k = poisson(L.g, lam, recip);
                    L.poisson++;
                }
            }
            if (k == 0) continue;
            fired += (uint64_t)k;
            const uint32_t h1 = CT_U32(p.Fptr)[j + 1];
            for (uint32_t h = CT_U32(p.Fptr)[j]; h < h1; ++h) {
                const uint32_t e = CT_U32(p.Fent)[h];
                const size_t s = (size_t)(e & 0xffffu) * V.st;
                const int v = (int)(int16_t)(e >> 16);
                const int64_t nx = (int64_t)X1[s] + (int64_t)v * (int64_t)k;
                if (nx < 0 || nx > 2147483647LL) { ok = false; break; }
                X1[s] = (int32_t)nx;
            }
        }
        if (ok) break;
        tau = tau * 0.5f;
    }
    // check_time_instant (T.cu:1507-1562): linear interpolation; a negative
    // delta saturates to 0 in the reference's float->unsigned conversion.
    {
        uint32_t i = L.cursor;
        const double tn = L.t + (double)tau;
        if (i < (uint32_t)p.K && tn >= CT_F64(p.ts)[i]) {
            uint32_t* out = p.storage + (size_t)L.sim * p.K * p.SS;
            while (i < (uint32_t)p.K && tn >= CT_F64(p.ts)[i]) {
                const double ti = CT_F64(p.ts)[i];
                for (int sk = 0; sk < p.SS; ++sk) {
                    const size_t sp = (size_t)CT_I32(p.cs)[sk] * V.st;
                    const int32_t d = X1[sp] - X[sp];
                    uint32_t val = (uint32_t)X[sp];
                    if (d > 0) {
                        const double delta = ((double)d * (ti - L.t)) / (double)tau;
                        val += (uint32_t)rint(delta);
                    }
                    out[(size_t)i * p.SS + sk] = val;
                }
                ++i;
            }
            L.cursor = i;
        }
    }
    L.xb ^= 1u;                          // X <- X1 (T.cu:2068-2070)
    L.t = L.t + (double)tau;             // T.cu:2084
    L.firings += fired;
    L.leaps++;
    return LEAP_OK;
}

// Up to `steps` direct-method SSA steps of the current burst
// (EventualSSASteps T.cu:2099-2273, LightweightSSA T.cu:2280-2454).  The
// reference recomputes every propensity each step; here only the reactions
// that read a species the fired reaction changed are recomputed (dependency
// graph), which yields the same values, and a0 is re-summed in reaction
// order each step.  Returns true when the simulation absorbed.
__device__ __forceinline__ bool ssa_steps(const KParams& p, Lane& L, const View& V, int steps) {
    int32_t* X = xcur(V, L);
    float dummy;
    float a0 = propensities(p, V, X, &dummy);
    for (int st = 0; st < steps; ++st) {
        if (a0 == 0.0f) { absorb(p, L, V, X); return true; }
        const float r1 = rng_oc(L.g) * a0;
        int j = 0;
        float acc = aval(V.A[0]);
        while (r1 > acc && j < p.R - 1) {
            ++j;
            acc = acc + aval(V.A[(size_t)j * V.st]);
        }
        const float dt = exp_rv(L.g) / a0;
        L.tacc = L.tacc + (double)dt;
        const double tn = L.t + L.tacc;
        uint32_t i = L.cursor;
        if (i < (uint32_t)p.K && tn > CT_F64(p.ts)[i]) {
            uint32_t* out = p.storage + (size_t)L.sim * p.K * p.SS;
            while (i < (uint32_t)p.K && tn > CT_F64(p.ts)[i]) {
                for (int sk = 0; sk < p.SS; ++sk)
                    out[(size_t)i * p.SS + sk] = (uint32_t)X[(size_t)CT_I32(p.cs)[sk] * V.st];
                ++i;
            }
            L.cursor = i;
        }
        const uint32_t h1 = CT_U32(p.Fptr)[j + 1];
        for (uint32_t h = CT_U32(p.Fptr)[j]; h < h1; ++h) {
            const uint32_t e = CT_U32(p.Fent)[h];
            const size_t s = (size_t)(e & 0xffffu) * V.st;
            const int v = (int)(int16_t)(e >> 16);
            X[s] = (int32_t)((uint32_t)X[s] + (uint32_t)v);
        }
        L.ssa_steps++;
        if (tn <= p.tmax) L.firings++;
        if (st + 1 < steps) {
            // j is per lane: walk its dependency list (not wave-uniform)
            const uint32_t d1 = p.Dptr[j + 1];
            for (uint32_t d = p.Dptr[j]; d < d1; ++d) {
                const int r = (int)p.Dent[d];
                float a;
                V.A[(size_t)r * V.st] = prop_one(p, V, X, r, &a);
            }
            a0 = 0.0f;
            for (int r = 0; r < p.R; ++r) a0 = a0 + aval(V.A[(size_t)r * V.st]);
        }
    }
    return false;
}

__device__ __forceinline__ void load_lane(const KParams& p, Lane& L) {
    const int s = L.slot;
    L.sim = p.st.sim[s];
    L.g.k0 = p.key0;
    L.g.k1 = p.key1;
    L.absorbed = 0;
    if (L.sim < 0) return;               // idle slot: fetch() initialises the lane
    L.t = p.st.t[s];
    L.tacc = p.st.tacc[s];
    L.cursor = p.st.cursor[s];
    L.mode = p.st.mode[s];
    L.evc = p.st.evc[s];
    L.ssa_left = p.st.ssa_left[s];
    L.xb = p.st.xb[s];
    L.firings = p.st.firings[s];
    L.ssa_steps = p.st.ssa_steps[s];
    L.poisson = p.st.poisson[s];
    L.gaps = p.st.gaps[s];
    L.leaps = p.st.leaps[s];
    L.rejections = p.st.rejections[s];
    L.switches = p.st.switches[s];
    L.g.sim = p.sim_base + (uint32_t)L.sim;
    L.g.ev = p.st.ev[s];
    L.g.blk = p.st.blk[s];
    L.g.pos = p.st.pos[s];
    if (L.g.pos < 4u) {   // regenerate the partially consumed block of an SSA burst
        philox_block(L.g.blk - 1u, L.g.ev, L.g.sim, 0u, L.g.k0, L.g.k1, L.g.b0, L.g.b1, L.g.b2, L.g.b3);
        for (uint32_t i = 0; i < L.g.pos; ++i) {   // drop the words already consumed
            L.g.b0 = L.g.b1;
            L.g.b1 = L.g.b2;
            L.g.b2 = L.g.b3;
        }
    }
}

__device__ __forceinline__ void store_lane(const KParams& p, const Lane& L) {
    const int s = L.slot;
    p.st.sim[s] = L.sim;
    if (L.sim < 0) return;
    p.st.t[s] = L.t;
    p.st.tacc[s] = L.tacc;
    p.st.cursor[s] = L.cursor;
    p.st.mode[s] = L.mode;
    p.st.evc[s] = L.evc;
    p.st.ssa_left[s] = L.ssa_left;
    p.st.xb[s] = L.xb;
    p.st.firings[s] = L.firings;
    p.st.ssa_steps[s] = L.ssa_steps;
    p.st.poisson[s] = L.poisson;
    p.st.gaps[s] = L.gaps;
    p.st.leaps[s] = L.leaps;
    p.st.rejections[s] = L.rejections;
    p.st.switches[s] = L.switches;
    p.st.ev[s] = L.g.ev;
    p.st.blk[s] = L.g.blk;
    p.st.pos[s] = L.g.pos;
}

// Per-simulation rates into the lane's rate array (rows that differ).
__device__ __forceinline__ void load_rates(const KParams& p, const View& V, int sim) {
    if (!p.c_bcast)
        for (int j = 0; j < p.R; ++j) V.Cs[(size_t)j * V.st] = p.c_sims[(size_t)sim * p.R + j];
}

// Takes the next simulation from the queue and initialises the lane
// (PrepareVariables, T.cu:1269-1297).  Returns false when the queue is empty.
__device__ __forceinline__ bool fetch(const KParams& p, Lane& L, const View& V) {
    const uint32_t q = atomicAdd(p.next_sim, 1u);
    if (q >= (uint32_t)p.n) { L.sim = -2; return false; }
    L.sim = (int)q;
    L.t = 0.0;
    L.tacc = 0.0;
    L.cursor = 0;
    L.evc = 0;
    L.ssa_left = 0;
    L.xb = 0;
    L.firings = L.ssa_steps = L.poisson = L.gaps = 0;
    L.leaps = L.rejections = L.switches = 0;
    L.absorbed = 0;
    L.mode = p.force_ssa ? MODE_SSA : MODE_LEAP;
    L.g.sim = p.sim_base + q;
    L.g.pos = 4;
    int32_t* X = xcur(V, L);
    for (int s = 0; s < p.S; ++s)
        X[(size_t)s * V.st] = (int32_t)(p.x0_bcast ? CT_U32(p.x0_bcast)[s] : p.x0_sims[(size_t)q * p.S + s]);
    load_rates(p, V, (int)q);
    return true;
}

__device__ __forceinline__ void finish(const KParams& p, Lane& L, const View& V) {
    const int32_t* X = xcur(V, L);
    ctl_sim_stat st;
    st.firings = L.firings;
    st.ssa_steps = L.ssa_steps;
    st.poisson = L.poisson;
    st.gaps = L.gaps;
    st.t_end = L.t;
    st.leaps = L.leaps;
    st.rejections = L.rejections;
    st.switches = L.switches;
    st.events = L.evc;
    st.absorbed = L.absorbed;
    st.reserved = 0;
    p.simstat[L.sim] = st;
    p.tfinal[L.sim] = L.t;
    for (int s = 0; s < p.S; ++s) p.xfinal[(size_t)L.sim * p.S + s] = X[(size_t)s * V.st];
    atomicAdd(p.done, 1u);
    L.sim = -1;
}

// One slot per lane; kLDS keeps the lane's state, trial state, propensities
// (and per-simulation rates) in LDS for the duration of the launch, copied
// from / to the global slot arrays at entry / exit.
template <bool kLDS>
__global__ __launch_bounds__(256) void tau_leap_kernel(KParams p) {
    extern __shared__ uint32_t smem[];
    __shared__ float recip[64];
    for (int i = threadIdx.x; i < 64; i += blockDim.x) recip[i] = c_recip[i];
    __syncthreads();
    Lane L;
    L.slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (L.slot >= p.ns) return;
    View V;
    if (kLDS) {
        const size_t B = blockDim.x;
        int32_t* base = (int32_t*)smem + threadIdx.x;
        V.X = base;
        V.xbuf = (size_t)p.S * B;
        V.A = (uint32_t*)(base + 2 * V.xbuf);
        V.Cs = (float*)(V.A + (size_t)p.R * B);
        V.st = B;
    } else {
        V.X = p.X + L.slot;
        V.xbuf = (size_t)p.S * p.ns;
        V.A = p.A + L.slot;
        V.Cs = p.Cs ? p.Cs + L.slot : nullptr;
        V.st = (size_t)p.ns;
    }
    load_lane(p, L);
    if (L.sim == -2) return;             // queue was exhausted for this slot
    if (kLDS && L.sim >= 0) {            // resume: state from the global slot arrays
        const int32_t* G = p.X + (L.xb ? (size_t)p.S * p.ns : 0) + L.slot;
        for (int s = 0; s < p.S; ++s) V.X[(size_t)s * V.st] = G[(size_t)s * p.ns];
        L.xb = 0;
        load_rates(p, V, L.sim);
    }
    for (int it = 0; it < p.iters; ++it) {
        if (L.sim < 0 && !fetch(p, L, V)) break;
        if (L.mode == MODE_LEAP) {
            const int r = leap(p, L, V, recip);
            if (r == LEAP_DONE || r == LEAP_ABSORB) {
                finish(p, L, V);
            } else if (r == LEAP_SWITCH) {
                L.mode = MODE_SSA;
            }
        } else {
            if (L.ssa_left == 0) {       // open a burst
                if (p.force_ssa && L.t >= p.tmax) { finish(p, L, V); continue; }
                rng_event(L.g, L.evc++);
                L.tacc = 0.0;
                L.ssa_left = p.ssa_steps;
            }
            const int n = L.ssa_left < p.ssa_chunk ? L.ssa_left : p.ssa_chunk;
            if (ssa_steps(p, L, V, n)) {
                L.ssa_left = 0;
                finish(p, L, V);
                continue;
            }
            L.ssa_left -= n;
            if (L.ssa_left == 0) {
                L.t = L.t + L.tacc;
                L.tacc = 0.0;
                if (!p.force_ssa) L.mode = MODE_LEAP;
            }
        }
    }
    if (kLDS && L.sim >= 0) {            // suspend: current state to the global slot arrays
        const int32_t* X = xcur(V, L);
        int32_t* G = p.X + L.slot;
        for (int s = 0; s < p.S; ++s) G[(size_t)s * p.ns] = X[(size_t)s * V.st];
        L.xb = 0;
    }
    store_lane(p, L);
}

// Reference sample layout [i][sk][sim] from the kernel's [sim][i][sk].
__global__ void to_reference_layout(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int n, int KS) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)n * KS;
    if (idx >= total) return;
    const size_t sim = idx / KS, ks = idx % KS;
    dst[ks * n + sim] = src[idx];
}

// calculateFitness (T.cu:3324-3408): group g of P parallel repetitions,
// thread index within the group; mean over the P simulations at
// (g*P + par)*threads + tid, then sum over samples, sampled species and
// target repetitions of |tgt - mean| / tgt (tgt != 0), divided by K*SS.
__global__ void fitness_kernel(const uint32_t* __restrict__ samples, const uint32_t* __restrict__ target,
                               const int8_t* __restrict__ t2e, double* __restrict__ fit, int threads,
                               int groups, int P, int K, int SS, int D, int E) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= threads * groups) return;
    const int g = tid / threads, th = tid % threads;
    const int experiment = t2e[tid];
    float sub = 0.0f;
    for (int i = 0; i < K; ++i)
        for (int s = 0; s < SS; ++s) {
            double acc = 0.0;
            for (int par = 0; par < P; ++par) {
                const size_t sim = (size_t)(g * P + par) * threads + th;
                acc += samples[(sim * K + i) * SS + s];
            }
            acc /= P;
            for (int rep = 0; rep < E; ++rep) {
                const int tgt = (int)target[(size_t)i * (SS * D * E) + (size_t)experiment * (SS * E) + rep * SS + s];
                double dist = fabs((double)tgt - acc);
                if (tgt != 0) dist /= tgt;
                sub += dist;
            }
        }
    fit[tid] = (1.0 / K) * (1.0 / SS) * sub;
}

}  // namespace ctl

// ===========================================================================
// Host runtime
// ===========================================================================
using namespace ctl;

struct ctl_sim {
    std::vector<double> ts_host;     // t_vector (host copy for the writers)
    ctl::Fitness fit;                // fitness inputs captured at creation
    ctl_model_info info{};
    ctl_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int n = 0, ns = 0, block = 256, grid = 0;
    bool lds = false;
    size_t smem = 0;
    KParams kp{};
    std::vector<void*> allocs;
    bool ran = false;
    ctl_stats stats{};
    // fitness
    uint32_t* d_target = nullptr;
    int8_t* d_t2e = nullptr;
};

namespace {

int hip_fail(hipError_t e, const char* what) {
    char b[256];
    std::snprintf(b, sizeof b, "%s: %s", what, hipGetErrorString(e));
    return ctl::fail(e == hipErrorOutOfMemory ? CTL_E_NOMEM : CTL_E_DEVICE, b);
}

#define HIPCHK(call, what)                         \
    do {                                           \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

template <typename T>
int dalloc(ctl_sim* s, T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    s->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

template <typename T>
int upload(ctl_sim* s, T** p, const T* src, size_t count) {
    int rc = dalloc(s, p, count);
    if (rc) return rc;
    HIPCHK(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return 0;
}

void release(ctl_sim* s) {
    for (void* q : s->allocs) (void)hipFree(q);
    s->allocs.clear();
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->ev0 = s->ev1 = nullptr;
    s->stream = nullptr;
}

}  // namespace

int ctl::sim_dims(const ctl_sim* s, int* n, int* K, int* SS, const double** ts) {
    if (!s) return ctl::fail(CTL_E_ARGS, "null sim");
    *n = s->n;
    *K = s->kp.K;
    *SS = s->kp.SS;
    *ts = s->ts_host.data();
    return 0;
}

extern "C" {

int ctl_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ctl_sim_create(const ctl_model* mh, const ctl_opts* o, ctl_sim** out) {
    ctl::clear_error();
    if (!mh || !o || !out) return ctl::fail(CTL_E_ARGS, "null argument");
    *out = nullptr;
    const ctl::Model& m = mh->m;
    if (o->threads <= 0 || o->blocks <= 0) return ctl::fail(CTL_E_ARGS, "threads and blocks must be positive");
    const int64_t n64 = (int64_t)o->threads * o->blocks;
    if (n64 > 0x7fffffffLL) return ctl::fail(CTL_E_LIMIT, "too many simulations");
    const int n = (int)n64;
    if (m.x0_rows != 1 && m.x0_rows < n) return ctl::fail(CTL_E_ROWS, "MX_0 has fewer rows than simulations");
    if (m.c_rows != 1 && m.c_rows < n) return ctl::fail(CTL_E_ROWS, "c_matrix has fewer rows than simulations");
    if (o->nc < 0 || o->ssa_steps <= 0 || !(o->eps > 0.0f)) return ctl::fail(CTL_E_ARGS, "bad tunables");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return ctl::fail(CTL_E_DEVICE, "no HIP device available: libctl needs an AMD GPU (gfx950)");
    if (o->gpu < 0 || o->gpu >= ndev) return ctl::fail(CTL_E_DEVICE, "gpu ordinal out of range");

    ctl_sim* s = new ctl_sim();
    s->ts_host = m.ts;
    s->fit = m.fit;
    ctl_model_info_get(mh, &s->info);
    s->opts = *o;
    s->device = o->gpu;
    s->n = n;
    auto bail = [&](int rc) { release(s); delete s; return rc; };
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) return bail(hip_fail(e, "hipSetDevice"));
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreate(&s->ev0)) != hipSuccess || (e = hipEventCreate(&s->ev1)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));

    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, s->device);
    if (e != hipSuccess) return bail(hip_fail(e, "hipGetDeviceProperties"));
    const int S = m.S, R = m.R, K = (int)m.ts.size(), SS = (int)m.cs.size();
    // Launch geometry.  Per-lane working set (state, trial state,
    // propensities, per-simulation rates) in LDS when a 64-lane wavefront's
    // share stays within 16 KB (>= 8 wavefronts per CU); otherwise in the
    // global slot arrays with 256-lane workgroups.
    const size_t lane_words = (size_t)2 * S + R + (m.c_rows == 1 ? 0 : R);
    const bool lds_auto = lane_words * 4 * 64 <= 16384;
    s->lds = o->block_size < 0 ? false : lds_auto;
    s->block = o->block_size > 0 ? o->block_size : (s->lds ? 64 : 256);
    if (s->block % 64 != 0 || s->block > 256) return bail(ctl::fail(CTL_E_ARGS, "block_size must be 64..256, multiple of 64"));
    s->smem = s->lds ? lane_words * 4 * s->block : 0;
    if (s->smem > 65536) { s->lds = false; s->smem = 0; }
    const void* kfn = s->lds ? (const void*)tau_leap_kernel<true> : (const void*)tau_leap_kernel<false>;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, s->block, s->smem);
    if (e != hipSuccess || per_cu <= 0) per_cu = 1;
    int ns = o->slots > 0 ? o->slots : prop.multiProcessorCount * per_cu * s->block;
    if (ns > n) ns = ((n + s->block - 1) / s->block) * s->block;
    s->ns = ns;
    s->grid = (ns + s->block - 1) / s->block;

    KParams& p = s->kp;
    p.S = S; p.R = R; p.K = K; p.SS = SS;
    int rc = 0;
    std::vector<uint32_t> horw(S);
    for (int i = 0; i < S; ++i) horw[i] = (uint32_t)m.hor[i] | ((uint32_t)m.hort[i] << 8);
    std::vector<float> cf(m.c.size());
    for (size_t i = 0; i < m.c.size(); ++i) cf[i] = (float)m.c[i];
    // SSA dependency graph: for each reaction j, the reactions with a
    // reactant among the species j changes (its fear entries), ascending.
    std::vector<uint32_t> Dptr(R + 1, 0), Dent;
    {
        std::vector<std::vector<uint32_t>> readers(S);
        for (int r = 0; r < R; ++r)
            for (uint32_t h = m.Lptr[r]; h < m.Lptr[r + 1]; ++h) readers[m.Lent[h] & 0xffffu].push_back((uint32_t)r);
        std::vector<char> mark(R, 0);
        std::vector<uint32_t> list;
        for (int j = 0; j < R; ++j) {
            Dptr[j] = (uint32_t)Dent.size();
            list.clear();
            for (uint32_t h = m.Fptr[j]; h < m.Fptr[j + 1]; ++h)
                for (uint32_t r : readers[m.Fent[h] & 0xffffu])
                    if (!mark[r]) { mark[r] = 1; list.push_back(r); }
            std::sort(list.begin(), list.end());
            for (uint32_t r : list) { mark[r] = 0; Dent.push_back(r); }
        }
        Dptr[R] = (uint32_t)Dent.size();
    }
    uint32_t *Lptr, *Lent, *Fptr, *Fent, *Vptr, *Vent, *Dp, *De, *hw;
    int32_t* cs;
    double* ts;
    if ((rc = upload(s, &Lptr, m.Lptr.data(), m.Lptr.size())) || (rc = upload(s, &Lent, m.Lent.data(), m.Lent.size())) ||
        (rc = upload(s, &Fptr, m.Fptr.data(), m.Fptr.size())) || (rc = upload(s, &Fent, m.Fent.data(), m.Fent.size())) ||
        (rc = upload(s, &Vptr, m.Vptr.data(), m.Vptr.size())) || (rc = upload(s, &Vent, m.Vent.data(), m.Vent.size())) ||
        (rc = upload(s, &Dp, Dptr.data(), Dptr.size())) || (rc = upload(s, &De, Dent.data(), Dent.size())) ||
        (rc = upload(s, &hw, horw.data(), horw.size())) || (rc = upload(s, &cs, m.cs.data(), m.cs.size())) ||
        (rc = upload(s, &ts, m.ts.data(), m.ts.size())))
        return bail(rc);
    p.Lptr = Lptr; p.Lent = Lent; p.Fptr = Fptr; p.Fent = Fent; p.Vptr = Vptr; p.Vent = Vent;
    p.Dptr = Dp; p.Dent = De;
    p.horw = hw; p.cs = cs; p.ts = ts; p.tmax = m.ts.back();
    float* cdev;
    uint32_t* xdev;
    if ((rc = upload(s, &cdev, cf.data(), cf.size())) || (rc = upload(s, &xdev, m.x0.data(), m.x0.size())))
        return bail(rc);
    p.c_bcast = m.c_rows == 1 ? cdev : nullptr;
    p.c_sims = m.c_rows == 1 ? nullptr : cdev;
    p.x0_bcast = m.x0_rows == 1 ? xdev : nullptr;
    p.x0_sims = m.x0_rows == 1 ? nullptr : xdev;
    p.n = n;
    p.sim_base = (uint32_t)o->offset;
    p.key0 = (uint32_t)o->seed;
    p.key1 = (uint32_t)(o->seed >> 32);
    p.nc = o->nc;
    p.eps = o->eps;
    p.ssa_steps = o->ssa_steps;
    p.ssa_factor = o->ssa_factor;
    p.force_ssa = o->force_ssa ? 1 : 0;
    p.ssa_chunk = o->ssa_chunk > 0 ? o->ssa_chunk : 10;
    p.iters = o->iters_per_launch > 0 ? o->iters_per_launch : 4096;
    p.ns = ns;
    if ((rc = dalloc(s, &p.X, (size_t)2 * S * ns)) || (rc = dalloc(s, &p.A, (size_t)R * ns))) return bail(rc);
    p.Cs = nullptr;
    if (!p.c_bcast && (rc = dalloc(s, &p.Cs, (size_t)R * ns))) return bail(rc);
    SlotState& st = p.st;
    if ((rc = dalloc(s, &st.sim, ns)) || (rc = dalloc(s, &st.t, ns)) || (rc = dalloc(s, &st.tacc, ns)) ||
        (rc = dalloc(s, &st.cursor, ns)) || (rc = dalloc(s, &st.mode, ns)) || (rc = dalloc(s, &st.evc, ns)) ||
        (rc = dalloc(s, &st.ev, ns)) || (rc = dalloc(s, &st.blk, ns)) || (rc = dalloc(s, &st.pos, ns)) ||
        (rc = dalloc(s, &st.ssa_left, ns)) || (rc = dalloc(s, &st.xb, ns)) || (rc = dalloc(s, &st.firings, ns)) ||
        (rc = dalloc(s, &st.ssa_steps, ns)) || (rc = dalloc(s, &st.poisson, ns)) || (rc = dalloc(s, &st.gaps, ns)) ||
        (rc = dalloc(s, &st.leaps, ns)) ||
        (rc = dalloc(s, &st.rejections, ns)) || (rc = dalloc(s, &st.switches, ns)))
        return bail(rc);
    if ((rc = dalloc(s, &p.storage, (size_t)n * K * SS)) || (rc = dalloc(s, &p.simstat, n)) ||
        (rc = dalloc(s, &p.xfinal, (size_t)n * S)) || (rc = dalloc(s, &p.tfinal, n)) ||
        (rc = dalloc(s, &p.next_sim, 2)))
        return bail(rc);
    p.done = p.next_sim + 1;
    s->stats.slots = ns;
    s->stats.block_size = s->block;
    s->stats.lds_state = s->lds ? 1 : 0;
    *out = s;
    return 0;
}

int ctl_sim_run(ctl_sim* s) {
    ctl::clear_error();
    if (!s) return ctl::fail(CTL_E_ARGS, "null sim");
    KParams& p = s->kp;
    HIPCHK(hipSetDevice(s->device), "hipSetDevice");
    const auto w0 = std::chrono::steady_clock::now();
    // PrepareVariables: zero the samples, reset slots and the queue.
    HIPCHK(hipMemsetAsync(p.storage, 0, sizeof(uint32_t) * (size_t)s->n * p.K * p.SS, s->stream), "memset");
    HIPCHK(hipMemsetAsync(p.st.sim, 0xff, sizeof(int32_t) * s->ns, s->stream), "memset");
    HIPCHK(hipMemsetAsync(p.next_sim, 0, 2 * sizeof(uint32_t), s->stream), "memset");
    ctl_stats& S = s->stats;
    S.kernel_ms = 0.0;
    S.launches = 0;
    uint32_t done = 0;
    const int max_launches = 1 << 20;
    float total_ms = 0.0f;
    while (done < (uint32_t)s->n) {
        HIPCHK(hipEventRecord(s->ev0, s->stream), "hipEventRecord");
        if (s->lds)
            hipLaunchKernelGGL(tau_leap_kernel<true>, dim3(s->grid), dim3(s->block), s->smem, s->stream, p);
        else
            hipLaunchKernelGGL(tau_leap_kernel<false>, dim3(s->grid), dim3(s->block), 0, s->stream, p);
        HIPCHK(hipGetLastError(), "tau_leap_kernel launch");
        HIPCHK(hipEventRecord(s->ev1, s->stream), "hipEventRecord");
        HIPCHK(hipMemcpyAsync(&done, p.done, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream), "hipMemcpy D2H");
        HIPCHK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1), "hipEventElapsedTime");
        total_ms += ms;
        S.launches++;
        if (s->opts.verbose)
            std::printf(" * launch %d: %u/%d simulations done (%.3f ms)\n", S.launches, done, s->n, ms);
        if (S.launches >= max_launches) return ctl::fail(CTL_E_STATE, "simulation did not finish");
    }
    S.kernel_ms = total_ms;
    // totals
    std::vector<ctl_sim_stat> st(s->n);
    HIPCHK(hipMemcpy(st.data(), p.simstat, sizeof(ctl_sim_stat) * s->n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    S.firings = S.leaps = S.rejections = S.switches = S.ssa_steps = S.poisson = S.gaps = S.absorbed = 0;
    for (const auto& x : st) {
        S.firings += x.firings; S.leaps += x.leaps; S.rejections += x.rejections; S.switches += x.switches;
        S.ssa_steps += x.ssa_steps; S.poisson += x.poisson; S.gaps += x.gaps;
        S.absorbed += x.absorbed;
    }
    S.simulations = (uint64_t)s->n;
    S.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    s->ran = true;
    return 0;
}

int ctl_sim_samples_device(const ctl_sim* s, void** dev_ptr, size_t* count) {
    if (!s || !dev_ptr || !count) return ctl::fail(CTL_E_ARGS, "null argument");
    *dev_ptr = s->kp.storage;
    *count = (size_t)s->n * s->kp.K * s->kp.SS;
    return 0;
}

int ctl_sim_samples(const ctl_sim* s, uint32_t* out, size_t count, int layout) {
    ctl::clear_error();
    if (!s || !out) return ctl::fail(CTL_E_ARGS, "null argument");
    if (!s->ran) return ctl::fail(CTL_E_STATE, "ctl_sim_run has not completed");
    const size_t total = (size_t)s->n * s->kp.K * s->kp.SS;
    if (count != total) return ctl::fail(CTL_E_ARGS, "sample buffer size must be n*K*SS");
    HIPCHK(hipSetDevice(s->device), "hipSetDevice");
    if (layout == 1) {
        HIPCHK(hipMemcpy(out, s->kp.storage, total * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return 0;
    }
    if (layout != 0) return ctl::fail(CTL_E_ARGS, "layout must be 0 or 1");
    uint32_t* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, total * sizeof(uint32_t)), "hipMalloc");
    const int KS = s->kp.K * s->kp.SS;
    const size_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(to_reference_layout, dim3((unsigned)blocks), dim3(256), 0, s->stream, s->kp.storage, tmp, s->n, KS);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(out, tmp, total * sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "to_reference_layout");
    return 0;
}

int ctl_sim_stats(const ctl_sim* s, ctl_stats* out) {
    if (!s || !out) return ctl::fail(CTL_E_ARGS, "null argument");
    *out = s->stats;
    return 0;
}

int ctl_sim_sim_stats(const ctl_sim* s, ctl_sim_stat* out, size_t n) {
    if (!s || !out) return ctl::fail(CTL_E_ARGS, "null argument");
    if (!s->ran) return ctl::fail(CTL_E_STATE, "ctl_sim_run has not completed");
    if (n != (size_t)s->n) return ctl::fail(CTL_E_ARGS, "n must equal the simulation count");
    HIPCHK(hipSetDevice(s->device), "hipSetDevice");
    HIPCHK(hipMemcpy(out, s->kp.simstat, sizeof(ctl_sim_stat) * n, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return 0;
}

int ctl_sim_final_state(const ctl_sim* s, int32_t* x, double* t) {
    if (!s) return ctl::fail(CTL_E_ARGS, "null argument");
    if (!s->ran) return ctl::fail(CTL_E_STATE, "ctl_sim_run has not completed");
    HIPCHK(hipSetDevice(s->device), "hipSetDevice");
    if (x) HIPCHK(hipMemcpy(x, s->kp.xfinal, sizeof(int32_t) * (size_t)s->n * s->kp.S, hipMemcpyDeviceToHost), "D2H");
    if (t) HIPCHK(hipMemcpy(t, s->kp.tfinal, sizeof(double) * (size_t)s->n, hipMemcpyDeviceToHost), "D2H");
    return 0;
}

int ctl_fitness(const ctl_sim* sc, double* out, size_t count) {
    ctl::clear_error();
    ctl_sim* s = const_cast<ctl_sim*>(sc);
    if (!s || !out) return ctl::fail(CTL_E_ARGS, "null argument");
    if (!s->ran) return ctl::fail(CTL_E_STATE, "ctl_sim_run has not completed");
    const ctl::Fitness& f = s->fit;
    if (!f.loaded) return ctl::fail(CTL_E_STATE, "fitness inputs not loaded");
    const int threads = s->opts.threads, P = (int)f.par_repetitions;
    const int groups = s->opts.blocks / P;
    const size_t nf = (size_t)threads * groups;
    if (count != nf) return ctl::fail(CTL_E_ARGS, "fitness buffer size must be threads * (blocks / p_rep)");
    if ((int)f.target_quantities != s->kp.SS)
        return ctl::fail(CTL_E_ARGS, "ts_numtgt must equal the number of sampled species");
    HIPCHK(hipSetDevice(s->device), "hipSetDevice");
    int rc;
    if (!s->d_target) {
        if ((rc = upload(s, &s->d_target, f.series.data(), f.series.size()))) return rc;
        if ((rc = upload(s, &s->d_t2e, f.thread2experiment.data(), f.thread2experiment.size()))) return rc;
    }
    double* d_fit;
    if ((rc = dalloc(s, &d_fit, nf))) return rc;
    hipLaunchKernelGGL(fitness_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s->stream, s->kp.storage,
                       s->d_target, s->d_t2e, d_fit, threads, groups, P, s->kp.K, s->kp.SS, (int)f.experiments,
                       (int)f.repetitions);
    HIPCHK(hipGetLastError(), "fitness_kernel launch");
    HIPCHK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    HIPCHK(hipMemcpy(out, d_fit, sizeof(double) * nf, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return 0;
}

void ctl_sim_destroy(ctl_sim* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    release(s);
    delete s;
}

}  // extern "C"
