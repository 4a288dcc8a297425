// ctl_kernels.h -- device side of the simulation kernels (gfx950): per-slot
// state, the compiled-model tables, propensities, the adaptive leap, SSA
// steps and the lane life cycle.  Included by the single-model runtime
// (ctl_sim.hip) and the batched runtime (ctl_batch.hip).
//
// Layout in HBM (ns = number of slots):
//   per slot, structure of arrays, coalesced across a wavefront:
//     X  int32 [2][S][ns]  state and trial state (parity bit swaps them)
//     A  uint32[R][ns]     propensity (FP32 bits) | critical flag << 31
//     Cs float [R][ns]     rate constants when rows differ between sims
//   per simulation: samples uint32 [sim][K][SS], stats, final X [sim][S]
// Small models keep X, X1, A (and Cs) in LDS instead, same [index][lane]
// layout.  The compiled model (flats, HOR) is read through scalar loads:
// every lane of a wavefront walks the same reaction / species loops, so each
// table entry is a wave-uniform constant-cache broadcast.
#pragma once

#include <hip/hip_runtime.h>

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
    uint32_t* blk;          // one lane: next Philox block; group kernels: words of the event consumed
    uint32_t* pos;          // one lane: words of the buffered block used (group kernels: unused)
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
    const uint32_t* FSptr;    // fear entries species-major, ascending reaction
    const uint32_t* FSent;    //   (group kernel's per-species update)
    // The u32 tables packed in one buffer, each at a 32-bit word offset: the
    // group kernels address every table from one base pointer, which keeps
    // their scalar-register footprint small.
    const uint32_t* T;
    int oLp, oLe, oVp, oVe, oHw, oFSp, oFSe, oFp, oFe, oDp, oDe, oCs;
    // Lane-interleaved round streams of the group kernel, built for its G:
    // round r covers reactions (or species) r*G .. r*G+G-1, lane l takes the
    // r*G+l-th; the round holds len = T[o*m + r] entries per lane, entry k of
    // lane l at T[o*e + base_r + k*G + l] (base_r = G * sum of the earlier
    // lengths).  Shorter lists are padded with 0 (multiplicity / change 0);
    // var and fear round lengths are multiples of 4 (loads issue 4 at a time).
    // Reactant entries (species | multiplicity << 16; with pfast the index of
    // the reactant's factor in the group's factor table, padding 2S), var
    // entries of the tau' sums and species-major fear entries (reaction |
    // change << 16).
    // Var and fear rounds list species in the order T[oSPI + r*G + l]
    // (species | HOR << 16 | HORtype << 24; 0xffffffff: none), longest lists
    // first.  gsimple: no species has HOR 3 or HOR 2 of type 2, so g is 1 or 2.
    int oPRm, oPRe, oVRm, oVRe, oFRm, oFRe, oSPI, gsimple;
    int vhalf;                // var entries as FP16 coefficient << 16 | byte offset of A[reaction] (every |v| <= 2048)
    int oDR;                  // pfast 1 / 2: SSA dependent records {r, entry 0, entry 1, 0} (uint4, Dptr order),
                              //   the entries as factor indices (1) or species | multiplicity << 16 (2)
    int pfast;                // propensity mode: 0 from the reactant entries; with <= 2 reactant entries of
                              //   multiplicity <= 2 per reaction (exactly 2 entries per round): 1 factor
                              //   indices into the LDS factor table, 2 entries read from the state
    const uint32_t* horw;     // hor | hort << 8
    const int32_t* cs;
    const double* ts;
    double tmax;
    const float* c_bcast;     // R (rates shared by every simulation) or null
    const float* c_sims;      // [row][R] or null; row = sim / c_div
    const uint32_t* x0_bcast; // S or null
    const uint32_t* x0_sims;  // [row][S] or null; row = sim / x0_div
    int c_div, x0_div;        // simulations per input row (1: one row each)
    int nfs;                  // per-simulation feeds: species fs[u] reset to
    const int32_t* fs;        //   fv[sim / fv_div][u] after every accepted leap
    const uint32_t* fv;
    int fv_div;
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
    int compact;              // iterations between two regroupings of a workgroup's
                              //   simulations by mode (0: never)
    // slots (global per-lane storage, [.][ns])
    int ns;
    int32_t* X;               // [2][S][ns]
    uint32_t* A;              // [R][ns]
    float* Cs;                // [R][ns] when rates differ between simulations
    const SlotState* st;      // device copy of the slot arrays' pointers (read at launch start / end)
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
typedef const __attribute__((address_space(4))) int64_t* ctab_i64;
#define CT_U32(ptr) ((ctab_u32)(ptr))
#define CT_I32(ptr) ((ctab_i32)(ptr))
#define CT_F32(ptr) ((ctab_f32)(ptr))
#define CT_F64(ptr) ((ctab_f64)(ptr))
#define CT_I64(ptr) ((ctab_i64)(ptr))

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
        // 0 < floor(x / v) < nc, as compares: v <= x < nc * v
        crit = crit || ((uint32_t)x >= (uint32_t)v && (uint64_t)(uint32_t)x < (uint64_t)nc * (uint32_t)v);
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
    if (p.nfs > 0) {                     // feeds (T.cu:2072-2079)
        int32_t* Xn = xcur(V, L);
        const uint32_t* fv = p.fv + (size_t)(L.sim / p.fv_div) * p.nfs;
        for (int u = 0; u < p.nfs; ++u) Xn[(size_t)CT_I32(p.fs)[u] * V.st] = (int32_t)fv[u];
    }
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
    L.sim = p.st->sim[s];
    L.g.k0 = p.key0;
    L.g.k1 = p.key1;
    L.absorbed = 0;
    if (L.sim < 0) return;               // idle slot: fetch() initialises the lane
    L.t = p.st->t[s];
    L.tacc = p.st->tacc[s];
    L.cursor = p.st->cursor[s];
    L.mode = p.st->mode[s];
    L.evc = p.st->evc[s];
    L.ssa_left = p.st->ssa_left[s];
    L.xb = p.st->xb[s];
    L.firings = p.st->firings[s];
    L.ssa_steps = p.st->ssa_steps[s];
    L.poisson = p.st->poisson[s];
    L.gaps = p.st->gaps[s];
    L.leaps = p.st->leaps[s];
    L.rejections = p.st->rejections[s];
    L.switches = p.st->switches[s];
    L.g.sim = p.sim_base + (uint32_t)L.sim;
    L.g.ev = p.st->ev[s];
    L.g.blk = p.st->blk[s];
    L.g.pos = p.st->pos[s];
    if (L.g.pos < 4u) {   // regenerate the partially consumed block of an SSA burst
        philox_block(L.g.blk - 1u, L.g.ev, L.g.sim, L.g.tag, L.g.k0, L.g.k1, L.g.b0, L.g.b1, L.g.b2, L.g.b3);
        for (uint32_t i = 0; i < L.g.pos; ++i) {   // drop the words already consumed
            L.g.b0 = L.g.b1;
            L.g.b1 = L.g.b2;
            L.g.b2 = L.g.b3;
        }
    }
}

__device__ __forceinline__ void store_lane(const KParams& p, const Lane& L) {
    const int s = L.slot;
    p.st->sim[s] = L.sim;
    if (L.sim < 0) return;
    p.st->t[s] = L.t;
    p.st->tacc[s] = L.tacc;
    p.st->cursor[s] = L.cursor;
    p.st->mode[s] = L.mode;
    p.st->evc[s] = L.evc;
    p.st->ssa_left[s] = L.ssa_left;
    p.st->xb[s] = L.xb;
    p.st->firings[s] = L.firings;
    p.st->ssa_steps[s] = L.ssa_steps;
    p.st->poisson[s] = L.poisson;
    p.st->gaps[s] = L.gaps;
    p.st->leaps[s] = L.leaps;
    p.st->rejections[s] = L.rejections;
    p.st->switches[s] = L.switches;
    p.st->ev[s] = L.g.ev;
    p.st->blk[s] = L.g.blk;
    p.st->pos[s] = L.g.pos;
}

// Per-simulation rates into the lane's rate array (rows that differ).
__device__ __forceinline__ void load_rates(const KParams& p, const View& V, int sim) {
    if (!p.c_bcast)
        for (int j = 0; j < p.R; ++j) V.Cs[(size_t)j * V.st] = p.c_sims[(size_t)(sim / p.c_div) * p.R + j];
}

// Takes the next simulation from the queue and initialises the lane
// (PrepareVariables, T.cu:1269-1297).  Returns false when the queue is empty.
__device__ __forceinline__ void start(const KParams& p, Lane& L, const View& V, uint32_t q) {
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
        X[(size_t)s * V.st] = (int32_t)(p.x0_bcast ? CT_U32(p.x0_bcast)[s] : p.x0_sims[(size_t)(q / (uint32_t)p.x0_div) * p.S + s]);
    load_rates(p, V, (int)q);
}

__device__ __forceinline__ bool fetch(const KParams& p, Lane& L, const View& V) {
    const uint32_t q = atomicAdd(p.next_sim, 1u);
    if (q >= (uint32_t)p.n) { L.sim = -2; return false; }
    start(p, L, V, q);
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


// ---------------------------------------------------------------------------
// Regrouping of a workgroup's slots by mode (one-lane kernel).  A lane runs
// whatever slot it holds: a wavefront whose lanes hold slots in both modes
// executes a leap and an SSA chunk one after the other, each with part of its
// lanes idle.  Every p.compact iterations the lanes of a workgroup hand their
// slots over through LDS so that leap-mode slots fill the first wavefronts,
// SSA-mode slots the next ones and idle slots the last: each wavefront then
// runs one mode.  A slot keeps its column (its state arrays, its global slot
// index): only the lane that runs it changes, so no trajectory does.
// ---------------------------------------------------------------------------
constexpr int LANE_WORDS = 30;      // words of a Lane's per-slot state

__device__ __forceinline__ void lane_put(const Lane& L, uint32_t* W, int col, int B) {
    uint32_t* w = W + col;
    int k = 0;
    auto put32 = [&](uint32_t x) { w[(size_t)(k++) * B] = x; };
    auto put64 = [&](uint64_t x) { put32((uint32_t)x); put32((uint32_t)(x >> 32)); };
    put32((uint32_t)L.sim);
    put64((uint64_t)__double_as_longlong(L.t));
    put64((uint64_t)__double_as_longlong(L.tacc));
    put32(L.cursor); put32((uint32_t)L.mode); put32(L.evc); put32((uint32_t)L.ssa_left); put32(L.xb);
    put64(L.firings); put64(L.ssa_steps); put64(L.poisson); put64(L.gaps);
    put32(L.leaps); put32(L.rejections); put32(L.switches); put32(L.absorbed);
    put32(L.g.sim); put32(L.g.ev); put32(L.g.blk);
    put32(L.g.b0); put32(L.g.b1); put32(L.g.b2); put32(L.g.b3); put32(L.g.pos);
}

__device__ __forceinline__ void lane_get(Lane& L, const uint32_t* W, int col, int B) {
    const uint32_t* w = W + col;
    int k = 0;
    auto get32 = [&]() { return w[(size_t)(k++) * B]; };
    auto get64 = [&]() { const uint64_t lo = get32(); return lo | ((uint64_t)get32() << 32); };
    L.sim = (int)get32();
    L.t = __longlong_as_double((long long)get64());
    L.tacc = __longlong_as_double((long long)get64());
    L.cursor = get32(); L.mode = (int)get32(); L.evc = get32(); L.ssa_left = (int)get32(); L.xb = get32();
    L.firings = get64(); L.ssa_steps = get64(); L.poisson = get64(); L.gaps = get64();
    L.leaps = get32(); L.rejections = get32(); L.switches = get32(); L.absorbed = get32();
    L.g.sim = get32(); L.g.ev = get32(); L.g.blk = get32();
    L.g.b0 = get32(); L.g.b1 = get32(); L.g.b2 = get32(); L.g.b3 = get32(); L.g.pos = get32();
}

// New column of this lane after a stable partition of the workgroup's
// columns by key (0 leap, 1 SSA, 2 idle); cnt is LDS scratch of 3 words per
// wavefront, perm of B words.  Every lane of the workgroup must call it.
__device__ __forceinline__ int regroup(int key, int col, uint32_t* cnt, uint32_t* perm) {
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));      // lanes below this one
    uint32_t rank = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint64_t b = __ballot(key == k);
        if (key == k) rank = (uint32_t)__popcll(b & lt);
        if (lane == 0) cnt[wave * 3 + k] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    uint32_t pos = 0;
    for (int k = 0; k < key; ++k)
        for (int w = 0; w < nw; ++w) pos += cnt[w * 3 + k];
    for (int w = 0; w < wave; ++w) pos += cnt[w * 3 + key];
    perm[pos + rank] = (uint32_t)col;
    __syncthreads();
    return (int)perm[threadIdx.x];
}

}  // namespace ctl
