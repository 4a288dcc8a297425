// ctl_group.h -- the group kernel: G lanes of a wavefront (G = 2..64, a
// power of two) cooperate on one simulation, for networks whose per-lane
// working set does not fit on chip.  Included by ctl_sim.hip.
//
// Work split: reaction j belongs to lane j mod G; species go to lanes in
// rounds of G, longest var/fear lists first (the round streams).  A group's
// simulation state lives in LDS for the whole launch:
//     X   int32 [2][S]   state and trial state (parity bit swaps them)
//     A   uint32[Rp]     propensity (FP32 bits) | critical flag << 31; zero
//                        past R (Rp = R + 1 rounded up to a multiple of G)
//     Kc  int32 [Rp]     counts drawn for the current leap attempt
//     Q   uint16[R]      reactions queued for a Poisson draw (lambda >= 1), in
//                        the trial-state half of X while it is free
//     F   uint32[2S+1]   reactant factors of the current state (reactions of
//                        at most two reactant entries of multiplicity <= 2):
//                        F[s] = x_s, F[S+s] = x_s (x_s - 1) / 2 as FP32 bits,
//                        each with its criticality in bit 31, F[2S] = 1.0; in
//                        the Kc region, which is free outside the leap's
//                        count and update phases
// so every table-driven access is an LDS access instead of a dependent HBM
// load.  Group-wide values (a0, tau, the chosen reaction) are computed with
// DPP row operations and lane reads and are identical in every lane of the
// group, so the group's control flow is uniform.
//
// Algorithm (restated lane by lane by the oracle's group mode):
//  * propensities and criticality per reaction as in the one-lane kernel;
//  * a0 and every cumulative selection use the lane partial sums (each
//    lane's reactions in ascending order) combined by an inclusive scan over
//    the G lanes (Hillis-Steele inside rows of 16 lanes, then the row totals
//    carried across rows; see gscan); a selection takes the first lane with
//    a positive partial whose inclusive sum reaches the threshold (else the
//    last lane with a positive partial), then walks that lane's reactions;
//  * tau' per species as in the one-lane kernel, minimum across lanes;
//  * group-wide draws (tau'', critical pick, SSA uniforms) come from the
//    event's Philox stream with counter word 3 = 0; lane l's small-lambda
//    Poisson process from the stream with word 3 = 1 + l (first gap drawn at
//    the start of the attempt); the count of a reaction with lambda >= 1
//    from the reaction's own stream (word 3 = 2^31 | j), so those draws can
//    be queued and spread over the group's lanes;
//  * the leap update and its negativity / int32 test run per species over
//    the species-major fear entries in ascending reaction order, which
//    rejects exactly when the reaction-major order of test_update
//    (2phase_n-tau-leaping.cu:1300-1351) would.
#pragma once

#include "ctl_kernels.h"

#pragma clang fp contract(off)

namespace ctl {

// LDS ordering between lanes of one wavefront (a group never spans two).
__device__ __forceinline__ void gsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Phase clock, compiled only into the profiling build of the library
// (CTL_PHASE_CLOCK=1: libctl_phases.so, ctl_phase_clock): the time the
// wavefronts spend in each phase of the group kernel, from s_memtime read at
// phase boundaries and summed over the launch's wavefronts.  The segment
// between two consecutive marks of a wavefront is added to the phase of the
// later mark by the wavefront's lowest active lane (the last mark's time is
// kept per wavefront, so the two sides of a divergent branch count once
// each).  The normal build compiles the marks to nothing.
// ---------------------------------------------------------------------------
enum : int {
    PH_PROPS = 0, PH_TAU1, PH_PICK, PH_COUNTS, PH_POISSON, PH_UPDATE, PH_SAMPLE,
    PH_SSA_OPEN, PH_SSA_SELECT, PH_SSA_UPDATE, PH_SSA_DEPS,
    PH_PRE_LEAP, PH_PRE_SSA, PH_LOOP, PH_REGROUP, PH_PARK, PH_COUNT
};
#if CTL_PHASE_CLOCK
static __device__ unsigned long long phase_total[PH_COUNT];
static __shared__ unsigned long long phase_row[4][PH_COUNT + 1];    // per wavefront; last: time of the last mark
__device__ __forceinline__ bool phase_leader() {
    return (int)(threadIdx.x & 63u) == __builtin_ffsll((long long)__ballot(1)) - 1;
}
// The clock read is a volatile asm with a memory clobber, so that the
// compiler neither merges the reads of two branches nor moves the row
// updates across them (a predicated if/else would interleave both sides).
__device__ __forceinline__ unsigned long long phase_now() {
    unsigned long long t;
    asm volatile("s_memtime %0\n\ts_waitcnt lgkmcnt(0)" : "=s"(t) : : "memory");
    return t;
}
__device__ __forceinline__ void phase_start() {
    if (phase_leader()) {
        unsigned long long* r = phase_row[threadIdx.x >> 6];
        for (int k = 0; k < PH_COUNT; ++k) r[k] = 0ull;
        r[PH_COUNT] = phase_now();
    }
}
__device__ __forceinline__ void phase_mark(int k) {
    const unsigned long long t = phase_now();
    if (phase_leader()) {
        unsigned long long* r = phase_row[threadIdx.x >> 6];
        r[k] += t - r[PH_COUNT];
        r[PH_COUNT] = t;
    }
}
__device__ __forceinline__ void phase_flush() {
    if (phase_leader()) {
        unsigned long long* r = phase_row[threadIdx.x >> 6];
        for (int k = 0; k < PH_COUNT; ++k)
            if (r[k]) {
                atomicAdd(&phase_total[k], r[k]);
                r[k] = 0ull;
            }
    }
}
#define CTL_PHASE_START() phase_start()
#define CTL_PHASE(k) phase_mark(k)
#define CTL_PHASE_FLUSH() phase_flush()
#else
#define CTL_PHASE_START() ((void)0)
#define CTL_PHASE(k) ((void)0)
#define CTL_PHASE_FLUSH() ((void)0)
#endif

// How propensities are computed (KParams::pfast): 0 from the reactant
// entries (any multiplicity), 1 as two entries of the LDS factor table, 2 as
// two entries computed from the state (models with at most two reactant
// entries of multiplicity <= 2, 1 when the table fits in the Kc region).
// PFM 0..2 fixes the mode at compile time (the simulation kernels: one
// instance per mode, so none keeps another's registers live); PFM 3 reads
// p.pfast (the batch kernels, whose models differ).
template <int PFM>
__device__ __forceinline__ int pmode(int pfast) {
    if constexpr (PFM == 3) return pfast;
    else return PFM;
}

// Two reactant entries (multiplicity 0, 1 or 2) from the state: the
// arithmetic and the criticality test of greactant, in the same order.
__device__ __forceinline__ void gtwo(float& a, bool& c, int32_t x0, uint32_t v0, int32_t x1, uint32_t v1,
                                     uint32_t nc) {
    const int32_t xs[2] = {x0, x1};
    const uint32_t vs[2] = {v0, v1};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int32_t x = xs[k];
        const uint32_t v = vs[k];
        const float xf = (float)x;
        const float f2 = (x < 2) ? 0.0f : (0.5f * xf) * (float)(x - 1);
        a = a * (v == 1u ? xf : (v == 2u ? f2 : 1.0f));
        c = c || ((uint32_t)x >= v && (uint32_t)x < nc * v);
    }
}

// A kernel parameter read at its point of use: the compiler may not hoist a
// test of it out of the main loop, where it would hold a 64-bit lane mask in
// scalar registers for the whole kernel (and spill others).
static __device__ __forceinline__ int kuse(int v) {
#if CTL_UNIFORM_PARAMS
    v = __builtin_amdgcn_readfirstlane(v);         // uniform; a no-op when already scalar
    asm volatile("" : "+s"(v));
#endif
    return v;
}

template <int G>
__device__ __forceinline__ uint64_t gballot(bool pred) {
    const uint64_t b = __ballot(pred);
    if (G == 64) return b;
    const int base = (int)(threadIdx.x & 63u) & ~(G - 1);
    return (b >> base) & ((1ull << G) - 1ull);
}

// Cross-lane moves inside a row of 16 lanes through DPP (a VALU operand
// modifier, no LDS round trip).  Lanes whose source is outside the row read 0.
template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
enum : int {
    DPP_QP_1133 = 0xF5,       // quad_perm [1,1,3,3]
    DPP_QP_3333 = 0xFF,       // quad_perm [3,3,3,3]
    DPP_QP_1032 = 0xB1,       // quad_perm [1,0,3,2]  (xor 1)
    DPP_QP_2301 = 0x4E,       // quad_perm [2,3,0,1]  (xor 2)
    DPP_ROW_SHL4 = 0x104,
    DPP_ROW_SHR1 = 0x111,
    DPP_ROW_SHR2 = 0x112,
    DPP_ROW_SHR4 = 0x114,
    DPP_ROW_SHR8 = 0x118,
    DPP_ROW_MIRROR = 0x140,
    DPP_ROW_HALF_MIRROR = 0x141,
    DPP_ROW_NEWBCAST15 = 0x15F,   // lane 15 of each row to the whole row
    DPP_WAVE_SHR1 = 0x138,        // lane l reads lane l - 1 across rows
    DPP_ROW_BCAST15 = 0x142,      // lane 15 of each row to the next row
    DPP_ROW_BCAST31 = 0x143,      // lane 31 to rows 2 and 3
};

// Inclusive scan over the group.  Inside each DPP row of 16 lanes (the whole
// group when G <= 16) a Hillis-Steele scan, v_l <- v_{l-d} + v_l for
// d = 1, 2, 4, 8; wider groups then add the previous row's total to rows 1
// and 3 (row_bcast:15) and, for G = 64, the first two rows' total to rows 2
// and 3 (row_bcast:31).
template <int G>
__device__ __forceinline__ float gscan(float v, int l) {
    float t;
    if (G == 16) {
        // the group is the DPP row: a shifted-in lane reads 0, and 0 + v = v
        // (v is a sum of non-negative terms started from +0, never -0)
        v = dppf<DPP_ROW_SHR1>(v) + v;
        v = dppf<DPP_ROW_SHR2>(v) + v;
        v = dppf<DPP_ROW_SHR4>(v) + v;
        v = dppf<DPP_ROW_SHR8>(v) + v;
    } else if (G < 16) {
        if (G >= 2) { t = dppf<DPP_ROW_SHR1>(v); if (l >= 1) v = t + v; }
        if (G >= 4) { t = dppf<DPP_ROW_SHR2>(v); if (l >= 2) v = t + v; }
        if (G >= 8) { t = dppf<DPP_ROW_SHR4>(v); if (l >= 4) v = t + v; }
        if (G >= 16) { t = dppf<DPP_ROW_SHR8>(v); if (l >= 8) v = t + v; }
    } else {
        v = dppf<DPP_ROW_SHR1>(v) + v;              // a shifted-in lane reads 0
        v = dppf<DPP_ROW_SHR2>(v) + v;
        v = dppf<DPP_ROW_SHR4>(v) + v;
        v = dppf<DPP_ROW_SHR8>(v) + v;
        v = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), DPP_ROW_BCAST15, 0xa, 0xf, false)) + v;
        if (G == 64)
            v = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), DPP_ROW_BCAST31, 0xc, 0xf, false)) + v;
    }
    return v;
}

__device__ __forceinline__ float rdlane(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// The value of the group's last lane, in every lane of the group.
template <int G>
__device__ __forceinline__ float gbcast_last(float v, int l) {
    if (G == 2) return dppf<DPP_QP_1133>(v);
    if (G == 4) return dppf<DPP_QP_3333>(v);
    if (G == 8) {
        const float t = dppf<DPP_QP_3333>(v);
        const float u = dppf<DPP_ROW_SHL4>(t);
        return l < 4 ? u : t;
    }
    const int wl = (int)(threadIdx.x & 63u);
    if (G == 16) return dppf<DPP_ROW_NEWBCAST15>(v);
    if (G == 32) {
        const float r0 = rdlane(v, 31), r1 = rdlane(v, 63);
        return wl < 32 ? r0 : r1;
    }
    return rdlane(v, 63);
}

// Minimum over the group (exact in any order).
template <int G>
__device__ __forceinline__ float gmin(float v) {
    float o;
    if (G >= 2) { o = dppf<DPP_QP_1032>(v); v = o < v ? o : v; }
    if (G >= 4) { o = dppf<DPP_QP_2301>(v); v = o < v ? o : v; }
    if (G >= 8) { o = dppf<DPP_ROW_HALF_MIRROR>(v); v = o < v ? o : v; }
    if (G >= 16) { o = dppf<DPP_ROW_MIRROR>(v); v = o < v ? o : v; }
    if (G >= 32) {                                  // combine the rows' minima
        const float r0 = rdlane(v, 15), r1 = rdlane(v, 31), r2 = rdlane(v, 47), r3 = rdlane(v, 63);
        const float m01 = r1 < r0 ? r1 : r0, m23 = r3 < r2 ? r3 : r2;
        if (G == 64) v = m23 < m01 ? m23 : m01;
        else v = (threadIdx.x & 32u) ? m23 : m01;
    }
    return v;
}

template <int CTRL>
__device__ __forceinline__ uint64_t dppu64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xf, 0xf, false);
    return ((uint64_t)hi << 32) | lo;
}

template <int G>
__device__ __forceinline__ uint64_t gsum64(uint64_t v) {
    if constexpr (G <= 16) {
        // butterfly inside the DPP row (the pairings of gmin); integer sums are exact in any order
        if (G >= 2) v += dppu64<DPP_QP_1032>(v);
        if (G >= 4) v += dppu64<DPP_QP_2301>(v);
        if (G >= 8) v += dppu64<DPP_ROW_HALF_MIRROR>(v);
        if (G >= 16) v += dppu64<DPP_ROW_MIRROR>(v);
        return v;
    }
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, m, G);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, G);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// Group state: the scalars are identical in every lane of the group; the
// Poisson / gap counters are per lane and summed when the simulation ends.
// The group's stream (counter word 3 = 0) with its Philox blocks spread over
// the group: lane i holds block cb + i of the current event, so 4G words (2G
// SSA steps) cost each lane one Philox evaluation where a per-lane buffer
// costs one per 4 words, and the groups of a wavefront, at different points
// of their streams, seldom make the wavefront evaluate Philox for one of them
// alone.  A word is read from the lane that holds its block; the words and
// their order are those of a per-lane stream (rng_u32).
struct GRng {
    uint32_t k0, k1;             // key
    uint32_t sim;                // global simulation index
    uint32_t ev;                 // current event
    uint32_t wc;                 // words of the event consumed
    uint32_t cb;                 // first block held by the group (GRNG_NONE: none)
    uint32_t c0, c1, c2, c3;     // this lane's block, cb + lane
};
constexpr uint32_t GRNG_NONE = 0x80000000u;   // b - cb >= 2^31 for every block b of an event

__device__ __forceinline__ void grng_event(GRng& g, uint32_t ev) {
    g.ev = ev;
    g.wc = 0u;
    g.cb = GRNG_NONE;
}

template <int G>
__device__ __forceinline__ uint32_t grng_u32(GRng& g, int l) {
    const uint32_t b = g.wc >> 2;
    if (b - g.cb >= (uint32_t)G) {
        g.cb = b;
        philox_block(b + (uint32_t)l, g.ev, g.sim, 0u, g.k0, g.k1, g.c0, g.c1, g.c2, g.c3);
    }
    // the four words as values: a select between loads of the struct's fields
    // would become a load through a selected address, which keeps the whole
    // lane state in scratch memory
    uint32_t c0 = g.c0, c1 = g.c1, c2 = g.c2, c3 = g.c3;
    asm volatile("" : "+v"(c0), "+v"(c1), "+v"(c2), "+v"(c3));
    const uint32_t k = g.wc & 3u;
    const uint32_t lo = (k & 1u) ? c1 : c0;
    const uint32_t hi = (k & 1u) ? c3 : c2;
    const uint32_t mine = (k & 2u) ? hi : lo;
    const int src = (int)(b - g.cb);
    g.wc += 1u;
    return __shfl(mine, src, G);
}

struct GLane {
    int slot;                // group slot
    int sim;                 // local index, -1 none, -2 queue exhausted
    double t;
    uint32_t cursor, evc, xb;
    uint64_t firings, ssa_steps;
    uint32_t leaps, rejections, switches, absorbed;
    uint64_t poisson, gaps;
    const float* crate;      // the simulation's rate constants (HBM, cached)
    const uint32_t* fvrow;   // the simulation's feed values (when the model has feeds)
    GRng gg;                 // group stream (word 3 = 0)
    Rng lg;                  // this lane's stream (word 3 = 1 + lane)
    double tacc;             // clock advance of the SSA burst in progress
    int ssa_left;            // steps left in the SSA burst in progress (0: none)
    int ssa_mode;            // 1 once a leap attempt switched, until its burst has run
};

struct GView {
    int32_t* Xb;             // [2][S]
    uint32_t* A;             // [Rp] (entries R .. Rp-1 stay 0)
    int32_t* Kc;             // [Rp]
    uint16_t* Qx;            // [R] own queue region, or null: the queue lives in the trial state
    uint32_t* F;             // [2S+1] reactant factors, in the Kc region (used only when 2S + 1 <= Rp)
    uint32_t* qn;            // length of the Poisson queue (LDS atomic counter)
};

// Padded reaction count of the LDS arrays: R + 1 rounded up to a multiple of
// G.  The entries past R stay zero in A, so a loop over a lane's rounds needs
// no per-lane bounds test and A[R] is a zero propensity that padding entries
// of the tau' stream point at.
__host__ __device__ __forceinline__ int group_rpad(int R, int G) { return (R + G) / G * G; }

// LDS words of one group's simulation: X[2][S], A[Rp], Kc[Rp], then one
// region shared by the queue of lambda >= 1 reactions (16-bit indices; it
// lives in the trial-state half of X, free until the update writes it, when
// R/2 <= S) and the factor table F[2S+1] (in Kc when 2S+1 <= Rp): the queue
// is used only by the count phase, the table only before it and in SSA;
// last, the queue's length.
__host__ __device__ __forceinline__ size_t group_lds_words(int S, int R, int G) {
    const size_t Rp = (size_t)group_rpad(R, G);
    const size_t q = ((size_t)R + 1) / 2;
    const size_t qx = q > (size_t)S ? q : 0;
    return (size_t)2 * S + 2 * Rp + qx + 1;
}

__device__ __forceinline__ GView group_view(uint32_t* base, int S, int R, int G) {
    GView V;
    const int Rp = group_rpad(R, G);
    V.Xb = (int32_t*)base;
    V.A = (uint32_t*)(V.Xb + 2 * S);
    V.Kc = (int32_t*)(V.A + Rp);
    V.Qx = ((size_t)R + 1) / 2 > (size_t)S ? (uint16_t*)(V.Kc + Rp) : nullptr;
    V.F = (uint32_t*)V.Kc;              // factor table (only when 2S + 1 <= Rp: see pfast)
    V.qn = base + group_lds_words(S, R, G) - 1;
    return V;
}

// The reactant factors of species s from its count x (the arithmetic and the
// criticality test of greactant for multiplicities 1 and 2): x and
// x (x - 1) / 2 as FP32, bit 31 set when v <= x < nc v.
__device__ __forceinline__ void gfactor(uint32_t* F, int S, int s, int32_t x, uint32_t nc) {
    const float xf = (float)x;
    const float f2 = (x < 2) ? 0.0f : (0.5f * xf) * (float)(x - 1);
    const bool c1 = (uint32_t)x >= 1u && (uint32_t)x < nc;
    const bool c2 = (uint32_t)x >= 2u && (uint32_t)x < nc * 2u;
    F[s] = __float_as_uint(xf) | (c1 ? 0x80000000u : 0u);
    F[S + s] = __float_as_uint(f2) | (c2 ? 0x80000000u : 0u);
}

// The whole factor table of state X; F[2S] = 1.0 is the factor of a padding
// entry (multiplicity 0), never critical.
template <int G>
__device__ __forceinline__ void gfactors(const KParams& p, const int32_t* X, uint32_t* F, int l) {
    const uint32_t nc = (uint32_t)p.nc;
    for (int s = l; s < p.S; s += G) gfactor(F, p.S, s, X[s], nc);
    if (l == 0) F[2 * p.S] = 0x3f800000u;
}

// Zeroes a group's A padding (R .. Rp-1); done when the group's model is set.
template <int G>
__device__ __forceinline__ void gzero_pad(const GView& V, int R, int l) {
    for (int j = R + l; j < group_rpad(R, G); j += G) V.A[j] = 0u;
}

// Propensity of reaction j from the group's LDS state; the arithmetic of
// prop_one (CalculatePropensity, T.cu:1443-1470, in FP32).
__device__ __forceinline__ void greactant(float& a, bool& c, int32_t x, int v, uint32_t nc) {
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
    c = c || ((uint32_t)x >= (uint32_t)v && (uint64_t)(uint32_t)x < (uint64_t)nc * (uint32_t)v);
}

// Reaction j from its own reactant list (the SSA dependents).
__device__ __forceinline__ float gprop(const KParams& p, const int32_t* X, const float* crate, int j, bool& crit) {
    float a = crate[j];
    bool c = false;
    const uint32_t h1 = p.T[p.oLp + j + 1];
    for (uint32_t h = p.T[p.oLp + j]; h < h1; ++h) {
        const uint32_t e = p.T[p.oLe + h];
        greactant(a, c, X[e & 0xffffu], (int)(e >> 16), (uint32_t)p.nc);
    }
    crit = c;
    return a;
}

// Every reaction of lane l (j = l, l + G, ...) from the reactant round
// stream: the table addresses do not depend on loaded data, so the loads of a
// round issue together.  Writes A[j] (propensity bits | critical << 31) and
// returns the lane's partial sums of all and of the critical propensities.
// With pfast the two entries of a reaction are indices into the factor table
// F of the state (built by gfactors), so a reaction costs two multiplies:
// the same factors in the same order as greactant, hence the same bits.
template <int G, int PFM = 3>
__device__ __forceinline__ void gprops(const KParams& p, const int32_t* X, const uint32_t* F, const float* crate,
                                       uint32_t* A, int l, float& P, float& PC) {
    const int nr = (p.R + G - 1) / G;
    P = 0.0f;
    PC = 0.0f;
    const int pm = pmode<PFM>(kuse(p.pfast));
    if (pm == 2) {
        // two entries per reaction (species | multiplicity << 16) read from
        // the state; four rounds per batch of loads
        const uint32_t* E = p.T + p.oPRe + l;
        const uint32_t nc = (uint32_t)p.nc;
        for (int r0 = 0; r0 < nr; r0 += 4) {
            float cr[4];
            uint32_t e0[4], e1[4];
            int32_t x0[4], x1[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + q, j = l + r * G;
                cr[q] = 0.0f;
                e0[q] = e1[q] = 0u;
                if (r < nr) {
                    e0[q] = E[(2 * r) * G];
                    e1[q] = E[(2 * r + 1) * G];
                    if (j < p.R) cr[q] = crate[j];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                x0[q] = X[e0[q] & 0xffffu];
                x1[q] = X[e1[q] & 0xffffu];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + q, j = l + r * G;
                if (r < nr) {                          // past R: 0 (rate 0) into A's zero padding
                    float a = cr[q];
                    bool c = false;
                    gtwo(a, c, x0[q], e0[q] >> 16, x1[q], e1[q] >> 16, nc);
                    A[j] = __float_as_uint(a) | (c ? 0x80000000u : 0u);
                    P = P + a;
                    if (c) PC = PC + a;
                }
            }
        }
        return;
    }
    if (pm == 1) {
        // two entries per reaction; four rounds per batch of loads
        const uint32_t* E = p.T + p.oPRe + l;
        for (int r0 = 0; r0 < nr; r0 += 4) {
            float cr[4];
            uint32_t e0[4], e1[4];
            uint32_t u0[4], u1[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + q, j = l + r * G;
                cr[q] = 0.0f;
                e0[q] = e1[q] = (uint32_t)(2 * p.S);
                if (r < nr) {
                    e0[q] = E[(2 * r) * G];
                    e1[q] = E[(2 * r + 1) * G];
                    if (j < p.R) cr[q] = crate[j];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                u0[q] = F[e0[q]];
                u1[q] = F[e1[q]];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = r0 + q, j = l + r * G;
                if (r < nr) {                          // past R: 0 (rate 0) into A's zero padding
                    float a = cr[q];
                    a = a * __uint_as_float(u0[q] & 0x7fffffffu);
                    a = a * __uint_as_float(u1[q] & 0x7fffffffu);
                    const uint32_t c = (u0[q] | u1[q]) & 0x80000000u;
                    A[j] = __float_as_uint(a) | c;
                    P = P + a;
                    if (c) PC = PC + a;
                }
            }
        }
        return;
    }
    const ctab_u32 TC = CT_U32(p.T);
    const uint32_t* E = p.T + p.oPRe + l;
    for (int r = 0, j = l; r < nr; ++r, j += G) {
        const uint32_t len = TC[p.oPRm + r];
        if (j < p.R) {
            float a = crate[j];
            bool c = false;
            for (uint32_t k = 0; k < len; ++k) {
                const uint32_t e = E[k * G];
                const int v = (int)(e >> 16);
                if (v != 0) greactant(a, c, X[e & 0xffffu], v, (uint32_t)p.nc);
            }
            A[j] = __float_as_uint(a) | (c ? 0x80000000u : 0u);
            P = P + a;
            if (c) PC = PC + a;
        }
        E += len * G;
    }
}

// Selection: among the lanes with a positive own partial (Pown), the first
// whose inclusive sum P reaches w, else the last such lane; then the walk
// through that lane's reactions from the previous lane's inclusive sum
// (which = 0: all reactions; which = 2: the critical ones plus reaction 0,
// the reference's pick walk that starts from a_0, T.cu:1993-2006).  Lanes
// with a zero partial are skipped: the scan associates differently per lane,
// so such a lane's inclusive sum can round above its predecessor's.  Only a
// reaction with a positive propensity is chosen.  Returns the chosen
// reaction (group-uniform) or -1 when no lane has a positive partial.
template <int G>
__device__ __forceinline__ int gselect(const KParams& p, const uint32_t* A, float Pown, float P, float w, int l,
                                       int which) {
    const uint64_t cand = gballot<G>(Pown > 0.0f);
    if (cand == 0) return -1;
    const uint64_t hit = gballot<G>(Pown > 0.0f && P >= w);
    const int sel = hit ? __builtin_ctzll(hit) : 63 - __builtin_clzll(cand);
    const float excl = G <= 16 ? dppf<DPP_ROW_SHR1>(P) : dppf<DPP_WAVE_SHR1>(P);   // lane 0's value is unused
    // The walk through lane sel's reactions sel + i*G, spread over the first
    // W lanes of the group (a row of 16 at most): lane i holds the i-th
    // propensity and the sequential sums excl + a_0 + ... + a_i are built by
    // W-1 row shifts (after k shifts lanes 0..k hold their final sum; a
    // shifted-in 0 or a skipped entry adds an exact 0), so every sum is
    // rounded as in the one-lane walk.  Lane sel's reactions number at most
    // 16 per round; groups wider than 16 leave their other rows idle.
    constexpr int W = G <= 16 ? G : 16;
    const int nj = (p.R - sel + G - 1) / G;
    float carry = __shfl(excl, sel, G);
    if (sel == 0) carry = 0.0f;
    int last = -1;
    for (int c = 0; c < nj; c += W) {
        const int i = c + l;
        const bool in = l < W && i < nj;
        const int j = sel + i * G;
        const uint32_t u = in ? A[j] : 0u;
        const bool use = which != 2 || (u & 0x80000000u) || j == 0;
        const float a = use ? aval(u) : 0.0f;
        const float a1 = l == 0 ? carry + a : a;
        float x = a1;
#pragma unroll
        for (int k = 1; k < W; ++k) {
            const float t = dppf<DPP_ROW_SHR1>(x) + a1;
            x = (W == 16 || l > 0) ? t : a1;
        }
        const uint64_t pos = gballot<G>(in && a > 0.0f);
        const uint64_t hm = gballot<G>(in && a > 0.0f && w <= x);
        if (hm) return sel + (c + __builtin_ctzll(hm)) * G;
        if (pos) last = sel + (c + 63 - __builtin_clzll(pos)) * G;
        carry = G <= 16 ? gbcast_last<G>(x, l) : __shfl(x, W - 1, G);
    }
    return last;
}

template <int G>
__device__ __forceinline__ void gabsorb(const KParams& p, GLane& L, const int32_t* X, int l) {
    uint32_t* out = p.storage + (size_t)L.sim * p.K * p.SS;
    if (L.cursor == 0) {
        for (int sk = l; sk < p.SS; sk += G) out[sk] = (uint32_t)X[p.T[p.oCs + sk]];
        L.cursor = 1;
    }
    const uint32_t* last = out + (size_t)(L.cursor - 1) * p.SS;
    for (int i = (int)L.cursor; i < p.K; ++i)
        for (int sk = l; sk < p.SS; sk += G) out[(size_t)i * p.SS + sk] = last[sk];
    L.t = p.tmax + (double)0.1f;
    L.cursor = (uint32_t)p.K;
    L.absorbed = 1;
}

// One adaptive leap attempt (PreCalculations, T.cu:1568-2093).
template <int G, int PFM = 3>
__device__ __forceinline__ int gleap(const KParams& p, GLane& L, const GView& V, int l,
                                     const float* __restrict__ recip) {
    if (L.t >= p.tmax) return LEAP_DONE;
    int32_t* X = V.Xb + (L.xb ? p.S : 0);
    int32_t* X1 = V.Xb + (L.xb ? 0 : p.S);
    float P, PC;
    if (pmode<PFM>(kuse(p.pfast)) == 1) {
        gfactors<G>(p, X, V.F, l);
        gsync();
    }
    gprops<G, PFM>(p, X, V.F, L.crate, V.A, l, P, PC);
    gsync();
    P = gscan<G>(P, l);
    PC = gscan<G>(PC, l);
    const float a0 = gbcast_last<G>(P, l);
    const float a0c = gbcast_last<G>(PC, l);
    CTL_PHASE(PH_PROPS);
    if (a0 == 0.0f) { gabsorb<G>(p, L, X, l); return LEAP_ABSORB; }
    // tau' (T.cu:1860-1922): the one-lane kernel's per-species values, min over lanes
    float tau1 = 3.40282347e+38f;
    const ctab_u32 TC = CT_U32(p.T);
    const int nrs = (p.S + G - 1) / G;
    const uint32_t* E = p.T + p.oVRe + l;
    const uint32_t* SPI = p.T + p.oSPI + l;
    for (int r = 0; r < nrs; ++r) {
        const uint32_t len = TC[p.oVRm + r];
        const uint32_t spw = SPI[r * G];             // species | HOR << 16 | HORtype << 24
        float mu = 0.0f, sg = 0.0f;
        if (kuse(p.vhalf)) {
            // entries (coefficient as FP16 << 16 | byte offset of A[reaction]):
            // both halves are read as operand selects, and v * v in FP32 is exact
            const char* Ab = reinterpret_cast<const char*>(V.A);
            for (uint32_t k = 0; k < len; k += 4) {  // len is a multiple of 4
                uint32_t e[4];
                int32_t u[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) e[i] = E[(k + i) * G];
#pragma unroll
                for (int i = 0; i < 4; ++i) u[i] = *reinterpret_cast<const int32_t*>(Ab + (e[i] & 0xffffu));
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float v = (float)__builtin_bit_cast(_Float16, (uint16_t)(e[i] >> 16));
                    // a critical propensity has bit 31 set: as an integer it is
                    // negative and the maximum with 0 gives +0; padding (A[R] = 0)
                    // adds an exact zero too
                    const float a = __int_as_float(u[i] > 0 ? u[i] : 0);
                    mu = mu + v * a;
                    sg = sg + (v * v) * a;
                }
            }
        } else {
            for (uint32_t k = 0; k < len; k += 4) {  // len is a multiple of 4
                uint32_t e[4], u[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) e[i] = E[(k + i) * G];
#pragma unroll
                for (int i = 0; i < 4; ++i) u[i] = V.A[e[i] & 0xffffu];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int v = (int)(int16_t)(e[i] >> 16);
                    // critical reactions and padding (A[R] = 0) add an exact zero
                    const float a = (u[i] & 0x80000000u) ? 0.0f : __uint_as_float(u[i]);
                    mu = mu + (float)v * a;
                    sg = sg + (float)(v * v) * a;
                }
            }
        }
        E += len * G;
        if (mu == 0.0f || sg == 0.0f) continue;      // no var entries (no species included)
        const int32_t x = X[spw & 0xffffu];
        const float ex = p.eps * (float)x;
        // g is 1 or 2 unless some species has HOR 3 or HOR 2 of type 2: then
        // the division by 2 is the exact scaling by 0.5
        const float tmp2 = kuse(p.gsimple) ? (((spw >> 16) & 0xffu) == 2u ? ex * 0.5f : ex)
                                           : ex / gfun(spw >> 16, x);
        const float macs = tmp2 > 1.0f ? tmp2 : 1.0f;
        const float t3 = macs / fabsf(mu);
        const float t4 = (macs * macs) / sg;
        const float tmp = t3 < t4 ? t3 : t4;
        if (tmp < tau1) tau1 = tmp;
    }
    tau1 = gmin<G>(tau1);
    CTL_PHASE(PH_TAU1);
    if (tau1 < p.ssa_factor / a0) { L.switches++; return LEAP_SWITCH; }      // T.cu:1935-1941
    uint32_t e = L.evc++;
    grng_event(L.gg, e);
    float tau2 = 3.40282347e+38f;
    if (a0c > 0.0f) tau2 = u32_exp(grng_u32<G>(L.gg, l)) / a0c;                                // T.cu:1962-1965
    float tau = tau1 < tau2 ? tau1 : tau2;
    int pick = -1;
    if (tau1 > tau2) {                                                         // T.cu:1993-2006
        const float w = u32_oc(grng_u32<G>(L.gg, l)) * a0c;
        float Q = 0.0f;
        for (int j = l; j < p.R; j += G) {
            const uint32_t u = V.A[j];
            if ((u & 0x80000000u) || j == 0) Q = Q + aval(u);
        }
        const float Qown = Q;
        Q = gscan<G>(Q, l);
        pick = gselect<G>(p, V.A, Qown, Q, w, l, 2);
    }
    CTL_PHASE(PH_PICK);
    uint64_t fired = 0;
    for (int attempt = 0;; ++attempt) {                                        // T.cu:2011-2050
        if (attempt > 0) { e = L.evc++; L.rejections++; }
        // counts: critical reactions (the pick), the small-lambda process of
        // this lane (its first gap drawn up front), and a queue of the
        // reactions with lambda >= 1, drawn afterwards by the group's lanes in
        // parallel, each from the reaction's own stream (word 3 = 2^31 | j)
        rng_event(L.lg, e);
        uint32_t fk = 0;                         // counts of this loop: process events and the pick
        uint16_t* Q = V.Qx ? V.Qx : (uint16_t*)X1;
        if (l == 0) *V.qn = 0u;
        gsync();
        double cum = 0.0, pos = (double)exp_rv(L.lg);
        L.gaps++;
        // The lane's second event position is drawn now, with the first, by
        // every lane at once; drawn inside the loop below it would cost the
        // wavefront a divergent draw for almost every reaction round.  It is
        // the same value from the same stream (the lane stream serves only
        // these gaps), counted as a gap when it is used.
        double pos2 = pos + (double)exp_rv(L.lg);
        bool have2 = true;
        const int nit = (p.R + G - 1) / G;
        for (int i = 0; i < nit; ++i) {
            const int j = l + i * G;
            bool need = false;
            {                                          // past R: the zero padding of A and Kc
                const uint32_t u = V.A[j];
                const bool crit = (u & 0x80000000u) != 0u;
                const float lam = tau * __uint_as_float(u);
                need = !crit && !(lam < 1.0f);
                const bool small = !crit && lam < 1.0f && lam > 0.0f;
                int32_t k = (crit && attempt == 0 && j == pick) ? 1 : 0;
                cum = small ? cum + (double)lam : cum;
                // pos >= cum held before this reaction, so only a lane whose
                // process mass grew draws gaps
                while (pos < cum) {
                    ++k;
                    if (have2) {
                        pos = pos2;
                        have2 = false;
                    } else {
                        pos = pos + (double)exp_rv(L.lg);
                    }
                    L.gaps++;
                }
                V.Kc[j] = k;                         // a queued reaction's count is drawn below
                fk += (uint32_t)k;
            }
            // the queue's order is free: each count comes from its reaction's own stream
            if (need) Q[atomicAdd(V.qn, 1u)] = (uint16_t)j;
        }
        gsync();
        CTL_PHASE(PH_COUNTS);
        const uint32_t nq = *V.qn;
        uint64_t f = fk;
        for (uint32_t qi = (uint32_t)l; qi < nq; qi += G) {
            const int j = Q[qi];
            const float lam = tau * __uint_as_float(V.A[j]);
            Rng gr{};
            gr.k0 = L.gg.k0;
            gr.k1 = L.gg.k1;
            gr.sim = L.gg.sim;
            gr.tag = 0x80000000u | (uint32_t)j;
            rng_event(gr, e);
            const int32_t k = poisson(gr, lam, recip);
            V.Kc[j] = k;
            f += (uint64_t)k;
            L.poisson++;
        }
        gsync();
        CTL_PHASE(PH_POISSON);
        bool bad = false;
        const uint32_t* F = p.T + p.oFRe + l;
        for (int r = 0; r < nrs; ++r) {
            const uint32_t len = TC[p.oFRm + r];
            const uint32_t sp = SPI[r * G] & 0xffffu;
            if (sp < (uint32_t)p.S) {
                int64_t x = X[sp];
                bool bs = false;
                for (uint32_t k = 0; k < len; k += 4) {   // len is a multiple of 4; padding adds 0
                    uint32_t e[4];
                    int32_t kc[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) e[i] = F[(k + i) * G];
#pragma unroll
                    for (int i = 0; i < 4; ++i) kc[i] = V.Kc[e[i] & 0xffffu];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        x += (int64_t)(int16_t)(e[i] >> 16) * (int64_t)kc[i];
                        bs = bs || x < 0 || x > 2147483647LL;
                    }
                }
                X1[sp] = (int32_t)x;
                bad = bad || bs;
            }
            F += len * G;
        }
        fired = gsum64<G>(f);
        const bool any_bad = gballot<G>(bad) != 0;
        gsync();
        CTL_PHASE(PH_UPDATE);
        if (!any_bad) break;
        tau = tau * 0.5f;
    }
    // check_time_instant (T.cu:1507-1562), per sampled species
    {
        uint32_t i = L.cursor;
        const double tn = L.t + (double)tau;
        if (i < (uint32_t)p.K && tn >= p.ts[i]) {
            uint32_t* out = p.storage + (size_t)L.sim * p.K * p.SS;
            while (i < (uint32_t)p.K && tn >= p.ts[i]) {
                const double ti = p.ts[i];
                for (int sk = l; sk < p.SS; sk += G) {
                    const int sp = (int)p.T[p.oCs + sk];
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
    L.xb ^= 1u;
    if (kuse(p.nfs) > 0) {               // feeds (T.cu:2072-2079)
        int32_t* Xn = V.Xb + (L.xb ? p.S : 0);
        for (int u = l; u < p.nfs; u += G) Xn[p.fs[u]] = (int32_t)L.fvrow[u];
        gsync();
    }
    L.t = L.t + (double)tau;
    L.firings += fired;
    L.leaps++;
    CTL_PHASE(PH_SAMPLE);
    return LEAP_OK;
}

// Up to nmax steps of an SSA burst (EventualSSASteps T.cu:2099-2273 /
// LightweightSSA T.cu:2280-2454), opening the burst when none is in progress:
// a burst of p.ssa_steps steps may run in several calls, its clock advance,
// propensities and factor table kept meanwhile (L.tacc, L.ssa_left, the
// group's LDS), so the steps and their draws are those of one call.  After
// the fired reaction only its dependent propensities are recomputed (same
// values as recomputing all).  Returns true on absorption.
template <int G, int PFM = 3>
__device__ __forceinline__ bool gssa_run(const KParams& p, GLane& L, const GView& V, int l, int nmax) {
    int32_t* X = V.Xb + (L.xb ? p.S : 0);
    if (L.ssa_left == 0) {
        grng_event(L.gg, L.evc++);
        L.tacc = 0.0;
        L.ssa_left = p.ssa_steps;
        if (pmode<PFM>(kuse(p.pfast)) == 1) {
            gfactors<G>(p, X, V.F, l);
            gsync();
        }
        float P, PC;
        gprops<G, PFM>(p, X, V.F, L.crate, V.A, l, P, PC);
        gsync();
    }
    CTL_PHASE(PH_SSA_OPEN);
    double tacc = L.tacc;
    const int n = nmax < L.ssa_left ? nmax : L.ssa_left;
    for (int step = 0; step < n; ++step) {
        float P = 0.0f;
        const int nrp = group_rpad(p.R, G) / G;        // rounds over A with its zero padding
        for (int r0 = 0; r0 < nrp; r0 += 4) {          // padding adds an exact 0
            uint32_t u[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) u[i] = r0 + i < nrp ? V.A[l + (r0 + i) * G] : 0u;
#pragma unroll
            for (int i = 0; i < 4; ++i) P = P + aval(u[i]);
        }
        const float Pown = P;
        P = gscan<G>(P, l);
        const float a0 = gbcast_last<G>(P, l);
        if (a0 == 0.0f) { gabsorb<G>(p, L, X, l); L.ssa_left = 0; return true; }
        const float r1 = u32_oc(grng_u32<G>(L.gg, l)) * a0;
        const int j = gselect<G>(p, V.A, Pown, P, r1, l, 0);
        if (j < 0) { gabsorb<G>(p, L, X, l); L.ssa_left = 0; return true; }    // unreachable for finite a0
        const float dt = u32_exp(grng_u32<G>(L.gg, l)) / a0;
        CTL_PHASE(PH_SSA_SELECT);
        tacc = tacc + (double)dt;
        const double tn = L.t + tacc;
        uint32_t i = L.cursor;
        if (i < (uint32_t)p.K && tn > p.ts[i]) {                               // T.cu:2213-2236
            uint32_t* out = p.storage + (size_t)L.sim * p.K * p.SS;
            while (i < (uint32_t)p.K && tn > p.ts[i]) {
                for (int sk = l; sk < p.SS; sk += G) out[(size_t)i * p.SS + sk] = (uint32_t)X[p.T[p.oCs + sk]];
                ++i;
            }
            L.cursor = i;
        }
        gsync();
        const uint32_t f1 = p.T[p.oFp + j + 1];
        const int pm = pmode<PFM>(kuse(p.pfast));
        const bool fac = pm == 1;
        for (uint32_t h = p.T[p.oFp + j] + (uint32_t)l; h < f1; h += G) {
            const uint32_t ent = p.T[p.oFe + h];
            const int s = (int)(ent & 0xffffu);
            const int v = (int)(int16_t)(ent >> 16);
            const int32_t x = (int32_t)((uint32_t)X[s] + (uint32_t)v);
            X[s] = x;
            if (fac) gfactor(V.F, p.S, s, x, (uint32_t)p.nc);   // a species occurs once per reaction
        }
        gsync();
        CTL_PHASE(PH_SSA_UPDATE);
        L.ssa_steps++;
        if (tn <= p.tmax) L.firings++;
        const uint32_t d0 = p.T[p.oDp + j], d1 = p.T[p.oDp + j + 1];
        if (fac) {
            // dependent records {r, factor index 0, factor index 1, 0}: one 16-byte load each
            const uint4* DR = reinterpret_cast<const uint4*>(p.T + p.oDR);
            for (uint32_t d = d0 + (uint32_t)l; d < d1; d += G) {
                const uint4 rec = DR[d];
                const int r = (int)rec.x;
                const uint32_t u0 = V.F[rec.y], u1 = V.F[rec.z];
                float a = L.crate[r];
                a = a * __uint_as_float(u0 & 0x7fffffffu);
                a = a * __uint_as_float(u1 & 0x7fffffffu);
                V.A[r] = __float_as_uint(a) | ((u0 | u1) & 0x80000000u);
            }
        } else if (pm == 2) {
            // dependent records {r, reactant entry 0, entry 1, 0} read from the state
            const uint4* DR = reinterpret_cast<const uint4*>(p.T + p.oDR);
            const uint32_t nc = (uint32_t)p.nc;
            for (uint32_t d = d0 + (uint32_t)l; d < d1; d += G) {
                const uint4 rec = DR[d];
                const int r = (int)rec.x;
                float a = L.crate[r];
                bool c = false;
                gtwo(a, c, X[rec.y & 0xffffu], rec.y >> 16, X[rec.z & 0xffffu], rec.z >> 16, nc);
                V.A[r] = __float_as_uint(a) | (c ? 0x80000000u : 0u);
            }
        } else {
            for (uint32_t d = d0 + (uint32_t)l; d < d1; d += G) {
                const int r = (int)p.T[p.oDe + d];
                bool crit;
                const float a = gprop(p, X, L.crate, r, crit);
                V.A[r] = __float_as_uint(a) | (crit ? 0x80000000u : 0u);
            }
        }
        gsync();
        CTL_PHASE(PH_SSA_DEPS);
    }
    L.ssa_left -= n;
    L.tacc = tacc;
    if (L.ssa_left == 0) {
        L.t = L.t + tacc;
        L.tacc = 0.0;
    }
    return false;
}

// One whole SSA burst.
template <int G, int PFM = 3>
__device__ __forceinline__ bool gssa(const KParams& p, GLane& L, const GView& V, int l) {
    L.ssa_left = 0;
    return gssa_run<G, PFM>(p, L, V, l, p.ssa_steps);
}

__device__ __forceinline__ const float* grates(const KParams& p, int sim) {
    return p.c_bcast ? p.c_bcast : p.c_sims + (size_t)(sim / p.c_div) * p.R;
}

// Next simulation from the queue (PrepareVariables, T.cu:1269-1297).
template <int G>
__device__ __forceinline__ bool gfetch(const KParams& p, GLane& L, const GView& V, int l) {
    uint32_t q = 0;
    if (l == 0) q = atomicAdd(p.next_sim, 1u);
    q = __shfl(q, 0, G);
    if (q >= (uint32_t)p.n) { L.sim = -2; return false; }
    L.sim = (int)q;
    L.t = 0.0;
    L.tacc = 0.0;
    L.ssa_left = 0;
    L.ssa_mode = kuse(p.force_ssa);
    L.cursor = 0;
    L.evc = 0;
    L.xb = 0;
    L.firings = L.ssa_steps = 0;
    L.leaps = L.rejections = L.switches = L.absorbed = 0;
    L.poisson = L.gaps = 0;
    L.gg.sim = L.lg.sim = p.sim_base + q;
    L.gg.cb = GRNG_NONE;
    L.lg.pos = 4;
    for (int s = l; s < p.S; s += G)
        V.Xb[s] = (int32_t)(p.x0_bcast ? p.x0_bcast[s] : p.x0_sims[(size_t)(q / (uint32_t)p.x0_div) * p.S + s]);
    L.crate = grates(p, (int)q);
    L.fvrow = p.nfs > 0 ? p.fv + (size_t)(q / (uint32_t)p.fv_div) * p.nfs : nullptr;
    gsync();
    return true;
}

template <int G>
__device__ __forceinline__ void gfinish(const KParams& p, GLane& L, const GView& V, int l) {
    const int32_t* X = V.Xb + (L.xb ? p.S : 0);
    const uint64_t pois = gsum64<G>(L.poisson);
    const uint64_t gaps = gsum64<G>(L.gaps);
    for (int s = l; s < p.S; s += G) p.xfinal[(size_t)L.sim * p.S + s] = X[s];
    if (l == 0) {
        ctl_sim_stat st;
        st.firings = L.firings;
        st.ssa_steps = L.ssa_steps;
        st.poisson = pois;
        st.gaps = gaps;
        st.t_end = L.t;
        st.leaps = L.leaps;
        st.rejections = L.rejections;
        st.switches = L.switches;
        st.events = L.evc;
        st.absorbed = L.absorbed;
        st.reserved = 0;
        p.simstat[L.sim] = st;
        p.tfinal[L.sim] = L.t;
        atomicAdd(p.done, 1u);
    }
    L.sim = -1;
}

// Longest a switched group waits for a partner burst (leap iterations).
constexpr int GBURST_WAIT = 24;

// ---------------------------------------------------------------------------
// Regrouping groups by mode.  The groups of a wavefront run in lockstep: while
// one group runs SSA steps the wavefront's leaping groups wait, and the
// other way round.  With p.compact > 0 a workgroup's group slots ("regions":
// a simulation's LDS state and its global slot) are reassigned to the
// workgroup's wavefronts every p.compact iterations, leap-mode regions first,
// SSA-mode ones next, regions without a simulation last, so that each
// wavefront runs one mode.  SSA bursts then run in chunks of p.ssa_chunk steps
// per iteration, comparable to a leap.  A region's simulation keeps its LDS
// state; its registers travel
// through the slot arrays in global memory (the parking format).  No event of
// any simulation changes: only which lanes run it.
// ---------------------------------------------------------------------------

// Writes the group's registers to the slot arrays of its region's slot s0.
template <int G>
__device__ __forceinline__ void gstage_put(const KParams& p, const GLane& L, int s0, int l) {
    p.st->poisson[(size_t)s0 * G + l] = L.poisson;
    p.st->gaps[(size_t)s0 * G + l] = L.gaps;
    if (l == 0) {
        p.st->sim[s0] = L.sim;
        p.st->t[s0] = L.t;
        p.st->tacc[s0] = L.tacc;
        p.st->cursor[s0] = L.cursor;
        p.st->ssa_left[s0] = L.ssa_left;
        p.st->mode[s0] = L.ssa_mode;
        p.st->evc[s0] = L.evc;
        p.st->xb[s0] = L.xb;
        p.st->firings[s0] = L.firings;
        p.st->ssa_steps[s0] = L.ssa_steps;
        p.st->leaps[s0] = L.leaps;
        p.st->rejections[s0] = L.rejections;
        p.st->switches[s0] = L.switches;
        p.st->ev[s0] = L.gg.ev;
        p.st->blk[s0] = L.gg.wc;      // words of the event consumed
    }
}

// Takes over the simulation of slot s0: its registers, rate and feed rows and
// the position in its group stream.
template <int G>
__device__ __forceinline__ void gstage_get(const KParams& p, GLane& L, int s0, int l) {
    L.sim = p.st->sim[s0];
    L.poisson = p.st->poisson[(size_t)s0 * G + l];
    L.gaps = p.st->gaps[(size_t)s0 * G + l];
    L.absorbed = 0;
    if (L.sim < 0) return;
    L.t = p.st->t[s0];
    L.tacc = p.st->tacc[s0];
    L.cursor = p.st->cursor[s0];
    L.ssa_left = p.st->ssa_left[s0];
    L.ssa_mode = p.st->mode[s0];
    L.evc = p.st->evc[s0];
    L.xb = p.st->xb[s0];
    L.firings = p.st->firings[s0];
    L.ssa_steps = p.st->ssa_steps[s0];
    L.leaps = p.st->leaps[s0];
    L.rejections = p.st->rejections[s0];
    L.switches = p.st->switches[s0];
    L.gg.ev = p.st->ev[s0];
    L.gg.wc = p.st->blk[s0];
    L.gg.cb = GRNG_NONE;                 // the blocks are regenerated at the next word
    L.gg.sim = L.lg.sim = p.sim_base + (uint32_t)L.sim;
    L.crate = grates(p, L.sim);
    L.fvrow = p.nfs > 0 ? p.fv + (size_t)(L.sim / p.fv_div) * p.nfs : nullptr;
}

// Stable partition of the workgroup's regions by key (0 leap, 1 SSA, 2
// none); cnt: 3 words per wavefront, perm: one word per region.  Every lane
// of the workgroup must call it.  Returns false, with no other effect, when
// no region of the workgroup holds a simulation; else the new region of
// this lane's group in *next and whether any region moves to other lanes in
// *moves (the same for the whole workgroup).
template <int G>
__device__ __forceinline__ bool gregroup(int key, int reg, uint32_t* cnt, uint32_t* perm, int l, int* next,
                                         bool* moves) {
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t rank = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint64_t b = __ballot(l == 0 && key == k);      // one bit per group (its lane 0)
        if (key == k) rank = (uint32_t)__popcll(b & lt);
        if (lane == 0) cnt[wave * 3 + k] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    uint32_t active = 0, pos = 0;
    for (int w = 0; w < nw; ++w) active += cnt[w * 3] + cnt[w * 3 + 1];
    if (active == 0) return false;
    for (int k = 0; k < key; ++k)
        for (int w = 0; w < nw; ++w) pos += cnt[w * 3 + k];
    for (int w = 0; w < wave; ++w) pos += cnt[w * 3 + key];
    if (l == 0) perm[pos + rank] = (uint32_t)reg;
    __syncthreads();
    const int nreg = (int)(blockDim.x / G);
    bool mv = false;
    for (int q = 0; q < nreg; ++q) mv = mv || perm[q] != (uint32_t)q;
    *moves = mv;
    *next = (int)perm[threadIdx.x / G];
    return true;
}

// Persistent group kernel: p.ns lanes = p.ns / G group slots, each pulling
// simulations from the queue; a launch runs up to p.iters iterations per
// group (a leap attempt, or SSA steps), then parks the group's state in the
// slot arrays for the next launch.  Without regrouping (RG false: G = 64, or
// p.compact = 0) a group that switched to SSA holds its whole burst until a second
// group of the wavefront switches too (or it has waited GBURST_WAIT
// iterations, or every group still running is waiting), so that bursts run
// side by side.  Either way each simulation's events and draws are its own;
// only the order in which the lanes serve the simulations moves.
// G = 8, 16 and 32 are compiled for four waves per SIMD (128 VGPRs, a few
// bytes of scratch): measured +9 % at G = 16 and +4 % at G = 32 on the
// 100x200 network, +24 % at G = 8 on the 50x100 sweep network.  G = 4
// groups are bounded by LDS instead, and G = 64 would spill heavily.
template <int G, bool RG, int PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(G == 8 || G == 16 || G == 32 ? 4 : 1))) void group_kernel(KParams p) {
    extern __shared__ uint32_t smem[];
    __shared__ float recip[64];
    __shared__ uint32_t rg_cnt[12];
    __shared__ uint32_t rg_perm[256 / G];
    for (int i = threadIdx.x; i < 64; i += blockDim.x) recip[i] = c_recip[i];
    __syncthreads();
    const int l = (int)(threadIdx.x & (unsigned)(G - 1));
    const int nreg = (int)(blockDim.x / G);
    int gib = (int)(threadIdx.x / G);
    const size_t gw = group_lds_words(p.S, p.R, G);
    constexpr bool rg = RG && G < 64;     // the host launches RG only with p.compact > 0 and several wavefronts
    GLane L;
    L.slot = (int)blockIdx.x * nreg + gib;
    CTL_PHASE_START();
    if (L.slot * G >= p.ns) return;                  // whole workgroups when regrouping
    GView V = group_view(smem + (size_t)gib * gw, p.S, p.R, G);
    gzero_pad<G>(V, p.R, l);
    L.gg.k0 = L.lg.k0 = p.key0;
    L.gg.k1 = L.lg.k1 = p.key1;
    L.lg.tag = 1u + (uint32_t)l;
    L.tacc = 0.0;
    L.ssa_left = 0;
    L.ssa_mode = 0;
    L.sim = p.st->sim[L.slot];
    if (L.sim == -2 && !rg) {
        CTL_PHASE(PH_PARK);
        CTL_PHASE_FLUSH();
        return;
    }
    if (L.sim >= 0) {                    // resume a parked simulation
        const int s0 = L.slot;
        L.t = p.st->t[s0];
        L.cursor = p.st->cursor[s0];
        L.evc = p.st->evc[s0];
        L.xb = 0;
        L.firings = p.st->firings[s0];
        L.ssa_steps = p.st->ssa_steps[s0];
        L.leaps = p.st->leaps[s0];
        L.rejections = p.st->rejections[s0];
        L.switches = p.st->switches[s0];
        L.absorbed = 0;
        L.ssa_mode = kuse(p.force_ssa);
        L.poisson = p.st->poisson[(size_t)s0 * G + l];
        L.gaps = p.st->gaps[(size_t)s0 * G + l];
        L.gg.sim = L.lg.sim = p.sim_base + (uint32_t)L.sim;
        L.gg.cb = GRNG_NONE;
        L.lg.pos = 4;
        for (int s = l; s < p.S; s += G) V.Xb[s] = p.X[(size_t)s0 * p.S + s];
        L.crate = grates(p, L.sim);
        L.fvrow = p.nfs > 0 ? p.fv + (size_t)(L.sim / p.fv_div) * p.nfs : nullptr;
        gsync();
    }
    bool pend = false;
    int waited = 0;
    for (int it = 0; it < p.iters; ++it) {
        CTL_PHASE(PH_LOOP);
        if (rg) {
            if (it % p.compact == 0) {
                // 0: leap (or about to take a simulation), 1: SSA, 2: no simulation left
                const int key = L.sim == -2 ? 2 : (L.sim >= 0 ? L.ssa_mode : kuse(p.force_ssa));
                int ng;
                bool moves;
                if (!gregroup<G>(key, gib, rg_cnt, rg_perm, l, &ng, &moves)) break;
                // a group whose region stays on its lanes keeps its registers;
                // the others hand their state over through the slot arrays
                if (moves) {
                    const bool move = ng != gib;
                    if (move) gstage_put<G>(p, L, L.slot, l);
                    __syncthreads();
                    if (move) {
                        gib = ng;
                        L.slot = (int)blockIdx.x * nreg + gib;
                        gstage_get<G>(p, L, L.slot, l);
                        V = group_view(smem + (size_t)gib * gw, p.S, p.R, G);
                    }
                }
                CTL_PHASE(PH_REGROUP);
            }
            if (L.sim == -2) continue;
            if (L.sim < 0 && !gfetch<G>(p, L, V, l)) continue;
            if (L.ssa_mode) {
                CTL_PHASE(PH_PRE_SSA);
                if (kuse(p.force_ssa) && L.ssa_left == 0 && L.t >= p.tmax) {
                    gfinish<G>(p, L, V, l);
                } else if (gssa_run<G, PF>(p, L, V, l, p.ssa_chunk)) {
                    gfinish<G>(p, L, V, l);
                } else if (L.ssa_left == 0 && !kuse(p.force_ssa)) {
                    L.ssa_mode = 0;
                }
            } else {
                CTL_PHASE(PH_PRE_LEAP);
                const int r = gleap<G, PF>(p, L, V, l, recip);
                if (r == LEAP_DONE || r == LEAP_ABSORB) {
                    gfinish<G>(p, L, V, l);
                } else if (r == LEAP_SWITCH) {
                    L.ssa_mode = 1;
                }
            }
            continue;
        }
        if (L.sim < 0 && !gfetch<G>(p, L, V, l)) break;
        if (kuse(p.force_ssa)) {
            if (L.t >= p.tmax || gssa<G, PF>(p, L, V, l)) gfinish<G>(p, L, V, l);
            continue;
        }
        if (!pend) {
            CTL_PHASE(PH_PRE_LEAP);
            const int r = gleap<G, PF>(p, L, V, l, recip);
            if (r == LEAP_DONE || r == LEAP_ABSORB) {
                gfinish<G>(p, L, V, l);
            } else if (r == LEAP_SWITCH) {
                pend = true;
                waited = 0;
            }
        } else {
            ++waited;
        }
        bool go = pend;
        if (G < 64) {
            const uint64_t lead = __ballot(l == 0);            // groups still in the loop
            const uint64_t pm = __ballot(pend && l == 0);       // ... waiting for a burst
            go = pm != 0 && (__popcll(pm) >= 2 || pm == lead || __ballot(pend && waited >= GBURST_WAIT) != 0);
        }
        if (go && pend) {
            pend = false;
            CTL_PHASE(PH_PRE_SSA);
            if (gssa<G, PF>(p, L, V, l)) gfinish<G>(p, L, V, l);
        }
    }
    // before parking: the burst of a switched simulation runs to its end (a
    // parked simulation resumes with a leap attempt, or with a new burst in
    // force-SSA mode)
    if (rg) {
        if (L.sim >= 0 && (L.ssa_left > 0 || (L.ssa_mode && !kuse(p.force_ssa))) &&
            gssa_run<G, PF>(p, L, V, l, L.ssa_left > 0 ? L.ssa_left : p.ssa_steps))
            gfinish<G>(p, L, V, l);
    } else if (pend && gssa<G, PF>(p, L, V, l)) {
        gfinish<G>(p, L, V, l);
    }
    const int s0 = L.slot;
    if (L.sim >= 0) {                    // park: state to the slot arrays
        const int32_t* X = V.Xb + (L.xb ? p.S : 0);
        for (int s = l; s < p.S; s += G) p.X[(size_t)s0 * p.S + s] = X[s];
        p.st->poisson[(size_t)s0 * G + l] = L.poisson;
        p.st->gaps[(size_t)s0 * G + l] = L.gaps;
        if (l == 0) {
            p.st->t[s0] = L.t;
            p.st->cursor[s0] = L.cursor;
            p.st->evc[s0] = L.evc;
            p.st->firings[s0] = L.firings;
            p.st->ssa_steps[s0] = L.ssa_steps;
            p.st->leaps[s0] = L.leaps;
            p.st->rejections[s0] = L.rejections;
            p.st->switches[s0] = L.switches;
        }
    }
    if (l == 0) p.st->sim[s0] = L.sim;
    CTL_PHASE(PH_PARK);
    CTL_PHASE_FLUSH();
}

// Host: appends the lane-interleaved round streams (KParams::oPRm ..) of a
// group kernel with G lanes to the packed table buffer T.
inline void append_group_streams(std::vector<uint32_t>& T, KParams& p, int G, const Model& m,
                                 const std::vector<uint32_t>& FSptr, const std::vector<uint32_t>& FSent,
                                 const std::vector<uint32_t>& Dent) {
    // Item order[r * G + l] (-1: none) goes to lane l of round r.  fixed > 0:
    // every round has exactly that many entries per lane; else the round's
    // longest list rounded up to a multiple of quantum.
    auto rounds = [&](const std::vector<int>& order, const std::vector<uint32_t>& ptr,
                      const std::vector<uint32_t>& ent, int& om, int& oe, uint32_t quantum, uint32_t fixed,
                      uint32_t pad) {
        const int nr = (int)(order.size() / (size_t)G);
        auto cnt = [&](int i) { return i < 0 ? 0u : ptr[i + 1] - ptr[i]; };
        std::vector<uint32_t> len((size_t)nr, 0u), e;
        for (int r = 0; r < nr; ++r) {
            uint32_t mx = 0;
            for (int l = 0; l < G; ++l) mx = std::max(mx, cnt(order[(size_t)r * G + l]));
            mx = fixed ? fixed : (mx + quantum - 1) / quantum * quantum;
            len[r] = mx;
            for (uint32_t k = 0; k < mx; ++k)
                for (int l = 0; l < G; ++l) {
                    const int i = order[(size_t)r * G + l];
                    e.push_back(k < cnt(i) ? ent[ptr[i] + k] : pad);
                }
        }
        om = (int)T.size();
        T.insert(T.end(), len.begin(), len.end());
        oe = (int)T.size();
        T.insert(T.end(), e.begin(), e.end());
    };
    // propensity mode (pmode): with at most two reactant entries of
    // multiplicity <= 2 per reaction, two table entries per reaction -- factor
    // indices (1) when the factor table fits in the Kc region, free outside
    // the count and update phases, else species | multiplicity read from the
    // state (2): a separate table would enlarge every group's LDS and cost
    // occupancy (the enzyme cascade, 61 species and 60 reactions at G = 4,
    // ran 20 % slower with one)
    bool two = true;
    for (int j = 0; j < m.R; ++j) {
        if (m.Lptr[j + 1] - m.Lptr[j] > 2) two = false;
        for (uint32_t h = m.Lptr[j]; h < m.Lptr[j + 1]; ++h)
            if ((m.Lent[h] >> 16) > 2) two = false;
    }
    p.pfast = !two ? 0 : (2 * m.S + 1 <= group_rpad(m.R, G) ? 1 : 2);
    std::vector<int> rorder((size_t)((m.R + G - 1) / G) * G, -1);
    for (int j = 0; j < m.R; ++j) rorder[j] = j;          // reaction j: lane j mod G (the selection's partials)
    // pfast: a reactant entry becomes its index in the factor table F
    // (multiplicity 1: s, 2: S + s; padding: 2S, the constant 1.0)
    const uint32_t one = (uint32_t)(2 * m.S);
    auto findex = [&](uint32_t e) { const uint32_t v = e >> 16, sp = e & 0xffffu;
                                    return v == 1u ? sp : (v == 2u ? (uint32_t)m.S + sp : one); };
    if (p.pfast == 1) {
        std::vector<uint32_t> fent(m.Lent.size());
        for (size_t h = 0; h < m.Lent.size(); ++h) fent[h] = findex(m.Lent[h]);
        rounds(rorder, m.Lptr, fent, p.oPRm, p.oPRe, 1, 2, one);
    } else {
        rounds(rorder, m.Lptr, m.Lent, p.oPRm, p.oPRe, 1, p.pfast == 2 ? 2 : 0, 0u);
    }
    p.oDR = 0;
    if (p.pfast) {                        // dependent records, 16-byte aligned
        while (T.size() % 4) T.push_back(0u);
        p.oDR = (int)T.size();
        for (uint32_t r : Dent) {       // factor indices (mode 1) or the entries themselves (mode 2)
            const uint32_t n = m.Lptr[r + 1] - m.Lptr[r];
            const uint32_t e0 = n > 0 ? m.Lent[m.Lptr[r]] : 0u, e1 = n > 1 ? m.Lent[m.Lptr[r] + 1] : 0u;
            T.push_back(r);
            T.push_back(p.pfast == 1 ? (n > 0 ? findex(e0) : one) : e0);
            T.push_back(p.pfast == 1 ? (n > 1 ? findex(e1) : one) : e1);
            T.push_back(0u);
        }
    }
    // species by entry count, longest first, so that a round's lanes carry
    // lists of similar length (tau' is a minimum and the update is per
    // species, so the species-to-lane map is free)
    std::vector<int> sorder((size_t)((m.S + G - 1) / G) * G, -1);
    for (int i = 0; i < m.S; ++i) sorder[i] = i;
    std::stable_sort(sorder.begin(), sorder.begin() + m.S, [&](int a, int b) {
        return std::max(m.Vptr[a + 1] - m.Vptr[a], FSptr[a + 1] - FSptr[a]) >
               std::max(m.Vptr[b + 1] - m.Vptr[b], FSptr[b + 1] - FSptr[b]);
    });
    p.oSPI = (int)T.size();
    p.gsimple = 1;
    for (int i = 0; i < m.S; ++i)
        if (m.hor[i] >= 3 || (m.hor[i] == 2 && m.hort[i] == 2)) p.gsimple = 0;
    for (int i : sorder)
        T.push_back(i < 0 ? 0xffffffffu
                          : (uint32_t)i | ((uint32_t)m.hor[i] << 16) | ((uint32_t)m.hort[i] << 24));
    // var entries: with every |coefficient| <= 2048 (exact in FP16, v * v exact
    // in FP32) and A's byte offsets below 2^16, as FP16 coefficient << 16 |
    // byte offset of A[reaction]; else reaction | v << 16
    p.vhalf = 4 * (m.R + 1) <= 0xffff ? 1 : 0;
    for (uint32_t e : m.Vent) {
        const int v = (int)(int16_t)(e >> 16);
        if (v > 2048 || v < -2048) p.vhalf = 0;
    }
    if (p.vhalf) {
        std::vector<uint32_t> vent(m.Vent.size());
        for (size_t h = 0; h < m.Vent.size(); ++h) {
            const _Float16 hv = (_Float16)(float)(int16_t)(m.Vent[h] >> 16);
            vent[h] = (uint32_t)__builtin_bit_cast(uint16_t, hv) << 16 | (m.Vent[h] & 0xffffu) * 4u;
        }
        rounds(sorder, m.Vptr, vent, p.oVRm, p.oVRe, 4, 0, (uint32_t)m.R * 4u);   // padding: A[R] = 0
    } else {
        rounds(sorder, m.Vptr, m.Vent, p.oVRm, p.oVRe, 4, 0, (uint32_t)m.R);       // padding: the zero slot A[R]
    }
    rounds(sorder, FSptr, FSent, p.oFRm, p.oFRe, 4, 0, 0u);
}

}  // namespace ctl
