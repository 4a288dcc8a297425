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

#include "ctl_group.h"
#include "ctl_internal.h"
#include "ctl_kernels.h"

#pragma clang fp contract(off)

namespace ctl {

// One slot per lane; kLDS keeps the lane's state, trial state, propensities
// (and per-simulation rates) in LDS for the duration of the launch, copied
// from / to the global slot arrays at entry / exit.  With p.compact > 0 (and
// more than one wavefront per workgroup) the workgroup regroups its slots by
// mode every p.compact iterations (regroup): a lane then runs the slot of
// another column, whose state it takes over through the LDS staging area
// that follows the slot arrays; idle lanes stay in the loop until the whole
// workgroup is idle.
template <bool kLDS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void tau_leap_kernel(KParams p) {
    extern __shared__ uint32_t smem[];
    __shared__ float recip[64];
    __shared__ uint32_t rg_cnt[12];
    for (int i = threadIdx.x; i < 64; i += blockDim.x) recip[i] = c_recip[i];
    __syncthreads();
    const int B = (int)blockDim.x;
    int col = (int)threadIdx.x;
    Lane L;
    L.slot = blockIdx.x * B + col;
    if (L.slot >= p.ns) return;          // ns is a multiple of the workgroup size when regrouping
    const bool regrouping = p.compact > 0 && B > 64;
    const size_t lane_words = kLDS ? (size_t)2 * p.S + p.R + (p.c_bcast ? 0 : p.R) : 0;
    uint32_t* W = smem + lane_words * B;                     // staging: LANE_WORDS x B, then perm B
    uint32_t* perm = W + (size_t)LANE_WORDS * B;
    auto view_of = [&](int c) {
        View V;
        if (kLDS) {
            int32_t* base = (int32_t*)smem + c;
            V.X = base;
            V.xbuf = (size_t)p.S * B;
            V.A = (uint32_t*)(base + 2 * V.xbuf);
            V.Cs = (float*)(V.A + (size_t)p.R * B);
            V.st = (size_t)B;
        } else {
            const int slot = blockIdx.x * B + c;
            V.X = p.X + slot;
            V.xbuf = (size_t)p.S * p.ns;
            V.A = p.A + slot;
            V.Cs = p.Cs ? p.Cs + slot : nullptr;
            V.st = (size_t)p.ns;
        }
        return V;
    };
    View V = view_of(col);
    load_lane(p, L);
    if (L.sim == -2 && !regrouping) return;   // queue was exhausted for this slot
    if (kLDS && L.sim >= 0) {            // resume: state from the global slot arrays
        const int32_t* G = p.X + (L.xb ? (size_t)p.S * p.ns : 0) + L.slot;
        for (int s = 0; s < p.S; ++s) V.X[(size_t)s * V.st] = G[(size_t)s * p.ns];
        L.xb = 0;
        load_rates(p, V, L.sim);
    }
    for (int it = 0; it < p.iters; ++it) {
        if (regrouping && it % p.compact == 0) {
            // 0: leap (or about to take a simulation), 1: SSA, 2: no simulation left
            const int key = L.sim == -2 ? 2 : (L.sim >= 0 && L.mode == MODE_SSA) || (L.sim < 0 && p.force_ssa) ? 1 : 0;
            if (!__syncthreads_or(key != 2)) break;
            lane_put(L, W, col, B);
            col = regroup(key, col, rg_cnt, perm);
            lane_get(L, W, col, B);
            L.slot = blockIdx.x * B + col;
            V = view_of(col);
        }
        if (L.sim == -2) continue;
        if (L.sim < 0 && !fetch(p, L, V)) {
            if (regrouping) continue;
            break;
        }
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

// The group kernel instance for G lanes per simulation, with or without
// regrouping, for propensity mode pf (KParams::pfast: 0, 1 or 2).
template <int G, bool RG>
const void* group_kernel_fn_p(int pf) {
    switch (pf) {
        case 1: return (const void*)group_kernel<G, RG, 1>;
        case 2: return (const void*)group_kernel<G, RG, 2>;
    }
    return (const void*)group_kernel<G, RG, 0>;
}

template <int G>
const void* group_kernel_fn_g(bool rg, int pf) {
    if constexpr (G < 64) {
        if (rg) return group_kernel_fn_p<G, true>(pf);
    }
    return group_kernel_fn_p<G, false>(pf);
}

const void* group_kernel_fn(int G, bool rg, int pf) {
    switch (G) {
        case 2: return group_kernel_fn_g<2>(rg, pf);
        case 4: return group_kernel_fn_g<4>(rg, pf);
        case 8: return group_kernel_fn_g<8>(rg, pf);
        case 16: return group_kernel_fn_g<16>(rg, pf);
        case 32: return group_kernel_fn_g<32>(rg, pf);
        case 64: return group_kernel_fn_g<64>(false, pf);
    }
    return nullptr;
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

// Per-group moments of the samples in two passes.  Pass 1: the replicas of a
// group are cut into min(ceil(group / MOM_CHUNK), MOM_MAX_CHUNKS) equal
// chunks (the last one takes the remainder); a workgroup sums one chunk over a
// tile of 256 (sample,
// species) columns, one thread per column, so consecutive threads read
// consecutive words of each simulation's sample row and a single large group
// still spreads over many workgroups.  Partial sums (s1, s2) go to
// part[(g * nchunk + c) * 2 + {0, 1}][KS].  Pass 2 adds a column's chunks in
// chunk order and writes mean and population sd to out[0 or 1][g][KS].
// Every partial is written before it is read, so part needs no clearing.
constexpr int MOM_CHUNK = 256;
constexpr int MOM_MAX_CHUNKS = 64;
constexpr int MOM_TILE = 256;

__host__ __device__ inline int moment_chunks(int group) {
    const int c = (group + MOM_CHUNK - 1) / MOM_CHUNK;
    return c < MOM_MAX_CHUNKS ? c : MOM_MAX_CHUNKS;
}

// Doubles of device memory the two passes need beside the samples: the
// partial sums and the mean / sd staging.
inline size_t moment_words(int groups, int group, int KS) {
    return (size_t)groups * KS * 2 * (size_t)moment_chunks(group) + (size_t)2 * groups * KS;
}

__global__ void moments_partial(const uint32_t* __restrict__ samples, double* __restrict__ part, int group, int KS,
                                int nchunk, int ntiles, size_t nblocks) {
    const int per = group / nchunk;                     // replicas per chunk; the last chunk takes the rest
    for (size_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const int tile = (int)(b % (size_t)ntiles);
        const int c = (int)((b / (size_t)ntiles) % (size_t)nchunk);
        const size_t g = b / ((size_t)ntiles * nchunk);
        const int ks = tile * MOM_TILE + (int)threadIdx.x;
        if (ks >= KS) continue;
        const int r0 = c * per, r1 = c + 1 == nchunk ? group : r0 + per;
        const uint32_t* col = samples + (g * group + r0) * (size_t)KS + ks;
        double s1 = 0.0, s2 = 0.0;
        for (int r = r0; r < r1; ++r) {
            const double v = (double)col[(size_t)(r - r0) * KS];
            s1 += v;
            s2 += v * v;
        }
        double* o = part + ((g * nchunk + c) * 2) * (size_t)KS + ks;
        o[0] = s1;
        o[KS] = s2;
    }
}

__global__ void moments_final(const double* __restrict__ part, double* __restrict__ out, int groups, int group,
                              int KS, int nchunk) {
    const size_t total = (size_t)groups * KS;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const size_t g = idx / KS, ks = idx % KS;
        double s1 = 0.0, s2 = 0.0;
        for (int c = 0; c < nchunk; ++c) {
            const double* o = part + ((g * nchunk + c) * 2) * (size_t)KS + ks;
            s1 += o[0];
            s2 += o[KS];
        }
        const double m = s1 / group;
        double var = s2 / group - m * m;
        if (var < 0.0) var = 0.0;
        out[idx] = m;
        out[total + idx] = sqrt(var);
    }
}

// Device scratch of the allocation-free moment entry point
// (ctl_sim_moments_noalloc): the partial sums and the mean / sd staging of
// the two moment passes.  Each GPU the library runs on holds its own copy
// until the process ends.  Host code never uses this name as an address: it
// takes the per-device address from hipGetSymbolAddress (scratch_setup).
constexpr size_t SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

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
    int lanes = 1;                   // lanes per simulation (1: tau_leap_kernel, else group_kernel<lanes>)
    int compact = 0;                 // iterations between regroupings by mode (0: off)
    bool lds = false;
    size_t smem = 0;
    KParams kp{};
    SlotState sst{};                 // host copy of the slot arrays' device pointers
    std::vector<void*> allocs;
    bool ran = false;
    ctl_stats stats{};
    // fitness
    uint32_t* d_target = nullptr;
    int8_t* d_t2e = nullptr;
    double* d_fit = nullptr;
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

// Device address of `scratch` on each device, looked up once per device
// with that device current (the first ctl_sim_create on it).
constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return ctl::fail(CTL_E_DEVICE, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return 0;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = ptr;
    return 0;
}

void* scratch_ptr(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

}  // namespace

// Argument checks shared by the moment entry points.
static int moments_check(const ctl_sim* s, int32_t group, size_t count) {
    if (!s) return ctl::fail(CTL_E_ARGS, "null sim");
    if (!s->ran) return ctl::fail(CTL_E_STATE, "ctl_sim_run has not completed");
    if (group <= 0 || s->n % group != 0) return ctl::fail(CTL_E_ARGS, "group must divide the simulation count");
    if (count != (size_t)(s->n / group) * s->kp.K * s->kp.SS) return ctl::fail(CTL_E_ARGS, "count must be (n/group)*K*SS");
    return 0;
}

// The two moment passes on the sim's stream, with `work` (device memory of
// moment_words doubles) for the partial sums and the mean / sd staging, then
// the copies to the caller's buffers.  No allocation.
static int moments_run(ctl_sim* s, int32_t group, double* mean, double* sd, size_t count, double* work) {
    HIPCHK(hipSetDevice(s->device), "hipSetDevice");
    const int KS = s->kp.K * s->kp.SS;
    const int groups = s->n / group;
    const int nchunk = moment_chunks(group);
    const int ntiles = (KS + MOM_TILE - 1) / MOM_TILE;
    const size_t nblocks = (size_t)groups * nchunk * ntiles;
    double* part = work;
    double* out = work + (size_t)groups * KS * 2 * nchunk;
    const unsigned g1 = (unsigned)std::min<size_t>(nblocks, 65536);
    hipLaunchKernelGGL(moments_partial, dim3(g1), dim3(MOM_TILE), 0, s->stream, s->kp.storage, part, group, KS,
                       nchunk, ntiles, nblocks);
    HIPCHK(hipGetLastError(), "moments_partial launch");
    const unsigned g2 = (unsigned)std::min<size_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL(moments_final, dim3(g2), dim3(256), 0, s->stream, part, out, groups, group, KS, nchunk);
    HIPCHK(hipGetLastError(), "moments_final launch");
    HIPCHK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    if (mean) HIPCHK(hipMemcpy(mean, out, sizeof(double) * count, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    if (sd) HIPCHK(hipMemcpy(sd, out + count, sizeof(double) * count, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return 0;
}

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

int ctl_phase_clock(int32_t gpu, uint64_t* out, int32_t count, int32_t reset) {
    ctl::clear_error();
#if CTL_PHASE_CLOCK
    if (!out || count < ctl::PH_COUNT) return ctl::fail(CTL_E_ARGS, "need a buffer of CTL_PHASES entries");
    HIPCHK(hipSetDevice(gpu), "hipSetDevice");
    HIPCHK(hipMemcpyFromSymbol(out, HIP_SYMBOL(ctl::phase_total), sizeof(uint64_t) * ctl::PH_COUNT),
           "hipMemcpyFromSymbol");
    if (reset) {
        const uint64_t zero[ctl::PH_COUNT] = {};
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(ctl::phase_total), zero, sizeof zero), "hipMemcpyToSymbol");
    }
    return 0;
#else
    (void)gpu; (void)out; (void)count; (void)reset;
    return ctl::fail(CTL_E_STATE, "built without the phase clock: load libctl_phases.so");
#endif
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
    // rows: 1 = shared, >= n = one per simulation, otherwise n/rows
    // consecutive simulations (a replica group) per row
    if (m.x0_rows > 1 && m.x0_rows < n && n % m.x0_rows)
        return ctl::fail(CTL_E_ROWS, "MX_0 rows must be 1, at least the simulation count, or divide it");
    if (m.c_rows > 1 && m.c_rows < n && n % m.c_rows)
        return ctl::fail(CTL_E_ROWS, "c_matrix rows must be 1, at least the simulation count, or divide it");
    if (o->nc < 0 || o->ssa_steps <= 0 || !(o->eps > 0.0f)) return ctl::fail(CTL_E_ARGS, "bad tunables");
    // Philox stream indices are the 32-bit global simulation indices offset .. offset + n - 1:
    // a range past 2^32 would wrap and alias the streams of another shard
    if (o->offset < 0 || o->offset + n64 > 0x100000000LL)
        return ctl::fail(CTL_E_ARGS, "offset must be >= 0 and offset + threads * blocks <= 2^32");
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
    if (int rc0 = scratch_setup(s->device)) return bail(rc0);
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreate(&s->ev0)) != hipSuccess || (e = hipEventCreate(&s->ev1)) != hipSuccess)
        return bail(hip_fail(e, "hipEventCreate"));

    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, s->device);
    if (e != hipSuccess) return bail(hip_fail(e, "hipGetDeviceProperties"));
    const int S = m.S, R = m.R, K = (int)m.ts.size(), SS = (int)m.cs.size();
    // Launch geometry.  One lane per simulation when its working set (state,
    // trial state, propensities, per-simulation rates) fits a 64-lane
    // wavefront's 16 KB share of LDS, or when asked for; otherwise G lanes
    // per simulation with the simulation's state in LDS (group kernel).
    const size_t lane_words = (size_t)2 * S + R + (m.c_rows == 1 ? 0 : R);
    const bool lds_auto = lane_words * 4 * 64 <= 16384;
    const size_t gwords64 = group_lds_words(S, R, 64);     // largest group LDS words per simulation
    int G = o->lanes_per_sim;
    if (G == 0) {
        if (lds_auto || gwords64 * 4 > 65536) {
            G = 1;
        } else {
            G = 4;
            while (G < 64 && G * 16 < R) G *= 2;      // about 16 reactions per lane at most
        }
    }
    if (G < 1 || G > 64 || (G & (G - 1)))
        return bail(ctl::fail(CTL_E_ARGS, "lanes_per_sim must be 0 or a power of two up to 64"));
    s->lanes = G;
    int per_cu = 0;
    if (G > 1) {
        const size_t gwords = group_lds_words(S, R, G);    // group LDS words per simulation
        if (gwords * 4 > 65536) return bail(ctl::fail(CTL_E_LIMIT, "model too large for the group kernel (more than 64 KB of LDS)"));
        // regrouping simulations by mode (leap / SSA) every `compact` iterations;
        // by default for groups of 16 or more lanes (networks of more than 128
        // reactions, where a leap costs several SSA steps) and never in
        // force-SSA mode, where every simulation is in one mode (measured: the
        // 100x200 network, G = 16: +10 %; the 50x100 sweep network, G = 8:
        // -6 %; the SSA-dominated enzyme cascade, G = 4: -40 %)
        const int compact = o->compact_interval > 0 ? o->compact_interval
                            : (o->compact_interval < 0 || o->force_ssa || G < 16) ? 0 : 8;
        // groups per workgroup: block_size / G when given (64..256), else two
        // wavefronts when regrouping (fewer wavefronts wait for the slowest at
        // each regrouping: C1 +1.5 % over four) and four otherwise
        if (o->block_size > 0 && (o->block_size % 64 != 0 || o->block_size > 256))
            return bail(ctl::fail(CTL_E_ARGS, "block_size must be 64..256, multiple of 64"));
        int gpb = (o->block_size > 0 ? o->block_size : (compact > 0 && G < 64 ? 128 : 256)) / G;
        while (gpb > 64 / G && (size_t)gpb * gwords * 4 > 65536) gpb /= 2;
        if ((size_t)gpb * gwords * 4 > 65536) gpb = 64 / G;   // one wavefront
        if ((size_t)gpb * gwords * 4 > 65536)
            return bail(ctl::fail(CTL_E_LIMIT, "model too large for this lanes_per_sim (use 64)"));
        s->block = gpb * G;
        s->smem = (size_t)gpb * gwords * 4;
        s->lds = true;
        s->compact = compact > 0 && G < 64 && s->block > 64 ? compact : 0;
        const void* kfn = group_kernel_fn(G, s->compact > 0, 1);   // occupancy (the instances alike)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, s->block, s->smem);
        if (e != hipSuccess || per_cu <= 0) per_cu = 1;
        long long groups = o->slots > 0 ? (long long)o->slots : (long long)prop.multiProcessorCount * per_cu * gpb;
        if (groups > n) groups = ((n + gpb - 1) / gpb) * (long long)gpb;
        groups = ((groups + gpb - 1) / gpb) * gpb;
        s->ns = (int)(groups * G);
        s->grid = (int)(groups / gpb);
    } else {
        // regrouping slots by mode needs several wavefronts per workgroup
        const int compact = o->compact_interval > 0 ? o->compact_interval : 0;    // measured slower by default
        s->lds = o->block_size < 0 ? false : lds_auto;
        auto smem_of = [&](bool lds, int block) {
            return (lds ? lane_words * 4 * block : 0) +
                   (compact > 0 && block > 64 ? (size_t)(LANE_WORDS + 1) * 4 * block : 0);
        };
        auto block_of = [&](bool lds) {
            if (o->block_size > 0) return o->block_size;
            if (compact <= 0) return lds ? 64 : 256;
            return smem_of(lds, 256) <= 65536 ? 256 : 128;
        };
        s->block = block_of(s->lds);
        if (s->block % 64 != 0 || s->block > 256) return bail(ctl::fail(CTL_E_ARGS, "block_size must be 64..256, multiple of 64"));
        s->smem = smem_of(s->lds, s->block);
        if (s->smem > 65536) {
            s->lds = false;
            s->block = block_of(false);
            s->smem = smem_of(false, s->block);
        }
        s->compact = compact > 0 && s->block > 64 ? compact : 0;
        const void* kfn = s->lds ? (const void*)tau_leap_kernel<true> : (const void*)tau_leap_kernel<false>;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, s->block, s->smem);
        if (e != hipSuccess || per_cu <= 0) per_cu = 1;
        int ns = o->slots > 0 ? o->slots : prop.multiProcessorCount * per_cu * s->block;
        if (ns > n) ns = ((n + s->block - 1) / s->block) * s->block;
        if (s->compact) ns = ((ns + s->block - 1) / s->block) * s->block;   // whole workgroups
        s->ns = ns;
        s->grid = (ns + s->block - 1) / s->block;
    }
    const int ns = s->ns;

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
    // fear entries species-major, ascending reaction (group kernel update)
    std::vector<uint32_t> FSptr(S + 1, 0), FSent(m.Fent.size());
    {
        std::vector<uint32_t> cnt(S + 1, 0);
        for (uint32_t e : m.Fent) cnt[(e & 0xffffu) + 1]++;
        for (int i = 0; i < S; ++i) cnt[i + 1] += cnt[i];
        for (int i = 0; i <= S; ++i) FSptr[i] = cnt[i];
        for (int j = 0; j < R; ++j)
            for (uint32_t h = m.Fptr[j]; h < m.Fptr[j + 1]; ++h) {
                const uint32_t e = m.Fent[h];
                FSent[cnt[e & 0xffffu]++] = (uint32_t)j | (e & 0xffff0000u);
            }
    }
    uint32_t *Lptr, *Lent, *Fptr, *Fent, *Vptr, *Vent, *Dp, *De, *hw, *FSp, *FSe;
    int32_t* cs;
    double* ts;
    if ((rc = upload(s, &Lptr, m.Lptr.data(), m.Lptr.size())) || (rc = upload(s, &Lent, m.Lent.data(), m.Lent.size())) ||
        (rc = upload(s, &Fptr, m.Fptr.data(), m.Fptr.size())) || (rc = upload(s, &Fent, m.Fent.data(), m.Fent.size())) ||
        (rc = upload(s, &Vptr, m.Vptr.data(), m.Vptr.size())) || (rc = upload(s, &Vent, m.Vent.data(), m.Vent.size())) ||
        (rc = upload(s, &Dp, Dptr.data(), Dptr.size())) || (rc = upload(s, &De, Dent.data(), Dent.size())) ||
        (rc = upload(s, &hw, horw.data(), horw.size())) || (rc = upload(s, &cs, m.cs.data(), m.cs.size())) ||
        (rc = upload(s, &ts, m.ts.data(), m.ts.size())) ||
        (rc = upload(s, &FSp, FSptr.data(), FSptr.size())) || (rc = upload(s, &FSe, FSent.data(), FSent.size())))
        return bail(rc);
    p.FSptr = FSp;
    p.FSent = FSe;
    {   // the same tables packed behind one base pointer (group kernels)
        std::vector<uint32_t> T;
        auto put = [&](const uint32_t* d, size_t k) { const int o = (int)T.size(); T.insert(T.end(), d, d + k); return o; };
        p.oLp = put(m.Lptr.data(), m.Lptr.size());
        p.oLe = put(m.Lent.data(), m.Lent.size());
        p.oVp = put(m.Vptr.data(), m.Vptr.size());
        p.oVe = put(m.Vent.data(), m.Vent.size());
        p.oHw = put(horw.data(), horw.size());
        p.oFSp = put(FSptr.data(), FSptr.size());
        p.oFSe = put(FSent.data(), FSent.size());
        p.oFp = put(m.Fptr.data(), m.Fptr.size());
        p.oFe = put(m.Fent.data(), m.Fent.size());
        p.oDp = put(Dptr.data(), Dptr.size());
        p.oDe = put(Dent.data(), Dent.size());
        p.oCs = put(reinterpret_cast<const uint32_t*>(m.cs.data()), m.cs.size());
        if (s->lanes > 1) append_group_streams(T, p, s->lanes, m, FSptr, FSent, Dent);
        uint32_t* Td;
        if ((rc = upload(s, &Td, T.data(), T.size()))) return bail(rc);
        p.T = Td;
    }
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
    p.x0_div = (m.x0_rows <= 1 || m.x0_rows >= n) ? 1 : (int)(n / m.x0_rows);
    p.c_div = (m.c_rows <= 1 || m.c_rows >= n) ? 1 : (int)(n / m.c_rows);
    p.nfs = (int)m.fs.size();
    p.fs = nullptr;
    p.fv = nullptr;
    p.fv_div = 1;
    if (p.nfs > 0) {
        if (m.fv_rows > 1 && m.fv_rows < n && n % m.fv_rows)
            return bail(ctl::fail(CTL_E_ROWS, "MX_feed rows must be 1, at least the simulation count, or divide it"));
        int32_t* fsd;
        uint32_t* fvd;
        if ((rc = upload(s, &fsd, m.fs.data(), m.fs.size())) || (rc = upload(s, &fvd, m.fv.data(), m.fv.size())))
            return bail(rc);
        p.fs = fsd;
        p.fv = fvd;
        p.fv_div = m.fv_rows <= 1 ? 0x7fffffff : (m.fv_rows >= n ? 1 : (int)(n / m.fv_rows));
    }
    p.n = n;
    p.sim_base = (uint32_t)o->offset;
    p.key0 = (uint32_t)o->seed;
    p.key1 = (uint32_t)(o->seed >> 32);
    p.nc = o->nc;
    p.eps = o->eps;
    p.ssa_steps = o->ssa_steps;
    p.ssa_factor = o->ssa_factor;
    p.force_ssa = o->force_ssa ? 1 : 0;
    p.ssa_chunk = o->ssa_chunk > 0 ? o->ssa_chunk : (G > 1 ? 8 : 10);
    p.compact = s->compact;
    p.iters = o->iters_per_launch > 0 ? o->iters_per_launch : (G > 1 ? 65536 : 4096);
    p.ns = ns;
    p.Cs = nullptr;
    if (G > 1) {                         // parked group state only: X [group][S]
        if ((rc = dalloc(s, &p.X, (size_t)S * (ns / G))) || (rc = dalloc(s, &p.A, 1))) return bail(rc);
    } else {
        if ((rc = dalloc(s, &p.X, (size_t)2 * S * ns)) || (rc = dalloc(s, &p.A, (size_t)R * ns))) return bail(rc);
        if (!p.c_bcast && (rc = dalloc(s, &p.Cs, (size_t)R * ns))) return bail(rc);
    }
    SlotState& st = s->sst;
    if ((rc = dalloc(s, &st.sim, ns)) || (rc = dalloc(s, &st.t, ns)) || (rc = dalloc(s, &st.tacc, ns)) ||
        (rc = dalloc(s, &st.cursor, ns)) || (rc = dalloc(s, &st.mode, ns)) || (rc = dalloc(s, &st.evc, ns)) ||
        (rc = dalloc(s, &st.ev, ns)) || (rc = dalloc(s, &st.blk, ns)) || (rc = dalloc(s, &st.pos, ns)) ||
        (rc = dalloc(s, &st.ssa_left, ns)) || (rc = dalloc(s, &st.xb, ns)) || (rc = dalloc(s, &st.firings, ns)) ||
        (rc = dalloc(s, &st.ssa_steps, ns)) || (rc = dalloc(s, &st.poisson, ns)) || (rc = dalloc(s, &st.gaps, ns)) ||
        (rc = dalloc(s, &st.leaps, ns)) ||
        (rc = dalloc(s, &st.rejections, ns)) || (rc = dalloc(s, &st.switches, ns)))
        return bail(rc);
    {
        SlotState* dst;
        if ((rc = upload(s, &dst, &s->sst, 1))) return bail(rc);
        p.st = dst;
    }
    if ((rc = dalloc(s, &p.storage, (size_t)n * K * SS)) || (rc = dalloc(s, &p.simstat, n)) ||
        (rc = dalloc(s, &p.xfinal, (size_t)n * S)) || (rc = dalloc(s, &p.tfinal, n)) ||
        (rc = dalloc(s, &p.next_sim, 2)))
        return bail(rc);
    p.done = p.next_sim + 1;
    s->stats.slots = ns;
    s->stats.block_size = s->block;
    s->stats.lds_state = s->lds ? 1 : 0;
    s->stats.lanes_per_sim = G;
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
    HIPCHK(hipMemsetAsync(s->sst.sim, 0xff, sizeof(int32_t) * s->ns, s->stream), "memset");
    HIPCHK(hipMemsetAsync(p.next_sim, 0, 2 * sizeof(uint32_t), s->stream), "memset");
    ctl_stats& S = s->stats;
    S.kernel_ms = 0.0;
    S.launches = 0;
    uint32_t done = 0;
    const int max_launches = 1 << 20;
    float total_ms = 0.0f;
    while (done < (uint32_t)s->n) {
        HIPCHK(hipEventRecord(s->ev0, s->stream), "hipEventRecord");
        switch (s->lanes) {
            case 1:
                if (s->lds)
                    hipLaunchKernelGGL(tau_leap_kernel<true>, dim3(s->grid), dim3(s->block), s->smem, s->stream, p);
                else
                    hipLaunchKernelGGL(tau_leap_kernel<false>, dim3(s->grid), dim3(s->block), s->smem, s->stream, p);
                break;
            default: {
                const void* fn = group_kernel_fn(s->lanes, s->compact > 0, p.pfast);
                if (!fn) return ctl::fail(CTL_E_STATE, "bad lanes_per_sim");
                void* args[] = {&p};
                HIPCHK(hipLaunchKernel(fn, dim3(s->grid), dim3(s->block), args, s->smem, s->stream),
                       "group_kernel launch");
            }
        }
        HIPCHK(hipGetLastError(), "tau_leap_kernel launch");
        HIPCHK(hipEventRecord(s->ev1, s->stream), "hipEventRecord");
        HIPCHK(hipMemcpyAsync(&done, p.done, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream), "hipMemcpy D2H");
        HIPCHK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1), "hipEventElapsedTime");
        total_ms += ms;
        S.launches++;
        if (s->opts.verbose)   // progress on stderr (unbuffered): long runs report as they go
            std::fprintf(stderr, " * launch %d: %u/%d simulations done (%.3f ms)\n", S.launches, done, s->n, ms);
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
    if (!s->d_fit && (rc = dalloc(s, &s->d_fit, nf))) return rc;    // once per sim, reused by later calls
    double* d_fit = s->d_fit;
    hipLaunchKernelGGL(fitness_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s->stream, s->kp.storage,
                       s->d_target, s->d_t2e, d_fit, threads, groups, P, s->kp.K, s->kp.SS, (int)f.experiments,
                       (int)f.repetitions);
    HIPCHK(hipGetLastError(), "fitness_kernel launch");
    HIPCHK(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    HIPCHK(hipMemcpy(out, d_fit, sizeof(double) * nf, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return 0;
}

int ctl_sim_moments(const ctl_sim* sc, int32_t group, double* mean, double* sd, size_t count) {
    ctl::clear_error();
    ctl_sim* s = const_cast<ctl_sim*>(sc);
    int rc = moments_check(s, group, count);
    if (rc) return rc;
    const size_t words = moment_words(s->n / group, group, s->kp.K * s->kp.SS);
    double* dev = nullptr;
    HIPCHK(hipMalloc(&dev, sizeof(double) * words), "hipMalloc");
    rc = moments_run(s, group, mean, sd, count, dev);
    (void)hipFree(dev);
    return rc;
}

int ctl_sim_moments_noalloc(const ctl_sim* sc, int32_t group, double* mean, double* sd, size_t count) {
    ctl::clear_error();
    ctl_sim* s = const_cast<ctl_sim*>(sc);
    int rc = moments_check(s, group, count);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * moment_words(s->n / group, group, s->kp.K * s->kp.SS);
    if (bytes > SCRATCH_BYTES) {
        char b[160];
        std::snprintf(b, sizeof b, "moments need %zu bytes of device scratch; the allocation-free entry point has %zu",
                      bytes, SCRATCH_BYTES);
        return ctl::fail(CTL_E_LIMIT, b);
    }
    void* ptr = scratch_ptr(s->device);
    if (!ptr) return ctl::fail(CTL_E_STATE, "device scratch not set up");
    return moments_run(s, group, mean, sd, count, static_cast<double*>(ptr));
}

int ctl_write_samples_binary(const ctl_sim* s, const char* path) {
    ctl::clear_error();
    if (!s || !path) return ctl::fail(CTL_E_ARGS, "null argument");
    if (!s->ran) return ctl::fail(CTL_E_STATE, "ctl_sim_run has not completed");
    const size_t total = (size_t)s->n * s->kp.K * s->kp.SS;
    std::vector<uint32_t> buf(total);
    HIPCHK(hipSetDevice(s->device), "hipSetDevice");
    HIPCHK(hipMemcpy(buf.data(), s->kp.storage, total * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    FILE* f = std::fopen(path, "wb");
    if (!f) return ctl::fail(CTL_E_ARGS, std::string("cannot write ") + path);
    uint32_t hdr[8] = {0x534c5443u /* "CTLS" */, 1u, (uint32_t)s->n, (uint32_t)s->kp.K, (uint32_t)s->kp.SS, 0, 0, 0};
    bool ok = std::fwrite(hdr, sizeof hdr, 1, f) == 1 && std::fwrite(buf.data(), sizeof(uint32_t), total, f) == total;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return ctl::fail(CTL_E_ARGS, std::string("short write to ") + path);
    return 0;
}

void ctl_sim_destroy(ctl_sim* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    release(s);
    delete s;
}

}  // extern "C"
