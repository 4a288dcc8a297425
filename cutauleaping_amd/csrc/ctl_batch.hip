// ctl_batch.hip -- many independent models of uneven size in one persistent
// launch (no counterpart in the reference, which runs one model per process).
//
// Two kernels.  batch_kernel runs one simulation per lane (small models).
// gbatch_kernel runs the group algorithm with one 64-lane wavefront per
// simulation and the simulation's state in LDS (large models): a wavefront
// takes the next simulation from one global counter over all models in
// estimated-cost order (largest first), so its model is wave-uniform and its
// descriptor lives in scalar registers.
//
// batch_kernel: work is split into chunks of 64 replicas of one model, ordered by an
// estimated per-replica cost, largest first.  A wavefront takes the next
// chunk's model as its current model -- so the model tables stay
// wave-uniform scalar loads -- and each of its lanes then steals replicas of
// that model from a per-model counter until the model is exhausted; when all
// its lanes are idle the wavefront takes the next chunk.  Several wavefronts
// can work on one model, and a lane never waits for a whole chunk to finish.
// Per-lane state lives in the global slot arrays, sized for the largest
// model.  A replica's Philox stream index is the model's base (sum of the
// replicas of the models before it, plus the offset) plus its replica index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

// batch_kernel lanes read their own model's descriptor: parameters are not
// known to be uniform
#define CTL_UNIFORM_PARAMS 0
#include "ctl_gbatch.h"
#include "ctl_group.h"
#include "ctl_kernels.h"

#pragma clang fp contract(off)

namespace ctl {

struct BatchParams {
    const KParams* desc;        // per-model descriptors (device)
    const int32_t* chunk_model; // chunk -> model
    int nchunks;
    uint32_t* chunk_next;
    uint32_t* rep_next;         // per-model replica counters
    int32_t* slot_model;        // model of each slot's current replica, -1 none
    int ns;
    int iters;
    int32_t* X;
    uint32_t* A;
    float* Cs;
};

// The descriptor of model m, copied through a constant-address-space pointer
// (m is wave-uniform, so the copy is a burst of scalar loads).
__device__ __forceinline__ KParams desc_of(const BatchParams& bp, int m) {
    static_assert(sizeof(KParams) % 4 == 0, "descriptor copied as words");
    KParams p;
    const ctab_u32 src = CT_U32(reinterpret_cast<const uint32_t*>(bp.desc + m));
    uint32_t* dst = reinterpret_cast<uint32_t*>(&p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(KParams) / 4); ++i) dst[i] = src[i];
    return p;
}


__global__ __launch_bounds__(256) void batch_kernel(BatchParams bp) {
    __shared__ float recip[64];
    for (int i = threadIdx.x; i < 64; i += blockDim.x) recip[i] = c_recip[i];
    __syncthreads();
    Lane L;
    L.slot = blockIdx.x * blockDim.x + threadIdx.x;    // ns is a multiple of the workgroup size
    View V;
    V.X = bp.X + L.slot;
    V.A = bp.A + L.slot;
    V.Cs = bp.Cs ? bp.Cs + L.slot : nullptr;
    V.st = (size_t)bp.ns;
    // the wavefront's current model: that of any lane still holding a replica
    const int lm = bp.slot_model[L.slot];
    const unsigned long long held = __ballot(lm >= 0);
    int m = -1;
    if (held) m = __builtin_amdgcn_readfirstlane(__shfl(lm, __ffsll((long long)held) - 1));
    L.sim = -1;
    L.g.k0 = 0;
    L.g.k1 = 0;
    if (m >= 0 && lm >= 0) {             // resume the replica this slot holds
        const KParams p = desc_of(bp, m);
        V.xbuf = (size_t)p.S * bp.ns;
        load_lane(p, L);
    }
    bool exhausted = false;
    for (int it = 0; it < bp.iters; ++it) {
        if (m < 0) {
            uint32_t c = 0;
            if (__lane_id() == 0) c = atomicAdd(bp.chunk_next, 1u);
            c = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
            if (c >= (uint32_t)bp.nchunks) break;
            m = __builtin_amdgcn_readfirstlane(CT_I32(bp.chunk_model)[c]);
            exhausted = false;
        }
        const KParams p = desc_of(bp, m);
        V.xbuf = (size_t)p.S * bp.ns;
        if (L.sim < 0 && !exhausted) {
            L.g.k0 = p.key0;
            L.g.k1 = p.key1;
            const uint32_t r = atomicAdd(bp.rep_next + m, 1u);
            if (r < (uint32_t)p.n) start(p, L, V, r);
            else exhausted = true;
        }
        if (!__any(L.sim >= 0)) { m = -1; continue; }
        if (L.sim < 0) continue;
        if (L.mode == MODE_LEAP) {
            const int r = leap(p, L, V, recip);
            if (r == LEAP_DONE || r == LEAP_ABSORB) {
                finish(p, L, V);
            } else if (r == LEAP_SWITCH) {
                L.mode = MODE_SSA;
            }
        } else {
            if (L.ssa_left == 0) {
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
    bp.slot_model[L.slot] = L.sim >= 0 ? m : -1;
    if (L.sim >= 0) store_lane(desc_of(bp, m), L);
}


}  // namespace ctl

// ===========================================================================
// Host runtime
// ===========================================================================
using namespace ctl;

struct ctl_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nmodels = 0, ns = 0, block = 256, grid = 0, nchunks = 0;
    int64_t total = 0;
    std::vector<KParams> desc;                 // host copies (device pointers inside)
    std::vector<int64_t> reps, base;
    std::vector<int> K, SS;
    KParams* d_desc = nullptr;
    BatchParams bp{};
    int lanes = 1;                             // 1: batch_kernel; 64: gbatch_kernel
    int verbose = 0;
    GBatchParams gp{};
    size_t smem = 0;
    uint32_t* d_done = nullptr;
    std::vector<void*> allocs;
    bool ran = false;
    ctl_stats stats{};
};

namespace {

int bhip(hipError_t e, const char* what) {
    char b[256];
    std::snprintf(b, sizeof b, "%s: %s", what, hipGetErrorString(e));
    return ctl::fail(e == hipErrorOutOfMemory ? CTL_E_NOMEM : CTL_E_DEVICE, b);
}

#define BHIP(call, what)                            \
    do {                                            \
        hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return bhip(e_, what); \
    } while (0)

template <typename T>
int balloc(ctl_batch* b, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return bhip(e, "hipMalloc");
    b->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

template <typename T>
int bupload(ctl_batch* b, T** p, const T* src, size_t count) {
    int rc = balloc(b, p, count);
    if (rc) return rc;
    if (count) BHIP(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return 0;
}

void brelease(ctl_batch* b) {
    for (void* q : b->allocs) (void)hipFree(q);
    b->allocs.clear();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->ev0 = b->ev1 = nullptr;
    b->stream = nullptr;
}

// Per-model device tables (the same compile products as a ctl_sim).
int upload_model(ctl_batch* b, const ctl::Model& m, KParams& p) {
    const int S = m.S, R = m.R;
    std::vector<uint32_t> horw(S);
    for (int i = 0; i < S; ++i) horw[i] = (uint32_t)m.hor[i] | ((uint32_t)m.hort[i] << 8);
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
            for (int32_t fsp : m.fs)     // feeds reset these species after every SSA step
                for (uint32_t r : readers[fsp])
                    if (!mark[r]) { mark[r] = 1; list.push_back(r); }
            std::sort(list.begin(), list.end());
            for (uint32_t r : list) { mark[r] = 0; Dent.push_back(r); }
        }
        Dptr[R] = (uint32_t)Dent.size();
    }
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
    std::vector<float> cf(m.c.size());
    for (size_t i = 0; i < m.c.size(); ++i) cf[i] = (float)m.c[i];
    uint32_t *Lptr, *Lent, *Fptr, *Fent, *Vptr, *Vent, *Dp, *De, *hw, *x0, *FSp, *FSe;
    int32_t* cs;
    double* ts;
    float* cdev;
    int rc;
    if ((rc = bupload(b, &Lptr, m.Lptr.data(), m.Lptr.size())) || (rc = bupload(b, &Lent, m.Lent.data(), m.Lent.size())) ||
        (rc = bupload(b, &Fptr, m.Fptr.data(), m.Fptr.size())) || (rc = bupload(b, &Fent, m.Fent.data(), m.Fent.size())) ||
        (rc = bupload(b, &Vptr, m.Vptr.data(), m.Vptr.size())) || (rc = bupload(b, &Vent, m.Vent.data(), m.Vent.size())) ||
        (rc = bupload(b, &Dp, Dptr.data(), Dptr.size())) || (rc = bupload(b, &De, Dent.data(), Dent.size())) ||
        (rc = bupload(b, &hw, horw.data(), horw.size())) || (rc = bupload(b, &cs, m.cs.data(), m.cs.size())) ||
        (rc = bupload(b, &ts, m.ts.data(), m.ts.size())) || (rc = bupload(b, &cdev, cf.data(), cf.size())) ||
        (rc = bupload(b, &x0, m.x0.data(), m.x0.size())) ||
        (rc = bupload(b, &FSp, FSptr.data(), FSptr.size())) || (rc = bupload(b, &FSe, FSent.data(), FSent.size())))
        return rc;
    p.FSptr = FSp;
    p.FSent = FSe;
    p.nfs = (int)m.fs.size();
    p.fs = nullptr;
    p.fv = nullptr;
    if (p.nfs > 0) {
        int32_t* fsd;
        uint32_t* fvd;
        if ((rc = bupload(b, &fsd, m.fs.data(), m.fs.size())) || (rc = bupload(b, &fvd, m.fv.data(), m.fv.size())))
            return rc;
        p.fs = fsd;
        p.fv = fvd;
    }
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
        append_group_streams(T, p, 64, m, FSptr, FSent, Dent);
        uint32_t* Td;
        if ((rc = bupload(b, &Td, T.data(), T.size()))) return rc;
        p.T = Td;
    }
    p.S = S; p.R = R; p.K = (int)m.ts.size(); p.SS = (int)m.cs.size();
    p.Lptr = Lptr; p.Lent = Lent; p.Fptr = Fptr; p.Fent = Fent; p.Vptr = Vptr; p.Vent = Vent;
    p.Dptr = Dp; p.Dent = De; p.horw = hw; p.cs = cs; p.ts = ts; p.tmax = m.ts.back();
    p.c_bcast = m.c_rows == 1 ? cdev : nullptr;
    p.c_sims = m.c_rows == 1 ? nullptr : cdev;
    p.x0_bcast = m.x0_rows == 1 ? x0 : nullptr;
    p.x0_sims = m.x0_rows == 1 ? nullptr : x0;
    return 0;
}

}  // namespace

extern "C" {

int ctl_batch_create(const ctl_model* const* models, const int64_t* replicas, int32_t nmodels, const ctl_opts* o,
                     ctl_batch** out) {
    ctl::clear_error();
    if (!models || !replicas || !o || !out || nmodels <= 0) return ctl::fail(CTL_E_ARGS, "bad arguments");
    *out = nullptr;
    int maxS = 0, maxR = 0;
    bool any_rates = false;
    int64_t total = 0;
    for (int i = 0; i < nmodels; ++i) {
        if (!models[i] || replicas[i] <= 0) return ctl::fail(CTL_E_ARGS, "null model or non-positive replica count");
        const ctl::Model& m = models[i]->m;
        if ((!m.fs.empty() && m.fv_rows > 1 && m.fv_rows < replicas[i] && replicas[i] % m.fv_rows) ||
            (m.x0_rows > 1 && m.x0_rows < replicas[i] && replicas[i] % m.x0_rows) ||
            (m.c_rows > 1 && m.c_rows < replicas[i] && replicas[i] % m.c_rows))
            return ctl::fail(CTL_E_ROWS, "a model's input rows must be 1, at least its replicas, or divide them");
        maxS = std::max(maxS, m.S);
        maxR = std::max(maxR, m.R);
        any_rates = any_rates || m.c_rows != 1;
        total += replicas[i];
    }
    if (total > 0x7fffffffLL) return ctl::fail(CTL_E_LIMIT, "too many simulations");
    if (o->offset < 0 || o->offset + total > 0x100000000LL)    // 32-bit Philox stream indices must not wrap
        return ctl::fail(CTL_E_ARGS, "offset must be >= 0 and offset + total replicas <= 2^32");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return ctl::fail(CTL_E_DEVICE, "no HIP device available: libctl needs an AMD GPU (gfx950)");
    if (o->gpu < 0 || o->gpu >= ndev) return ctl::fail(CTL_E_DEVICE, "gpu ordinal out of range");
    ctl_batch* b = new ctl_batch();
    auto bail = [&](int rc) { brelease(b); delete b; return rc; };
    b->device = o->gpu;
    b->verbose = o->verbose;
    b->nmodels = nmodels;
    b->total = total;
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return bail(bhip(e, "hipSetDevice"));
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bail(bhip(e, "hipStreamCreate"));
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess)
        return bail(bhip(e, "hipEventCreate"));
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, b->device)) != hipSuccess) return bail(bhip(e, "hipGetDeviceProperties"));
    // kernel: one lane per simulation when every model's per-lane working set
    // fits 16 KB of LDS per wavefront, else one 64-lane wavefront per simulation
    size_t lane_words = 0, gwords = 0;
    for (int i = 0; i < nmodels; ++i) {
        const ctl::Model& m = models[i]->m;
        lane_words = std::max(lane_words, (size_t)2 * m.S + m.R + (m.c_rows == 1 ? 0 : m.R));
        gwords = std::max(gwords, ctl::group_lds_words(m.S, m.R, 64));
    }
    int G = o->lanes_per_sim;
    if (G == 0) G = lane_words * 4 * 64 <= 16384 ? 1 : 64;
    if (G != 1 && G != 64) return bail(ctl::fail(CTL_E_ARGS, "a batch runs with lanes_per_sim 0, 1 or 64"));
    if (G == 64 && gwords * 4 > 65536)
        return bail(ctl::fail(CTL_E_LIMIT, "a model is too large for the group kernel (more than 64 KB of LDS)"));
    b->lanes = G;
    int per_cu = 0;
    int rc;
    KParams shared{};
    SlotState st{};
    int32_t* X = nullptr;
    uint32_t* A = nullptr;
    float* Cs = nullptr;
    int ns = 0;
    if (G == 1) {
        b->block = o->block_size > 0 ? o->block_size : 256;
        if (b->block % 64 != 0 || b->block > 256) return bail(ctl::fail(CTL_E_ARGS, "block_size must be 64..256, multiple of 64"));
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)batch_kernel, b->block, 0);
        if (e != hipSuccess || per_cu <= 0) per_cu = 1;
        ns = o->slots > 0 ? o->slots : prop.multiProcessorCount * per_cu * b->block;
        ns = ((ns + b->block - 1) / b->block) * b->block;
        b->ns = ns;
        b->grid = ns / b->block;
        // shared slot arrays (sized for the largest model) and counters
        if ((rc = balloc(b, &X, (size_t)2 * maxS * ns)) || (rc = balloc(b, &A, (size_t)maxR * ns)) ||
            (any_rates && (rc = balloc(b, &Cs, (size_t)maxR * ns))))
            return bail(rc);
        if ((rc = balloc(b, &st.sim, ns)) || (rc = balloc(b, &st.t, ns)) || (rc = balloc(b, &st.tacc, ns)) ||
            (rc = balloc(b, &st.cursor, ns)) || (rc = balloc(b, &st.mode, ns)) || (rc = balloc(b, &st.evc, ns)) ||
            (rc = balloc(b, &st.ev, ns)) || (rc = balloc(b, &st.blk, ns)) || (rc = balloc(b, &st.pos, ns)) ||
            (rc = balloc(b, &st.ssa_left, ns)) || (rc = balloc(b, &st.xb, ns)) || (rc = balloc(b, &st.firings, ns)) ||
            (rc = balloc(b, &st.ssa_steps, ns)) || (rc = balloc(b, &st.poisson, ns)) || (rc = balloc(b, &st.gaps, ns)) ||
            (rc = balloc(b, &st.leaps, ns)) || (rc = balloc(b, &st.rejections, ns)) || (rc = balloc(b, &st.switches, ns)))
            return bail(rc);
        SlotState* dst;
        if ((rc = bupload(b, &dst, &st, 1))) return bail(rc);
        shared.st = dst;
    } else {
        b->block = 64;
        b->smem = gwords * 4;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)gbatch_kernel, 64, b->smem);
        if (e != hipSuccess || per_cu <= 0) per_cu = 1;
        long long groups = o->slots > 0 ? (long long)o->slots : (long long)prop.multiProcessorCount * per_cu;
        if (groups > total) groups = total;
        b->ns = (int)(groups * 64);
        b->grid = (int)groups;
        GBatchParams& g = b->gp;
        g.ngroups = (int)groups;
        g.maxS = maxS;
        g.total = total;
        if ((rc = balloc(b, &g.slot_model, groups)) || (rc = balloc(b, &g.slot_sim, groups)) ||
            (rc = balloc(b, &g.t, groups)) || (rc = balloc(b, &g.cursor, groups)) || (rc = balloc(b, &g.evc, groups)) ||
            (rc = balloc(b, &g.firings, groups)) || (rc = balloc(b, &g.ssa_steps, groups)) ||
            (rc = balloc(b, &g.leaps, groups)) || (rc = balloc(b, &g.rejections, groups)) ||
            (rc = balloc(b, &g.switches, groups)) || (rc = balloc(b, &g.poisson, (size_t)groups * 64)) ||
            (rc = balloc(b, &g.gaps, (size_t)groups * 64)) || (rc = balloc(b, &g.X, (size_t)groups * maxS)) ||
            (rc = balloc(b, &g.next, 1)))
            return bail(rc);
        g.iters = o->iters_per_launch > 0 ? o->iters_per_launch : 65536;
    }
    ns = b->ns;
    if ((rc = balloc(b, &b->d_done, 1))) return bail(rc);
    b->desc.resize(nmodels);
    b->reps.assign(replicas, replicas + nmodels);
    b->base.resize(nmodels);
    b->K.resize(nmodels);
    b->SS.resize(nmodels);
    std::vector<double> cost(nmodels);
    int64_t acc = 0;
    for (int i = 0; i < nmodels; ++i) {
        const ctl::Model& m = models[i]->m;
        KParams p = shared;
        if ((rc = upload_model(b, m, p))) return bail(rc);
        p.n = (int)replicas[i];
        p.x0_div = (m.x0_rows <= 1 || m.x0_rows >= p.n) ? 1 : (int)(p.n / m.x0_rows);
        p.c_div = (m.c_rows <= 1 || m.c_rows >= p.n) ? 1 : (int)(p.n / m.c_rows);
        p.fv_div = m.fv_rows <= 1 ? 0x7fffffff : (m.fv_rows >= p.n ? 1 : (int)(p.n / m.fv_rows));
        b->base[i] = acc;
        p.sim_base = (uint32_t)(o->offset + acc);
        acc += replicas[i];
        p.key0 = (uint32_t)o->seed;
        p.key1 = (uint32_t)(o->seed >> 32);
        p.nc = o->nc;
        p.eps = o->eps;
        p.ssa_steps = o->ssa_steps;
        p.ssa_factor = o->ssa_factor;
        p.force_ssa = o->force_ssa ? 1 : 0;
        p.ssa_chunk = o->ssa_chunk > 0 ? o->ssa_chunk : 10;
        p.iters = 0;
        p.ns = ns;
        p.X = X; p.A = A; p.Cs = Cs;
        if ((rc = balloc(b, &p.storage, (size_t)p.n * p.K * p.SS)) || (rc = balloc(b, &p.simstat, p.n)) ||
            (rc = balloc(b, &p.xfinal, (size_t)p.n * p.S)) || (rc = balloc(b, &p.tfinal, p.n)))
            return bail(rc);
        p.next_sim = nullptr;
        p.done = b->d_done;
        b->desc[i] = p;
        b->K[i] = p.K;
        b->SS[i] = p.SS;
        // per-replica cost estimate: the work of one leap attempt
        cost[i] = (double)(2 * m.R + 5 * m.nL + 5 * m.nV + m.nF + m.S);
    }
    if ((rc = bupload(b, &b->d_desc, b->desc.data(), b->desc.size()))) return bail(rc);
    std::vector<int> order(nmodels);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] * b->reps[a] > cost[c] * b->reps[c]; });
    b->stats.lanes_per_sim = G;
    if (G != 1) {                        // one global counter over the models in cost order
        std::vector<int64_t> cum(nmodels + 1, 0);
        for (int k = 0; k < nmodels; ++k) cum[k + 1] = cum[k] + b->reps[order[k]];
        GBatchParams& g = b->gp;
        int32_t* d_order;
        int64_t* d_cum;
        if ((rc = bupload(b, &d_order, order.data(), order.size())) || (rc = bupload(b, &d_cum, cum.data(), cum.size())))
            return bail(rc);
        g.desc = b->d_desc;
        g.order = d_order;
        g.cum = d_cum;
        g.nmodels = nmodels;
        b->stats.slots = b->ns;
        b->stats.block_size = 64;
        b->stats.lds_state = 1;
        *out = b;
        return 0;
    }
    std::vector<int32_t> chunks;
    for (int i : order)
        for (int64_t r = 0; r < b->reps[i]; r += 64) chunks.push_back(i);
    b->nchunks = (int)chunks.size();
    int32_t* d_chunks;
    uint32_t *d_next, *d_rep;
    int32_t* d_slot_model;
    if ((rc = bupload(b, &d_chunks, chunks.data(), chunks.size())) || (rc = balloc(b, &d_next, 1)) ||
        (rc = balloc(b, &d_rep, nmodels)) || (rc = balloc(b, &d_slot_model, ns)))
        return bail(rc);
    BatchParams& bp = b->bp;
    bp.desc = b->d_desc;
    bp.chunk_model = d_chunks;
    bp.nchunks = b->nchunks;
    bp.chunk_next = d_next;
    bp.rep_next = d_rep;
    bp.slot_model = d_slot_model;
    bp.ns = ns;
    bp.iters = o->iters_per_launch > 0 ? o->iters_per_launch : 4096;
    bp.X = X; bp.A = A; bp.Cs = Cs;
    b->stats.slots = ns;
    b->stats.block_size = b->block;
    *out = b;
    return 0;
}

int ctl_batch_run(ctl_batch* b) {
    ctl::clear_error();
    if (!b) return ctl::fail(CTL_E_ARGS, "null batch");
    BHIP(hipSetDevice(b->device), "hipSetDevice");
    const auto w0 = std::chrono::steady_clock::now();
    for (int i = 0; i < b->nmodels; ++i) {
        const KParams& p = b->desc[i];
        BHIP(hipMemsetAsync(p.storage, 0, sizeof(uint32_t) * (size_t)p.n * p.K * p.SS, b->stream), "memset");
    }
    if (b->lanes == 1) {
        BHIP(hipMemsetAsync(b->bp.slot_model, 0xff, sizeof(int32_t) * b->ns, b->stream), "memset");
        BHIP(hipMemsetAsync(b->bp.rep_next, 0, sizeof(uint32_t) * b->nmodels, b->stream), "memset");
        BHIP(hipMemsetAsync(b->bp.chunk_next, 0, sizeof(uint32_t), b->stream), "memset");
    } else {
        BHIP(hipMemsetAsync(b->gp.slot_sim, 0xff, sizeof(int32_t) * b->gp.ngroups, b->stream), "memset");
        BHIP(hipMemsetAsync(b->gp.slot_model, 0, sizeof(int32_t) * b->gp.ngroups, b->stream), "memset");
        BHIP(hipMemsetAsync(b->gp.next, 0, sizeof(uint32_t), b->stream), "memset");
    }
    BHIP(hipMemsetAsync(b->d_done, 0, sizeof(uint32_t), b->stream), "memset");
    ctl_stats& S = b->stats;
    S.launches = 0;
    float total_ms = 0.0f;
    uint32_t done = 0;
    while (done < (uint32_t)b->total) {
        BHIP(hipEventRecord(b->ev0, b->stream), "hipEventRecord");
        if (b->lanes == 1)
            hipLaunchKernelGGL(batch_kernel, dim3(b->grid), dim3(b->block), 0, b->stream, b->bp);
        else
            hipLaunchKernelGGL(gbatch_kernel, dim3(b->grid), dim3(64), b->smem, b->stream, b->gp);
        BHIP(hipGetLastError(), "batch kernel launch");
        BHIP(hipEventRecord(b->ev1, b->stream), "hipEventRecord");
        BHIP(hipMemcpyAsync(&done, b->d_done, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream), "hipMemcpy D2H");
        BHIP(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        float ms = 0.0f;
        BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1), "hipEventElapsedTime");
        total_ms += ms;
        if (++S.launches >= (1 << 20)) return ctl::fail(CTL_E_STATE, "batch did not finish");
        if (b->verbose)
            std::fprintf(stderr, " * launch %d: %u/%lld simulations done (%.3f ms)\n", S.launches, done,
                         (long long)b->total, ms);
    }
    S.kernel_ms = total_ms;
    S.firings = S.leaps = S.rejections = S.switches = S.ssa_steps = S.poisson = S.gaps = S.absorbed = 0;
    for (int i = 0; i < b->nmodels; ++i) {
        std::vector<ctl_sim_stat> st(b->desc[i].n);
        BHIP(hipMemcpy(st.data(), b->desc[i].simstat, sizeof(ctl_sim_stat) * st.size(), hipMemcpyDeviceToHost), "D2H");
        for (const auto& x : st) {
            S.firings += x.firings; S.leaps += x.leaps; S.rejections += x.rejections; S.switches += x.switches;
            S.ssa_steps += x.ssa_steps; S.poisson += x.poisson; S.gaps += x.gaps; S.absorbed += x.absorbed;
        }
    }
    S.simulations = (uint64_t)b->total;
    S.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    b->ran = true;
    return 0;
}

int ctl_batch_samples(const ctl_batch* b, int32_t model, uint32_t* out, size_t count) {
    ctl::clear_error();
    if (!b || !out || model < 0 || model >= b->nmodels) return ctl::fail(CTL_E_ARGS, "bad arguments");
    if (!b->ran) return ctl::fail(CTL_E_STATE, "ctl_batch_run has not completed");
    const KParams& p = b->desc[model];
    if (count != (size_t)p.n * p.K * p.SS) return ctl::fail(CTL_E_ARGS, "count must be replicas*K*SS");
    BHIP(hipSetDevice(b->device), "hipSetDevice");
    BHIP(hipMemcpy(out, p.storage, count * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return 0;
}

int ctl_batch_sim_stats(const ctl_batch* b, int32_t model, ctl_sim_stat* out, size_t count) {
    ctl::clear_error();
    if (!b || !out || model < 0 || model >= b->nmodels) return ctl::fail(CTL_E_ARGS, "bad arguments");
    if (!b->ran) return ctl::fail(CTL_E_STATE, "ctl_batch_run has not completed");
    const KParams& p = b->desc[model];
    if (count != (size_t)p.n) return ctl::fail(CTL_E_ARGS, "count must be the model's replica count");
    BHIP(hipSetDevice(b->device), "hipSetDevice");
    BHIP(hipMemcpy(out, p.simstat, count * sizeof(ctl_sim_stat), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return 0;
}

int ctl_batch_stats(const ctl_batch* b, ctl_stats* out) {
    if (!b || !out) return ctl::fail(CTL_E_ARGS, "null argument");
    *out = b->stats;
    return 0;
}

int ctl_batch_stream_base(const ctl_batch* b, int32_t model, int64_t* base) {
    if (!b || !base || model < 0 || model >= b->nmodels) return ctl::fail(CTL_E_ARGS, "bad arguments");
    *base = b->base[model];
    return 0;
}

void ctl_batch_destroy(ctl_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    brelease(b);
    delete b;
}

}  // extern "C"
