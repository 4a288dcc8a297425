// ctl_batch.hip -- many independent models of uneven size in one persistent
// launch sequence (no counterpart in the reference, which runs one model per
// process).
//
// Each model runs with its own number of lanes per simulation G (chosen per
// model, auto_lanes), and the models of one G form a class with its own
// persistent kernel on its own stream; the classes run concurrently:
//   G = 1      batch_kernel (below): one simulation per lane;
//   G = 2..32  gclass_kernel<G> (ctl_gbatch.hip): 64/G groups of a wavefront
//              run replicas of one model, the simulations' state in LDS;
//   G = 64     gbatch_kernel (ctl_gbatch.hip): one wavefront per simulation.
// Inside a class the work is ordered by an estimated cost, largest first.
//
// batch_kernel: work is split into chunks of 64 replicas of one model, ordered by an
// estimated per-replica cost, largest first.  A wavefront takes the next
// chunk's model as its current model -- so the model tables stay
// wave-uniform scalar loads -- and each of its lanes then steals replicas of
// that model from a per-model counter until the model is exhausted; when all
// its lanes are idle the wavefront takes the next chunk.  Several wavefronts
// can work on one model, and a lane never waits for a whole chunk to finish.
// Per-lane state lives in the global slot arrays, sized for the largest
// model of the class.  A replica's Philox stream index is the model's base
// (sum of the replicas of the models before it, plus the offset) plus its
// replica index, whatever its class.
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

// The models of a batch that run with the same lanes per simulation G form
// a class: one persistent kernel on its own stream (G = 1: batch_kernel,
// G = 2..32: gclass_kernel<G>, G = 64: gbatch_kernel).  The classes of a
// batch run concurrently.
struct BClass {
    int G = 1;
    std::vector<int> models;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int block = 64, grid = 0, ns = 0;        // ns: resident lanes (G = 1) or groups (G > 1)
    size_t smem = 0;
    int64_t total = 0;                        // replicas of the class's models
    uint32_t* d_done = nullptr;
    BatchParams bp{};
    GBatchParams gp{};
    GClassParams cp{};
    const SlotState* st = nullptr;           // G = 1: device copy of the slot arrays' pointers
    bool finished = false;
};

struct ctl_batch {
    int device = 0;
    int nmodels = 0;
    int64_t total = 0;
    std::vector<KParams> desc;                 // host copies (device pointers inside)
    std::vector<int64_t> reps, base;
    std::vector<int> K, SS, lanes;
    KParams* d_desc = nullptr;
    uint32_t* d_rep_next = nullptr;            // per-model replica counters
    std::vector<BClass> classes;
    int verbose = 0;
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
    for (BClass& c : b->classes) {
        if (c.ev0) (void)hipEventDestroy(c.ev0);
        if (c.ev1) (void)hipEventDestroy(c.ev1);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        c.ev0 = c.ev1 = nullptr;
        c.stream = nullptr;
    }
}

// Lanes per simulation the library picks for a model of a batch: one lane
// when the per-lane working set fits a wavefront's 16 KB share of LDS;
// otherwise about 16 reactions per lane (G = 4 .. 64), widened until a
// wavefront's groups need at most 32 KB of LDS.
int auto_lanes(const ctl::Model& m) {
    const size_t lane_words = (size_t)2 * m.S + m.R + (m.c_rows == 1 ? 0 : m.R);
    if (lane_words * 4 * 64 <= 16384) return 1;
    int G = 4;
    while (G < 64 && G * 16 < m.R) G *= 2;
    while (G < 64 && (size_t)(64 / G) * ctl::group_lds_words(m.S, m.R, G) * 4 > 32768) G *= 2;
    return G;
}

// Per-model device tables (the same compile products as a ctl_sim).
int upload_model(ctl_batch* b, const ctl::Model& m, KParams& p, int G) {
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
        if (G > 1) append_group_streams(T, p, G, m, FSptr, FSent, Dent);
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
    return ctl_batch_create_lanes(models, replicas, nullptr, nmodels, o, out);
}

int ctl_batch_create_lanes(const ctl_model* const* models, const int64_t* replicas, const int32_t* lanes_in,
                           int32_t nmodels, const ctl_opts* o, ctl_batch** out) {
    ctl::clear_error();
    if (!models || !replicas || !o || !out || nmodels <= 0) return ctl::fail(CTL_E_ARGS, "bad arguments");
    *out = nullptr;
    int64_t total = 0;
    std::vector<int> lanes(nmodels);
    for (int i = 0; i < nmodels; ++i) {
        if (!models[i] || replicas[i] <= 0) return ctl::fail(CTL_E_ARGS, "null model or non-positive replica count");
        const ctl::Model& m = models[i]->m;
        if ((!m.fs.empty() && m.fv_rows > 1 && m.fv_rows < replicas[i] && replicas[i] % m.fv_rows) ||
            (m.x0_rows > 1 && m.x0_rows < replicas[i] && replicas[i] % m.x0_rows) ||
            (m.c_rows > 1 && m.c_rows < replicas[i] && replicas[i] % m.c_rows))
            return ctl::fail(CTL_E_ROWS, "a model's input rows must be 1, at least its replicas, or divide them");
        total += replicas[i];
        const int G = lanes_in && lanes_in[i] > 0 ? lanes_in[i] : o->lanes_per_sim > 0 ? o->lanes_per_sim : auto_lanes(m);
        if (G < 1 || G > 64 || (G & (G - 1)))
            return ctl::fail(CTL_E_ARGS, "lanes_per_sim must be 0 or a power of two up to 64");
        if (G > 1 && (size_t)(64 / G) * ctl::group_lds_words(m.S, m.R, G) * 4 > 65536)
            return ctl::fail(CTL_E_LIMIT, G == 64 ? "a model is too large for the group kernel (more than 64 KB of LDS)"
                                                  : "a model is too large for this lanes_per_sim (more than 64 KB of LDS per wavefront)");
        lanes[i] = G;
    }
    if (total > 0x7fffffffLL) return ctl::fail(CTL_E_LIMIT, "too many simulations");
    if (o->offset < 0 || o->offset + total > 0x100000000LL)    // 32-bit Philox stream indices must not wrap
        return ctl::fail(CTL_E_ARGS, "offset must be >= 0 and offset + total replicas <= 2^32");
    if (o->block_size != 0 && (o->block_size % 64 != 0 || o->block_size < 0 || o->block_size > 256))
        return ctl::fail(CTL_E_ARGS, "block_size must be 64..256, multiple of 64");
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
    b->lanes = lanes;
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return bail(bhip(e, "hipSetDevice"));
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, b->device)) != hipSuccess) return bail(bhip(e, "hipGetDeviceProperties"));
    int rc;
    // classes, ascending G
    for (int G = 1; G <= 64; G *= 2) {
        BClass c;
        c.G = G;
        for (int i = 0; i < nmodels; ++i)
            if (lanes[i] == G) { c.models.push_back(i); c.total += replicas[i]; }
        if (!c.models.empty()) b->classes.push_back(c);
    }
    if ((rc = balloc(b, &b->d_rep_next, nmodels))) return bail(rc);
    std::vector<int> cls(nmodels, -1);
    for (size_t k = 0; k < b->classes.size(); ++k) {
        BClass& c = b->classes[k];
        for (int i : c.models) cls[i] = (int)k;
        if ((e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking)) != hipSuccess)
            return bail(bhip(e, "hipStreamCreate"));
        if ((e = hipEventCreate(&c.ev0)) != hipSuccess || (e = hipEventCreate(&c.ev1)) != hipSuccess)
            return bail(bhip(e, "hipEventCreate"));
        if ((rc = balloc(b, &c.d_done, 1))) return bail(rc);
        int maxS = 0, maxR = 0;
        size_t gwords = 0;
        bool any_rates = false;
        for (int i : c.models) {
            const ctl::Model& m = models[i]->m;
            maxS = std::max(maxS, m.S);
            maxR = std::max(maxR, m.R);
            any_rates = any_rates || m.c_rows != 1;
            if (c.G > 1) gwords = std::max(gwords, ctl::group_lds_words(m.S, m.R, c.G));
        }
        int per_cu = 0;
        if (c.G == 1) {
            c.block = o->block_size > 0 ? o->block_size : 256;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)batch_kernel, c.block, 0);
            if (e != hipSuccess || per_cu <= 0) per_cu = 1;
            int ns = o->slots > 0 ? o->slots : prop.multiProcessorCount * per_cu * c.block;
            ns = ((ns + c.block - 1) / c.block) * c.block;
            c.ns = ns;
            c.grid = ns / c.block;
            BatchParams& bp = c.bp;
            SlotState st{};
            if ((rc = balloc(b, &bp.X, (size_t)2 * maxS * ns)) || (rc = balloc(b, &bp.A, (size_t)maxR * ns)) ||
                (any_rates && (rc = balloc(b, &bp.Cs, (size_t)maxR * ns))))
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
            c.st = dst;
            bp.ns = ns;
            bp.iters = o->iters_per_launch > 0 ? o->iters_per_launch : 4096;
            bp.rep_next = b->d_rep_next;
            if ((rc = balloc(b, &bp.chunk_next, 1)) || (rc = balloc(b, &bp.slot_model, ns))) return bail(rc);
        } else if (c.G == 64) {
            c.block = 64;
            c.smem = gwords * 4;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)gbatch_kernel, 64, c.smem);
            if (e != hipSuccess || per_cu <= 0) per_cu = 1;
            long long groups = o->slots > 0 ? (long long)o->slots : (long long)prop.multiProcessorCount * per_cu;
            if (groups > c.total) groups = c.total;
            c.ns = (int)groups;
            c.grid = (int)groups;
            GBatchParams& g = c.gp;
            g.ngroups = (int)groups;
            g.maxS = maxS;
            g.total = c.total;
            if ((rc = balloc(b, &g.slot_model, groups)) || (rc = balloc(b, &g.slot_sim, groups)) ||
                (rc = balloc(b, &g.t, groups)) || (rc = balloc(b, &g.cursor, groups)) || (rc = balloc(b, &g.evc, groups)) ||
                (rc = balloc(b, &g.firings, groups)) || (rc = balloc(b, &g.ssa_steps, groups)) ||
                (rc = balloc(b, &g.leaps, groups)) || (rc = balloc(b, &g.rejections, groups)) ||
                (rc = balloc(b, &g.switches, groups)) || (rc = balloc(b, &g.poisson, (size_t)groups * 64)) ||
                (rc = balloc(b, &g.gaps, (size_t)groups * 64)) || (rc = balloc(b, &g.X, (size_t)groups * maxS)) ||
                (rc = balloc(b, &g.next, 1)))
                return bail(rc);
            g.iters = o->iters_per_launch > 0 ? o->iters_per_launch : 65536;
        } else {
            const int gpw = 64 / c.G;                       // groups per wavefront (= workgroup)
            c.block = 64;
            c.smem = (size_t)gpw * gwords * 4;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ctl::gclass_kernel_fn(c.G), 64, c.smem);
            if (e != hipSuccess || per_cu <= 0) per_cu = 1;
            long long waves = o->slots > 0 ? ((long long)o->slots + gpw - 1) / gpw
                                           : (long long)prop.multiProcessorCount * per_cu;
            const long long need = (c.total + gpw - 1) / gpw;
            if (waves > need) waves = need;
            c.grid = (int)waves;
            c.ns = (int)(waves * gpw);
            GClassParams& g = c.cp;
            g.nwaves = (int)waves;
            g.maxS = maxS;
            g.gwords = (int)gwords;
            const size_t ng = (size_t)c.ns;
            if ((rc = balloc(b, &g.wave_model, waves)) || (rc = balloc(b, &g.slot_sim, ng)) ||
                (rc = balloc(b, &g.t, ng)) || (rc = balloc(b, &g.cursor, ng)) || (rc = balloc(b, &g.evc, ng)) ||
                (rc = balloc(b, &g.firings, ng)) || (rc = balloc(b, &g.ssa_steps, ng)) ||
                (rc = balloc(b, &g.leaps, ng)) || (rc = balloc(b, &g.rejections, ng)) ||
                (rc = balloc(b, &g.switches, ng)) || (rc = balloc(b, &g.poisson, ng * c.G)) ||
                (rc = balloc(b, &g.gaps, ng * c.G)) || (rc = balloc(b, &g.X, ng * maxS)) ||
                (rc = balloc(b, &g.chunk_next, 1)))
                return bail(rc);
            g.iters = o->iters_per_launch > 0 ? o->iters_per_launch : 65536;
        }
    }
    b->desc.resize(nmodels);
    b->reps.assign(replicas, replicas + nmodels);
    b->base.resize(nmodels);
    b->K.resize(nmodels);
    b->SS.resize(nmodels);
    std::vector<double> cost(nmodels);
    int64_t acc = 0;
    for (int i = 0; i < nmodels; ++i) {
        const ctl::Model& m = models[i]->m;
        const BClass& c = b->classes[cls[i]];
        KParams p{};
        if ((rc = upload_model(b, m, p, lanes[i]))) return bail(rc);
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
        if (c.G == 1) {
            p.ns = c.ns;
            p.X = c.bp.X; p.A = c.bp.A; p.Cs = c.bp.Cs;
            p.st = c.st;
        }
        if ((rc = balloc(b, &p.storage, (size_t)p.n * p.K * p.SS)) || (rc = balloc(b, &p.simstat, p.n)) ||
            (rc = balloc(b, &p.xfinal, (size_t)p.n * p.S)) || (rc = balloc(b, &p.tfinal, p.n)))
            return bail(rc);
        p.next_sim = b->d_rep_next + i;       // the group classes pull replicas through gfetch
        p.done = c.d_done;
        b->desc[i] = p;
        b->K[i] = p.K;
        b->SS[i] = p.SS;
        // per-replica cost estimate: the work of one leap attempt
        cost[i] = (double)(2 * m.R + 5 * m.nL + 5 * m.nV + m.nF + m.S);
    }
    if ((rc = bupload(b, &b->d_desc, b->desc.data(), b->desc.size()))) return bail(rc);
    int maxG = 0;
    bool uniform = true;
    for (BClass& c : b->classes) {
        std::vector<int> order = c.models;     // largest work first
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return cost[x] * b->reps[x] > cost[y] * b->reps[y]; });
        if (c.G == 64) {                       // one counter over the class's models in cost order
            std::vector<int64_t> cum(order.size() + 1, 0);
            for (size_t k = 0; k < order.size(); ++k) cum[k + 1] = cum[k] + b->reps[order[k]];
            GBatchParams& g = c.gp;
            int32_t* d_order;
            int64_t* d_cum;
            if ((rc = bupload(b, &d_order, order.data(), order.size())) || (rc = bupload(b, &d_cum, cum.data(), cum.size())))
                return bail(rc);
            g.desc = b->d_desc;
            g.order = d_order;
            g.cum = d_cum;
            g.nmodels = (int)order.size();
        } else {                               // chunks of one wavefront's worth of replicas
            const int per = c.G == 1 ? 64 : 64 / c.G;
            std::vector<int32_t> chunks;
            for (int i : order)
                for (int64_t r = 0; r < b->reps[i]; r += per) chunks.push_back(i);
            int32_t* d_chunks;
            if ((rc = bupload(b, &d_chunks, chunks.data(), chunks.size()))) return bail(rc);
            if (c.G == 1) {
                c.bp.desc = b->d_desc;
                c.bp.chunk_model = d_chunks;
                c.bp.nchunks = (int)chunks.size();
            } else {
                c.cp.desc = b->d_desc;
                c.cp.chunk_model = d_chunks;
                c.cp.nchunks = (int)chunks.size();
            }
        }
        if (maxG && c.G != maxG) uniform = false;
        maxG = std::max(maxG, c.G);
    }
    b->stats.lanes_per_sim = uniform ? maxG : 0;
    b->stats.slots = 0;
    for (BClass& c : b->classes) b->stats.slots += c.ns;
    b->stats.block_size = b->classes.size() == 1 ? b->classes[0].block : 0;
    b->stats.lds_state = b->classes.size() == 1 && b->classes[0].G == 1 ? 0 : 1;
    *out = b;
    return 0;
}

int ctl_batch_run(ctl_batch* b) {
    ctl::clear_error();
    if (!b) return ctl::fail(CTL_E_ARGS, "null batch");
    BHIP(hipSetDevice(b->device), "hipSetDevice");
    const auto w0 = std::chrono::steady_clock::now();
    for (BClass& c : b->classes) {
        for (int i : c.models) {
            const KParams& p = b->desc[i];
            BHIP(hipMemsetAsync(p.storage, 0, sizeof(uint32_t) * (size_t)p.n * p.K * p.SS, c.stream), "memset");
        }
        if (c.G == 1) {
            BHIP(hipMemsetAsync(c.bp.slot_model, 0xff, sizeof(int32_t) * c.ns, c.stream), "memset");
            BHIP(hipMemsetAsync(c.bp.chunk_next, 0, sizeof(uint32_t), c.stream), "memset");
        } else if (c.G == 64) {
            BHIP(hipMemsetAsync(c.gp.slot_sim, 0xff, sizeof(int32_t) * c.gp.ngroups, c.stream), "memset");
            BHIP(hipMemsetAsync(c.gp.slot_model, 0, sizeof(int32_t) * c.gp.ngroups, c.stream), "memset");
            BHIP(hipMemsetAsync(c.gp.next, 0, sizeof(uint32_t), c.stream), "memset");
        } else {
            BHIP(hipMemsetAsync(c.cp.wave_model, 0xff, sizeof(int32_t) * c.cp.nwaves, c.stream), "memset");
            BHIP(hipMemsetAsync(c.cp.slot_sim, 0xff, sizeof(int32_t) * c.ns, c.stream), "memset");
            BHIP(hipMemsetAsync(c.cp.chunk_next, 0, sizeof(uint32_t), c.stream), "memset");
        }
        BHIP(hipMemsetAsync(c.d_done, 0, sizeof(uint32_t), c.stream), "memset");
        c.finished = false;
    }
    // replica counters of every model (before any class starts)
    BHIP(hipMemset(b->d_rep_next, 0, sizeof(uint32_t) * b->nmodels), "memset");
    ctl_stats& S = b->stats;
    S.launches = 0;
    // the classes run concurrently: each round launches every unfinished
    // class on its own stream, then waits for all of them
    const auto k0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> done(b->classes.size(), 0);
    for (;;) {
        bool any = false;
        for (size_t k = 0; k < b->classes.size(); ++k) {
            BClass& c = b->classes[k];
            if (c.finished) continue;
            any = true;
            BHIP(hipEventRecord(c.ev0, c.stream), "hipEventRecord");
            if (c.G == 1)
                hipLaunchKernelGGL(batch_kernel, dim3(c.grid), dim3(c.block), 0, c.stream, c.bp);
            else if (c.G == 64)
                hipLaunchKernelGGL(gbatch_kernel, dim3(c.grid), dim3(64), c.smem, c.stream, c.gp);
            else
                switch (c.G) {
                    case 2: hipLaunchKernelGGL(gclass_kernel<2>, dim3(c.grid), dim3(64), c.smem, c.stream, c.cp); break;
                    case 4: hipLaunchKernelGGL(gclass_kernel<4>, dim3(c.grid), dim3(64), c.smem, c.stream, c.cp); break;
                    case 8: hipLaunchKernelGGL(gclass_kernel<8>, dim3(c.grid), dim3(64), c.smem, c.stream, c.cp); break;
                    case 16: hipLaunchKernelGGL(gclass_kernel<16>, dim3(c.grid), dim3(64), c.smem, c.stream, c.cp); break;
                    default: hipLaunchKernelGGL(gclass_kernel<32>, dim3(c.grid), dim3(64), c.smem, c.stream, c.cp); break;
                }
            BHIP(hipGetLastError(), "batch kernel launch");
            BHIP(hipEventRecord(c.ev1, c.stream), "hipEventRecord");
            BHIP(hipMemcpyAsync(&done[k], c.d_done, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream), "hipMemcpy D2H");
            ++S.launches;
        }
        if (!any) break;
        for (size_t k = 0; k < b->classes.size(); ++k) {
            BClass& c = b->classes[k];
            if (c.finished) continue;
            BHIP(hipStreamSynchronize(c.stream), "hipStreamSynchronize");
            if (done[k] >= (uint32_t)c.total) c.finished = true;
            if (b->verbose) {
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, c.ev0, c.ev1);
                std::fprintf(stderr, " * class G=%d launch: %u/%lld simulations done (%.3f ms)\n", c.G, done[k],
                             (long long)c.total, ms);
            }
        }
        if (S.launches >= (1 << 20)) return ctl::fail(CTL_E_STATE, "batch did not finish");
    }
    S.kernel_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - k0).count();
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

int ctl_batch_model_lanes(const ctl_batch* b, int32_t model, int32_t* lanes) {
    if (!b || !lanes || model < 0 || model >= b->nmodels) return ctl::fail(CTL_E_ARGS, "bad arguments");
    *lanes = b->lanes[model];
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
