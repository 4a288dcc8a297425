// ctl_gbatch.hip -- gbatch_kernel: the group algorithm with one 64-lane
// wavefront per simulation for batches of many models (host runtime in
// ctl_batch.hip).  A wavefront's model index is read through
// readfirstlane, so its descriptor is uniform: this translation unit keeps
// the uniform-parameter treatment of the group code (CTL_UNIFORM_PARAMS).
#include <hip/hip_runtime.h>

#include "ctl_gbatch.h"
#include "ctl_group.h"

#pragma clang fp contract(off)

namespace ctl {



__global__ __launch_bounds__(64) void gbatch_kernel(GBatchParams bp) {
    constexpr int G = 64;
    extern __shared__ uint32_t smem[];
    __shared__ float recip[64];
    recip[threadIdx.x] = c_recip[threadIdx.x];
    __syncthreads();
    const int l = (int)threadIdx.x;
    const int slot = (int)blockIdx.x;
    if (slot >= bp.ngroups) return;
    GLane L;
    L.tacc = 0.0;
    L.ssa_left = 0;
    L.ssa_mode = 0;
    L.slot = slot;
    L.lg.tag = 1u + (uint32_t)l;
    int m = __builtin_amdgcn_readfirstlane(bp.slot_model[slot]);
    L.sim = __builtin_amdgcn_readfirstlane(bp.slot_sim[slot]);
    if (L.sim == -2) return;
    KParams p;
    GView V;
    if (L.sim >= 0) {                    // resume a parked simulation
        p = desc_of_b(bp.desc, m);
        V = group_view(smem, p.S, p.R, G);
        gzero_pad<G>(V, p.R, l);
        L.t = bp.t[slot];
        L.cursor = bp.cursor[slot];
        L.evc = bp.evc[slot];
        L.xb = 0;
        L.firings = bp.firings[slot];
        L.ssa_steps = bp.ssa_steps[slot];
        L.leaps = bp.leaps[slot];
        L.rejections = bp.rejections[slot];
        L.switches = bp.switches[slot];
        L.absorbed = 0;
        L.poisson = bp.poisson[(size_t)slot * G + l];
        L.gaps = bp.gaps[(size_t)slot * G + l];
        L.gg.k0 = L.lg.k0 = p.key0;
        L.gg.k1 = L.lg.k1 = p.key1;
        L.gg.sim = L.lg.sim = p.sim_base + (uint32_t)L.sim;
        L.gg.cb = GRNG_NONE;
        L.lg.pos = 4;
        for (int s = l; s < p.S; s += G) V.Xb[s] = bp.X[(size_t)slot * bp.maxS + s];
        L.crate = grates(p, L.sim);
        L.fvrow = p.nfs > 0 ? p.fv + (size_t)(L.sim / p.fv_div) * p.nfs : nullptr;
        gsync();
    }
    for (int it = 0; it < bp.iters; ++it) {
        if (L.sim < 0) {                 // next simulation of any model
            uint32_t q = 0;
            if (l == 0) q = atomicAdd(bp.next, 1u);
            q = __builtin_amdgcn_readfirstlane(q);
            if ((int64_t)q >= bp.total) { L.sim = -2; break; }
            int lo = 0, hi = bp.nmodels - 1;           // last k with cum[k] <= q
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if ((int64_t)q >= CT_I64(bp.cum)[mid]) lo = mid; else hi = mid - 1;
            }
            lo = __builtin_amdgcn_readfirstlane(lo);
            m = __builtin_amdgcn_readfirstlane(CT_I32(bp.order)[lo]);
            const uint32_t r = (uint32_t)((int64_t)q - CT_I64(bp.cum)[lo]);
            p = desc_of_b(bp.desc, m);
            V = group_view(smem, p.S, p.R, G);
            gzero_pad<G>(V, p.R, l);
            L.sim = (int)r;
            L.t = 0.0;
            L.cursor = 0;
            L.evc = 0;
            L.xb = 0;
            L.firings = L.ssa_steps = 0;
            L.leaps = L.rejections = L.switches = L.absorbed = 0;
            L.poisson = L.gaps = 0;
            L.gg.k0 = L.lg.k0 = p.key0;
            L.gg.k1 = L.lg.k1 = p.key1;
            L.gg.sim = L.lg.sim = p.sim_base + r;
            L.gg.cb = GRNG_NONE;
            L.lg.pos = 4;
            for (int s = l; s < p.S; s += G)
                V.Xb[s] = (int32_t)(p.x0_bcast ? p.x0_bcast[s] : p.x0_sims[(size_t)(r / (uint32_t)p.x0_div) * p.S + s]);
            L.crate = grates(p, (int)r);
            L.fvrow = p.nfs > 0 ? p.fv + (size_t)(r / (uint32_t)p.fv_div) * p.nfs : nullptr;
            gsync();
        }
        if (p.force_ssa) {
            if (L.t >= p.tmax || gssa<G>(p, L, V, l)) gfinish<G>(p, L, V, l);
            continue;
        }
        const int r = gleap<G>(p, L, V, l, recip);
        if (r == LEAP_DONE || r == LEAP_ABSORB) {
            gfinish<G>(p, L, V, l);
        } else if (r == LEAP_SWITCH) {
            if (gssa<G>(p, L, V, l)) gfinish<G>(p, L, V, l);
        }
    }
    if (L.sim >= 0) {                    // park
        const int32_t* X = V.Xb + (L.xb ? p.S : 0);
        for (int s = l; s < p.S; s += G) bp.X[(size_t)slot * bp.maxS + s] = X[s];
        bp.poisson[(size_t)slot * G + l] = L.poisson;
        bp.gaps[(size_t)slot * G + l] = L.gaps;
        if (l == 0) {
            bp.t[slot] = L.t;
            bp.cursor[slot] = L.cursor;
            bp.evc[slot] = L.evc;
            bp.firings[slot] = L.firings;
            bp.ssa_steps[slot] = L.ssa_steps;
            bp.leaps[slot] = L.leaps;
            bp.rejections[slot] = L.rejections;
            bp.switches[slot] = L.switches;
        }
    }
    if (l == 0) {
        bp.slot_model[slot] = m;
        bp.slot_sim[slot] = L.sim;
    }
}

// Group batch, G = 2..32 (see GClassParams).  The loop mirrors group_kernel:
// one leap attempt per group per iteration, SSA bursts held until a second
// group of the wavefront switches (at most GBURST_WAIT iterations, or until
// every group with a replica waits); a group without a replica idles.  Each
// replica's trajectory depends only on its own state and streams.
template <int G>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(G == 8 || G == 16 || G == 32 ? 4 : 1)))
void gclass_kernel(GClassParams bp) {
    extern __shared__ uint32_t smem[];
    __shared__ float recip[64];
    recip[threadIdx.x] = c_recip[threadIdx.x];
    __syncthreads();
    const int l = (int)(threadIdx.x & (unsigned)(G - 1));
    const int gib = (int)(threadIdx.x / G);
    const int wave = (int)blockIdx.x;
    if (wave >= bp.nwaves) return;
    uint32_t* gbase = smem + (size_t)gib * (size_t)bp.gwords;
    GLane L;
    L.tacc = 0.0;
    L.ssa_left = 0;
    L.ssa_mode = 0;
    L.slot = wave * (64 / G) + gib;
    L.lg.tag = 1u + (uint32_t)l;
    int m = __builtin_amdgcn_readfirstlane(bp.wave_model[wave]);
    L.sim = m >= 0 ? bp.slot_sim[L.slot] : -1;
    if (m >= 0) {
        const KParams p = desc_of_b(bp.desc, m);
        const GView V = group_view(gbase, p.S, p.R, G);
        gzero_pad<G>(V, p.R, l);
        L.gg.k0 = L.lg.k0 = p.key0;
        L.gg.k1 = L.lg.k1 = p.key1;
        if (L.sim >= 0) {                // resume a parked replica
            const int s0 = L.slot;
            L.t = bp.t[s0];
            L.cursor = bp.cursor[s0];
            L.evc = bp.evc[s0];
            L.xb = 0;
            L.firings = bp.firings[s0];
            L.ssa_steps = bp.ssa_steps[s0];
            L.leaps = bp.leaps[s0];
            L.rejections = bp.rejections[s0];
            L.switches = bp.switches[s0];
            L.absorbed = 0;
            L.poisson = bp.poisson[(size_t)s0 * G + l];
            L.gaps = bp.gaps[(size_t)s0 * G + l];
            L.gg.sim = L.lg.sim = p.sim_base + (uint32_t)L.sim;
            L.gg.cb = GRNG_NONE;
            L.lg.pos = 4;
            for (int s = l; s < p.S; s += G) V.Xb[s] = bp.X[(size_t)s0 * bp.maxS + s];
            L.crate = grates(p, L.sim);
            L.fvrow = p.nfs > 0 ? p.fv + (size_t)(L.sim / p.fv_div) * p.nfs : nullptr;
        }
        gsync();
    }
    bool pend = false;
    int waited = 0;
    for (int it = 0; it < bp.iters; ++it) {
        if (m < 0) {                     // the next chunk's model
            uint32_t c = 0;
            if (threadIdx.x == 0) c = atomicAdd(bp.chunk_next, 1u);
            c = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
            if (c >= (uint32_t)bp.nchunks) break;
            m = __builtin_amdgcn_readfirstlane(CT_I32(bp.chunk_model)[c]);
            const KParams p = desc_of_b(bp.desc, m);
            gzero_pad<G>(group_view(gbase, p.S, p.R, G), p.R, l);
            L.sim = -1;
        }
        const KParams p = desc_of_b(bp.desc, m);
        const GView V = group_view(gbase, p.S, p.R, G);
        if (L.sim == -1) {               // a replica of the model (-2 once it has none left)
            L.gg.k0 = L.lg.k0 = p.key0;
            L.gg.k1 = L.lg.k1 = p.key1;
            gfetch<G>(p, L, V, l);
        }
        if (__ballot(L.sim >= 0) == 0) { m = -1; continue; }
        const bool active = L.sim >= 0;
        if (active) {
            if (kuse(p.force_ssa)) {
                if (L.t >= p.tmax || gssa<G>(p, L, V, l)) gfinish<G>(p, L, V, l);
            } else if (!pend) {
                const int r = gleap<G>(p, L, V, l, recip);
                if (r == LEAP_DONE || r == LEAP_ABSORB) {
                    gfinish<G>(p, L, V, l);
                } else if (r == LEAP_SWITCH) {
                    pend = true;
                    waited = 0;
                }
            } else {
                ++waited;
            }
        }
        const uint64_t lead = __ballot(l == 0 && active);       // groups holding a replica
        const uint64_t pm = __ballot(pend && l == 0);           // ... waiting for a burst
        const bool go = pm != 0 && (__popcll(pm) >= 2 || pm == lead || __ballot(pend && waited >= GBURST_WAIT) != 0);
        if (go && pend) {
            pend = false;
            if (gssa<G>(p, L, V, l)) gfinish<G>(p, L, V, l);
        }
    }
    if (m >= 0) {
        const KParams p = desc_of_b(bp.desc, m);
        const GView V = group_view(gbase, p.S, p.R, G);
        if (pend && gssa<G>(p, L, V, l)) gfinish<G>(p, L, V, l);   // before parking
        if (L.sim >= 0) {                // park
            const int s0 = L.slot;
            const int32_t* X = V.Xb + (L.xb ? p.S : 0);
            for (int s = l; s < p.S; s += G) bp.X[(size_t)s0 * bp.maxS + s] = X[s];
            bp.poisson[(size_t)s0 * G + l] = L.poisson;
            bp.gaps[(size_t)s0 * G + l] = L.gaps;
            if (l == 0) {
                bp.t[s0] = L.t;
                bp.cursor[s0] = L.cursor;
                bp.evc[s0] = L.evc;
                bp.firings[s0] = L.firings;
                bp.ssa_steps[s0] = L.ssa_steps;
                bp.leaps[s0] = L.leaps;
                bp.rejections[s0] = L.rejections;
                bp.switches[s0] = L.switches;
            }
        }
    }
    if (l == 0) bp.slot_sim[L.slot] = L.sim;
    if (threadIdx.x == 0) bp.wave_model[wave] = m;
}

template __global__ void gclass_kernel<2>(GClassParams);
template __global__ void gclass_kernel<4>(GClassParams);
template __global__ void gclass_kernel<8>(GClassParams);
template __global__ void gclass_kernel<16>(GClassParams);
template __global__ void gclass_kernel<32>(GClassParams);

const void* gclass_kernel_fn(int G) {
    switch (G) {
        case 2: return (const void*)gclass_kernel<2>;
        case 4: return (const void*)gclass_kernel<4>;
        case 8: return (const void*)gclass_kernel<8>;
        case 16: return (const void*)gclass_kernel<16>;
        case 32: return (const void*)gclass_kernel<32>;
    }
    return nullptr;
}

}  // namespace ctl
