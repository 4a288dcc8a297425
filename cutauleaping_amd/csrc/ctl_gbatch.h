// ctl_gbatch.h -- the group batch kernel's parameters and declaration (the
// kernel is defined in ctl_gbatch.hip, the host runtime in ctl_batch.hip).
#pragma once

#include "ctl_kernels.h"

namespace ctl {

// Group batch: one simulation per 64-lane wavefront (group algorithm, G = 64).
struct GBatchParams {
    const KParams* desc;        // per-model descriptors (device)
    const int32_t* order;       // models in cost order
    const int64_t* cum;         // cumulative replicas in that order [nmodels + 1]
    int nmodels;
    int64_t total;
    uint32_t* next;             // global simulation counter
    int ngroups;                // resident group slots (one wavefront each)
    int iters;
    int maxS;
    // parked state per group slot
    int32_t* slot_model;
    int32_t* slot_sim;          // replica index, -1 none, -2 queue exhausted
    double* t;
    uint32_t* cursor;
    uint32_t* evc;
    uint64_t* firings;
    uint64_t* ssa_steps;
    uint32_t* leaps;
    uint32_t* rejections;
    uint32_t* switches;
    uint64_t* poisson;          // [slot][64]
    uint64_t* gaps;             // [slot][64]
    int32_t* X;                 // [slot][maxS]
};

// The descriptor of model m (wave-uniform), copied through a
// constant-address-space pointer: a burst of scalar loads.
static __device__ __forceinline__ KParams desc_of_b(const KParams* desc, int m) {
    KParams p;
    const ctab_u32 src = CT_U32(reinterpret_cast<const uint32_t*>(desc + m));
    uint32_t* dst = reinterpret_cast<uint32_t*>(&p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(KParams) / 4); ++i) dst[i] = src[i];
    return p;
}

__global__ void gbatch_kernel(GBatchParams bp);

// Group batch with G = 2..32 lanes per simulation: a 64-lane wavefront holds
// 64 / G groups that all run replicas of one model at a time (the model is
// wave-uniform, so its tables stay scalar loads).  The wavefront takes the
// next chunk of the class's chunk list (chunks name models, largest work
// first); its groups then pull replicas of that model from the model's own
// counter (KParams::next_sim) until it is exhausted, and when no group of
// the wavefront holds a replica the wavefront takes the next chunk.
struct GClassParams {
    const KParams* desc;        // per-model descriptors (device)
    const int32_t* chunk_model; // chunk -> model
    int nchunks;
    uint32_t* chunk_next;
    int nwaves;                 // resident wavefronts (one per workgroup)
    int iters;
    int maxS;
    int gwords;                 // LDS words per group (the class's largest model)
    // parked state
    int32_t* wave_model;        // [wave] model in progress, -1 none
    int32_t* slot_sim;          // [group slot] replica, -1 none, -2 the model is exhausted for it
    double* t;
    uint32_t* cursor;
    uint32_t* evc;
    uint64_t* firings;
    uint64_t* ssa_steps;
    uint32_t* leaps;
    uint32_t* rejections;
    uint32_t* switches;
    uint64_t* poisson;          // [slot][G]
    uint64_t* gaps;             // [slot][G]
    int32_t* X;                 // [slot][maxS]
};

template <int G>
__global__ void gclass_kernel(GClassParams bp);
const void* gclass_kernel_fn(int G);

}  // namespace ctl
