// ctl_mlp.hip -- the learned surrogate's MLP on gfx950: hand-written BF16
// MFMA GEMMs with FP32 accumulation for inference and training (forward,
// backward, Adam).  No counterpart in the reference (the surrogate is an
// addition of this build); the simulator is its data source, its fallback
// and the judge of its error.
//
// Network: widths w0 -> w1 -> ... -> wL, ReLU after every hidden layer,
// linear output.  Weights W_l are [w_{l-1}][w_l] row-major; activations are
// [batch][width] row-major.  Master weights, Adam moments and gradients are
// FP32; the GEMM operands (weights, activations, back-propagated errors) are
// BF16 (round to nearest even); every product accumulates in FP32.
//
// GEMM (gemm_tile<WT>): a 256-lane workgroup computes a (64 WT) x (64 WT)
// tile of C; its four wavefronts each own a (32 WT) x (32 WT) quarter as
// WT x WT accumulators of v_mfma_f32_32x32x16_bf16.  K advances 32 at a
// time: the A and B tiles are staged in LDS k-contiguous ([row][k], rows
// padded to 40 elements), filled from global memory with 16-byte loads
// along whichever dimension of the operand is contiguous, the next slab's
// loads in flight while the MFMAs consume the current one.  An operand is
// either k-contiguous (A[i*ld + k], B[j*ld + k]) or not (A[k*ld + i],
// B[k*ld + j]), so one kernel serves X*W (forward), dY*W^T (input error)
// and X^T*dY (weight gradient).  MFMA lane layout: lane l holds
// A[l%32][8*(l/32)+0..7] and B[8*(l/32)+0..7][l%32]; accumulator v of lane
// l is C[8*(v/4) + 4*(l/32) + v%4][l%32].  Workgroups that share rows of A
// are placed on one XCD (one L2).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ctl_internal.h"

namespace ctl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum : int { EPI_STORE_F32 = 0, EPI_BIAS_RELU_BF16 = 1, EPI_BIAS_F32 = 2, EPI_MASK_BF16 = 3 };

struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    long lda, ldb;
    int a_kc, b_kc;            // operand contiguous along k
    int a_vec, b_vec;          // 16-byte loads allowed (16-byte aligned base, ld a multiple of 8)
    int M, N, K;
    int tiles_n;               // column tiles
    int kchunk;                // k per split (split-K: blockIdx.y selects [y*kchunk, (y+1)*kchunk))
    int epi;
    const float* bias;         // [N] (EPI_BIAS_*)
    const uint16_t* mask;      // [M][N] bf16 activation whose sign gates (EPI_MASK_BF16)
    float* Cf;                 // [M][N] fp32 out
    uint16_t* Cb;              // [M][N] bf16 out
};

constexpr int GBK = 32;        // k per LDS slab
constexpr int GKP = 40;        // padded LDS row (elements)

// Eight consecutive elements of operand X along its contiguous dimension,
// starting at (row r, k): (r, k..k+7) when kc, else (r..r+7, k); zero beyond
// R rows / K columns.
__device__ __forceinline__ uint4 gload8(const uint16_t* __restrict__ X, long ld, bool kc, bool vec, int r, int k,
                                        int R, int K) {
    uint4 z = make_uint4(0u, 0u, 0u, 0u);
    const uint16_t* p;
    int lim;
    if (kc) {
        if (r >= R) return z;
        p = X + (long)r * ld + k;
        lim = K - k;
    } else {
        if (k >= K) return z;
        p = X + (long)k * ld + r;
        lim = R - r;
    }
    if (vec && lim >= 8) return *reinterpret_cast<const uint4*>(p);
    uint16_t e[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) e[q] = q < lim ? p[q] : (uint16_t)0;
    z.x = e[0] | ((uint32_t)e[1] << 16);
    z.y = e[2] | ((uint32_t)e[3] << 16);
    z.z = e[4] | ((uint32_t)e[5] << 16);
    z.w = e[6] | ((uint32_t)e[7] << 16);
    return z;
}

// Rows [r0, r0 + BR) x k [k0, k0 + GBK) of an operand: BR*GBK/8 vectors,
// BR/64 per lane.  Vector v covers (row v/4, k 8*(v%4)..) when kc; else a
// lane loads the two k rows of a pair, vectors v and v + 1 = (rows 8*u..,
// k 2*(w%16) and 2*(w%16) + 1) with w = v/2, u = w/16, so that the store
// can write each row's two k values as one 32-bit word and the 16 lanes of
// a row block write 16 consecutive words.
template <int BR>
__device__ __forceinline__ void gstage_load(uint4 (&reg)[BR / 64], const uint16_t* __restrict__ X, long ld, bool kc,
                                            bool vec, int r0, int k0, int R, int K) {
#pragma unroll
    for (int q = 0; q < BR / 64; ++q) {
        const int v = (int)threadIdx.x * (BR / 64) + q;
        if (kc) {
            reg[q] = gload8(X, ld, true, vec, r0 + (v >> 2), k0 + (v & 3) * 8, R, K);
        } else if constexpr (BR / 64 >= 2) {
            const int w = v >> 1, u = w >> 4;
            reg[q] = gload8(X, ld, false, vec, r0 + u * 8, k0 + 2 * (w & 15) + (v & 1), R, K);
        } else {                         // one vector per lane: (rows 8*(v % (BR/8)).., k v / (BR/8))
            reg[q] = gload8(X, ld, false, vec, r0 + (v % (BR / 8)) * 8, k0 + v / (BR / 8), R, K);
        }
    }
}

template <int BR>
__device__ __forceinline__ void gstage_store(const uint4 (&reg)[BR / 64], uint16_t* S, bool kc) {
    if (kc) {
#pragma unroll
        for (int q = 0; q < BR / 64; ++q) {
            const int v = (int)threadIdx.x * (BR / 64) + q;
            *reinterpret_cast<uint4*>(S + (v >> 2) * GKP + (v & 3) * 8) = reg[q];
        }
    } else if constexpr (BR / 64 < 2) {
        const int v = (int)threadIdx.x, r = (v % (BR / 8)) * 8, k = v / (BR / 8);
        const uint32_t w[4] = {reg[0].x, reg[0].y, reg[0].z, reg[0].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) S[(r + e) * GKP + k] = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
    } else {
#pragma unroll
        for (int q = 0; q < BR / 64; q += 2) {
            const int w = ((int)threadIdx.x * (BR / 64) + q) >> 1, u = w >> 4;
            const uint32_t lo[4] = {reg[q].x, reg[q].y, reg[q].z, reg[q].w};          // k even
            const uint32_t hi[4] = {reg[q + 1].x, reg[q + 1].y, reg[q + 1].z, reg[q + 1].w};   // k odd
            uint32_t* row = reinterpret_cast<uint32_t*>(S + (u * 8) * GKP + 2 * (w & 15));
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t a = (lo[e >> 1] >> (16 * (e & 1))) & 0xffffu;
                const uint32_t b = (hi[e >> 1] >> (16 * (e & 1))) & 0xffffu;
                row[e * (GKP / 2)] = a | (b << 16);
            }
        }
    }
}

template <int WT>
__global__ __launch_bounds__(256) void gemm_tile(GemmArgs g) {
    constexpr int BM = 64 * WT;
    __shared__ __align__(16) uint16_t As[BM * GKP];
    __shared__ __align__(16) uint16_t Bs[BM * GKP];
    // XCD-aware order: consecutive workgroup ids go to different XCDs; give
    // each XCD a run of consecutive tiles (same rows of A, several columns).
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int full = nb / 8 * 8;
    const int id = b < full ? (b % 8) * (full / 8) + b / 8 : b;
    const int m0 = (id / g.tiles_n) * BM, n0 = (id % g.tiles_n) * BM;
    const int kb = (int)blockIdx.y * g.kchunk;                      // this split's k range [kb, ke)
    const int ke = kb + g.kchunk < g.K ? kb + g.kchunk : g.K;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32 * WT, wn = (wave & 1) * 32 * WT;
    f32x16 acc[WT][WT];
#pragma unroll
    for (int a = 0; a < WT; ++a)
#pragma unroll
        for (int c = 0; c < WT; ++c)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][c][v] = 0.0f;
    uint4 ra[BM / 64], rb[BM / 64];
    gstage_load<BM>(ra, g.A, g.lda, g.a_kc, g.a_vec, m0, kb, g.M, ke);
    gstage_load<BM>(rb, g.B, g.ldb, g.b_kc, g.b_vec, n0, kb, g.N, ke);
    gstage_store<BM>(ra, As, g.a_kc);
    gstage_store<BM>(rb, Bs, g.b_kc);
    __syncthreads();
    for (int k0 = kb; k0 < ke; k0 += GBK) {
        const bool more = k0 + GBK < ke;
        if (more) {
            gstage_load<BM>(ra, g.A, g.lda, g.a_kc, g.a_vec, m0, k0 + GBK, g.M, ke);
            gstage_load<BM>(rb, g.B, g.ldb, g.b_kc, g.b_vec, n0, k0 + GBK, g.N, ke);
        }
#pragma unroll
        for (int ks = 0; ks < GBK; ks += 16) {
            const int kk = ks + 8 * (lane >> 5);
            bf16x8 fa[WT], fb[WT];
#pragma unroll
            for (int t = 0; t < WT; ++t) {
                fa[t] = *reinterpret_cast<const bf16x8*>(As + (wm + 32 * t + (lane & 31)) * GKP + kk);
                fb[t] = *reinterpret_cast<const bf16x8*>(Bs + (wn + 32 * t + (lane & 31)) * GKP + kk);
            }
#pragma unroll
            for (int a = 0; a < WT; ++a)
#pragma unroll
                for (int c = 0; c < WT; ++c)
                    acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[c], acc[a][c], 0, 0, 0);
        }
        __syncthreads();
        if (more) {
            gstage_store<BM>(ra, As, g.a_kc);
            gstage_store<BM>(rb, Bs, g.b_kc);
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < WT; ++a)
#pragma unroll
        for (int c = 0; c < WT; ++c) {
            const int col = n0 + wn + 32 * c + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = m0 + wm + 32 * a + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row >= g.M || col >= g.N) continue;
                const long o = (long)row * g.N + col;
                float x = acc[a][c][v];
                switch (g.epi) {
                    case EPI_STORE_F32: g.Cf[(long)blockIdx.y * g.M * g.N + o] = x; break;   // split y's partial
                    case EPI_BIAS_RELU_BF16: x += g.bias[col]; g.Cb[o] = f2bf(x > 0.0f ? x : 0.0f); break;
                    case EPI_BIAS_F32: g.Cf[o] = x + g.bias[col]; break;
                    case EPI_MASK_BF16: g.Cb[o] = f2bf(bf2f(g.mask[o]) > 0.0f ? x : 0.0f); break;
                }
            }
        }
}

// Split-K partial sums, added in split order: C[i] = sum_s part[s][i].
__global__ void ksplit_sum(const float* __restrict__ part, float* __restrict__ C, int splits, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[(long)k * n + i];
    C[i] = s;
}

// Loss gradient of the mean squared error over [B][N]: d = 2 (y - t) / (B N),
// stored as BF16 for the backward GEMMs; per-block partial sums of squared
// errors in FP32 (loss_part), or added to a device FP64 total (loss_acc).
__global__ void mse_grad(const float* __restrict__ y, const float* __restrict__ t, uint16_t* __restrict__ d,
                         float* __restrict__ dscale_out, float* __restrict__ loss_part, double* __restrict__ loss_acc,
                         long total, float scale) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const float e = y[i] - t[i];
        acc += e * e;
        d[i] = f2bf(2.0f * e * scale);
        if (dscale_out) dscale_out[i] = 2.0f * e * scale;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (loss_part) loss_part[blockIdx.x] = red[0];
        if (loss_acc) atomicAdd(loss_acc, (double)red[0]);
    }
}

// Column sums db[j] = sum_b d[b][j] (d in BF16), FP32, in two passes with a
// fixed order: col_sum_part sums rows [c*CS_ROWS, (c+1)*CS_ROWS) of 64
// columns per workgroup (4 row lanes per column, combined in LDS), col_sum
// adds the chunks in order.  Coalesced: a wavefront reads 64 consecutive
// columns of one row.
constexpr int CS_ROWS = 512;
__global__ __launch_bounds__(256) void col_sum_part(const uint16_t* __restrict__ d, float* __restrict__ part, int B,
                                                    int N) {
    __shared__ float red[4][64];
    const int j = blockIdx.x * 64 + (int)(threadIdx.x & 63u);
    const int q = (int)(threadIdx.x >> 6);
    const int b0 = blockIdx.y * CS_ROWS;
    const int b1 = b0 + CS_ROWS < B ? b0 + CS_ROWS : B;
    float s = 0.0f;
    if (j < N)
        for (int b = b0 + q; b < b1; b += 4) s += bf2f(d[(long)b * N + j]);
    red[q][threadIdx.x & 63u] = s;
    __syncthreads();
    if (q == 0 && j < N)
        part[(long)blockIdx.y * N + j] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

__global__ void col_sum(const float* __restrict__ part, float* __restrict__ db, int chunks, int N) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[(long)c * N + j];
    db[j] = s;
}

__global__ void to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// Rows idx[0..B) of a [n][w] FP32 array, as BF16 (inputs) or FP32 (targets).
__global__ void gather_rows(const float* __restrict__ src, const int32_t* __restrict__ idx, uint16_t* __restrict__ dstb,
                            float* __restrict__ dstf, int B, int w) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)B * w) return;
    const long b = i / w, k = i % w;
    const float v = src[(long)idx[b] * w + k];
    if (dstb) dstb[i] = f2bf(v);
    if (dstf) dstf[i] = v;
}

// Adam (Kingma & Ba 2015) on the FP32 master copy, refreshing the BF16 copy.
__global__ void adam(float* __restrict__ w, uint16_t* __restrict__ wb, const float* __restrict__ gr,
                     float* __restrict__ m, float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                     float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

}  // namespace ctl

using namespace ctl;

struct ctl_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int> widths;
    int L = 0;                          // layers
    std::vector<long> woff, boff;       // offsets into the flat parameter vector
    long nparam = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // flat fp32: W_1, b_1, ..., W_L, b_L
    uint16_t* wb = nullptr;             // bf16 copy (biases unused)
    long step = 0;
    // work buffers for a batch of up to cap rows
    int cap = 0;
    std::vector<uint16_t*> act;         // act[l]: [cap][w_l] bf16, l = 0..L-1
    uint16_t *err0 = nullptr, *err1 = nullptr;   // [cap][max w] bf16
    float* yout = nullptr;              // [cap][w_L]
    float* tgt = nullptr;
    float* part = nullptr;
    int32_t* idx = nullptr;             // [ndata] epoch permutation
    double* lacc = nullptr;             // epoch loss total
    float* cpart = nullptr;             // [ceil(cap / CS_ROWS)][max w] column-sum partials
    float* kpart = nullptr;             // split-K partial tiles of the weight gradients
    long kpart_words = 0;
    // training data
    float *dX = nullptr, *dY = nullptr;
    long ndata = 0;
    std::vector<void*> allocs;
};

namespace {

int mfail(hipError_t e, const char* what) {
    char b[256];
    std::snprintf(b, sizeof b, "%s: %s", what, hipGetErrorString(e));
    return ctl::fail(e == hipErrorOutOfMemory ? CTL_E_NOMEM : CTL_E_DEVICE, b);
}

#define MHIP(call, what)                             \
    do {                                             \
        hipError_t e_ = (call);                      \
        if (e_ != hipSuccess) return mfail(e_, what); \
    } while (0)

template <typename T>
int malloc_dev(ctl_mlp* n, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return mfail(e, "hipMalloc");
    n->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

// C[M][N] = A * B with element (i, k) of A at A[i*ai + k*ak] and (k, j) of
// B at B[k*bk + j*bj]; one of each operand's strides is 1.
void gemm(ctl_mlp* n, const uint16_t* A, long ai, long ak, const uint16_t* B, long bk, long bj, int M, int N, int K,
          int epi, const float* bias, const uint16_t* mask, float* Cf, uint16_t* Cb) {
    GemmArgs g{};
    g.A = A;
    g.B = B;
    g.a_kc = ak == 1;
    g.lda = g.a_kc ? ai : ak;
    g.b_kc = bk == 1;
    g.ldb = g.b_kc ? bj : bk;
    g.a_vec = (g.lda % 8 == 0) && ((uintptr_t)A % 16 == 0);
    g.b_vec = (g.ldb % 8 == 0) && ((uintptr_t)B % 16 == 0);
    g.M = M; g.N = N; g.K = K;
    g.epi = epi; g.bias = bias; g.mask = mask; g.Cf = Cf; g.Cb = Cb;
    g.kchunk = K;
    // 128 x 128 tiles when they still give every CU a workgroup, else 64 x 64
    const long big = (long)((M + 127) / 128) * ((N + 127) / 128);
    if (big >= 256) {
        g.tiles_n = (N + 127) / 128;
        hipLaunchKernelGGL(gemm_tile<2>, dim3((unsigned)big), dim3(256), 0, n->stream, g);
        return;
    }
    g.tiles_n = (N + 63) / 64;
    const long tiles = (long)((M + 63) / 64) * g.tiles_n;
    // a long k loop over few tiles (weight gradients over a large batch):
    // split k so that about 1024 workgroups run, partial tiles summed in order
    int splits = 1;
    if (epi == EPI_STORE_F32 && n->kpart)
        while (splits < 64 && tiles * splits * 2 <= 1024 && K / (splits * 2) >= 1024 &&
               (long)splits * 2 * M * N <= n->kpart_words)
            splits *= 2;
    if (splits == 1) {
        hipLaunchKernelGGL(gemm_tile<1>, dim3((unsigned)tiles), dim3(256), 0, n->stream, g);
        return;
    }
    g.kchunk = (K + splits - 1) / splits;
    g.kchunk = (g.kchunk + GBK - 1) / GBK * GBK;
    g.Cf = n->kpart;
    const long mn = (long)M * N;
    hipLaunchKernelGGL(gemm_tile<1>, dim3((unsigned)tiles, (unsigned)splits), dim3(256), 0, n->stream, g);
    hipLaunchKernelGGL(ksplit_sum, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, n->stream, n->kpart, Cf, splits,
                       mn);
}

int ensure_cap(ctl_mlp* n, int B) {
    if (B <= n->cap) return 0;
    int wmax = 0;
    for (int w : n->widths) wmax = std::max(wmax, w);
    n->act.assign(n->L, nullptr);
    int rc;
    for (int l = 0; l < n->L; ++l)
        if ((rc = malloc_dev(n, &n->act[l], (size_t)B * n->widths[l]))) return rc;
    if ((rc = malloc_dev(n, &n->err0, (size_t)B * wmax)) || (rc = malloc_dev(n, &n->err1, (size_t)B * wmax)) ||
        (rc = malloc_dev(n, &n->yout, (size_t)B * n->widths[n->L])) ||
        (rc = malloc_dev(n, &n->tgt, (size_t)B * n->widths[n->L])) || (rc = malloc_dev(n, &n->part, 1024)) ||
        (rc = malloc_dev(n, &n->lacc, 1)) ||
        (rc = malloc_dev(n, &n->cpart, (size_t)((B + CS_ROWS - 1) / CS_ROWS) * wmax)))
        return rc;
    if (B >= 2048) {                    // large batches: room for 16 splits of the largest weight block
        long wmn = 0;
        for (int l = 0; l < n->L; ++l) wmn = std::max(wmn, (long)n->widths[l] * n->widths[l + 1]);
        if ((rc = malloc_dev(n, &n->kpart, (size_t)16 * wmn))) return rc;
        n->kpart_words = 16 * wmn;
    }
    n->cap = B;
    return 0;
}

// Forward pass of B rows already in act[0]; output FP32 in out (default yout).
void forward(ctl_mlp* n, int B, float* out = nullptr) {
    for (int l = 0; l < n->L; ++l) {
        const int K = n->widths[l], N = n->widths[l + 1];
        const uint16_t* W = n->wb + n->woff[l];
        const float* b = n->w + n->boff[l];
        if (l + 1 < n->L)
            gemm(n, n->act[l], K, 1, W, N, 1, B, N, K, EPI_BIAS_RELU_BF16, b, nullptr, nullptr, n->act[l + 1]);
        else
            gemm(n, n->act[l], K, 1, W, N, 1, B, N, K, EPI_BIAS_F32, b, nullptr, out ? out : n->yout, nullptr);
    }
}

// Backward pass from the BF16 output error in err0: fills n->g.
void backward(ctl_mlp* n, int B) {
    uint16_t* e = n->err0;
    uint16_t* e2 = n->err1;
    for (int l = n->L - 1; l >= 0; --l) {
        const int K = n->widths[l], N = n->widths[l + 1];
        // dW_l = act[l]^T * e : (K x B) * (B x N)
        gemm(n, n->act[l], 1, K, e, N, 1, K, N, B, EPI_STORE_F32, nullptr, nullptr, n->g + n->woff[l], nullptr);
        const int chunks = (B + CS_ROWS - 1) / CS_ROWS;
        hipLaunchKernelGGL(col_sum_part, dim3((N + 63) / 64, chunks), dim3(256), 0, n->stream, e, n->cpart, B, N);
        hipLaunchKernelGGL(col_sum, dim3((N + 255) / 256), dim3(256), 0, n->stream, n->cpart, n->g + n->boff[l],
                           chunks, N);
        if (l > 0) {
            // e_{l} = (e * W_l^T) gated by act[l] > 0 : (B x N) * (N x K)
            const uint16_t* W = n->wb + n->woff[l];
            gemm(n, e, N, 1, W, 1, N, B, K, N, EPI_MASK_BF16, nullptr, n->act[l], nullptr, e2);
            std::swap(e, e2);
        }
    }
}

float loss_and_grad(ctl_mlp* n, int B, const float* t_dev, float* dscale_dbg) {
    const long total = (long)B * n->widths[n->L];
    const int blocks = (int)std::min<long>(1024, (total + 255) / 256);
    hipLaunchKernelGGL(mse_grad, dim3(blocks), dim3(256), 0, n->stream, n->yout, t_dev, n->err0, dscale_dbg, n->part,
                       (double*)nullptr,
                       total, 1.0f / (float)total);
    std::vector<float> part(blocks);
    if (hipMemcpyAsync(part.data(), n->part, sizeof(float) * blocks, hipMemcpyDeviceToHost, n->stream) != hipSuccess)
        return NAN;
    if (hipStreamSynchronize(n->stream) != hipSuccess) return NAN;
    double s = 0.0;
    for (float p : part) s += p;
    return (float)(s / (double)total);
}

void refresh_bf16(ctl_mlp* n) {
    hipLaunchKernelGGL(to_bf16, dim3((unsigned)((n->nparam + 255) / 256)), dim3(256), 0, n->stream, n->w, n->wb,
                       n->nparam);
}

}  // namespace

extern "C" {

int ctl_mlp_create(const int32_t* widths, int32_t nlayers, int32_t gpu, uint64_t seed, ctl_mlp** out) {
    ctl::clear_error();
    if (!widths || nlayers < 1 || !out) return ctl::fail(CTL_E_ARGS, "bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return ctl::fail(CTL_E_DEVICE, "no HIP device available: libctl needs an AMD GPU (gfx950)");
    for (int l = 0; l <= nlayers; ++l)
        if (widths[l] <= 0) return ctl::fail(CTL_E_ARGS, "layer widths must be positive");
    ctl_mlp* n = new ctl_mlp();
    auto bail = [&](int rc) { ctl_mlp_destroy(n); return rc; };
    n->device = gpu;
    n->L = nlayers;
    n->widths.assign(widths, widths + nlayers + 1);
    hipError_t e = hipSetDevice(gpu);
    if (e != hipSuccess) return bail(mfail(e, "hipSetDevice"));
    if ((e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking)) != hipSuccess) return bail(mfail(e, "stream"));
    long off = 0;
    for (int l = 0; l < nlayers; ++l) {
        n->woff.push_back(off);
        off += (long)widths[l] * widths[l + 1];
        n->boff.push_back(off);
        off += widths[l + 1];
    }
    n->nparam = off;
    int rc;
    if ((rc = malloc_dev(n, &n->w, off)) || (rc = malloc_dev(n, &n->g, off)) || (rc = malloc_dev(n, &n->m, off)) ||
        (rc = malloc_dev(n, &n->v, off)) || (rc = malloc_dev(n, &n->wb, off)))
        return bail(rc);
    // He-uniform weights, zero biases (host RNG, seeded)
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < nlayers; ++l) {
        const float a = std::sqrt(6.0f / (float)widths[l]);
        std::uniform_real_distribution<float> U(-a, a);
        for (long i = 0; i < (long)widths[l] * widths[l + 1]; ++i) h[n->woff[l] + i] = U(rng);
    }
    if ((e = hipMemcpy(n->w, h.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return bail(mfail(e, "H2D"));
    if ((e = hipMemset(n->m, 0, sizeof(float) * off)) != hipSuccess || (e = hipMemset(n->v, 0, sizeof(float) * off)) != hipSuccess)
        return bail(mfail(e, "memset"));
    refresh_bf16(n);
    if ((e = hipStreamSynchronize(n->stream)) != hipSuccess) return bail(mfail(e, "init"));
    *out = n;
    return 0;
}

int ctl_mlp_num_params(const ctl_mlp* n, int64_t* count) {
    if (!n || !count) return ctl::fail(CTL_E_ARGS, "null argument");
    *count = n->nparam;
    return 0;
}

int ctl_mlp_params(ctl_mlp* n, float* host, size_t count, int32_t set) {
    ctl::clear_error();
    if (!n || !host || count != (size_t)n->nparam) return ctl::fail(CTL_E_ARGS, "count must equal the parameter count");
    MHIP(hipSetDevice(n->device), "hipSetDevice");
    if (set) {
        MHIP(hipMemcpy(n->w, host, sizeof(float) * count, hipMemcpyHostToDevice), "H2D");
        refresh_bf16(n);
        MHIP(hipStreamSynchronize(n->stream), "sync");
    } else {
        MHIP(hipMemcpy(host, n->w, sizeof(float) * count, hipMemcpyDeviceToHost), "D2H");
    }
    return 0;
}

int ctl_mlp_forward(ctl_mlp* n, const float* x, int64_t rows, float* y) {
    ctl::clear_error();
    if (!n || !x || !y || rows <= 0) return ctl::fail(CTL_E_ARGS, "bad arguments");
    MHIP(hipSetDevice(n->device), "hipSetDevice");
    const int chunk = 8192;
    int rc = ensure_cap(n, (int)std::min<int64_t>(rows, chunk));
    if (rc) return rc;
    float* xin = nullptr;
    MHIP(hipMalloc(&xin, sizeof(float) * (size_t)std::min<int64_t>(rows, chunk) * n->widths[0]), "hipMalloc");
    for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
        const int B = (int)std::min<int64_t>(chunk, rows - r0);
        const long nin = (long)B * n->widths[0];
        hipError_t e = hipMemcpyAsync(xin, x + r0 * n->widths[0], sizeof(float) * nin, hipMemcpyHostToDevice, n->stream);
        if (e != hipSuccess) { (void)hipFree(xin); return mfail(e, "H2D"); }
        hipLaunchKernelGGL(to_bf16, dim3((unsigned)((nin + 255) / 256)), dim3(256), 0, n->stream, xin, n->act[0], nin);
        forward(n, B);
        e = hipMemcpyAsync(y + r0 * n->widths[n->L], n->yout, sizeof(float) * (size_t)B * n->widths[n->L],
                           hipMemcpyDeviceToHost, n->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
        if (e != hipSuccess) { (void)hipFree(xin); return mfail(e, "forward"); }
    }
    (void)hipFree(xin);
    return 0;
}

int ctl_mlp_forward_device(ctl_mlp* n, const float* x_dev, int64_t rows, float* y_dev, float* ms) {
    ctl::clear_error();
    if (!n || !x_dev || !y_dev || rows <= 0) return ctl::fail(CTL_E_ARGS, "bad arguments");
    MHIP(hipSetDevice(n->device), "hipSetDevice");
    const int chunk = 65536;
    int rc = ensure_cap(n, (int)std::min<int64_t>(rows, chunk));
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms) {
        MHIP(hipEventCreate(&e0), "event");
        MHIP(hipEventCreate(&e1), "event");
        MHIP(hipEventRecord(e0, n->stream), "event");
    }
    const int w0 = n->widths[0], wL = n->widths[n->L];
    for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
        const int B = (int)std::min<int64_t>(chunk, rows - r0);
        const long nin = (long)B * w0;
        hipLaunchKernelGGL(to_bf16, dim3((unsigned)((nin + 255) / 256)), dim3(256), 0, n->stream, x_dev + r0 * w0,
                           n->act[0], nin);
        forward(n, B, y_dev + r0 * wL);              // the output layer writes the caller's rows
    }
    hipError_t e = hipSuccess;
    if (ms) e = hipEventRecord(e1, n->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    if (e == hipSuccess && ms) e = hipEventElapsedTime(ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return mfail(e, "forward");
    return 0;
}

// One forward + backward on rows (x, t); returns the loss and the FP32
// gradient of the mean squared error with respect to every parameter.
int ctl_mlp_gradient(ctl_mlp* n, const float* x, const float* t, int32_t rows, float* grad, size_t count, float* loss) {
    ctl::clear_error();
    if (!n || !x || !t || rows <= 0 || rows > 65536 || (grad && count != (size_t)n->nparam)) return ctl::fail(CTL_E_ARGS, "bad arguments");
    MHIP(hipSetDevice(n->device), "hipSetDevice");
    int rc = ensure_cap(n, rows);
    if (rc) return rc;
    const long nin = (long)rows * n->widths[0], nout = (long)rows * n->widths[n->L];
    float* xin = nullptr;
    MHIP(hipMalloc(&xin, sizeof(float) * nin), "hipMalloc");
    hipError_t e = hipMemcpy(xin, x, sizeof(float) * nin, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n->tgt, t, sizeof(float) * nout, hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(xin); return mfail(e, "H2D"); }
    hipLaunchKernelGGL(to_bf16, dim3((unsigned)((nin + 255) / 256)), dim3(256), 0, n->stream, xin, n->act[0], nin);
    forward(n, rows);
    const float l = loss_and_grad(n, rows, n->tgt, nullptr);
    backward(n, rows);
    e = hipStreamSynchronize(n->stream);
    if (e == hipSuccess && grad) e = hipMemcpy(grad, n->g, sizeof(float) * count, hipMemcpyDeviceToHost);
    (void)hipFree(xin);
    if (e != hipSuccess) return mfail(e, "gradient");
    if (loss) *loss = l;
    return 0;
}

// Training data (normalised inputs [n][w0], targets [n][wL]) kept on the device.
int ctl_mlp_set_data(ctl_mlp* n, const float* x, const float* t, int64_t rows) {
    ctl::clear_error();
    if (!n || !x || !t || rows <= 0) return ctl::fail(CTL_E_ARGS, "bad arguments");
    MHIP(hipSetDevice(n->device), "hipSetDevice");
    if (n->dX) { (void)hipFree(n->dX); n->dX = nullptr; }
    if (n->dY) { (void)hipFree(n->dY); n->dY = nullptr; }
    MHIP(hipMalloc(&n->dX, sizeof(float) * rows * n->widths[0]), "hipMalloc");
    MHIP(hipMalloc(&n->dY, sizeof(float) * rows * n->widths[n->L]), "hipMalloc");
    MHIP(hipMemcpy(n->dX, x, sizeof(float) * rows * n->widths[0], hipMemcpyHostToDevice), "H2D");
    MHIP(hipMemcpy(n->dY, t, sizeof(float) * rows * n->widths[n->L], hipMemcpyHostToDevice), "H2D");
    if (n->idx) { (void)hipFree(n->idx); n->idx = nullptr; }
    MHIP(hipMalloc(&n->idx, sizeof(int32_t) * rows), "hipMalloc");
    n->ndata = rows;
    return 0;
}

// Mini-batch Adam over the device data: epochs of shuffled batches.
// losses[e] receives the mean training loss of epoch e.
int ctl_mlp_train(ctl_mlp* n, int32_t epochs, int32_t batch, float lr, uint64_t seed, float* losses) {
    ctl::clear_error();
    if (!n || epochs <= 0 || batch <= 0) return ctl::fail(CTL_E_ARGS, "bad arguments");
    if (!n->ndata) return ctl::fail(CTL_E_STATE, "no training data");
    MHIP(hipSetDevice(n->device), "hipSetDevice");
    int rc = ensure_cap(n, batch);
    if (rc) return rc;
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm(n->ndata);
    for (long i = 0; i < n->ndata; ++i) perm[i] = (int32_t)i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int w0 = n->widths[0], wL = n->widths[n->L];
    const long total = (long)batch * wL;
    const int blocks = (int)std::min<long>(1024, (total + 255) / 256);
    for (int ep = 0; ep < epochs; ++ep) {
        // one upload of the epoch's permutation; the loss stays on the device
        // until the epoch ends, so the batches run without host round trips
        std::shuffle(perm.begin(), perm.end(), rng);
        MHIP(hipMemcpyAsync(n->idx, perm.data(), sizeof(int32_t) * n->ndata, hipMemcpyHostToDevice, n->stream), "H2D");
        MHIP(hipMemsetAsync(n->lacc, 0, sizeof(double), n->stream), "memset");
        long nb = 0;
        for (long s0 = 0; s0 + batch <= n->ndata; s0 += batch) {
            hipLaunchKernelGGL(gather_rows, dim3((unsigned)(((long)batch * w0 + 255) / 256)), dim3(256), 0, n->stream,
                               n->dX, n->idx + s0, n->act[0], (float*)nullptr, batch, w0);
            hipLaunchKernelGGL(gather_rows, dim3((unsigned)(((long)batch * wL + 255) / 256)), dim3(256), 0, n->stream,
                               n->dY, n->idx + s0, (uint16_t*)nullptr, n->tgt, batch, wL);
            forward(n, batch);
            hipLaunchKernelGGL(mse_grad, dim3(blocks), dim3(256), 0, n->stream, n->yout, n->tgt, n->err0,
                               (float*)nullptr, (float*)nullptr, n->lacc, total, 1.0f / (float)total);
            backward(n, batch);
            n->step += 1;
            const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
            hipLaunchKernelGGL(adam, dim3((unsigned)((n->nparam + 255) / 256)), dim3(256), 0, n->stream, n->w, n->wb,
                               n->g, n->m, n->v, n->nparam, lr, b1, b2, eps, c1, c2);
            ++nb;
        }
        double sum = 0.0;
        MHIP(hipMemcpyAsync(&sum, n->lacc, sizeof(double), hipMemcpyDeviceToHost, n->stream), "D2H");
        MHIP(hipStreamSynchronize(n->stream), "train");
        if (losses) losses[ep] = nb ? (float)(sum / ((double)total * nb)) : NAN;
    }
    return 0;
}

void ctl_mlp_destroy(ctl_mlp* n) {
    if (!n) return;
    (void)hipSetDevice(n->device);
    for (void* q : n->allocs) (void)hipFree(q);
    if (n->dX) (void)hipFree(n->dX);
    if (n->dY) (void)hipFree(n->dY);
    if (n->idx) (void)hipFree(n->idx);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

}  // extern "C"
