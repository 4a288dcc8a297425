// ctl_device.h -- device-side random numbers and FP32 math of libctl (gfx950).
//
// Everything here is built with -ffp-contract=off: a fused multiply-add only
// happens where fmaf() is written, so the op order is fixed and a host
// replay of the same algorithm reproduces every value bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctl {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Per simulation the
// counter is {block, event, global simulation index, 0}; the key is the
// 64-bit seed.  An "event" is one leap attempt, one tau-halving retry or one
// SSA burst; its uniforms are consumed in order, four per Philox block.
// ---------------------------------------------------------------------------
// 1 when the kernels of the translation unit read their parameters (the
// Philox key included) from kernel arguments, the same in every lane; a unit
// whose lanes may read different parameter sets defines it 0 before
// including this header.
#ifndef CTL_UNIFORM_PARAMS
#define CTL_UNIFORM_PARAMS 1
#endif

struct Rng {
    uint32_t k0, k1;     // key
    uint32_t sim;        // global simulation index
    uint32_t ev;         // current event
    uint32_t blk;        // next block of the event
    uint32_t b0, b1, b2, b3;
    uint32_t pos;        // 4 = buffer empty
    uint32_t tag = 0u;   // counter word 3: 0, or 1 + lane for a group's lane streams
};

static __device__ __forceinline__ void philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                             uint32_t k1, uint32_t& o0, uint32_t& o1, uint32_t& o2,
                                             uint32_t& o3) {
#if CTL_UNIFORM_PARAMS
    // The key is kernel-uniform: without this barrier the compiler hoists all
    // ten round keys into scalar registers for the whole kernel, which pushes
    // other uniform values into spill slots.  After it, the round keys are two
    // scalar adds per round, issued beside the vector work.
    k0 = __builtin_amdgcn_readfirstlane(k0);      // uniform; a no-op when already scalar
    k1 = __builtin_amdgcn_readfirstlane(k1);
    asm volatile("" : "+s"(k0), "+s"(k1));
#endif
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32 -> 64-bit multiply each (v_mad_u64_u32), not a lo and a hi
        const uint64_t p0 = (uint64_t)c0 * 0xD2511F53u;
        const uint64_t p1 = (uint64_t)c2 * 0xCD9E8D57u;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

__device__ __forceinline__ void rng_event(Rng& g, uint32_t ev) {
    g.ev = ev;
    g.blk = 0;
    g.pos = 4;
}

// The block is consumed front to back by rotating the buffer registers (b0 is
// always the next word), which keeps the whole state in VGPRs.
__device__ __forceinline__ uint32_t rng_u32(Rng& g) {
    if (g.pos >= 4) {
        philox_block(g.blk, g.ev, g.sim, g.tag, g.k0, g.k1, g.b0, g.b1, g.b2, g.b3);
        g.blk += 1;
        g.pos = 0;
    }
    const uint32_t v = g.b0;
    g.b0 = g.b1;
    g.b1 = g.b2;
    g.b2 = g.b3;
    g.pos += 1;
    return v;
}

// U(0,1] and U[0,1) on the 2^-24 grid.
__device__ __forceinline__ float u32_oc(uint32_t r) { return (float)((r >> 8) + 1u) * 0x1p-24f; }
__device__ __forceinline__ float rng_oc(Rng& g) { return u32_oc(rng_u32(g)); }
__device__ __forceinline__ float rng_co(Rng& g) { return (float)(rng_u32(g) >> 8) * 0x1p-24f; }

// ---------------------------------------------------------------------------
// FP32 log and exp: Cody-Waite reduction with ln 2 split into a 16-bit head
// (exact products with the exponent) and a tail; log(1+f) from the atanh
// series (coefficients 2/3, 2/5, 2/7, 2/9), exp(r) from its degree-8 Taylor
// polynomial.  Only +, *, / and explicit fmaf: identical on host and device.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dlogf(float x) {
    const uint32_t ix = __float_as_uint(x);
    int e = (int)(ix >> 23) - 127;
    uint32_t mb = (ix & 0x007fffffu) | 0x3f800000u;
    if (mb > 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
    const float f = __uint_as_float(mb) - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float R = fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f);
    R = fmaf(z, R, 0x1.99999ap-2f);
    R = fmaf(z, R, 0x1.555556p-1f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float ef = (float)e;
    return ef * 0x1.62e4p-1f + (f - (hfsq - (s * (hfsq + R) + ef * 0x1.7f7d1cp-20f)));
}

__device__ __forceinline__ float dexpf(float x) {
    x = x < -87.0f ? -87.0f : x;
    x = x > 88.0f ? 88.0f : x;
    const float kf = rintf(x * 0x1.715476p+0f);
    const float r = (x - kf * 0x1.62e4p-1f) - kf * 0x1.7f7d1cp-20f;
    float p = fmaf(r, 0x1.a01a02p-16f, 0x1.a01a02p-13f);
    p = fmaf(r, p, 0x1.6c16c2p-10f);
    p = fmaf(r, p, 0x1.111112p-7f);
    p = fmaf(r, p, 0x1.555556p-5f);
    p = fmaf(r, p, 0x1.555556p-3f);
    p = fmaf(r, p, 0.5f);
    p = fmaf(r, p, 1.0f);
    p = fmaf(r, p, 1.0f);
    const int k = (int)kf;
    return p * __uint_as_float((uint32_t)(k + 127) << 23);
}

// Exponential(1) variate -log(1 - V): V = r 2^-32 below 2^-8 (exact) and
// (r >> 8) 2^-24 above, so short gaps keep 32-bit resolution and V < 1.
// log(1 - V) is dlogf's series with f = -V exactly while 1 - V > sqrt(1/2)
// (small arguments keep their full relative precision), else dlogf(1 - V).
// The two cases share one series and one division without a branch (the
// lanes of a wavefront draw different V): f and the exponent are selected,
// and each case ends with exactly the additions of its own formula.
__device__ __forceinline__ float u32_exp(uint32_t r) {
    const float v = r < 0x01000000u ? (float)r * 0x1p-32f : (float)(r >> 8) * 0x1p-24f;
    const bool small = v < 0x1.2bec34p-2f;
    // dlogf's reduction of x = 1 - v (exact on the large branch)
    const uint32_t ix = __float_as_uint(1.0f - v);
    int e = (int)(ix >> 23) - 127;
    uint32_t mb = (ix & 0x007fffffu) | 0x3f800000u;
    if (mb > 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
    const float f = small ? -v : __uint_as_float(mb) - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float R = fmaf(z, 0x1.c71c72p-3f, 0x1.24924ap-2f);
    R = fmaf(z, R, 0x1.99999ap-2f);
    R = fmaf(z, R, 0x1.555556p-1f);
    R = R * z;
    const float hfsq = 0.5f * f * f;
    const float t = s * (hfsq + R);
    const float ef = (float)e;
    const float lsmall = f - (hfsq - t);
    const float llarge = ef * 0x1.62e4p-1f + (f - (hfsq - (t + ef * 0x1.7f7d1cp-20f)));
    return -(small ? lsmall : llarge);
}
__device__ __forceinline__ float exp_rv(Rng& g) { return u32_exp(rng_u32(g)); }

// ---------------------------------------------------------------------------
// Poisson variates (the reference calls curand_poisson, T.cu:2018, 2040).
//   lam < 10         inversion by sequential search, one U[0,1); a search
//                    longer than 63 terms (cumulative stalled under u in
//                    float) restarts with a fresh uniform
//   10 <= lam < 2^23 PTRS (Hormann 1993) with the log-pmf from Loader's bd0
//                    deviance and the Stirling tail, stable in FP32
//   lam >= 2^23      normal approximation (Marsaglia polar) in FP64
// recip[k] = 1/k (k < 64) lives in LDS.
// ---------------------------------------------------------------------------
__constant__ float c_logfact[10] = {0.0f, 0.0f, 0x1.62e43p-1f, 0x1.cab0cp+0f, 0x1.96ca78p+1f,
                                     0x1.326644p+2f, 0x1.a51274p+2f, 0x1.10ce2p+3f, 0x1.5358e8p+3f,
                                     0x1.99a892p+3f};

__device__ __forceinline__ float bd0(float kf, float lam, float diff) {
    const float sum = kf + lam;
    if (fabsf(diff) < 0.1f * sum) {
        const float v = diff / sum;
        float s = diff * v;
        float ej = 2.0f * kf * v;
        const float v2 = v * v;
        ej = ej * v2; s = s + ej * 0x1.555556p-2f;
        ej = ej * v2; s = s + ej * 0x1.99999ap-3f;
        ej = ej * v2; s = s + ej * 0x1.24924ap-3f;
        ej = ej * v2; s = s + ej * 0x1.c71c72p-4f;
        ej = ej * v2; s = s + ej * 0x1.745d18p-4f;
        return s;
    }
    return (kf * dlogf(kf / lam) + lam) - kf;
}

__device__ __forceinline__ int32_t poisson_large(Rng& g, float lam) {
    if (lam >= 8388608.0f) {
        if (!(lam < 2147483648.0f)) return 2147483647;
        float u, v, s;
        do {
            u = 2.0f * rng_oc(g) - 1.0f;
            v = 2.0f * rng_oc(g) - 1.0f;
            s = u * u + v * v;
        } while (s >= 1.0f || s == 0.0f);
        const float z = u * sqrtf(-2.0f * dlogf(s) / s);
        const double kd = floor((double)lam + sqrt((double)lam) * (double)z + 0.5);
        if (kd < 0.0) return 0;
        if (kd > 2147483647.0) return 2147483647;
        return (int32_t)kd;
    }
    const float slam = sqrtf(lam);
    const float loglam = dlogf(lam);
    const float b = 0.931f + 2.53f * slam;
    const float a = -0.059f + 0.02483f * b;
    const float invalpha = 1.1239f + 1.1328f / (b - 3.4f);
    const float vr = 0.9277f - 3.6224f / (b - 2.0f);
    const float li = floorf(lam);
    const float lf = lam - li;
    for (;;) {
        const float U = rng_oc(g) - 0.5f;
        const float V = rng_oc(g);
        const float us = 0.5f - fabsf(U);
        if (us < 0.013f && V > us) continue;
        const float xk = (2.0f * a / us + b) * U + 0.43f;
        const float d = floorf(xk + lf);
        if (us >= 0.07f && V <= vr) return (int32_t)li + (int32_t)d;
        if (d < -li) continue;
        if (d > 1073741824.0f) continue;
        const int32_t k = (int32_t)li + (int32_t)d;
        const float lhs = dlogf(V * invalpha / (a / (us * us) + b));
        float rhs;
        if (k < 10) {
            rhs = (-lam + (float)k * loglam) - c_logfact[k];
        } else {
            const float kf = (float)k;
            const float kinv = 1.0f / kf;
            const float k2 = kinv * kinv;
            const float stirl = (0x1.555556p-4f - (0x1.6c16c2p-9f - 0x1.a01a02p-11f * k2) * k2) * kinv;
            rhs = (-bd0(kf, lam, d - lf) - 0.5f * dlogf(0x1.921fb6p+2f * kf)) - stirl;
        }
        if (lhs <= rhs) return k;
    }
}

__device__ __forceinline__ int32_t poisson(Rng& g, float lam, const float* __restrict__ recip) {
    if (!(lam > 0.0f)) return 0;
    if (lam < 10.0f) {
        const float p0 = dexpf(-lam);
        for (;;) {
            const float u = rng_co(g);
            float p = p0, F = p0;
            int k = 0;
            while (u >= F) {
                ++k;
                if (k > 63) break;
                p = (p * lam) * recip[k];
                F = F + p;
            }
            if (k <= 63) return k;
        }
    }
    return poisson_large(g, lam);
}

}  // namespace ctl
