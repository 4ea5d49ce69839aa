// closure_net.hpp - the learned closure's network evaluation, shared by the
// closure kernels and the fused gather-scatter: BF16 matrix-core products on
// wave64 (v_mfma_f32_16x16x16_bf16, FP32 accumulation), the network's
// operands packed in the order the lanes read them, and the inputs and the
// stress of a snow particle.
//
// With l the lane, A (16 x 16): lane l holds A[l & 15][4 (l >> 4) + i];
// B (16 x 16): lane l holds B[4 (l >> 4) + i][l & 15]; D: lane l holds
// D[4 (l >> 4) + i][l & 15], i < 4.  A wave evaluates 16 samples at a time,
// every layer transposed (D^T = W X^T, rows = units, columns = samples).
#pragma once

#include "device_math.hpp"
#include "internal.hpp"

namespace svamd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_round(float f) { return __uint_as_float(((uint32_t)(uint16_t)to_bf16(f)) << 16); }
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// p[0], p[s], p[2 s], p[3 s] as one BF16 operand.
__device__ __forceinline__ bf16x4 op4(const float *p, int s) {
    bf16x4 r;
    r.x = to_bf16(p[0]);
    r.y = to_bf16(p[s]);
    r.z = to_bf16(p[2 * s]);
    r.w = to_bf16(p[3 * s]);
    return r;
}
__device__ __forceinline__ bf16x4 pack(const f32x4 v) {
    bf16x4 r;
    r.x = to_bf16(v.x);
    r.y = to_bf16(v.y);
    r.z = to_bf16(v.z);
    r.w = to_bf16(v.w);
    return r;
}
__device__ __forceinline__ f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }

// The network's operands in the order the lanes read them (staged in LDS by
// the closure kernels, packed once in global memory for the fused kernel):
// op[o][l] is lane l's BF16 operand o (o = t for W1 tile t, 4 + 4 t + kk for
// W2 tile (t, kk), 20 + kk for W3 tile kk), so that every operand is one
// conflict-free 8-byte read; the biases follow as floats.
struct StagedNet {
    bf16x4 op[24][64];
    float b1[MLP_H], b2[MLP_H], b3[16];
};

__device__ __forceinline__ void stage_net(const float *__restrict__ P, StagedNet &S) {
    for (int idx = (int)threadIdx.x; idx < 24 * 64; idx += (int)blockDim.x) {
        const int o = idx >> 6, ln = idx & 63, r0 = 4 * (ln >> 4), c = ln & 15;
        const float *src;
        if (o < 4)
            src = P + MLP_OFF_W1 + (16 * o + c) * MLP_KIN + r0;
        else if (o < 20)
            src = P + MLP_OFF_W2 + (16 * ((o - 4) >> 2) + c) * MLP_H + 16 * ((o - 4) & 3) + r0;
        else
            src = P + MLP_OFF_W3 + c * MLP_H + 16 * (o - 20) + r0;
        S.op[o][ln] = op4(src, 1);
    }
    for (int idx = (int)threadIdx.x; idx < 2 * MLP_H + 16; idx += (int)blockDim.x) {
        if (idx < MLP_H)
            S.b1[idx] = P[MLP_OFF_B1 + idx];
        else if (idx < 2 * MLP_H)
            S.b2[idx - MLP_H] = P[MLP_OFF_B2 + idx - MLP_H];
        else
            S.b3[idx - 2 * MLP_H] = P[MLP_OFF_B3 + idx - 2 * MLP_H];
    }
}

// forward16 with the operands from a StagedNet (in LDS or in global memory).
__device__ __forceinline__ void forward16_staged(const StagedNet &S, bf16x4 xop, f32x4 &y) {
    const int l = (int)(threadIdx.x & 63u), r0 = 4 * (l >> 4);
    bf16x4 h1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const f32x4 acc = mfma(S.op[t][l], xop, zero4());
        f32x4 h;
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = fmaxf(acc[i] + S.b1[16 * t + r0 + i], 0.f);
        h1[t] = pack(h);
    }
    bf16x4 h2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = zero4();
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) acc = mfma(S.op[4 + 4 * t + kk][l], h1[kk], acc);
        f32x4 h;
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = fmaxf(acc[i] + S.b2[16 * t + r0 + i], 0.f);
        h2[t] = pack(h);
    }
    f32x4 acc = zero4();
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) acc = mfma(S.op[20 + kk][l], h2[kk], acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = acc[i] + S.b3[r0 + i];
}

// The network's inputs for a snow particle (singular values s of its F, its
// Lame parameters and Jp): (s - 1) / sigma scale, log10(E / 1e5),
// 10 (nu - 0.3), (Jp - 1) / Jp scale; E is returned for the stress unit.
__device__ __forceinline__ void closure_features(const float s[3], float mu, float lam, float jp, float x[MLP_IN],
                                                 float &E) {
    E = mu * (3.f * lam + 2.f * mu) / (lam + mu);
    const float nu = lam / (2.f * (lam + mu));
#pragma unroll
    for (int q = 0; q < 3; ++q) x[q] = (s[q] - 1.0f) * (1.0f / CLOSURE_SIGMA_SCALE);
    x[3] = log10f(E * 1e-5f);
    x[4] = (nu - 0.3f) * 10.0f;
    x[5] = (jp - 1.0f) * (1.0f / CLOSURE_JP_SCALE);
}

// The learned Kirchhoff stress: the exact linear part of the hardened
// fixed-corotated principal stresses plus the network's remainder y (in
// units of E h CLOSURE_STRESS_SCALE m), recomposed as U diag(d) U^T.
__device__ __forceinline__ void closure_tau(const float U[9], const float s[3], float mu, float lam, float hard,
                                            float E, const float y[3], float tau[9]) {
    const float unit = E * hard * CLOSURE_STRESS_SCALE * closure_strain_scale(s);
    float lin[3], rem[3], d[3];
    fcr_principal_parts(mu * hard, lam * hard, s, lin, rem);
#pragma unroll
    for (int q = 0; q < 3; ++q) d[q] = lin[q] + y[q] * unit;
    recompose(U, d, U, tau);
}

// The network's outputs for the wave's 64 samples (lane l: its sample's
// inputs x, and want: whether it needs them).  Every lane of the wave must
// run this (the products take all 64 lanes); lanes that do not want an
// output pass any finite inputs.  The features reach the operand layout and
// the outputs their owners through lane shuffles.
__device__ __forceinline__ void closure_outputs(const StagedNet &net, const float x[MLP_IN], bool want, float y[3]) {
    const int l = (int)(threadIdx.x & 63u), r0 = 4 * (l >> 4), c = l & 15;
    y[0] = y[1] = y[2] = 0.f;
    if (__ballot(want) == 0ull) return;
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
        const int src = 16 * g + c;
        float f[MLP_IN];
#pragma unroll
        for (int q = 0; q < MLP_IN; ++q) f[q] = __shfl(x[q], src, 64);
        bf16x4 xop;
        xop.x = r0 == 0 ? to_bf16(f[0]) : (r0 == 4 ? to_bf16(f[4]) : (short)0);
        xop.y = r0 == 0 ? to_bf16(f[1]) : (r0 == 4 ? to_bf16(f[5]) : (short)0);
        xop.z = r0 == 0 ? to_bf16(f[2]) : (short)0;
        xop.w = r0 == 0 ? to_bf16(f[3]) : (short)0;
        f32x4 out;
        // A fresh pointer per group keeps the operand loads inside the group:
        // hoisted out of the loops they would hold 48 registers for the
        // whole particle loop of the fused kernel.
        const StagedNet *S = &net;
        asm volatile("" : "+s"(S));
        forward16_staged(*S, xop, out);
        // lanes 0..15 hold samples 16 g .. 16 g + 15; lane 16 g + j takes lane j's
        const int from = (l - 16 * g) & 63;
        const bool mine = l >= 16 * g && l < 16 * g + 16;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float v = __shfl(out[q], from, 64);
            if (mine) y[q] = v;
        }
    }
}

}  // namespace svamd
