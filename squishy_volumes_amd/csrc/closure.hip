// closure.hip - the learned constitutive closure: a 6-64-64-3 perceptron
// with BF16 matrix-core products and FP32 accumulation, its training
// kernels (loss gradient, Adam) and its evaluation as the stress of snow
// particles.  The reference has no learned stress model; the analytic model
// it replaces is the fixed-corotated snow stress of the engine
// (tau_fixed_corotated_svd in device_math.hpp).
//
// Products use v_mfma_f32_16x16x16_bf16 on wave64.  With l the lane,
//   A (16 x 16): lane l holds A[l & 15][4 (l >> 4) + i], i < 4,
//   B (16 x 16): lane l holds B[4 (l >> 4) + i][l & 15],
//   D (16 x 16): lane l holds D[4 (l >> 4) + i][l & 15].
// A wave takes 16 samples at a time and evaluates every layer transposed,
// D^T = W X^T (rows = units, columns = samples): the tile a lane holds of one
// layer's output is then exactly the B operand the next layer wants from it,
// so the forward pass keeps its activations in registers.  The backward pass
// stages the activations in LDS, since its weight gradients contract over
// the samples.
//
// Rounding points (the CPU oracle rounds at the same places): weights,
// inputs, hidden activations and back-propagated errors are rounded to BF16
// (nearest even) where they enter a product; biases, accumulations and the
// outputs are FP32.
#include "closure_net.hpp"
#include "internal.hpp"

namespace svamd {
namespace {

// One layer tile set: tile t holds rows 16 t + 4 (l >> 4) + i, column l & 15.
struct Tiles {
    f32x4 t[4];
};

// Forward pass of 16 samples whose features the lanes hold as the B operand
// xop (lane l: features 4 (l >> 4) .. + 3 of sample l & 15).  h1, h2 get the
// BF16-rounded activations; lanes l < 16 get sample l's outputs in y.x..y.z.
__device__ __forceinline__ void forward16(const float *__restrict__ P, bf16x4 xop, Tiles &h1, Tiles &h2, f32x4 &y) {
    const int l = (int)(threadIdx.x & 63u), r0 = 4 * (l >> 4), c = l & 15;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = mfma(op4(P + MLP_OFF_W1 + (16 * t + c) * MLP_KIN + r0, 1), xop, zero4());
#pragma unroll
        for (int i = 0; i < 4; ++i) h1.t[t][i] = bf16_round(fmaxf(acc[i] + P[MLP_OFF_B1 + 16 * t + r0 + i], 0.f));
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = zero4();
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
            acc = mfma(op4(P + MLP_OFF_W2 + (16 * t + c) * MLP_H + 16 * kk + r0, 1), pack(h1.t[kk]), acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) h2.t[t][i] = bf16_round(fmaxf(acc[i] + P[MLP_OFF_B2 + 16 * t + r0 + i], 0.f));
    }
    f32x4 acc = zero4();
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) acc = mfma(op4(P + MLP_OFF_W3 + c * MLP_H + 16 * kk + r0, 1), pack(h2.t[kk]), acc);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = acc[i] + P[MLP_OFF_B3 + r0 + i];
}

// The B operand of 16 samples' features, from row-major x[n][MLP_IN].
__device__ __forceinline__ bf16x4 feature_op(const float *__restrict__ x, uint64_t s0, uint64_t n) {
    const int l = (int)(threadIdx.x & 63u), r0 = 4 * (l >> 4), c = l & 15;
    const uint64_t s = s0 + (uint64_t)c;
    float f[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) f[i] = (s < n && r0 + i < MLP_IN) ? x[s * MLP_IN + r0 + i] : 0.f;
    bf16x4 r;
    r.x = to_bf16(f[0]);
    r.y = to_bf16(f[1]);
    r.z = to_bf16(f[2]);
    r.w = to_bf16(f[3]);
    return r;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ x,
                                                    float *__restrict__ y, uint64_t n) {
    // The weights are staged in LDS once per work-group, as BF16 MFMA operands.
    __shared__ StagedNet net;
    stage_net(P, net);
    __syncthreads();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const int l = (int)(threadIdx.x & 63u);
    for (uint64_t g = wave; g * 16 < n; g += nwaves) {
        f32x4 out;
        forward16_staged(net, feature_op(x, g * 16, n), out);
        const uint64_t s = g * 16 + (uint64_t)l;
        if (l < 16 && s < n) {
#pragma unroll
            for (int i = 0; i < MLP_OUT; ++i) y[s * MLP_OUT + i] = out[i];
        }
    }
}

// Loss sum((y - t)^2) and its gradient scale * d/dP over n samples, added to
// loss and grad.  scale = 1 / N makes grad the gradient of the mean over N.
// Each wave keeps its weight-gradient tiles in registers over all of its
// sample groups and adds them to grad once.
constexpr int GRAD_WAVES = 4;
// Rows are padded (16 -> 20 and 64 -> 68 floats) so that the 16 lanes that
// write one column of a 16-sample tile, and the four lane groups that read
// four rows of it, fall on different LDS banks.
constexpr int XS = 20, HS = 68;
struct WaveStage {
    float x[16 * XS];   // [sample][feature]
    float h1[16 * HS];  // [sample][unit]
    float h2[16 * HS];
    float dy[16 * XS];  // [sample][output], BF16-rounded
    float d2[16 * HS];  // dL/dz2, BF16-rounded
    float d1[16 * HS];  // dL/dz1, BF16-rounded
};

// Two waves per SIMD: the register budget this sets (256 per lane, the
// gradient tiles included) costs a few spilled registers, and the second wave
// hides the first one's LDS and weight-load latency (1.27 -> 0.94 ms over
// 4.2M samples against one wave per SIMD).
__global__ __launch_bounds__(64 * GRAD_WAVES, 2) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ x,
                                                            const float *__restrict__ tgt, uint64_t n, float scale,
                                                            float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ WaveStage stage[GRAD_WAVES];
    WaveStage &S = stage[threadIdx.x >> 6];
    const int l = (int)(threadIdx.x & 63u), r0 = 4 * (l >> 4), c = l & 15;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;

    f32x4 g1[4], g2[4][4], g3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        g1[t] = zero4();
        g3[t] = zero4();
#pragma unroll
        for (int u = 0; u < 4; ++u) g2[t][u] = zero4();
    }
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;

    for (uint64_t g = wave; g * 16 < n; g += nwaves) {
        const bf16x4 xop = feature_op(x, g * 16, n);
#pragma unroll
        for (int i = 0; i < 4; ++i) S.x[c * XS + r0 + i] = __uint_as_float(((uint32_t)(uint16_t)xop[i]) << 16);
        Tiles h1, h2;
        f32x4 y;
        forward16(P, xop, h1, h2, y);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                S.h1[c * HS + 16 * t + r0 + i] = h1.t[t][i];
                S.h2[c * HS + 16 * t + r0 + i] = h2.t[t][i];
            }
        // Output error: lanes l < 16 hold sample l's outputs (rows r0 + i = i).
        const uint64_t s = g * 16 + (uint64_t)c;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float d = 0.f;
            if (r0 == 0 && i < MLP_OUT && s < n) {
                d = y[i] - tgt[s * MLP_OUT + i];
                lsum += d * d;
            }
            S.dy[c * XS + r0 + i] = bf16_round(2.f * scale * d);
        }
        __builtin_amdgcn_wave_barrier();
        // dz2^T = (W3^T dy^T) masked by relu'(z2).
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const f32x4 acc = mfma(op4(P + MLP_OFF_W3 + r0 * MLP_H + 16 * t + c, MLP_H), op4(S.dy + c * XS + r0, 1), zero4());
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = c * HS + 16 * t + r0 + i;
                S.d2[k] = S.h2[k] > 0.f ? bf16_round(acc[i]) : 0.f;
            }
        }
        __builtin_amdgcn_wave_barrier();
        // dz1^T = (W2^T dz2^T) masked by relu'(z1).
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = zero4();
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                acc = mfma(op4(P + MLP_OFF_W2 + (16 * kk + r0) * MLP_H + 16 * t + c, MLP_H), op4(S.d2 + c * HS + 16 * kk + r0, 1),
                           acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = c * HS + 16 * t + r0 + i;
                S.d1[k] = S.h1[k] > 0.f ? bf16_round(acc[i]) : 0.f;
            }
        }
        __builtin_amdgcn_wave_barrier();
        // Weight gradients, contracted over the 16 samples.
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            g3[t] = mfma(op4(S.dy + r0 * XS + c, XS), op4(S.h2 + r0 * HS + 16 * t + c, HS), g3[t]);
            g1[t] = mfma(op4(S.d1 + r0 * HS + 16 * t + c, HS), op4(S.x + r0 * XS + c, XS), g1[t]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                g2[t][u] = mfma(op4(S.d2 + r0 * HS + 16 * t + c, HS), op4(S.h1 + r0 * HS + 16 * u + c, HS), g2[t][u]);
        }
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            gb1 += S.d1[q * HS + l];
            gb2 += S.d2[q * HS + l];
            if (l < MLP_OUT) gb3 += S.dy[q * XS + l];
        }
        __builtin_amdgcn_wave_barrier();
    }

    // Tile (t, u) of dW2 holds rows 16 t + r0 + i (unit of layer 2), columns
    // 16 u + c (unit of layer 1); dW1 rows 16 t + r0 + i, columns c (feature);
    // dW3 rows r0 + i (output), columns 16 t + c (unit).
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = 16 * t + r0 + i;
            if (c < MLP_IN) atomicAdd(grad + MLP_OFF_W1 + row * MLP_KIN + c, g1[t][i]);
            if (r0 + i < MLP_OUT) atomicAdd(grad + MLP_OFF_W3 + (r0 + i) * MLP_H + 16 * t + c, g3[t][i]);
#pragma unroll
            for (int u = 0; u < 4; ++u) atomicAdd(grad + MLP_OFF_W2 + row * MLP_H + 16 * u + c, g2[t][u][i]);
        }
    atomicAdd(grad + MLP_OFF_B1 + l, gb1);
    atomicAdd(grad + MLP_OFF_B2 + l, gb2);
    if (l < MLP_OUT) atomicAdd(grad + MLP_OFF_B3 + l, gb3);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (l == 0) atomicAdd(loss, lsum);
}

// Adam on the FP32 master parameters; clears the gradient.  bc1, bc2 are
// 1 / (1 - beta^step).
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ G, float *__restrict__ M, float *__restrict__ V,
                           int n, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const float g = G[k];
    const float m = b1 * M[k] + (1.f - b1) * g;
    const float v = b2 * V[k] + (1.f - b2) * g * g;
    M[k] = m;
    V[k] = v;
    P[k] -= lr * (m * bc1) / (sqrtf(v * bc2) + eps);
    G[k] = 0.f;
}

// ------------------------------------------------- the closure in the substep
// Kirchhoff stress of every live snow particle from the network, written
// by slot (row k at tau[k * stride + slot]) for the scatter that follows.
// Per lane: the SVD of F, the network inputs and the training-hull test; per
// wave: the network on its 64 particles, 16 at a time, with the work-group's
// copy of the operands in LDS.  A particle whose
// inputs leave the hull takes the analytic stress; the particles with
// id % judge_stride == judge_phase take both, and the difference is summed
// into the statistics.
struct ClosureArgs {
    ParticleArrays P;
    const uint32_t *order;  // slot i reads P at order[i] (the scatter permutes), or null
    AuxArrays X;
    const float *params;
    float *tau;
    uint64_t stride, n;
    float hull_lo[MLP_IN], hull_hi[MLP_IN];
    float hardening;
    uint32_t judge_stride, judge_phase;
    ClosureStats *stats;
    const float *svd;  // U, s by P slot (12 rows of stride), or null
    const DeviceScalars *ds;  // halt flag, or null
};

__global__ __launch_bounds__(256) void k_closure(ClosureArgs a) {
    if (a.ds && a.ds->halt) return;  // a substep past the end of an adaptive call's loop
    __shared__ StagedNet net;
    __shared__ float feat[4][64 * 8];
    __shared__ float outs[4][64 * 4];
    stage_net(a.params, net);
    __syncthreads();
    const int l = (int)(threadIdx.x & 63u), w = (int)(threadIdx.x >> 6), r0 = 4 * (l >> 4), c = l & 15;
    uint32_t evaluated = 0, fallbacks = 0, judged = 0;
    double err2 = 0.0, ref2 = 0.0;
    for (uint64_t base = ((uint64_t)blockIdx.x * 4 + (uint64_t)w) * 64; base < a.n; base += (uint64_t)gridDim.x * 256) {
        const uint64_t i = base + (uint64_t)l;
        bool snow = false, inside = false;
        float U[9], s[3] = {1.f, 1.f, 1.f}, x[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float mu = 0.f, lam = 0.f, E = 0.f, hard = 1.f;
        uint32_t id = 0;
        if (i < a.n) {
            const uint64_t src = a.order ? a.order[i] : i;
            const uint32_t flags = a.P.flags(src);
            snow = (flags & F_SNOW) && !(flags & F_TOMB);
            if (snow) {
                if (a.svd) {
                    // the factors the gather formed F from
#pragma unroll
                    for (int k = 0; k < 9; ++k) U[k] = a.svd[(uint64_t)k * a.stride + src];
#pragma unroll
                    for (int q = 0; q < 3; ++q) s[q] = a.svd[(uint64_t)(9 + q) * a.stride + src];
                } else {
                    float F[9], V[9];
#pragma unroll
                    for (int k = 0; k < 9; ++k) F[k] = a.P.F(k, src);
                    svd3(F, U, s, V);
                }
                mu = a.P.p0(src);
                lam = a.P.p1(src);
                id = a.P.id(src);
                const float jp = a.X.jp[id];
                hard = expf(a.hardening * (1.0f - jp));
                closure_features(s, mu, lam, jp, x, E);
                inside = true;
#pragma unroll
                for (int q = 0; q < MLP_IN; ++q) inside = inside && x[q] >= a.hull_lo[q] && x[q] <= a.hull_hi[q];
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) feat[w][l * 8 + q] = q < MLP_IN ? x[q] : 0.f;
        __builtin_amdgcn_wave_barrier();
        if (__ballot(inside)) {
#pragma unroll 1
            for (int g = 0; g < 4; ++g) {
                // Rows 8..15 of the 16-row feature operand are zero padding
                // (the address stays inside the sample's 8 staged values).
                bf16x4 xop;
                const float *fp = &feat[w][(16 * g + c) * 8 + (r0 & 7)];
                const bool real = r0 < 8;
                xop.x = real ? to_bf16(fp[0]) : (short)0;
                xop.y = real ? to_bf16(fp[1]) : (short)0;
                xop.z = real ? to_bf16(fp[2]) : (short)0;
                xop.w = real ? to_bf16(fp[3]) : (short)0;
                f32x4 y;
                forward16_staged(net, xop, y);
                if (l < 16) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) outs[w][(16 * g + l) * 4 + q] = y[q];
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (snow) {
            float tau[9];
            if (inside) {
                // linear-elastic part (FP32, exact in s) plus the learned remainder
                const float y[3] = {outs[w][l * 4], outs[w][l * 4 + 1], outs[w][l * 4 + 2]};
                closure_tau(U, s, mu, lam, hard, E, y, tau);
                ++evaluated;
            } else {
                tau_fixed_corotated_svd(mu * hard, lam * hard, U, s, tau);
                ++fallbacks;
            }
            if (a.judge_stride && id % a.judge_stride == a.judge_phase) {
                float ref[9];
                tau_fixed_corotated_svd(mu * hard, lam * hard, U, s, ref);
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    err2 += (double)(tau[k] - ref[k]) * (double)(tau[k] - ref[k]);
                    ref2 += (double)ref[k] * (double)ref[k];
                }
                ++judged;
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) a.tau[(uint64_t)k * a.stride + i] = tau[k];
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        evaluated += __shfl_xor(evaluated, o, 64);
        fallbacks += __shfl_xor(fallbacks, o, 64);
        judged += __shfl_xor(judged, o, 64);
        err2 += __shfl_xor(err2, o, 64);
        ref2 += __shfl_xor(ref2, o, 64);
    }
    // Work-group w adds its sums to slot w with a plain read-add-write: one
    // launch touches a slot once, and launches on the stream are ordered
    // (atomics on five common addresses would serialize in memory).
    __shared__ ClosureStats part[4];
    if (l == 0) part[w] = ClosureStats{evaluated, fallbacks, judged, 0ull, err2, ref2};
    __syncthreads();
    if (threadIdx.x == 0) {
        ClosureStats &S = a.stats[blockIdx.x];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            S.evaluated += part[q].evaluated;
            S.fallbacks += part[q].fallbacks;
            S.judged += part[q].judged;
            S.err2 += part[q].err2;
            S.ref2 += part[q].ref2;
        }
    }
}

// Training pairs from the current state: every live snow particle with
// id % stride == phase contributes its network inputs (as the closure kernel
// forms them) and, as the target, its analytic principal Kirchhoff stresses
// 2 mu h (s - 1) s + lambda h (J - 1) J minus their linear part
// h (2 mu (s - 1) + lambda tr(s - 1)), in units of E h CLOSURE_STRESS_SCALE m
// (h the hardening factor, m the strain scale): what the network adds in the
// closure.  A wave reserves its
// rows with one atomic; rows past the capacity are counted, not written.
__global__ __launch_bounds__(256) void k_closure_pairs(ParticleArrays P, AuxArrays X, uint64_t n, float hardening,
                                                       uint32_t stride, uint32_t phase, float *__restrict__ xo,
                                                       float *__restrict__ to, uint64_t capacity,
                                                       unsigned long long *count) {
    const int l = (int)(threadIdx.x & 63u);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i - (uint64_t)l < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        bool take = false;
        float x[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, t[3] = {0.f, 0.f, 0.f};
        if (i < n) {
            const uint32_t flags = P.flags(i);
            const uint32_t id = P.id(i);
            if ((flags & F_SNOW) && !(flags & F_TOMB) && id % stride == phase) {
                float F[9], U[9], s[3], V[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) F[k] = P.F(k, i);
                svd3(F, U, s, V);
                const float mu = P.p0(i), lam = P.p1(i);
                const float jp = X.jp[id];
                const float hard = expf(hardening * (1.0f - jp));
                const float E = mu * (3.f * lam + 2.f * mu) / (lam + mu);
                const float nu = lam / (2.f * (lam + mu));
#pragma unroll
                for (int q = 0; q < 3; ++q) x[q] = (s[q] - 1.0f) * (1.0f / CLOSURE_SIGMA_SCALE);
                x[3] = log10f(E * 1e-5f);
                x[4] = (nu - 0.3f) * 10.0f;
                x[5] = (jp - 1.0f) * (1.0f / CLOSURE_JP_SCALE);
                const float unit = E * hard * CLOSURE_STRESS_SCALE * closure_strain_scale(s);
                float lin[3], rem[3];
                fcr_principal_parts(mu * hard, lam * hard, s, lin, rem);
#pragma unroll
                for (int q = 0; q < 3; ++q) t[q] = rem[q] / unit;
                take = true;
            }
        }
        const uint64_t mask = __ballot(take);
        if (!mask) continue;
        unsigned long long base = 0;
        if (l == __ffsll((unsigned long long)mask) - 1) base = atomicAdd(count, (unsigned long long)__popcll(mask));
        base = __shfl(base, __ffsll((unsigned long long)mask) - 1, 64);
        const uint64_t row = base + (uint64_t)__popcll(mask & ((1ull << l) - 1ull));
        if (take && row < capacity) {
#pragma unroll
            for (int q = 0; q < MLP_IN; ++q) xo[row * MLP_IN + q] = x[q];
#pragma unroll
            for (int q = 0; q < 3; ++q) to[row * 3 + q] = t[q];
        }
    }
}

__global__ __launch_bounds__(256) void k_pack_net(const float *__restrict__ P, StagedNet *__restrict__ out) {
    stage_net(P, *out);
}

}  // namespace

hipError_t launch_pack_net(const float *params, StagedNet *out, hipStream_t stream) {
    hipLaunchKernelGGL(k_pack_net, dim3(1), dim3(256), 0, stream, params, out);
    return hipGetLastError();
}

hipError_t launch_closure_pairs(const ParticleArrays &P, const AuxArrays &X, uint64_t n, float hardening,
                                uint32_t stride, uint32_t phase, float *x, float *t, uint64_t capacity,
                                unsigned long long *count, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_closure_pairs, dim3(blocks), dim3(256), 0, stream, P, X, n, hardening, stride, phase, x, t,
                       capacity, count);
    return hipGetLastError();
}

hipError_t launch_closure(const ClosureLaunch &c, hipStream_t stream) {
    if (c.n == 0) return hipSuccess;
    ClosureArgs a;
    a.P = c.P;
    a.order = c.order;
    a.X = c.X;
    a.params = c.params;
    a.tau = c.tau;
    a.stride = c.stride;
    a.n = c.n;
    for (int q = 0; q < MLP_IN; ++q) {
        a.hull_lo[q] = c.hull_lo[q];
        a.hull_hi[q] = c.hull_hi[q];
    }
    a.hardening = c.hardening;
    a.judge_stride = c.judge_stride;
    a.judge_phase = c.judge_phase;
    a.stats = c.stats;
    a.svd = c.svd;
    a.ds = c.ds;
    // Each work-group stages the network once and then takes 256 particles
    // at a time: 2048 work-groups (8 per CU) amortize the staging (one
    // work-group per 256 particles measured 10 % slower on 1M particles;
    // 512 to 2048 measure alike).
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((c.n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_closure, dim3(blocks), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float *P, const float *x, float *y, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t groups = (n + 15) / 16;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((groups + 3) / 4, 4096);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, stream, P, x, y, n);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float *P, const float *x, const float *t, uint64_t n, float scale, float *grad,
                           float *loss, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    // Each wave takes several sample groups so that its gradient tiles are
    // added to grad a few hundred times per call, not once per group.
    const uint64_t groups = (n + 15) / 16;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((groups + 8 * GRAD_WAVES - 1) / (8 * GRAD_WAVES), 512);
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64 * GRAD_WAVES), 0, stream, P, x, t, n, scale, grad, loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *P, float *G, float *M, float *V, float lr, float b1, float b2, float eps, int step,
                           hipStream_t stream) {
    const float bc1 = 1.f / (1.f - powf(b1, (float)step)), bc2 = 1.f / (1.f - powf(b2, (float)step));
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, stream, P, G, M, V, MLP_NPARAM, lr, b1,
                       b2, eps, bc1, bc2);
    return hipGetLastError();
}

}  // namespace svamd
