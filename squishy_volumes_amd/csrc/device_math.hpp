// device_math.hpp - per-particle math shared by the substep kernels (gfx950).
//
// Matrices are 3x3 float arrays in column-major order, a[3*col + row], the
// layout of IoState's [[f32;3];3] (rust/crates/file_frame/src/particles.rs:100-106).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SV_DEV __device__ __forceinline__
#define SV_HD __host__ __device__ __forceinline__

namespace svamd {

// ------------------------------------------------------------ Morton codes
SV_HD uint32_t part1by2(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
SV_HD uint32_t compact1by2(uint32_t x) {
    x &= 0x09249249u;
    x = (x ^ (x >> 2)) & 0x030c30c3u;
    x = (x ^ (x >> 4)) & 0x0300f00fu;
    x = (x ^ (x >> 8)) & 0xff0000ffu;
    x = (x ^ (x >> 16)) & 0x000003ffu;
    return x;
}
SV_HD uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}

// x / h from y = 1/h rounded to nearest: the product and one correction by
// its exact residual.  By Markstein's theorem on division by reciprocal
// this is the correctly rounded quotient, the same as the IEEE division,
// for finite operands without overflow or underflow on the way (three
// operations instead of the division's ten).
SV_DEV float div_by(float x, float h, float y) {
    const float q = x * y;
    const float r = fmaf(-q, h, x);
    return fmaf(r, y, q);
}

// ------------------------------------------------- quadratic B-spline weights
// kernel_quadratic (rust/crates/cpu/src/kernels.rs:17-26) evaluated at the
// three nodes base, base+1, base+2 of a particle whose offset from its base
// node is fx = x/h - floor(x/h - 1/2), fx in [1/2, 3/2).
SV_DEV void bspline_weights(float fx, float w[3]) {
    float a = 1.5f - fx;
    float b = fx - 1.0f;
    float c = fx - 0.5f;
    w[0] = 0.5f * a * a;
    w[1] = 0.75f - b * b;
    w[2] = 0.5f * c * c;
}

// ------------------------------------------------------------ small algebra
SV_HD float det3(const float F[9]) {
    return F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
           F[6] * (F[1] * F[5] - F[4] * F[2]);
}
// C = A * B (column-major)
SV_HD void matmul3(const float A[9], const float B[9], float C[9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            C[3 * c + r] = A[r] * B[3 * c] + A[3 + r] * B[3 * c + 1] + A[6 + r] * B[3 * c + 2];
}
// C = A * B^T
SV_HD void matmul3_bt(const float A[9], const float B[9], float C[9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            C[3 * c + r] = A[r] * B[c] + A[3 + r] * B[3 + c] + A[6 + r] * B[6 + c];
}

SV_DEV float powi_dev(float x, int n) {
    // f32::powi for the small exponents of the inviscid model.
    int e = n < 0 ? -n : n;
    float r = 1.0f, b = x;
    while (e) {
        if (e & 1) r *= b;
        b *= b;
        e >>= 1;
    }
    return n < 0 ? 1.0f / r : r;
}

// ------------------------------------------------------------ 3x3 SVD
// Signed SVD F = U diag(s) V^T with U, V proper rotations, s0 >= s1 >= |s2|
// and s2 < 0 iff det F < 0.  Eigen-analysis of F^T F by cyclic Jacobi
// sweeps with exact rotations, then a Givens QR of F V (the structure of
// McAdams et al., "Computing the SVD of 3x3 matrices with minimal branching
// and elementary floating point operations", UW-Madison TR1690, 2011).
template <class T>
SV_DEV void jacobi_pq(T S[3][3], T V[3][3], int p, int q) {
    T apq = S[p][q];
    T app = S[p][p], aqq = S[q][q];
    // tan of the rotation angle, t = sign(theta) / (|theta| + sqrt(theta^2 + 1)).
    bool zero = apq == T(0);
    T theta = (aqq - app) / (T(2) * (zero ? T(1) : apq));
    T t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
    t = zero ? T(0) : t;
    T c = T(1) / sqrt(t * t + T(1));
    T s = t * c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T skp = S[k][p], skq = S[k][q];
        S[k][p] = c * skp - s * skq;
        S[k][q] = s * skp + c * skq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T spk = S[p][k], sqk = S[q][k];
        S[p][k] = c * spk - s * sqk;
        S[q][k] = s * spk + c * sqk;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T vkp = V[k][p], vkq = V[k][q];
        V[k][p] = c * vkp - s * vkq;
        V[k][q] = s * vkp + c * vkq;
    }
}

template <class T, int SWEEPS>
SV_DEV void svd3_impl(const T F[3][3], T U[3][3], T s[3], T V[3][3]) {
    T S[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
            V[i][j] = i == j ? T(1) : T(0);
        }
#pragma unroll 1
    for (int it = 0; it < SWEEPS; ++it) {
        jacobi_pq<T>(S, V, 0, 1);
        jacobi_pq<T>(S, V, 0, 2);
        jacobi_pq<T>(S, V, 1, 2);
    }
    T dv = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) - V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
           V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
    if (dv < T(0)) {
#pragma unroll
        for (int k = 0; k < 3; ++k) V[k][2] = -V[k][2];
    }
    T B[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) B[i][j] = F[i][0] * V[0][j] + F[i][1] * V[1][j] + F[i][2] * V[2][j];
    T n[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) n[j] = B[0][j] * B[0][j] + B[1][j] * B[1][j] + B[2][j] * B[2][j];
    // Order columns by decreasing norm; a swap negates one column so V stays a rotation.
    auto swap_cols = [&](int a, int b) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            T t = B[k][a]; B[k][a] = B[k][b]; B[k][b] = -t;
            t = V[k][a]; V[k][a] = V[k][b]; V[k][b] = -t;
        }
        T t = n[a]; n[a] = n[b]; n[b] = t;
    };
    if (n[0] < n[1]) swap_cols(0, 1);
    if (n[0] < n[2]) swap_cols(0, 2);
    if (n[1] < n[2]) swap_cols(1, 2);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) U[i][j] = i == j ? T(1) : T(0);
    auto givens = [&](int i, int j) {
        T a = B[i][i], b = B[j][i];
        T r = sqrt(a * a + b * b);
        bool ok = r > T(0);
        T c = ok ? a / r : T(1), sn = ok ? b / r : T(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            T bi = B[i][k], bj = B[j][k];
            B[i][k] = c * bi + sn * bj;
            B[j][k] = -sn * bi + c * bj;
            T ui = U[k][i], uj = U[k][j];
            U[k][i] = c * ui + sn * uj;
            U[k][j] = -sn * ui + c * uj;
        }
    };
    givens(0, 1);
    givens(0, 2);
    givens(1, 2);
    s[0] = B[0][0];
    s[1] = B[1][1];
    s[2] = B[2][2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (s[j] < T(0)) {
            s[j] = -s[j];
            s[2] = -s[2];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                U[k][j] = -U[k][j];
                U[k][2] = -U[k][2];
            }
        }
}

// The singular values of F alone, non-negative and in descending order: the
// Jacobi sweeps of svd3_impl give V, and the singular values are the column
// norms of F V (taken from F itself, not from the squared F^T F).  In the
// float form a lane stops sweeping once the off-diagonal of its rotated F^T F
// is below 1e-7 of its smallest diagonal entry, which bounds the change of
// every eigenvalue (Weyl) at about that fraction of the smallest; the wave
// stops once all its lanes have.
template <class T, int SWEEPS>
SV_DEV void singular_values_impl(const T F[3][3], T s[3]) {
    T S[3][3], V[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            S[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
            V[i][j] = i == j ? T(1) : T(0);
        }
#pragma unroll 1
    for (int it = 0; it < SWEEPS; ++it) {
        if constexpr (sizeof(T) == 4) {
            const T off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
            const T small = T(1e-7) * fmin(fmin(S[0][0], S[1][1]), S[2][2]);
            const bool done = off <= small * small;
            if (__all(done)) break;
            if (done) continue;
        }
        jacobi_pq<T>(S, V, 0, 1);
        jacobi_pq<T>(S, V, 0, 2);
        jacobi_pq<T>(S, V, 1, 2);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T n = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T b = F[i][0] * V[0][j] + F[i][1] * V[1][j] + F[i][2] * V[2][j];
            n += b * b;
        }
        s[j] = sqrt(n);
    }
    T t;
    if (s[0] < s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; }
    if (s[1] < s[2]) { t = s[1]; s[1] = s[2]; s[2] = t; }
    if (s[0] < s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; }
}

// Singular values of F (column-major), non-negative and descending, with the
// FP64 recomputation of svd3 when the smallest is below 1e-3 of the largest.
SV_DEV void singular_values3(const float Fc[9], float s[3]) {
    float F[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) F[r][c] = Fc[3 * c + r];
    singular_values_impl<float, 5>(F, s);
    if (s[2] < 1e-3f * s[0] || !(s[0] == s[0])) {
        double Fd[3][3], sd[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Fd[i][j] = (double)F[i][j];
        singular_values_impl<double, 12>(Fd, sd);
#pragma unroll
        for (int i = 0; i < 3; ++i) s[i] = (float)sd[i];
    }
}

// Float SVD with an FP64 recomputation when F is near degenerate: when the
// smallest singular value is below 1e-3 of the largest, or when the float
// factors reconstruct F with a relative error above 1e-4.  Returns true when
// the FP64 path ran.
SV_DEV bool svd3(const float Fc[9], float Uc[9], float s[3], float Vc[9]) {
    float F[3][3], U[3][3], V[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) F[r][c] = Fc[3 * c + r];
    svd3_impl<float, 5>(F, U, s, V);
    float fn = 0.f, err = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float rec = U[i][0] * s[0] * V[j][0] + U[i][1] * s[1] * V[j][1] + U[i][2] * s[2] * V[j][2];
            float d = rec - F[i][j];
            err += d * d;
            fn += F[i][j] * F[i][j];
        }
    bool degenerate = fabsf(s[2]) < 1e-3f * s[0] || err > 1e-8f * fn;
    if (degenerate) {
        double Fd[3][3], Ud[3][3], Vd[3][3], sd[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Fd[i][j] = (double)F[i][j];
        svd3_impl<double, 12>(Fd, Ud, sd, Vd);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            s[i] = (float)sd[i];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                U[i][j] = (float)Ud[i][j];
                V[i][j] = (float)Vd[i][j];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            Uc[3 * c + r] = U[r][c];
            Vc[3 * c + r] = V[r][c];
        }
    return degenerate;
}

// M = U diag(d) V^T
SV_DEV void recompose(const float U[9], const float d[3], const float V[9], float M[9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            M[3 * c + r] = U[r] * d[0] * V[c] + U[3 + r] * d[1] * V[3 + c] + U[6 + r] * d[2] * V[6 + c];
}

// ------------------------------------------------------- constitutive models
// Material selector carried in the particle flags.
// Model sets the substep kernels are specialised on.
enum : uint32_t { M_NH = 1, M_FLUID = 2, M_FCR = 4, M_SNOW = 8, M_SAND = 16, M_VISC = 32 };
// Not a model: the variant for batched scenes (the cull in each particle's
// own copy of the domain).
constexpr uint32_t M_BATCH = 64;
constexpr uint32_t M_COLL = 128;  // mesh colliders (collide phase, twin nodes)
constexpr uint32_t M_LEARN = 256; // the learned snow stress inside the fused gather-scatter
constexpr uint32_t SET_POLAR = M_NH | M_FLUID | M_FCR | M_VISC;  // no SVD in the hot path
constexpr uint32_t SET_ALL = M_NH | M_FLUID | M_FCR | M_SNOW | M_SAND | M_VISC;

enum : uint32_t {
    F_SOLID = 1u << 0,
    F_FLUID = 1u << 1,
    F_VISC = 1u << 2,
    F_SAND = 1u << 3,
    F_GOAL = 1u << 4,
    F_TOMB = 1u << 5,
    F_FAILED = 1u << 6,
    F_FCR = 1u << 8,
    F_SNOW = 1u << 9
};

// Kirchhoff stress tau = P F^T.
// Neo-Hookean, P of rust/crates/util/src/elastic.rs:299-311: since
// cof(F) F^T = J I, tau = mu (F F^T - I) + lambda ln(J) I.
SV_DEV void tau_neo_hookean(float mu, float lambda, const float F[9], float J, float tau[9]) {
    float FFt[9];
    matmul3_bt(F, F, FFt);
    float lj = lambda * logf(J);
#pragma unroll
    for (int k = 0; k < 9; ++k) tau[k] = mu * FFt[k];
    tau[0] += lj - mu;
    tau[4] += lj - mu;
    tau[8] += lj - mu;
}

// Inviscid fluid, P = K (1 - J^-gamma) cof F (elastic.rs:582-592): tau = K (1 - J^-gamma) J I.
SV_DEV void tau_inviscid(float K, int exponent, float J, float tau[9]) {
    float p = K * (1.0f - 1.0f / powi_dev(J, exponent)) * J;
#pragma unroll
    for (int k = 0; k < 9; ++k) tau[k] = 0.f;
    tau[0] = tau[4] = tau[8] = p;
}

// Polar factor R of F (det F > 0) by the scaled Newton iteration
// X <- (g X + X^-T / g) / 2, g = det(X)^(-1/3) (Higham, "Computing the polar
// decomposition with applications", SIAM J. Sci. Stat. Comput. 7, 1986).
// X^-T = cof(X) / det(X) with the cofactor columns of partial_invariant_3.
// Returns false when det F <= 0 or the iteration has not converged.
SV_DEV bool polar3(const float F[9], float R[9]) {
    float X[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) X[k] = F[k];
    bool ok = false;
#pragma unroll 1
    for (int it = 0; it < 10; ++it) {
        float cof[9];
        // columns: x1 x x2, x2 x x0, x0 x x1
        cof[0] = X[4] * X[8] - X[5] * X[7];
        cof[1] = X[5] * X[6] - X[3] * X[8];
        cof[2] = X[3] * X[7] - X[4] * X[6];
        cof[3] = X[7] * X[2] - X[8] * X[1];
        cof[4] = X[8] * X[0] - X[6] * X[2];
        cof[5] = X[6] * X[1] - X[7] * X[0];
        cof[6] = X[1] * X[5] - X[2] * X[4];
        cof[7] = X[2] * X[3] - X[0] * X[5];
        cof[8] = X[0] * X[4] - X[1] * X[3];
        const float det = X[0] * cof[0] + X[1] * cof[1] + X[2] * cof[2];
        if (!(det > 0.f)) return false;
        const float g = 1.0f / cbrtf(det);
        const float a = 0.5f * g, bb = 0.5f / (g * det);
        float delta = 0.f;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const float nx = a * X[k] + bb * cof[k];
            delta = fmaxf(delta, fabsf(nx - X[k]));
            X[k] = nx;
        }
        if (delta < 1e-6f) {
            ok = true;
            break;
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = X[k];
    return ok;
}

// Fixed-corotated (no reference counterpart): tau = 2 mu (F - R) F^T + lambda (J - 1) J I.
// From the SVD: (F - R) F^T = U diag(s^2 - s) U^T.
// The principal stresses 2 mu (s - 1) s + lambda (J - 1) J split into the
// linear part 2 mu e + lambda tr(e) and the second-order remainder
// 2 mu e^2 + lambda (e0 e1 + e0 e2 + e1 e2 + e0 e1 e2 + (J - 1)^2), e = s - 1
// (exact in FP32 near 1), J - 1 summed from the same products.  No term
// cancels, so both parts keep their relative precision however small the
// strain; s s - s and s0 s1 s2 - 1 carry absolute errors of about 1e-7,
// which dominate strains below 1e-5.
SV_DEV void fcr_principal_parts(float mu, float lambda, const float s[3], float lin[3], float rem[3]) {
    const float e0 = s[0] - 1.0f, e1 = s[1] - 1.0f, e2 = s[2] - 1.0f;
    const float tr = e0 + e1 + e2;
    const float p23 = (e0 * e1 + e0 * e2 + e1 * e2) + e0 * e1 * e2;
    const float jm1 = tr + p23;
    const float vol2 = lambda * (p23 + jm1 * jm1);
    const float e[3] = {e0, e1, e2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        lin[q] = 2.0f * mu * e[q] + lambda * tr;
        rem[q] = 2.0f * mu * e[q] * e[q] + vol2;
    }
}

SV_DEV void tau_fixed_corotated_svd(float mu, float lambda, const float U[9], const float s[3], float tau[9]) {
    // (s - 1) s and (J - 1) J from e = s - 1, without cancellation
    const float e0 = s[0] - 1.0f, e1 = s[1] - 1.0f, e2 = s[2] - 1.0f;
    float d[3] = {e0 + e0 * e0, e1 + e1 * e1, e2 + e2 * e2};
    recompose(U, d, U, tau);
    const float jm1 = (e0 + e1 + e2) + ((e0 * e1 + e0 * e2 + e1 * e2) + e0 * e1 * e2);
    float vol = lambda * (jm1 + jm1 * jm1);
#pragma unroll
    for (int k = 0; k < 9; ++k) tau[k] *= 2.0f * mu;
    tau[0] += vol;
    tau[4] += vol;
    tau[8] += vol;
}

// The same from the polar factor: (F - R) F^T = F F^T - R F^T.
SV_DEV void tau_fixed_corotated_polar(float mu, float lambda, const float F[9], const float R[9], float J,
                                      float tau[9]) {
    float FFt[9], RFt[9];
    matmul3_bt(F, F, FFt);
    matmul3_bt(R, F, RFt);
    const float vol = lambda * (J - 1.0f) * J;
#pragma unroll
    for (int k = 0; k < 9; ++k) tau[k] = 2.0f * mu * (FFt[k] - RFt[k]);
    tau[0] += vol;
    tau[4] += vol;
    tau[8] += vol;
}

SV_DEV float energy_neo_hookean(float mu, float lambda, const float F[9], float J) {
    float i2 = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) i2 += F[k] * F[k];
    float lj = logf(J);
    return mu / 2.f * (i2 - 3.f) - mu * lj + lambda / 2.f * lj * lj;
}
SV_DEV float energy_inviscid(float K, int exponent, float J) {
    float at_rest = K * (1.f - 1.f / (1.f - (float)exponent));
    return K * (J - powi_dev(J, 1 - exponent) / (1.f - (float)exponent)) - at_rest;
}
SV_DEV float energy_fixed_corotated_polar(float mu, float lambda, const float F[9], const float R[9], float J) {
    float d = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) d += (F[k] - R[k]) * (F[k] - R[k]);
    return mu * d + lambda / 2.f * (J - 1.f) * (J - 1.f);
}
SV_DEV float energy_fixed_corotated(float mu, float lambda, const float s[3]) {
    float J = s[0] * s[1] * s[2];
    float a = s[0] - 1.f, b = s[1] - 1.f, c = s[2] - 1.f;
    return mu * (a * a + b * b + c * c) + lambda / 2.f * (J - 1.f) * (J - 1.f);
}

// cauchy_stress_general_viscosity, rust/crates/util/src/elastic.rs:670-677.
SV_DEV void cauchy_viscosity(float dyn, float bulk, const float C[9], float out[9]) {
    float div = C[0] + C[4] + C[8];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) out[3 * c + r] = dyn * (C[3 * c + r] + C[3 * r + c]);
    out[0] += bulk * div;
    out[4] += bulk * div;
    out[8] += bulk * div;
}

}  // namespace svamd
