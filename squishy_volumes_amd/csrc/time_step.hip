// time_step.hip - the adaptive time-step limits of the reference
// (rust/crates/cpu/src/phase/limit_time_step.rs:25-223) as device
// reductions over the live particles of the engine's current state:
//   by_sound     min h / c, c = sqrt(kappa / rho), kappa the largest
//                diagonal-space stiffness of the stress over J (:36-130)
//   by_isolated  min of the isolated-particle bounds (:132-185)
//   by_velocity  0.5 h / max |v| (:199-208)
//   by_deformation min 0.2 / max(|C_ij|, 1e-8) (:210-222)
// Solids take the Neo-Hookean formulas and fluids the inviscid ones, as the
// reference does for its two parameter variants (the engine's fixed-corotated
// and snow solids are Solid there too).  A limit with no particle is +inf
// (the reference's None).
#include <cfloat>

#include "device_math.hpp"
#include "internal.hpp"

namespace svamd {
namespace {

constexpr float SEPARATION = 1e-5f;  // SINGULAR_VALUE_SEPARATION, rust/crates/util/src/consts.rs:14

// h / c for one particle, or +inf when kappa / rho is not a positive number.
__device__ float sound_limit(bool fluid, float p0, float p1, float m, float vol, const float F[9], float h) {
    // nalgebra's singular values: non-negative, in descending order
    float s[3];
    singular_values3(F, s);
    const float x = s[0], y = s[1], z = s[2];
    const float J = x * y * z;
    const float g[3] = {y * z, x * z, x * y};  // dJ/ds
    float d1, d2, diag;                          // dPsi/dJ, d2Psi/dJ2, the identity's coefficient
    if (fluid) {
        // inviscid: Psi'(J) = K (1 - J^-gamma), Psi''(J) = gamma K / J^(gamma + 1) (elastic.rs:555-570)
        const float K = p0;
        const int gamma = (int)p1;
        float Jg = 1.f;
        for (int k = 0; k < gamma; ++k) Jg *= J;
        d1 = K * (1.f - 1.f / Jg);
        d2 = (float)gamma * K / (Jg * J);
        diag = 0.f;
    } else {
        // Neo-Hookean: dPsi/dI2 = mu / 2, dPsi/dJ = (lambda ln J - mu) / J,
        // d2Psi/dJ2 = (lambda (1 - ln J) + mu) / J^2 (elastic.rs:276-295)
        const float mu = p0, lambda = p1;
        const float lnJ = logf(J);
        d1 = (lambda * lnJ - mu) / J;
        d2 = (lambda * (1.f - lnJ) + mu) / (J * J);
        diag = mu;  // (mu / 2) * d2I2/ds2 = (mu / 2) * 2 I
    }
    // first[i] = dPsi/ds_i; M = second derivative in diagonal space
    float P[3];
    const float sv[3] = {x, y, z};
#pragma unroll
    for (int i = 0; i < 3; ++i) P[i] = (fluid ? 0.f : diag * sv[i]) + d1 * g[i];
    // d2J/ds2 = [[0, z, y], [z, 0, x], [y, x, 0]]
    const float M11 = diag + d2 * g[0] * g[0], M22 = diag + d2 * g[1] * g[1], M33 = diag + d2 * g[2] * g[2];
    const float M21 = d2 * g[1] * g[0] + d1 * z, M32 = d2 * g[2] * g[1] + d1 * x, M13 = d2 * g[0] * g[2] + d1 * y;
    const bool xy = fabsf(x - y) < SEPARATION, yz = fabsf(y - z) < SEPARATION, zx = fabsf(z - x) < SEPARATION;
    float k[6];
    k[0] = x * x * M11;
    k[1] = y * y * M22;
    k[2] = z * z * M33;
    k[3] = y * y * (xy ? (P[0] + x * M11 - y * M21) / 2.f / x : (x * P[0] - y * P[1]) / (x * x - y * y));
    k[4] = y * z * (yz ? (P[1] + y * M22 - z * M32) / 2.f / y : (y * P[1] - z * P[2]) / (y * y - z * z));
    k[5] = z * x * (zx ? (P[2] + x * M33 - x * M13) / 2.f / z : (z * P[2] - x * P[0]) / (z * z - x * x));
    float kappa = k[0];
#pragma unroll
    for (int i = 1; i < 6; ++i) kappa = fmaxf(kappa, k[i]);
    kappa /= J;
    const float rho = (m / vol) / J;
    const float c2 = kappa / rho;
    if (!(c2 > 0.f) || !isfinite(c2)) return INFINITY;
    return h / sqrtf(c2);
}

__device__ float isolated_limit(bool fluid, float p0, float p1, float m, float vol, const float F[9], float h) {
    if (!fluid) {
        // Stability analysis of explicit MPM 3.12: xi = 3 / h^2, R = 1, K = 1, D = 3
        const float xi = 3.f / h / h;
        const float r = m / (vol * xi * (1.f - 0.5f) * (p0 + 1.5f * p1));
        return r > 0.f ? sqrtf(r) : INFINITY;
    }
    const float rho0 = m / vol;
    const float j = det3(F);
    const float K = p0;
    const int gamma = (int)p1;
    float jg = 1.f;
    for (int k = 0; k < gamma; ++k) jg *= j;
    const float first = K * (1.f - 1.f / jg);
    float r;
    if (fabsf(j - 1.f) > SEPARATION) {
        r = h / j * sqrtf(rho0 * (j - 1.f) / (6.f * first * 3.f));
    } else {
        const float second = (float)gamma * K / (jg * j);
        r = h * sqrtf(rho0 / (6.f * second * 3.f));
    }
    return r == r ? r : INFINITY;
}

// Per work-group: {min by_sound, min by_isolated, max |v|, min by_deformation};
// without motion the stored v and C are not read (0 and +inf in their place).
__global__ __launch_bounds__(256) void k_step_limits(ParticleArrays P, uint64_t n, float h, bool motion,
                                                    float4 *__restrict__ part, const DeviceScalars *ds) {
    if (ds && ds->halt) return;
    float snd = INFINITY, iso = INFINITY, vmax = 0.f, def = INFINITY;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t flags = P.flags(i);
        if (flags & F_TOMB) continue;
        float F[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = P.F(k, i);
        const bool fluid = (flags & F_FLUID) != 0;
        const float m = P.m(i), vol = P.vol(i), p0 = P.p0(i), p1 = P.p1(i);
        snd = fminf(snd, sound_limit(fluid, p0, p1, m, vol, F, h));
        iso = fminf(iso, isolated_limit(fluid, p0, p1, m, vol, F, h));
        if (motion) {
            const float v0 = P.v(0, i), v1 = P.v(1, i), v2 = P.v(2, i);
            vmax = fmaxf(vmax, sqrtf(v0 * v0 + v1 * v1 + v2 * v2));
#pragma unroll
            for (int k = 0; k < 9; ++k) def = fminf(def, 0.2f / fmaxf(fabsf(P.at(i)[(FA + k) * TILE]), 1e-8f));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        snd = fminf(snd, __shfl_xor(snd, o, 64));
        iso = fminf(iso, __shfl_xor(iso, o, 64));
        vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
        def = fminf(def, __shfl_xor(def, o, 64));
    }
    __shared__ float4 w[4];
    if ((threadIdx.x & 63u) == 0) w[threadIdx.x >> 6] = make_float4(snd, iso, vmax, def);
    __syncthreads();
    if (threadIdx.x == 0) {
        float4 r = w[0];
        for (int k = 1; k < 4; ++k) {
            r.x = fminf(r.x, w[k].x);
            r.y = fminf(r.y, w[k].y);
            r.z = fmaxf(r.z, w[k].z);
            r.w = fminf(r.w, w[k].w);
        }
        part[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void k_step_limits_final(const float4 *__restrict__ part, uint32_t groups,
                                                           float4 *__restrict__ out, const DeviceScalars *ds) {
    if (ds && ds->halt) return;
    float snd = INFINITY, iso = INFINITY, vmax = 0.f, def = INFINITY;
    for (uint32_t i = threadIdx.x; i < groups; i += blockDim.x) {
        const float4 r = part[i];
        snd = fminf(snd, r.x);
        iso = fminf(iso, r.y);
        vmax = fmaxf(vmax, r.z);
        def = fminf(def, r.w);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        snd = fminf(snd, __shfl_xor(snd, o, 64));
        iso = fminf(iso, __shfl_xor(iso, o, 64));
        vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
        def = fminf(def, __shfl_xor(def, o, 64));
    }
    __shared__ float4 w[4];
    if ((threadIdx.x & 63u) == 0) w[threadIdx.x >> 6] = make_float4(snd, iso, vmax, def);
    __syncthreads();
    if (threadIdx.x == 0) {
        float4 r = w[0];
        for (int k = 1; k < 4; ++k) {
            r.x = fminf(r.x, w[k].x);
            r.y = fminf(r.y, w[k].y);
            r.z = fmaxf(r.z, w[k].z);
            r.w = fminf(r.w, w[k].w);
        }
        *out = r;
    }
}

// The bounding box of the live particles' positions and their largest
// speed, for the grid window: per work-group {min x, max x, max |v|}, then one
// work-group over the partials.  NaN positions are skipped (fminf/fmaxf).
struct Bounds {
    float lo[3], hi[3], vmax, live;
};
__device__ __forceinline__ void bounds_merge(Bounds &r, const Bounds &o) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = fminf(r.lo[a], o.lo[a]);
        r.hi[a] = fmaxf(r.hi[a], o.hi[a]);
    }
    r.vmax = fmaxf(r.vmax, o.vmax);
    r.live = fmaxf(r.live, o.live);
}
__device__ __forceinline__ Bounds bounds_empty() {
    Bounds r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = INFINITY;
        r.hi[a] = -INFINITY;
    }
    r.vmax = 0.f;
    r.live = 0.f;
    return r;
}
__device__ Bounds bounds_block_reduce(Bounds r) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Bounds q;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            q.lo[a] = __shfl_xor(r.lo[a], o, 64);
            q.hi[a] = __shfl_xor(r.hi[a], o, 64);
        }
        q.vmax = __shfl_xor(r.vmax, o, 64);
        q.live = __shfl_xor(r.live, o, 64);
        bounds_merge(r, q);
    }
    __shared__ Bounds w[4];
    if ((threadIdx.x & 63u) == 0) w[threadIdx.x >> 6] = r;
    __syncthreads();
    Bounds t = w[0];
    for (int k = 1; k < 4; ++k) bounds_merge(t, w[k]);
    return t;
}

__global__ __launch_bounds__(256) void k_bounds(ParticleArrays P, uint64_t n, Bounds *__restrict__ part) {
    Bounds r = bounds_empty();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (P.flags(i) & F_TOMB) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            r.lo[a] = fminf(r.lo[a], P.x(a, i));
            r.hi[a] = fmaxf(r.hi[a], P.x(a, i));
        }
        const float v0 = P.v(0, i), v1 = P.v(1, i), v2 = P.v(2, i);
        r.vmax = fmaxf(r.vmax, sqrtf(v0 * v0 + v1 * v1 + v2 * v2));
        r.live = 1.f;
    }
    const Bounds t = bounds_block_reduce(r);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_bounds_final(const Bounds *__restrict__ part, uint32_t groups,
                                                      float *__restrict__ out) {
    Bounds r = bounds_empty();
    for (uint32_t i = threadIdx.x; i < groups; i += blockDim.x) bounds_merge(r, part[i]);
    const Bounds t = bounds_block_reduce(r);
    if (threadIdx.x < 8u) {
        const float v[8] = {t.lo[0], t.lo[1], t.lo[2], t.hi[0], t.hi[1], t.hi[2], t.vmax, t.live};
        out[threadIdx.x] = v[threadIdx.x];
    }
}

}  // namespace

hipError_t launch_particle_bounds(const ParticleArrays &P, uint64_t n, float4 *part, float *out, hipStream_t stream) {
    static_assert(sizeof(Bounds) == 2 * sizeof(float4), "two float4 per partial");
    const uint32_t groups = step_limit_groups(n);
    hipLaunchKernelGGL(k_bounds, dim3(groups), dim3(256), 0, stream, P, n, reinterpret_cast<Bounds *>(part));
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, stream, reinterpret_cast<const Bounds *>(part), groups,
                       out);
    return hipGetLastError();
}

uint32_t step_limit_groups(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 2048)); }

hipError_t launch_step_limits(const ParticleArrays &P, uint64_t n, float h, bool motion, float4 *part, float4 *out,
                              const DeviceScalars *ds, hipStream_t stream) {
    const uint32_t groups = step_limit_groups(n);
    hipLaunchKernelGGL(k_step_limits, dim3(groups), dim3(256), 0, stream, P, n, h, motion, part, ds);
    hipLaunchKernelGGL(k_step_limits_final, dim3(1), dim3(256), 0, stream, part, groups, out, ds);
    return hipGetLastError();
}

namespace {

// ---- AdaptiveTimeStepState on the device (adaptive_time_step_state.rs:13-70)

constexpr uint32_t FAIL_BITS = MPM_STATUS_ENERGY_ERROR | MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED;

// A limit of the reductions: +inf (no particle) or NaN is the reference's
// None, which the state counts as f32::MAX.
__device__ __forceinline__ float as_limit(float v) { return isinf(v) || isnan(v) ? FLT_MAX : v; }

__device__ __forceinline__ float without_prior(const DeviceAdaptive &A) {
    return fminf(fminf(fminf(A.max_time_step, A.by_velocity), fminf(A.by_deformation, A.by_sound)), A.by_isolated);
}
__device__ __forceinline__ float allowed(const DeviceAdaptive &A) {
    float a = without_prior(A);
    for (uint32_t k = 0; k < A.nprior; ++k) a = fminf(a, A.prior[k]);
    return a;
}
// push_current_limit: the history keeps the last 11 pushes (pop_front past 10).
__device__ __forceinline__ void push(DeviceAdaptive &A) {
    const float v = without_prior(A);
    if (A.nprior < 11u) {
        A.prior[(A.head + A.nprior) % 11u] = v;
        ++A.nprior;
    } else {
        A.prior[A.head] = v;
        A.head = (A.head + 1u) % 11u;
    }
}
__device__ __forceinline__ void halt(DeviceAdaptive &A, DeviceScalars *ds, uint32_t why) {
    A.stop = why;
    ds->halt = 1u;
}

// The start of a substep: the call ends when the last gather failed a
// particle (an energy error leaves the time at the start of the failing
// substep, cpu_state.rs:178-190), when the target time is reached, or with
// a zero step or a time outside frame a; else entry 0 gets the frame
// factor, the interpolated gravity (interpolate_input.rs:24-28, in the same
// f64 -> f32 order as the host) and the step of the external force.
__global__ void k_adapt_begin(DeviceAdaptive *__restrict__ Ap, SubstepParams *__restrict__ table,
                              DeviceScalars *__restrict__ ds) {
    DeviceAdaptive &A = *Ap;
    if (ds->halt) return;
    if (ds->status & FAIL_BITS) {
        if ((ds->status & MPM_STATUS_ENERGY_ERROR) && A.substeps > 0) A.time -= (double)A.last_advance;
        halt(A, ds, ADAPT_FAILED);
        return;
    }
    if (!(A.time < A.target)) {
        halt(A, ds, ADAPT_DONE);
        return;
    }
    const float dt = allowed(A);
    if (dt == 0.f) {
        halt(A, ds, ADAPT_ZERO_STEP);
        return;
    }
    const double ft = A.time * A.fps;
    if ((uint64_t)floor(ft) != A.frame) {
        halt(A, ds, ADAPT_WRONG_FRAME);
        return;
    }
    SubstepParams &P = table[0];
    P.factor_b = (float)fmod(ft, 1.0);
    const float fa = 1.0f - P.factor_b;
    for (int a = 0; a < 3; ++a) P.gravity[a] = fa * A.gravity_a[a] + P.factor_b * A.gravity_b[a];
    P.dt_force = dt;
    P.epoch = A.epoch0 + A.substeps;
    P.index = 0u;
    A.dt_force = dt;
    ds->param_index = 0u;
    ds->grid_param = 0u;
}

// limit_time_step_before_force (limit_time_step.rs:25-34): the sound-speed
// and isolated-particle limits of the state, pushed; the allowed step is the
// step of the scatter and the grid update (entry 0) and, until the second
// push, of the advance (entry 1, a copy of entry 0 otherwise).
__global__ void k_adapt_force(DeviceAdaptive *__restrict__ Ap, SubstepParams *__restrict__ table,
                              DeviceScalars *__restrict__ ds, const float4 *__restrict__ limits) {
    DeviceAdaptive &A = *Ap;
    if (ds->halt) return;
    A.by_sound = as_limit(limits->x);
    A.by_isolated = as_limit(limits->y);
    push(A);
    const float dt = allowed(A);
    if (dt == 0.f) {
        halt(A, ds, ADAPT_ZERO_STEP);
        return;
    }
    table[0].dt = dt;
    table[1] = table[0];
    table[1].index = 1u;
    table[1].epoch = table[0].epoch + 1u;
    for (int k = 0; k < 3; ++k) ds->limit_bits[k] = 0u;
}

// limit_time_step_before_integrate (limit_time_step.rs:187-222): the
// velocity and deformation limits of the gathered v and C, pushed; the
// allowed step is the step of the advance and of the time (cpu_state.rs:
// 187-190).  A particle that failed in the scatter (a stored state too close
// to inversion) ends the call before the advance; so does an overflow of the
// collider twins, after which the call runs again with a larger table.
__global__ void k_adapt_integrate(DeviceAdaptive *__restrict__ Ap, SubstepParams *__restrict__ table,
                                  DeviceScalars *__restrict__ ds) {
    DeviceAdaptive &A = *Ap;
    if (ds->halt) return;
    if (ds->status & FAIL_BITS) {
        halt(A, ds, ADAPT_FAILED);
        return;
    }
    if (ds->status & MPM_STATUS_TABLE_TRIES_EXCEEDED) {
        halt(A, ds, ADAPT_TWIN_OVERFLOW);
        return;
    }
    const float vmax = __uint_as_float(ds->limit_bits[0]), cmax = __uint_as_float(ds->limit_bits[1]);
    const bool any = ds->limit_bits[2] != 0u;
    A.by_velocity = any && vmax != 0.f ? as_limit(0.5f * A.h / vmax) : FLT_MAX;
    A.by_deformation = any ? as_limit(0.2f / fmaxf(cmax, 1e-8f)) : FLT_MAX;
    push(A);
    const float dt = allowed(A);
    if (dt == 0.f) {
        halt(A, ds, ADAPT_ZERO_STEP);
        return;
    }
    table[1].dt_force = dt;
    A.time += (double)dt;
    A.last_advance = dt;
    A.substeps += 1u;
}

}  // namespace

hipError_t launch_adapt_begin(DeviceAdaptive *A, SubstepParams *table, DeviceScalars *ds, hipStream_t stream) {
    hipLaunchKernelGGL(k_adapt_begin, dim3(1), dim3(1), 0, stream, A, table, ds);
    return hipGetLastError();
}
hipError_t launch_adapt_force(DeviceAdaptive *A, SubstepParams *table, DeviceScalars *ds, const float4 *limits,
                              hipStream_t stream) {
    hipLaunchKernelGGL(k_adapt_force, dim3(1), dim3(1), 0, stream, A, table, ds, limits);
    return hipGetLastError();
}
hipError_t launch_adapt_integrate(DeviceAdaptive *A, SubstepParams *table, DeviceScalars *ds, hipStream_t stream) {
    hipLaunchKernelGGL(k_adapt_integrate, dim3(1), dim3(1), 0, stream, A, table, ds);
    return hipGetLastError();
}

}  // namespace svamd
