// transfer.hip - P2G, grid update and G2P of the MLS-MPM substep (gfx950).
//
// Dataflow of one substep (after the block sort):
//   P2G   per occupied 4^3 block, 4 waves: the block's particles are loaded
//         once (coalesced through the sort order), reduced to a 16-float
//         payload, and grouped by cell in LDS.  Lane l of every wave owns cell
//         l and accumulates its particles' stencil one 3x3 plane at a time in
//         registers; each wave adds its planes into its own 6^3-node float4
//         arena with plain LDS read-add-write (within one wave no two lanes
//         touch the same node, and the arena layout puts them on distinct
//         banks).  The summed arenas are flushed to the dense node grid with
//         global float atomics.  Mirrors scatter_momentum
//         (rust/crates/cpu/src/phase/scatter_momentum.rs:22-93).
//   grid  v = p / m per node of every touched block (meld_grid without
//         collider twins, rust/crates/cpu/src/phase/meld_grid.rs:61-66) plus the
//         optional grid gravity and boundary conditions; clears the
//         accumulation grid.
//   G2P   per occupied block: the 6^3 node velocities are staged in LDS, each
//         thread gathers v and C of one particle (collect_velocity,
//         rust/crates/cpu/src/phase/collect_velocity.rs:19-75), advances x and
//         F (advance_particles, rust/crates/cpu/src/phase/advance_particles.rs:17-93),
//         culls (rust/crates/cpu/src/phase/cull_particles.rs:17-41), and writes
//         the fused affine momentum A = m C - dt 4/h^2 V0 (tau + J sigma_visc)
//         for the next P2G, permuting the particle arrays into block order.
//
// The kernels are specialised on the set of constitutive models present: the
// common set (Neo-Hookean, fluid, fixed-corotated, viscosity) needs only a
// polar decomposition; snow and sand need the full SVD.
#include <algorithm>

#include "device_math.hpp"
#include "internal.hpp"

namespace svamd {

namespace {

constexpr int P2G_THREADS = 256;
constexpr int P2G_WAVES = P2G_THREADS / 64;
constexpr int P2G_CAP = 256;  // particles staged per pass
constexpr int P2G_PER_THREAD = P2G_CAP / P2G_THREADS;
constexpr int ARENA = 512;    // float4 slots of the 6^3 arena (8 rows x 64 banks)
constexpr int G2P_NODES = 216;
constexpr uint32_t NO_BLOCK = 0xffffffffu;

// LDS slot of arena node (x, y, z), 0 <= x, y, z < 6.  The low six bits are
// (x&3) | (y&3)<<2 | (z&3)<<4: for one stencil offset the 64 lanes (cells)
// of a wave address 64 distinct slots, and any 16 consecutive lanes (one
// ds_read_b128 phase) address 16 distinct 16-byte bank groups.
__device__ __forceinline__ int arena_slot(int x, int y, int z) {
    return (((x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2)) << 6) | (x & 3) | ((y & 3) << 2) | ((z & 3) << 4);
}

__device__ __forceinline__ uint32_t block_linear(const GridGeom &G, uint32_t bx, uint32_t by, uint32_t bz) {
    return bx + (uint32_t)G.nb[0] * (by + (uint32_t)G.nb[1] * bz);
}

__device__ __forceinline__ void decode_block(uint32_t key, uint32_t b[3]) {
    b[0] = compact1by2(key);
    b[1] = compact1by2(key >> 1);
    b[2] = compact1by2(key >> 2);
}

// Dense node index of local node (x, y, z).
__device__ __forceinline__ uint32_t node_index(const GridGeom &G, uint32_t x, uint32_t y, uint32_t z) {
    return block_linear(G, x >> 2, y >> 2, z >> 2) * 64u + (x & 3u) + ((y & 3u) << 2) + ((z & 3u) << 4);
}

// The block-list range of this work-group.  Work-group w runs on XCD w % 8;
// the Morton-ordered block list is cut into 8 contiguous parts, one per XCD,
// so that neighbouring blocks, which share halo nodes, share an L2.  The
// work-groups of an XCD stride through its part.  gridDim.x is a multiple of 8.
struct BlockRange {
    uint32_t begin, end, step;
};
__device__ __forceinline__ BlockRange xcd_block_range(uint32_t nblocks) {
    const uint32_t x = blockIdx.x & 7u, r = blockIdx.x >> 3, per = gridDim.x >> 3;
    const uint32_t chunk = (nblocks + 7u) >> 3;
    const uint32_t b0 = min(nblocks, x * chunk), b1 = min(nblocks, b0 + chunk);
    return {b0 + r, b1, per};
}

__device__ __noinline__ bool svd3_call(const float F[9], float U[9], float s[3], float V[9]) {
    return svd3(F, U, s, V);
}

// Kirchhoff stress tau = P F^T of a particle.  R must hold the polar factor
// of F for fixed-corotated particles; U, s hold the SVD for snow particles.
template <uint32_t SET>
__device__ __forceinline__ void kirchhoff(uint32_t flags, float p0, float p1, float hard, const float F[9], float J,
                                          const float R[9], const float U[9], const float s[3], float tau[9]) {
    if ((SET & M_FLUID) && (flags & F_FLUID)) {
        tau_inviscid(p0, (int)p1, J, tau);
    } else if ((SET & M_SNOW) && (flags & F_SNOW)) {
        tau_fixed_corotated_svd(p0 * hard, p1 * hard, U, s, tau);
    } else if ((SET & M_FCR) && (flags & F_FCR)) {
        tau_fixed_corotated_polar(p0, p1, F, R, J, tau);
    } else {
        tau_neo_hookean(p0, p1, F, J, tau);
    }
}

// A = m C - s V0 tau - s J V0 sigma_visc(C)  (scatter_momentum.rs:46-84 regrouped).
template <uint32_t SET>
__device__ __forceinline__ void fused_affine(float m, float vol, float scaling, const float C[9], const float tau[9],
                                             float J, bool visc, float vd, float vb, float A[9]) {
    const float sv = scaling * vol;
#pragma unroll
    for (int k = 0; k < 9; ++k) A[k] = m * C[k] - sv * tau[k];
    if ((SET & M_VISC) && visc) {
        float sig[9];
        cauchy_viscosity(vd, vb, C, sig);
        const float svj = sv * J;
#pragma unroll
        for (int k = 0; k < 9; ++k) A[k] -= svj * sig[k];
    }
}

// Stress inputs of a particle at P2G time (first substep of a call: the
// state holds C, not A).
template <uint32_t SET>
__device__ __forceinline__ void affine_from_state(const ParticleArrays &P, const AuxArrays &X, uint32_t p, float m,
                                                  float scaling, float hardening, DeviceScalars *ds, float A[9]) {
    const uint32_t flags = P.flags(p);
    float F[9], C[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        F[k] = P.F(k, p);
        C[k] = P.A(k, p);
    }
    float J = det3(F);
    float R[9], U[9], s[3] = {1.f, 1.f, 1.f};
    float hard = 1.0f;
    if ((SET & M_SNOW) && (flags & F_SNOW)) {
        float V[9];
        if (svd3_call(F, U, s, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
        hard = expf(hardening * (1.0f - X.jp[P.id(p)]));
    } else if ((SET & M_FCR) && (flags & F_FCR)) {
        if (!polar3(F, R)) {
            float V[9];
            svd3_call(F, U, s, V);
            matmul3_bt(U, V, R);
        }
    }
    if (!(J >= 1e-7f) && !(flags & F_FLUID))
        atomicOr(&ds->status, MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED | (KID_P2G << 16));
    float tau[9];
    kirchhoff<SET>(flags, P.p0(p), P.p1(p), hard, F, J, R, U, s, tau);
    const bool visc = flags & F_VISC;
    float vd = 0.f, vb = 0.f;
    if ((SET & M_VISC) && visc) {
        vd = X.visc_dyn[P.id(p)];
        vb = X.visc_bulk[P.id(p)];
    }
    fused_affine<SET>(m, P.vol(p), scaling, C, tau, J, visc, vd, vb, A);
}

// ------------------------------------------------------------------- P2G
template <bool FROM_C, bool GOALS, uint32_t SET>
__global__ __launch_bounds__(P2G_THREADS) void k_p2g(ParticleArrays P, AuxArrays X, const uint32_t *__restrict__ order,
                                                     const uint32_t *__restrict__ blk_start,
                                                     const uint32_t *__restrict__ blk_key, GridGeom G, SubstepParams S,
                                                     float4 *__restrict__ grid, uint32_t *__restrict__ blk_flag,
                                                     uint32_t *__restrict__ cand, DeviceScalars *ds) {
    __shared__ float4 arena[P2G_WAVES][ARENA];  // one arena per wave: no barriers between their adds
    __shared__ float4 pay[P2G_CAP][4];
    __shared__ uint32_t cell_count[64];
    __shared__ uint32_t cell_start[64];
    __shared__ uint32_t cell_fill[64];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const uint32_t nblocks = ds->num_blocks;
    const uint32_t live = ds->live_count;
    const float h = S.h;
    const float scaling = S.dt * 4.0f / (h * h);

    const BlockRange range = xcd_block_range(nblocks);
    for (uint32_t bi = range.begin; bi < range.end; bi += range.step) {
        const uint32_t start = blk_start[bi];
        const uint32_t end = bi + 1 < nblocks ? blk_start[bi + 1] : live;
        uint32_t b[3];
        decode_block(blk_key[bi], b);
        int cell0[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) cell0[a] = G.origin[a] + 4 * (int)b[a];
        // Claim the 8 blocks the arena covers (first claimant lists the block
        // for the grid update); issued first so the latency overlaps the staging.
        if (tid < 8) {
            const uint32_t T = block_linear(G, b[0] + (tid & 1u), b[1] + ((tid >> 1) & 1u), b[2] + (tid >> 2));
            const uint32_t old = atomicExch(&blk_flag[T], 1u);
            cand[8u * bi + tid] = old == 0u ? T : NO_BLOCK;
        }
        for (int k = lane; k < ARENA; k += 64) arena[wave][k] = make_float4(0.f, 0.f, 0.f, 0.f);

        for (uint32_t cs = start; cs < end; cs += P2G_CAP) {
            const uint32_t ce = min(end, cs + (uint32_t)P2G_CAP);
            if (tid < 64) {
                cell_count[tid] = 0;
                cell_fill[tid] = 0;
            }
            __syncthreads();
            // Stage: each thread reduces its particles to payloads
            //   (fx, m | q | h A columns), contribution at stencil node (i,j,k):
            //   w_ijk (m v + A d) = w_ijk (q + i c0 + j c1 + k c2).
            float pl[P2G_PER_THREAD][16];
            uint32_t mycell[P2G_PER_THREAD];
#pragma unroll
            for (int k = 0; k < P2G_PER_THREAD; ++k) {
                const uint32_t i = cs + tid + (uint32_t)(k * P2G_THREADS);
                mycell[k] = 64u;
                if (i >= ce) continue;
                const uint32_t p = order[i];
                float xp[3], fx[3];
                int c[3];
                bool inside = true;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    xp[a] = P.x(a, p);
                    const float xn = xp[a] / h;
                    const float fl = floorf(xn - 0.5f);
                    fx[a] = xn - fl;
                    c[a] = (int)fl - cell0[a];
                    inside = inside && c[a] >= 0 && c[a] < 4;
                }
                if (!inside) {
                    atomicOr(&ds->status, MPM_STATUS_OUT_OF_GRID | (KID_P2G << 16));
                    continue;
                }
                const float m = P.m(p);
                float v[3] = {P.v(0, p), P.v(1, p), P.v(2, p)};
                // external_force (rust/crates/cpu/src/phase/external_force.rs:36-47).
                bool goal = false;
                if constexpr (GOALS) {
                    const uint32_t o = P.id(p);
                    const uint32_t fb = S.has_b ? X.goal_flags_b[o] : X.goal_flags_a[o];
                    if ((X.goal_flags_a[o] & F_GOAL) && (fb & F_GOAL)) {
                        goal = true;
                        const float fa = 1.0f - S.factor_b;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const float gb = S.has_b ? X.goal_b[a][o] : X.goal_a[a][o];
                            const float gp = fa * X.goal_a[a][o] + S.factor_b * gb;
                            v[a] = (gp - xp[a]) / S.dt;
                        }
                    }
                }
                if (!goal && !(S.options & MPM_OPT_GRID_GRAVITY)) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) v[a] += S.dt * S.gravity[a];
                }
                float A[9];
                if constexpr (FROM_C) {
                    affine_from_state<SET>(P, X, p, m, scaling, S.snow_hardening, ds, A);
                } else {
#pragma unroll
                    for (int q = 0; q < 9; ++q) A[q] = P.A(q, p);
                }
                float *o = pl[k];
                o[0] = fx[0];
                o[1] = fx[1];
                o[2] = fx[2];
                o[3] = m;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    o[4 + r] = m * v[r] - h * (A[r] * fx[0] + A[3 + r] * fx[1] + A[6 + r] * fx[2]);
                    o[7 + r] = h * A[r];
                    o[10 + r] = h * A[3 + r];
                    o[13 + r] = h * A[6 + r];
                }
                const uint32_t cell = (uint32_t)(c[0] | (c[1] << 2) | (c[2] << 4));
                atomicAdd(&cell_count[cell], 1u);
                mycell[k] = cell;
            }
            __syncthreads();
            if (tid < 64) {
                const uint32_t cnt = cell_count[tid];
                uint32_t incl = cnt;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t y = __shfl_up(incl, o, 64);
                    if (lane >= (uint32_t)o) incl += y;
                }
                cell_start[tid] = incl - cnt;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < P2G_PER_THREAD; ++k) {
                if (mycell[k] < 64u) {
                    const uint32_t slot = cell_start[mycell[k]] + atomicAdd(&cell_fill[mycell[k]], 1u);
                    const float *o = pl[k];
                    pay[slot][0] = make_float4(o[0], o[1], o[2], o[3]);
                    pay[slot][1] = make_float4(o[4], o[5], o[6], o[7]);
                    pay[slot][2] = make_float4(o[8], o[9], o[10], o[11]);
                    pay[slot][3] = make_float4(o[12], o[13], o[14], o[15]);
                }
            }
            __syncthreads();

            // Accumulate: lane = cell, wave w takes every P2G_WAVES-th particle
            // of the cell; one stencil plane (fixed i) at a time, 36 registers.
            const uint32_t cnt = cell_count[lane];
            const uint32_t st = cell_start[lane];
            const int cx = (int)(lane & 3u), cy = (int)((lane >> 2) & 3u), cz = (int)(lane >> 4);
            if (cnt > wave) {
#pragma unroll 1
                for (int i = 0; i < 3; ++i) {
                    float acc[9][4];
#pragma unroll
                    for (int n = 0; n < 9; ++n)
#pragma unroll
                        for (int c = 0; c < 4; ++c) acc[n][c] = 0.f;
                    for (uint32_t r = wave; r < cnt; r += P2G_WAVES) {
                        const float4 a0 = pay[st + r][0], a1 = pay[st + r][1], a2 = pay[st + r][2], a3 = pay[st + r][3];
                        const float m = a0.w;
                        const float c0[3] = {a1.w, a2.x, a2.y};
                        const float c1[3] = {a2.z, a2.w, a3.x};
                        const float c2[3] = {a3.y, a3.z, a3.w};
                        float wx[3], wy[3], wz[3];
                        bspline_weights(a0.x, wx);
                        bspline_weights(a0.y, wy);
                        bspline_weights(a0.z, wz);
                        const float wi = i == 0 ? wx[0] : (i == 1 ? wx[1] : wx[2]);
                        const float fi = (float)i;
                        const float ti[3] = {a1.x + fi * c0[0], a1.y + fi * c0[1], a1.z + fi * c0[2]};
#pragma unroll
                        for (int j = 0; j < 3; ++j) {
                            const float wij = wi * wy[j];
#pragma unroll
                            for (int k = 0; k < 3; ++k) {
                                const float wt = wij * wz[k];
                                const int n = 3 * j + k;
                                acc[n][0] += wt * m;
#pragma unroll
                                for (int r3 = 0; r3 < 3; ++r3)
                                    acc[n][1 + r3] += wt * (ti[r3] + (float)j * c1[r3] + (float)k * c2[r3]);
                            }
                        }
                    }
                    // Add the plane into this wave's arena, one stencil offset per
                    // read-add-write round: for a fixed offset the 64 lanes hit 64
                    // distinct nodes, while two different offsets of two lanes can
                    // hit the same node, so a round's write must land before the
                    // next round's read.
#pragma unroll
                    for (int n = 0; n < 9; ++n) {
                        const int slot = arena_slot(cx + i, cy + n / 3, cz + n % 3);
                        float4 g = arena[wave][slot];
                        g.x += acc[n][0];
                        g.y += acc[n][1];
                        g.z += acc[n][2];
                        g.w += acc[n][3];
                        arena[wave][slot] = g;
                        // Keeps the compiler from moving the next round's read
                        // above this write (LDS ops of one wave execute in order).
                        __builtin_amdgcn_wave_barrier();
                    }
                }
            }
            __syncthreads();
        }

        // Flush the sum of the wave arenas (6^3 nodes) to the grid.
        for (int nidx = tid; nidx < 216; nidx += P2G_THREADS) {
            const int ax = nidx % 6, ay = (nidx / 6) % 6, az = nidx / 36;
            const int slot = arena_slot(ax, ay, az);
            float4 g4 = arena[0][slot];
#pragma unroll
            for (int w = 1; w < P2G_WAVES; ++w) {
                const float4 o = arena[w][slot];
                g4.x += o.x;
                g4.y += o.y;
                g4.z += o.z;
                g4.w += o.w;
            }
            if (g4.x != 0.f) {
                const uint32_t node = node_index(G, 4u * b[0] + ax, 4u * b[1] + ay, 4u * b[2] + az);
                float *g = reinterpret_cast<float *>(grid + node);
                unsafeAtomicAdd(g + 0, g4.x);
                unsafeAtomicAdd(g + 1, g4.y);
                unsafeAtomicAdd(g + 2, g4.z);
                unsafeAtomicAdd(g + 3, g4.w);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ grid update
// One wave per touched block: the candidate slots written by P2G are read 64
// at a time and the claimed blocks processed in turn.
__global__ __launch_bounds__(256) void k_grid_update(const uint32_t *__restrict__ cand, GridGeom G, SubstepParams S,
                                                     float4 *__restrict__ grid, float4 *__restrict__ vgrid,
                                                     uint32_t *__restrict__ blk_flag, uint32_t *__restrict__ blk_epoch,
                                                     uint32_t *__restrict__ partial, DeviceScalars *ds) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint32_t nw = gridDim.x * (blockDim.x / 64);
    const uint32_t ncand = ds->num_blocks * 8u;
    uint32_t active = 0;
    for (uint32_t chunk = gw; chunk * 64u < ncand; chunk += nw) {
        const uint32_t c = chunk * 64u + lane;
        const uint32_t mine = c < ncand ? cand[c] : NO_BLOCK;
        uint64_t todo = __ballot(mine != NO_BLOCK);
        while (todo) {
            const int l = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1ull;
            const uint32_t B = __shfl(mine, l, 64);
            const uint32_t node = B * 64u + lane;
            const float4 g = grid[node];
            const float m = g.x;
            float v[3] = {0.f, 0.f, 0.f};
            if (m > 0.f) {
                v[0] = g.y / m;
                v[1] = g.z / m;
                v[2] = g.w / m;
                // Extensions without a reference counterpart.
                if (S.options & MPM_OPT_GRID_GRAVITY) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) v[a] += S.gravity[a] * S.dt;
                }
                if (S.options & (MPM_OPT_STICKY_PLANE | MPM_OPT_BOX_WALLS)) {
                    const uint32_t bx = B % (uint32_t)G.nb[0];
                    const uint32_t by = (B / (uint32_t)G.nb[0]) % (uint32_t)G.nb[1];
                    const uint32_t bz = B / ((uint32_t)G.nb[0] * (uint32_t)G.nb[1]);
                    const int lx = (int)(4u * bx + (lane & 3u)), ly = (int)(4u * by + ((lane >> 2) & 3u)),
                              lz = (int)(4u * bz + (lane >> 4));
                    const float xi[3] = {(float)(lx + G.origin[0]) * S.h, (float)(ly + G.origin[1]) * S.h,
                                         (float)(lz + G.origin[2]) * S.h};
                    if ((S.options & MPM_OPT_STICKY_PLANE) && xi[2] < S.plane_z) v[0] = v[1] = v[2] = 0.f;
                    if (S.options & MPM_OPT_BOX_WALLS) {
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            if (xi[a] < S.wall_min[a] && v[a] < 0.f) v[a] = 0.f;
                            if (xi[a] > S.wall_max[a] && v[a] > 0.f) v[a] = 0.f;
                        }
                    }
                }
                ++active;
            }
            vgrid[node] = make_float4(m, v[0], v[1], v[2]);
            grid[node] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lane == 0) {
                blk_flag[B] = 0u;
                blk_epoch[B] = S.epoch;
            }
        }
    }
    // Per-wave count of active nodes (summed once per step_to call).
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) active += __shfl_xor(active, o, 64);
    if (lane == 0) partial[gw] = active;
}

__global__ __launch_bounds__(256) void k_sum_partials(const uint32_t *__restrict__ partial, uint32_t n,
                                                      DeviceScalars *ds) {
    __shared__ uint32_t ws[4];
    uint32_t s = 0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63u) == 0) ws[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) ds->active_nodes = ws[0] + ws[1] + ws[2] + ws[3];
}

// ------------------------------------------------------------------- G2P
template <bool LAST, uint32_t SET>
__global__ __launch_bounds__(256) void k_g2p(ParticleArrays Pin, ParticleArrays Pout, AuxArrays X,
                                             const uint32_t *__restrict__ order, const uint32_t *__restrict__ blk_start,
                                             const uint32_t *__restrict__ blk_key, GridGeom G, SubstepParams S,
                                             const float4 *__restrict__ vgrid, uint32_t *__restrict__ keys_out,
                                             DeviceScalars *ds) {
    __shared__ float4 vloc[G2P_NODES];
    const uint32_t tid = threadIdx.x;
    const uint32_t NT = blockDim.x;
    const uint32_t nblocks = ds->num_blocks;
    const uint32_t live = ds->live_count;
    const float h = S.h;
    const float dt = S.dt;
    const float scaling = dt * 4.0f / (h * h);
    const float apic = 4.0f / h;

    const BlockRange range = xcd_block_range(nblocks);
    for (uint32_t bi = range.begin; bi < range.end; bi += range.step) {
        const uint32_t start = blk_start[bi];
        const uint32_t end = bi + 1 < nblocks ? blk_start[bi + 1] : live;
        uint32_t b[3];
        decode_block(blk_key[bi], b);
        for (uint32_t nidx = tid; nidx < (uint32_t)G2P_NODES; nidx += NT) {
            const uint32_t ax = nidx % 6u, ay = (nidx / 6u) % 6u, az = nidx / 36u;
            vloc[nidx] = vgrid[node_index(G, 4u * b[0] + ax, 4u * b[1] + ay, 4u * b[2] + az)];
        }
        __syncthreads();
        int cell0[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) cell0[a] = G.origin[a] + 4 * (int)b[a];

        for (uint32_t i = start + tid; i < end; i += NT) {
            const uint32_t p = order[i];
            float xp[3], fx[3], w[3][3];
            int c[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                xp[a] = Pin.x(a, p);
                const float xn = xp[a] / h;
                const float fl = floorf(xn - 0.5f);
                fx[a] = xn - fl;
                c[a] = min(max((int)fl - cell0[a], 0), 3);
                bspline_weights(fx[a], w[a]);
            }
            // collect_velocity: v = sum w v_i, C = 4/h^2 sum w v_i (x_i - x_p)^T
            float v[3] = {0.f, 0.f, 0.f};
            float C[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) C[k] = 0.f;
            // Planes one at a time: keeps the LDS reads of a single plane in flight.
#pragma unroll 1
            for (int i3 = 0; i3 < 3; ++i3) {
                const float dx = (float)i3 - fx[0];
#pragma unroll
                for (int j3 = 0; j3 < 3; ++j3) {
                    const float dy = (float)j3 - fx[1];
                    const float wij = w[0][i3] * w[1][j3];
#pragma unroll
                    for (int k3 = 0; k3 < 3; ++k3) {
                        const float dz = (float)k3 - fx[2];
                        const float wt = wij * w[2][k3];
                        const float4 gv = vloc[(c[0] + i3) + 6 * ((c[1] + j3) + 6 * (c[2] + k3))];
                        const float wv[3] = {wt * gv.y, wt * gv.z, wt * gv.w};
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            v[r] += wv[r];
                            C[r] += wv[r] * dx;
                            C[3 + r] += wv[r] * dy;
                            C[6 + r] += wv[r] * dz;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) C[k] *= apic;

            // advance_particles: x += dt v, F += (C F) dt, plasticity, energy.
            uint32_t flags = Pin.flags(p);
            const uint32_t id = Pin.id(p);
            const float m = Pin.m(p), vol = Pin.vol(p), p0 = Pin.p0(p), p1 = Pin.p1(p);
            float F[9], CF[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] = Pin.F(k, p);
            matmul3(C, F, CF);
#pragma unroll
            for (int k = 0; k < 9; ++k) F[k] += CF[k] * dt;
            float xnew[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) xnew[a] = xp[a] + v[a] * dt;

            float R[9], U[9], s[3] = {1.f, 1.f, 1.f};
            float hard = 1.0f;
            float energy = 0.f;
            float J = det3(F);
            if ((SET & M_FLUID) && (flags & F_FLUID)) {
                // advance_particles.rs:76-86: singular values <- (s1 s2 s3)^(1/3),
                // i.e. F <- J^(1/3) R with R the polar factor of F.
                if (!(J > 0.f) || !polar3(F, R)) {
                    float V[9];
                    if (svd3_call(F, U, s, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
                    matmul3_bt(U, V, R);
                    J = s[0] * s[1] * s[2];
                }
                const float g = cbrtf(J);
#pragma unroll
                for (int k = 0; k < 9; ++k) F[k] = g * R[k];
                J = det3(F);
                if (LAST) energy = energy_inviscid(p0, (int)p1, J);
            } else if ((SET & M_SNOW) && (flags & F_SNOW)) {
                float V[9];
                if (svd3_call(F, U, s, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
                float jp = X.jp[id];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float cl = fminf(fmaxf(s[a], S.snow_lo), S.snow_hi);
                    jp *= s[a] / cl;
                    s[a] = cl;
                }
                X.jp[id] = jp;
                recompose(U, s, V, F);
                J = s[0] * s[1] * s[2];
                hard = expf(S.snow_hardening * (1.0f - jp));
                if (LAST) energy = energy_fixed_corotated(p0 * hard, p1 * hard, s);
            } else {
                if ((SET & M_SAND) && (flags & F_SAND)) {
                    // advance_particles.rs:50-71: Drucker-Prager return map in log strain.
                    float V[9];
                    if (svd3_call(F, U, s, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
                    const float alpha = X.sand_alpha[id];
                    const float e[3] = {logf(s[0]), logf(s[1]), logf(s[2])};
                    const float etr = e[0] + e[1] + e[2];
                    const float eh[3] = {e[0] - etr / 3.f, e[1] - etr / 3.f, e[2] - etr / 3.f};
                    const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
                    if (etr < 0.f && ehn > 0.f) {
                        const float dg = ehn + (3.f * p1 + 2.f * p0) / 2.f / p0 * etr * alpha;
                        if (dg > 0.f) {
                            float d[3];
#pragma unroll
                            for (int a = 0; a < 3; ++a) d[a] = expf(e[a] - dg / ehn * eh[a]);
                            recompose(U, d, V, F);
                        }
                    } else {
                        const float one[3] = {1.f, 1.f, 1.f};
                        recompose(U, one, V, F);
                    }
                    J = det3(F);
                }
                if ((SET & M_FCR) && (flags & F_FCR)) {
                    if (!polar3(F, R)) {
                        float V[9];
                        svd3_call(F, U, s, V);
                        matmul3_bt(U, V, R);
                    }
                    if (LAST) energy = energy_fixed_corotated_polar(p0, p1, F, R, J);
                } else {
                    if (!(J > 0.f)) {
                        // try_elastic_energy_neo_hookean fails: EnergyError, particle FAILED.
                        flags |= F_FAILED;
                        atomicOr(&ds->status, MPM_STATUS_ENERGY_ERROR | (KID_G2P << 16));
                    } else if (LAST) {
                        energy = energy_neo_hookean(p0, p1, F, J);
                    }
                }
            }

            // cull_particles: strict inequalities.
            bool inside = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) inside = inside && xnew[a] > S.domain_min[a] && xnew[a] < S.domain_max[a];
            if (!inside) flags |= F_TOMB;

            float Aout[9];
            if (LAST || !inside) {
#pragma unroll
                for (int k = 0; k < 9; ++k) Aout[k] = C[k];
            } else {
                float tau[9];
                kirchhoff<SET>(flags, p0, p1, hard, F, J, R, U, s, tau);
                const bool visc = flags & F_VISC;
                float vd = 0.f, vb = 0.f;
                if ((SET & M_VISC) && visc) {
                    vd = X.visc_dyn[id];
                    vb = X.visc_bulk[id];
                }
                fused_affine<SET>(m, vol, scaling, C, tau, J, visc, vd, vb, Aout);
            }
            if (LAST) X.energy[id] = energy;
            if (keys_out) keys_out[i] = inside ? particle_block_key(xnew, G, ds) : G.tomb_key;

#pragma unroll
            for (int a = 0; a < 3; ++a) {
                Pout.x(a, i) = xnew[a];
                Pout.v(a, i) = v[a];
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                Pout.F(k, i) = F[k];
                Pout.A(k, i) = Aout[k];
            }
            Pout.m(i) = m;
            Pout.vol(i) = vol;
            Pout.p0(i) = p0;
            Pout.p1(i) = p1;
            Pout.flags(i) = flags;
            Pout.id(i) = id;
        }
        __syncthreads();
    }
}

// Particles that are not in any block (tombstoned, sorted to the end) are
// copied to their sorted slot unchanged.
__global__ __launch_bounds__(256) void k_copy_dead(ParticleArrays Pin, ParticleArrays Pout,
                                                   const uint32_t *__restrict__ order, uint64_t n,
                                                   uint32_t *__restrict__ keys_out, uint32_t tomb, DeviceScalars *ds) {
    const uint64_t live = ds->live_count;
    for (uint64_t i = live + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = order[i];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            Pout.x(a, i) = Pin.x(a, p);
            Pout.v(a, i) = Pin.v(a, p);
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            Pout.F(k, i) = Pin.F(k, p);
            Pout.A(k, i) = Pin.A(k, p);
        }
        Pout.m(i) = Pin.m(p);
        Pout.vol(i) = Pin.vol(p);
        Pout.p0(i) = Pin.p0(p);
        Pout.p1(i) = Pin.p1(p);
        Pout.flags(i) = Pin.flags(p);
        Pout.id(i) = Pin.id(p);
        if (keys_out) keys_out[i] = tomb;
    }
}

// to_io_state's reverse permutation (rust/crates/cpu/src/cpu_state.rs:84-93):
// out[id] = state[i], vectors as xyz triples, matrices column-major.
__global__ __launch_bounds__(256) void k_unpermute(ParticleArrays P, uint64_t n, float *__restrict__ x,
                                                   float *__restrict__ v, float *__restrict__ F,
                                                   float *__restrict__ C, uint32_t *__restrict__ flags) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t o = P.id(i);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            x[3ull * o + a] = P.x(a, i);
            v[3ull * o + a] = P.v(a, i);
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            F[9ull * o + k] = P.F(k, i);
            C[9ull * o + k] = P.A(k, i);
        }
        flags[o] = P.flags(i);
    }
}

__global__ __launch_bounds__(256) void k_stress(const uint32_t *__restrict__ flags,
                                                const mpm_particle_parameters *__restrict__ params,
                                                const float *__restrict__ Fin, const float *__restrict__ jp,
                                                float hardening, float *__restrict__ tau_out,
                                                float *__restrict__ energy_out, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = flags[i];
    const mpm_particle_parameters pp = params[i];
    const bool fluid = f & F_FLUID;
    const float p0 = fluid ? pp.bulk_modulus : pp.mu;
    const float p1 = fluid ? (float)pp.exponent : pp.lambda;
    float F[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = Fin[9 * i + k];
    float R[9], U[9], s[3] = {1.f, 1.f, 1.f}, V[9];
    float hard = 1.f;
    const float J = det3(F);
    float energy;
    if (!fluid && (f & F_SNOW)) {
        svd3_call(F, U, s, V);
        hard = expf(hardening * (1.f - (jp ? jp[i] : 1.f)));
        energy = energy_fixed_corotated(p0 * hard, p1 * hard, s);
    } else if (!fluid && (f & F_FCR)) {
        if (!polar3(F, R)) {
            svd3_call(F, U, s, V);
            matmul3_bt(U, V, R);
        }
        energy = energy_fixed_corotated_polar(p0, p1, F, R, J);
    } else if (fluid) {
        energy = energy_inviscid(p0, (int)p1, J);
    } else {
        energy = J > 0.f ? energy_neo_hookean(p0, p1, F, J) : __builtin_nanf("");
    }
    float tau[9];
    kirchhoff<SET_ALL>(f, p0, p1, hard, F, J, R, U, s, tau);
#pragma unroll
    for (int k = 0; k < 9; ++k) tau_out[9 * i + k] = tau[k];
    if (energy_out) energy_out[i] = energy;
}

__global__ __launch_bounds__(256) void k_svd(const float *__restrict__ Fin, float *__restrict__ Uo,
                                             float *__restrict__ so, float *__restrict__ Vo, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float F[9], U[9], s[3], V[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) F[k] = Fin[9 * i + k];
    svd3(F, U, s, V);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        Uo[9 * i + k] = U[k];
        Vo[9 * i + k] = V[k];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) so[3 * i + a] = s[a];
}

// Work-groups of a persistent launch: all resident at once (the occupancy of
// the kernel times the CU count), a multiple of 8 for the XCD mapping.
int resident_groups(const void *kernel, int threads, const TransferLaunch &L) {
    if (L.persistent_blocks > 0) return std::max(8, L.persistent_blocks & ~7);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return std::max(8, (per_cu * L.cu_count) & ~7);
}

template <uint32_t SET>
hipError_t p2g_dispatch(const TransferLaunch &L, bool from_c, bool goals, const ParticleArrays &P, const AuxArrays &X,
                        const uint32_t *order, const uint32_t *blk_start, const uint32_t *blk_key, const GridGeom &G,
                        const SubstepParams &S, float4 *grid, uint32_t *blk_flag, uint32_t *cand, DeviceScalars *ds,
                        hipStream_t stream) {
    const dim3 blk(P2G_THREADS);
#define SV_P2G(FC, GO)                                                                                                  \
    do {                                                                                                                \
        static const int groups = resident_groups(reinterpret_cast<const void *>(&k_p2g<FC, GO, SET>), P2G_THREADS, L); \
        hipLaunchKernelGGL((k_p2g<FC, GO, SET>), dim3(groups), blk, 0, stream, P, X, order, blk_start, blk_key, G, S,    \
                           grid, blk_flag, cand, ds);                                                                   \
    } while (0)
    if (from_c) {
        if (goals) SV_P2G(true, true);
        else SV_P2G(true, false);
    } else {
        if (goals) SV_P2G(false, true);
        else SV_P2G(false, false);
    }
#undef SV_P2G
    return hipGetLastError();
}

}  // namespace

hipError_t launch_p2g(const TransferLaunch &L, bool from_c, const ParticleArrays &P, const AuxArrays &X,
                      const uint32_t *order, const uint32_t *blk_start, const uint32_t *blk_key, const GridGeom &G,
                      const SubstepParams &S, float4 *grid, uint32_t *blk_flag, uint32_t *cand, DeviceScalars *ds,
                      hipStream_t stream) {
    const bool goals = S.has_goals != 0;
    if (L.model_set == SET_POLAR)
        return p2g_dispatch<SET_POLAR>(L, from_c, goals, P, X, order, blk_start, blk_key, G, S, grid, blk_flag, cand, ds,
                                       stream);
    return p2g_dispatch<SET_ALL>(L, from_c, goals, P, X, order, blk_start, blk_key, G, S, grid, blk_flag, cand, ds,
                                 stream);
}

// The grid update runs on a fixed grid so that its per-wave partial counts
// have a fixed size.
uint32_t grid_update_groups(const TransferLaunch &L) { return (uint32_t)std::max(8, 4 * L.cu_count); }
uint32_t grid_update_waves(const TransferLaunch &L) { return grid_update_groups(L) * 4u; }

hipError_t launch_grid_update(const TransferLaunch &L, const uint32_t *cand, const GridGeom &G, const SubstepParams &S,
                              float4 *grid, float4 *vgrid, uint32_t *blk_flag, uint32_t *blk_epoch, uint32_t *partial,
                              DeviceScalars *ds, hipStream_t stream) {
    hipLaunchKernelGGL(k_grid_update, dim3(grid_update_groups(L)), dim3(256), 0, stream, cand, G, S, grid, vgrid, blk_flag,
                       blk_epoch, partial, ds);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const TransferLaunch &L, const uint32_t *partial, DeviceScalars *ds,
                               hipStream_t stream) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, stream, partial, grid_update_waves(L), ds);
    return hipGetLastError();
}

hipError_t launch_g2p(const TransferLaunch &L, bool last, const ParticleArrays &Pin, const ParticleArrays &Pout,
                      const AuxArrays &X, const uint32_t *order, const uint32_t *blk_start, const uint32_t *blk_key,
                      const GridGeom &G, const SubstepParams &S, const float4 *vgrid, uint32_t *keys_out,
                      DeviceScalars *ds, hipStream_t stream) {
    const dim3 blk(L.g2p_threads);
#define SV_G2P(LA, SE)                                                                                              \
    do {                                                                                                            \
        const int groups = resident_groups(reinterpret_cast<const void *>(&k_g2p<LA, SE>), L.g2p_threads, L);       \
        hipLaunchKernelGGL((k_g2p<LA, SE>), dim3(groups), blk, 0, stream, Pin, Pout, X, order, blk_start, blk_key, G, \
                           S, vgrid, keys_out, ds);                                                                 \
    } while (0)
    if (L.model_set == SET_POLAR) {
        if (last) SV_G2P(true, SET_POLAR);
        else SV_G2P(false, SET_POLAR);
    } else {
        if (last) SV_G2P(true, SET_ALL);
        else SV_G2P(false, SET_ALL);
    }
#undef SV_G2P
    return hipGetLastError();
}

hipError_t launch_copy_dead(const ParticleArrays &Pin, const ParticleArrays &Pout, const uint32_t *order, uint64_t n,
                            uint32_t *keys_out, uint32_t tomb, DeviceScalars *ds, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_copy_dead, dim3(grid), dim3(256), 0, stream, Pin, Pout, order, n, keys_out, tomb, ds);
    return hipGetLastError();
}

hipError_t launch_unpermute(const ParticleArrays &P, uint64_t n, float *x, float *v, float *F, float *C,
                            uint32_t *flags, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_unpermute, dim3(grid), dim3(256), 0, stream, P, n, x, v, F, C, flags);
    return hipGetLastError();
}

hipError_t launch_stress(const uint32_t *flags, const mpm_particle_parameters *params, const float *F, const float *jp,
                         float snow_hardening, float *tau, float *energy, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stress, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, flags, params, F, jp,
                       snow_hardening, tau, energy, n);
    return hipGetLastError();
}

hipError_t launch_svd(const float *F, float *U, float *s, float *V, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_svd, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, F, U, s, V, n);
    return hipGetLastError();
}

}  // namespace svamd
