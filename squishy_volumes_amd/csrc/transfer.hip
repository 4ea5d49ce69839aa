// transfer.hip - particle/grid transfers of the MLS-MPM substep (gfx950).
//
// One persistent kernel, k_transfer<MODE, PERMUTE, GOALS, SET>, moves data
// between particles and grid.  It walks the occupied 4^3 blocks of the sorted
// block list (XCD-aware contiguous ranges) with one 4-wave work-group per
// block at a time.  Its modes:
//
//   XM_P2G    first substep of a call: the stress from the stored F and C, the
//             external force and the scatter (scatter_momentum,
//             rust/crates/cpu/src/phase/scatter_momentum.rs:22-93).
//   XM_G2P2G  every further substep: the gather of substep n (collect_velocity,
//             rust/crates/cpu/src/phase/collect_velocity.rs:19-75), the
//             advance and cull of substep n (advance_particles.rs:17-93,
//             cull_particles.rs:17-41), then the external force and scatter of
//             substep n+1 from the registers.  The fused affine momentum
//             A = m C - dt 4/h^2 V0 (tau + J sigma_visc) and the particle
//             velocity never leave the chip.
//   XM_LIMITS the gather alone, reduced to the largest |v| and |C_ij| for the
//             adaptive step's limits before the advance (limit_time_step.rs:187-222).
//   XM_G2P    last substep of a call: the gather, advance and cull; writes the
//             reference's state x, v, F, C and the elastic energies.
//
// Block frame.  Between sorts a particle keeps its slot in its block's range
// while it moves, so the kernel accepts base cells one cell outside the
// block: cells -1..4 per axis, stencil nodes -1..6 (an 8^3 node region).  The
// gather reads the region's node velocities from LDS; the scatter buckets the
// particles by cell (216 cells) and accumulates with lane = cell: for one
// stencil offset the lanes of a wave then touch distinct nodes, so the
// region's arena takes conflict-free LDS float atomics.  Waves 0-2 take the 64
// cells inside the block, one stencil plane each; wave 3 takes the sparse
// shell of cells around it.  A particle outside the region (it moved more
// than a cell since the last sort) takes a global-memory gather and scatter.
// The arena is flushed with global float atomics; the blocks it touched are
// listed for the grid update (first claimant only).
//
// The kernels are specialised on the set of constitutive models present: the
// common set (Neo-Hookean, fluid, fixed-corotated, viscosity) needs only a
// polar decomposition; snow and sand need the full SVD.
#include <algorithm>

#include "closure_net.hpp"
#include "collide.hpp"
#include "device_math.hpp"
#include "internal.hpp"

namespace svamd {

namespace {

constexpr int XF_THREADS = 256;
constexpr int XF_CAP = 768;            // particles per pass of a block
constexpr int XF_PER_THREAD = XF_CAP / XF_THREADS;
constexpr int NSLOT = 512;             // 8^3 node region
constexpr int NCELL = 216;             // 6^3 cells
constexpr int XF_RANKS = 12;           // core payloads per cell and pass
constexpr int XF_OVF = 96;             // overflow payloads per pass
constexpr int PSLOT = 256;             // nodes one stencil plane of the core cells reaches
// Node velocities of the region in LDS: node (x, y, z) at x + VS_Y y + VS_Z z.
// The strides make the 2x2x3 base cells that 16 consecutive particles of a
// jittered 8-per-cell lattice share fall on distinct 16-byte bank groups.
constexpr int VS_Y = 10;
constexpr int VS_Z = 84;
constexpr int VS_SLOTS = 7 + 7 * VS_Y + 7 * VS_Z + 1;
constexpr uint32_t NO_BLOCK = 0xffffffffu;

enum : int { XM_P2G = 0, XM_G2P2G = 1, XM_G2P = 2, XM_LIMITS = 3 };

// Arena slot of region node (x, y, z), 0 <= x, y, z < 8 (region node = block
// node + 1).  The low six bits are (x&3) | (y&3)<<2 | (z&3)<<4: for one
// stencil offset the 64 cells of a 4^3 cube address 64 distinct slots, and 16
// consecutive lanes 16 distinct 16-byte bank groups.
__device__ __forceinline__ int arena_slot(int x, int y, int z) {
    return (((x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2)) << 6) | (x & 3) | ((y & 3) << 2) | ((z & 3) << 4);
}

// Slot of region node (x, y, z) in the arena of core stencil plane i: the 64
// core cells reach x in [1 + i, 4 + i] only, so u = x - 1 - i in [0, 3] and
// the plane needs 4 x 8 x 8 = 256 slots, with the same bank property.
__device__ __forceinline__ int plane_slot(int u, int y, int z) {
    return (((y >> 2) | ((z >> 2) << 1)) << 6) | u | ((y & 3) << 2) | ((z & 3) << 4);
}

__device__ __forceinline__ uint32_t block_linear(const GridGeom &G, uint32_t bx, uint32_t by, uint32_t bz) {
    return bx + (uint32_t)G.nb[0] * (by + (uint32_t)G.nb[1] * bz);
}

__device__ __forceinline__ void decode_block(uint32_t key, uint32_t b[3]) {
    b[0] = compact1by2(key);
    b[1] = compact1by2(key >> 1);
    b[2] = compact1by2(key >> 2);
}

// Dense node index of local node (x, y, z) (all >= 0).
__device__ __forceinline__ uint32_t node_index(const GridGeom &G, uint32_t x, uint32_t y, uint32_t z) {
    return block_linear(G, x >> 2, y >> 2, z >> 2) * 64u + (x & 3u) + ((y & 3u) << 2) + ((z & 3u) << 4);
}

__device__ __forceinline__ bool node_in_grid(const GridGeom &G, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < 4 * G.nb[0] && y < 4 * G.nb[1] && z < 4 * G.nb[2];
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

// The SVD is out of line (it is rare and large); its operands travel by value
// so that no array of the caller has to live in private memory.
struct Mat9 {
    float v[9];
};
struct Svd3 {
    float U[9], s[3], V[9];
    bool fp64;
};
__device__ __noinline__ Svd3 svd3_value(Mat9 F) {
    Svd3 r;
    r.fp64 = svd3(F.v, r.U, r.s, r.V);
    return r;
}
__device__ __forceinline__ bool svd3_call(const float F[9], float U[9], float s[3], float V[9]) {
    Mat9 m;
#pragma unroll
    for (int k = 0; k < 9; ++k) m.v[k] = F[k];
    const Svd3 r = svd3_value(m);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        U[k] = r.U[k];
        V[k] = r.V[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = r.s[k];
    return r.fp64;
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

// Stress inputs of a particle whose state holds F and C (first substep of a
// call): polar factor / SVD, A.  With tau_ext, a snow particle's Kirchhoff
// stress is read from it (row k at tau_ext[k * stride + slot]) instead of
// being evaluated.  Returns false, with A untouched, for a particle too close
// to inversion (det F < INVERSE_EPS): it takes no part in the scatter
// (prepare_tmp.wesl:74-78).
template <uint32_t SET>
__device__ __forceinline__ bool affine_from_state(uint32_t flags, const float F[9], const float C[9], float m, float vol,
                                                  float p0, float p1, uint32_t id, const AuxArrays &X, float scaling,
                                                  float hardening, DeviceScalars *ds, const float *tau_ext,
                                                  uint64_t stride, uint64_t slot, float A[9]) {
    float J = det3(F);
    if (!(J >= INVERSE_EPS)) return false;
    float R[9], U[9], s[3] = {1.f, 1.f, 1.f};
    float hard = 1.0f;
    const bool ext = (SET & M_SNOW) && (flags & F_SNOW) && tau_ext;
    if (ext) {
    } else if ((SET & M_SNOW) && (flags & F_SNOW)) {
        float V[9];
        if (svd3_call(F, U, s, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
        hard = expf(hardening * (1.0f - X.jp[id]));
    } else if ((SET & M_FCR) && (flags & F_FCR)) {
        if (!polar3(F, R)) {
            float V[9];
            svd3_call(F, U, s, V);
            matmul3_bt(U, V, R);
        }
    }
    float tau[9];
    if (ext) {
#pragma unroll
        for (int k = 0; k < 9; ++k) tau[k] = tau_ext[(uint64_t)k * stride + slot];
    } else {
        kirchhoff<SET>(flags, p0, p1, hard, F, J, R, U, s, tau);
    }
    const bool visc = flags & F_VISC;
    float vd = 0.f, vb = 0.f;
    if ((SET & M_VISC) && visc) {
        vd = X.visc_dyn[id];
        vb = X.visc_bulk[id];
    }
    fused_affine<SET>(m, vol, scaling, C, tau, J, visc, vd, vb, A);
    return true;
}

// A particle flagged FAILED in this transfer: the status bits, and the
// transfer's parameter-table entry as the call's first failing transfer if
// no earlier one failed.
__device__ __forceinline__ void report_failure(DeviceScalars *ds, uint32_t bits, uint32_t transfer) {
    atomicOr(&ds->status, bits);
    atomicMin(&ds->fail_index, transfer);
}

// Separable gather of v = sum w v_i and C = 4/h sum w v_i (i - f)^T over the
// 3x3x3 stencil (collect_velocity.rs:19-75), with w_ijk = wx_i wy_j wz_k:
// rows over k first, then j, then the planes i.  node(i, j, k) returns the
// node velocity as a float4 (m, v).
template <class Node>
__device__ __forceinline__ void gather_stencil(const float w[3][3], const float f[3], float apic, Node node, float v[3],
                                               float C[9]) {
    float T[3][3];  // T[a] = sum w v (node offset along axis a)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        v[r] = 0.f;
        T[0][r] = T[1][r] = T[2][r] = 0.f;
    }
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {
        float S[3] = {0.f, 0.f, 0.f}, Sy[3] = {0.f, 0.f, 0.f}, Sz[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float R[3] = {0.f, 0.f, 0.f}, Rz[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float4 g = node(i, j, k);
                const float wz = w[2][k];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const float gr = r == 0 ? g.y : (r == 1 ? g.z : g.w);
                    R[r] += wz * gr;
                    if (k > 0) Rz[r] += ((float)k * wz) * gr;
                }
            }
            const float wy = w[1][j];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                S[r] += wy * R[r];
                if (j > 0) Sy[r] += ((float)j * wy) * R[r];
                Sz[r] += wy * Rz[r];
            }
        }
        const float wx = i == 0 ? w[0][0] : (i == 1 ? w[0][1] : w[0][2]);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            v[r] += wx * S[r];
            T[0][r] += ((float)i * wx) * S[r];
            T[1][r] += wx * Sy[r];
            T[2][r] += wx * Sz[r];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int r = 0; r < 3; ++r) C[3 * a + r] = apic * (T[a][r] - f[a] * v[r]);
}

// external_force (rust/crates/cpu/src/phase/external_force.rs:36-47): goal
// velocity when the particle has a goal in frames a and b, else gravity
// (unless gravity is applied on the grid).
template <bool GOALS>
__device__ __forceinline__ void external_velocity(const SubstepParams &S, const AuxArrays &X, uint32_t id,
                                                  const float xp[3], float v[3]) {
    if constexpr (GOALS) {
        const uint32_t fb = S.has_b ? X.goal_flags_b[id] : X.goal_flags_a[id];
        if ((X.goal_flags_a[id] & F_GOAL) && (fb & F_GOAL)) {
            const float fa = 1.0f - S.factor_b;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float gb = S.has_b ? X.goal_b[a][id] : X.goal_a[a][id];
                const float gp = fa * X.goal_a[a][id] + S.factor_b * gb;
                v[a] = (gp - xp[a]) / S.dt_force;
            }
            return;
        }
    }
    if (!(S.options & MPM_OPT_GRID_GRAVITY)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] += S.dt_force * S.gravity[a];
    }
}

struct XferArgs {
    ParticleArrays Pin;          // read; sorted order when PERMUTE, else the current order
    ParticleArrays Pout;         // written; equals Pin when not PERMUTE
    AuxArrays X;
    const uint32_t *order;       // PERMUTE: slot i of the sorted order holds Pin[order[i]]
    const uint32_t *blk_start;
    const uint32_t *blk_key;
    GridGeom G;

    const SubstepParams *params; // parameters of substep ds->param_index (the substep whose
                                 // scatter runs, P2G modes)
    GridBuffers B;               // B.ovf[epoch & 1]: the overflow list for that substep's grid
    uint32_t *keys_out;          // block keys of the new positions, or null
    uint32_t mover_bit;          // or'ed into a key that differs from the key of the block
                                 // the particle was processed in (BUCKET_MOVER or 0)
    DeviceScalars *ds;
    const float *tau_ext;        // P2G mode: Kirchhoff stress of snow particles by slot (9 rows
    uint64_t tau_stride;         // of tau_stride floats), from the learned closure; or null
    float *svd_out;              // G2P mode: U and s of snow particles' new F by slot (12 rows of
                                 // tau_stride floats), for the learned closure; or null
    uint32_t pass_cap;           // particles per pass of a block: XF_THREADS x 1..3
    ColliderCall coll;           // M_COLL kernels: the mesh and the twin tables
    LearnedCall learn;           // M_LEARN kernels: the network and its judge
};

// The substep parameters do not change while a kernel runs: read through the
// constant address space, the compiler may load them again where it needs
// them (as it does kernel arguments) instead of holding them in registers.
__device__ __forceinline__ SubstepParams load_invariant(const SubstepParams *p) {
    static_assert(sizeof(SubstepParams) % 4 == 0, "whole words");
    SubstepParams s;
    const __attribute__((address_space(4))) uint32_t *w = (const __attribute__((address_space(4))) uint32_t *)p;
    uint32_t *d = reinterpret_cast<uint32_t *>(&s);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(SubstepParams) / 4); ++i) d[i] = w[i];
    return s;
}

// Scatter of one particle straight into the global accumulation grid, for a
// particle outside its block's region; claims the blocks it touches for the
// grid update (overflow list of the epoch's parity).
__device__ __forceinline__ void scatter_global(const GridBuffers &B, const GridGeom &G, uint32_t epoch,
                                               DeviceScalars *ds, const int cell[3], const float pl[16]) {
    if (!node_in_grid(G, cell[0], cell[1], cell[2]) || !node_in_grid(G, cell[0] + 2, cell[1] + 2, cell[2] + 2)) {
        atomicOr(&ds->status, MPM_STATUS_OUT_OF_GRID | (KID_P2G << 16));
        return;
    }
    float wx[3], wy[3], wz[3];
    bspline_weights(pl[0], wx);
    bspline_weights(pl[1], wy);
    bspline_weights(pl[2], wz);
#pragma unroll 1
    for (int i = 0; i < 3; ++i)
#pragma unroll 1
        for (int j = 0; j < 3; ++j)
#pragma unroll 1
            for (int k = 0; k < 3; ++k) {
                const float wt = (i == 0 ? wx[0] : (i == 1 ? wx[1] : wx[2])) *
                                 (j == 0 ? wy[0] : (j == 1 ? wy[1] : wy[2])) *
                                 (k == 0 ? wz[0] : (k == 1 ? wz[1] : wz[2]));
                const uint32_t x = (uint32_t)(cell[0] + i), y = (uint32_t)(cell[1] + j), z = (uint32_t)(cell[2] + k);
                float *g = reinterpret_cast<float *>(B.grid + node_index(G, x, y, z));
                unsafeAtomicAdd(g + 0, wt * pl[3]);
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    unsafeAtomicAdd(g + 1 + r, wt * (pl[4 + r] + (float)i * pl[7 + r] + (float)j * pl[10 + r] +
                                                     (float)k * pl[13 + r]));
            }
#pragma unroll 1
    for (int bz = cell[2] >> 2; bz <= (cell[2] + 2) >> 2; ++bz)
#pragma unroll 1
        for (int by = cell[1] >> 2; by <= (cell[1] + 2) >> 2; ++by)
#pragma unroll 1
            for (int bx = cell[0] >> 2; bx <= (cell[0] + 2) >> 2; ++bx) {
                const uint32_t T = block_linear(G, (uint32_t)bx, (uint32_t)by, (uint32_t)bz);
                if (atomicExch(&B.blk_flag[T], 1u) == 0u) {
                    const uint32_t slot = atomicAdd(&ds->ovf_count[epoch & 1u], 1u);
                    if (slot < G.num_blocks_total) B.ovf[epoch & 1u][slot] = T;
                }
            }
}

// The blocks that the stencil of a particle in region cell c (relative to the
// block's first cell, each component in [-1, 4]) reaches, as a 27-bit mask:
// bit (ox + 1) + 3 (oy + 1) + 9 (oz + 1) for block offset (ox, oy, oz).  Its
// nodes c .. c + 2 lie in block offsets floor(c / 4) .. floor((c + 2) / 4).
__device__ __forceinline__ uint32_t touched_blocks(const int c[3]) {
    uint32_t m[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int lo = (c[q] >> 2) + 1, hi = ((c[q] + 2) >> 2) + 1;
        m[q] = ((2u << hi) - 1u) & ~((1u << lo) - 1u);
    }
    const uint32_t X = m[0] * 0x1249249u;  // bit i at i + 3t, t < 9
    const uint32_t y3 = ((m[1] & 1u) ? 07u : 0u) | ((m[1] & 2u) ? 070u : 0u) | ((m[1] & 4u) ? 0700u : 0u);
    const uint32_t Y = y3 * 0x40201u;      // repeated at 0, 9, 18
    const uint32_t Z = ((m[2] & 1u) ? 0x1ffu : 0u) | ((m[2] & 2u) ? 0x1ffu << 9 : 0u) | ((m[2] & 4u) ? 0x1ffu << 18 : 0u);
    return X & Y & Z;
}

// The fields of one particle that the transfer kernel reads in every mode.
struct ParticleIn {
    float x[3], F[9], m, vol, p0, p1;
    uint32_t id, flags, src;
};

// Issues all loads of slot i's particle (its source slot is order[i] when
// permuting) at once; nothing here waits on a loaded value except the source
// slot.
template <bool PERMUTE>
__device__ __forceinline__ void load_particle(const ParticleArrays &P, const uint32_t *order, uint32_t i,
                                              ParticleIn &r) {
    r.src = PERMUTE ? order[i] : i;
    const float *p = P.at(r.src);
#pragma unroll
    for (int q = 0; q < 3; ++q) r.x[q] = p[(FX + q) * TILE];
#pragma unroll
    for (int k = 0; k < 9; ++k) r.F[k] = p[(FF + k) * TILE];
    r.m = p[FM * TILE];
    r.vol = p[FVOL * TILE];
    r.p0 = p[FP0 * TILE];
    r.p1 = p[FP1 * TILE];
    r.id = __float_as_uint(p[FID * TILE]);
    r.flags = __float_as_uint(p[FFLAGS * TILE]);
}

// Payload of one particle for the scatter, 16 floats in four float4:
//   (f, m | q | h A columns); its contribution at stencil node (i, j, k) is
//   w_ijk (m v + A d) = w_ijk (q + i c0 + j c1 + k c2).
// Contributions of stencil plane i of the cnt particles of one cell (one
// lane, region cell (cx, cy, cz)): fetch(r, a0..a3) returns the r-th
// payload.  The 9 nodes of the plane are accumulated in registers, then added
// into the wave's own arena with a plain read-add-write, one stencil offset
// per round: for one offset the lanes (distinct cells) touch distinct nodes,
// whose arena slots differ in their low six bits (no bank conflict).
// PLANE: arena is core plane i's (plane_slot), else the full region arena.
template <bool PLANE, class Fetch>
__device__ __forceinline__ void accumulate_plane(int i, uint32_t cnt, int cx, int cy, int cz, Fetch fetch,
                                                 float4 *arena) {
    float acc[9][4];
#pragma unroll
    for (int n = 0; n < 9; ++n)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[n][c] = 0.f;
    const float fi = (float)i;
    for (uint32_t r = 0; r < cnt; ++r) {
        float4 a0, a1, a2, a3;
        fetch(r, a0, a1, a2, a3);
        const float m = a0.w;
        const float c0[3] = {a1.w, a2.x, a2.y};
        const float c1[3] = {a2.z, a2.w, a3.x};
        const float c2[3] = {a3.y, a3.z, a3.w};
        float wy[3], wz[3];
        bspline_weights(a0.y, wy);
        bspline_weights(a0.z, wz);
        const float d = i == 0 ? 1.5f - a0.x : (i == 1 ? a0.x - 1.0f : a0.x - 0.5f);
        const float wi = i == 1 ? 0.75f - d * d : 0.5f * d * d;
        // Node values built by running sums (no multiplications by the
        // offsets 0, which IEEE arithmetic cannot fold away).
        float tj[3] = {a1.x + fi * c0[0], a1.y + fi * c0[1], a1.z + fi * c0[2]};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float wij = wi * wy[j];
            float tk[3] = {tj[0], tj[1], tj[2]};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float wt = wij * wz[k];
                const int n = 3 * j + k;
                acc[n][0] += wt * m;
#pragma unroll
                for (int r3 = 0; r3 < 3; ++r3) acc[n][1 + r3] += wt * tk[r3];
                if (k < 2)
#pragma unroll
                    for (int r3 = 0; r3 < 3; ++r3) tk[r3] += c2[r3];
            }
            if (j < 2)
#pragma unroll
                for (int r3 = 0; r3 < 3; ++r3) tj[r3] += c1[r3];
        }
    }
    if (cnt == 0) return;
#pragma unroll
    for (int n = 0; n < 9; ++n) {
        float4 &g = arena[PLANE ? plane_slot(cx - 1, cy + n / 3, cz + n % 3) : arena_slot(cx + i, cy + n / 3, cz + n % 3)];
        float4 t = g;
        t.x += acc[n][0];
        t.y += acc[n][1];
        t.z += acc[n][2];
        t.w += acc[n][3];
        g = t;
        // The next round's read must not move above this write (two rounds
        // of two lanes can hit one node; LDS ops of a wave run in order).
        __builtin_amdgcn_wave_barrier();
    }
}

// The learned kernel asks for two waves per SIMD, which its network's
// registers would otherwise cost it (a few values then live in scratch).
// The gather-only mode has no scatter arenas (10 KB of LDS): asking for three
// waves per SIMD fits it into 168 registers (no spills without the permute),
// which takes the learned substep's gather from 0.148 to 0.124 ms per 1M
// particles.  The other modes stay at the compiler's choice (2 waves, bound
// by their 78 KB of LDS per work-group anyway).
template <int MODE, bool PERMUTE, bool GOALS, uint32_t SET>
__global__ __launch_bounds__(XF_THREADS) __attribute__((amdgpu_waves_per_eu(MODE == XM_G2P || MODE == XM_LIMITS ? 3 : ((SET & M_LEARN) && MODE == XM_G2P2G ? 2 : 1)))) void k_transfer(XferArgs a) {
    constexpr bool HAS_G2P = MODE != XM_P2G;
    constexpr bool HAS_P2G = MODE == XM_P2G || MODE == XM_G2P2G;
    constexpr bool LAST = MODE == XM_G2P;
    constexpr bool LIMITS = MODE == XM_LIMITS;
    float lim_v = 0.f, lim_c = 0.f;  // LIMITS: largest |v| and |C_ij| of this thread's particles
    uint32_t lim_any = 0;
    constexpr bool COLL = (SET & M_COLL) != 0;
    // The learned snow stress: the network runs on a wave's 64 particles at
    // once, so a wave keeps all its lanes in every iteration of the particle
    // loop (a lane with nothing to do skips its work instead of leaving).
    constexpr bool LEARN = (SET & M_LEARN) != 0 && MODE == XM_G2P2G;
    // kernels without the SVD, mesh colliders or the network: registers to spare
    constexpr bool LIGHT = (SET & (M_SNOW | M_SAND | M_COLL | M_LEARN)) == 0;
    uint32_t n_eval = 0, n_fallback = 0, n_judged = 0;
    double err2 = 0.0, ref2 = 0.0;
    // The region's node velocities (gather) and the arenas (scatter: one
    // 256-slot arena per core plane wave, one 512-slot arena for the overflow
    // wave) share LDS; a pass restages the velocities after the previous
    // pass's flush.
    constexpr int ARENA4 = 3 * PSLOT + NSLOT;
    constexpr int SHARED4 = (HAS_P2G ? ARENA4 : 0) > (HAS_G2P ? VS_SLOTS : 0) ? (HAS_P2G ? ARENA4 : 0)
                                                                              : (HAS_G2P ? VS_SLOTS : 1);
    __shared__ float4 shared4[SHARED4];
    float4 *vstage = shared4;                                    // node (x, y, z) at x + VS_Y y + VS_Z z: (v, m)
    float4 *planes = shared4;                                    // [wave][plane_slot], waves 0-2
    float4 *shell = shared4 + 3 * PSLOT;                         // [arena_slot], wave 3
    __shared__ float4 payc[HAS_P2G ? 4 * XF_RANKS * 64 : 1];     // core payloads [quarter][rank][cell]
    __shared__ float4 payx[HAS_P2G ? 4 * XF_OVF : 1];            // overflow payloads [quarter][entry]
    __shared__ uint8_t xcell[HAS_P2G ? XF_OVF : 1];              // region cell of each overflow entry
    __shared__ uint8_t xrank[HAS_P2G ? XF_OVF : 1];              // rank of the entry in its cell
    __shared__ uint8_t xperm[HAS_P2G ? XF_OVF : 1];              // entries in cell order
    __shared__ uint8_t xocc[HAS_P2G ? NCELL : 1];                // occupied overflow cells
    __shared__ uint32_t xcount[HAS_P2G ? NCELL : 1];
    __shared__ uint32_t xstart[HAS_P2G ? NCELL : 1];
    __shared__ uint32_t core_count[64];
    __shared__ uint32_t nx;
    __shared__ uint32_t blk_touch;  // blocks reached by the block's particles (touched_blocks)

    // a substep launched past the end of an adaptive call's loop
    if (a.ds->halt) return;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t wave = tid >> 6;
    const GridGeom &G = a.G;
    DeviceScalars *ds = a.ds;
    const uint32_t nblocks = ds->num_blocks;
    const uint32_t live = ds->live_count;
    const float h = G.h;
    const uint32_t pidx = ds->param_index;
    const SubstepParams S1 = load_invariant(a.params + pidx);
    // The grid update that follows reads the same entry.  Each of the two
    // kernels hands the index on to the other in a slot the other kernel
    // alone reads, so no work-group races a reader of its own launch.
    if (blockIdx.x == 0 && tid == 0) ds->grid_param = pidx;
    // The step of the advance that the gather precedes, and of the external
    // force and collide of the substep that starts (the same value: the
    // adaptive step does not change between them), and of the scatter (after
    // the sound and isolated-particle limits).  Equal with a fixed step.
    const float dt = S1.dt_force;
    const float scaling = S1.dt * 4.0f / (h * h);
    const float apic = 4.0f / h;

    // The work-group visits its block-list entries j = 0, 1, ... (list index
    // range.begin + j * range.step), each in passes of at most XF_CAP
    // particles.  The loads that start a pass (the region's node velocities,
    // two per thread, and each thread's first particle) are issued as soon
    // as the previous pass's gather is done, so that they travel while it
    // accumulates and flushes.  Lane l of every wave holds entry
    // meta_base + l's particle range and key, refilled 64 entries at a time.
    const BlockRange range = xcd_block_range(nblocks);
    uint32_t meta_base = 0, m_start = 0, m_end = 0, m_key = 0;
    auto load_meta = [&](uint32_t base) {
        meta_base = base;
        const uint32_t bj = range.begin + (base + lane) * range.step;
        if (bj < range.end) {
            m_start = a.blk_start[bj];
            m_end = bj + 1 < nblocks ? a.blk_start[bj + 1] : live;
            m_key = a.blk_key[bj];
        }
    };
    struct Item {
        uint32_t j, bi, start, end, cs;
        uint32_t b[3];
    };
    auto get_block = [&](uint32_t j, Item &w) -> bool {
        w.bi = range.begin + j * range.step;
        if (w.bi >= range.end) return false;
        if (j - meta_base >= 64u) load_meta(j);
        const int l = (int)(j - meta_base);
        w.j = j;
        w.start = (uint32_t)__builtin_amdgcn_readlane((int)m_start, l);
        w.end = (uint32_t)__builtin_amdgcn_readlane((int)m_end, l);
        w.cs = w.start;
        decode_block((uint32_t)__builtin_amdgcn_readlane((int)m_key, l), w.b);
        return true;
    };
    static_assert(NSLOT == 2 * XF_THREADS, "two staged nodes per thread");
    int so[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const uint32_t s = tid + (uint32_t)(t * XF_THREADS);
        so[t] = (int)(s & 7u) + VS_Y * (int)((s >> 3) & 7u) + VS_Z * (int)(s >> 6);
    }
    // The fields of a thread's next particle are loaded while the current
    // one computes (3 particles per thread, 256 apart).  The copy into cur at
    // the top of an iteration is the one wait for them: afterwards no load
    // that cur depends on is outstanding, so the conditional prefetch leaves
    // the compiler no wait to guess.  All loads that start a pass are
    // unconditional (clamped indices).
    ParticleIn cur, nxt;
    float4 sv[2];
    bool sin[2];
    // Thread t takes slot t ^ flip of each group of XF_THREADS particles, with
    // flip alternating between 0 and 255 from pass to pass: the short last
    // group of a pass falls on waves 0-1 and on waves 2-3 in turn, so that the
    // SIMDs (wave w of every work-group runs on SIMD w) share the load.
    uint32_t npass = 0;
    auto issue = [&](const Item &w, uint32_t flip) {
        const uint32_t ce = min(w.end, w.cs + a.pass_cap);
        if constexpr (HAS_G2P) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const uint32_t s = tid + (uint32_t)(t * XF_THREADS);
                const int x = 4 * (int)w.b[0] + (int)(s & 7u) - 1, y = 4 * (int)w.b[1] + (int)((s >> 3) & 7u) - 1,
                          z = 4 * (int)w.b[2] + (int)(s >> 6) - 1;
                sin[t] = node_in_grid(G, x, y, z);
                sv[t] = a.B.vgrid[sin[t] ? node_index(G, (uint32_t)x, (uint32_t)y, (uint32_t)z) : 0u];
            }
            // Keeps the node loads ahead of the particle loads (the in-order
            // load counter then lets the staging wait for them alone).
            __builtin_amdgcn_sched_barrier(0);
        }
        // (the learned kernel loads each particle when it starts it: the network
        // needs the registers a prefetched particle would hold)
        if constexpr (!LEARN) load_particle<PERMUTE>(a.Pin, a.order, min(w.cs + (tid ^ flip), ce - 1u), nxt);
    };

    load_meta(0);
    Item w;
    bool have = get_block(0, w);
    if (have) issue(w, 0u);
    if constexpr (HAS_P2G) {
        if (tid < 64u) core_count[tid] = 0u;
        if (tid == 0) nx = 0u;
    }
    while (have) {
        const uint32_t bi = w.bi, start = w.start, end = w.end, cs = w.cs;
        const uint32_t ce = min(end, cs + a.pass_cap);
        const uint32_t flip = npass & 1u ? 255u : 0u;
        const uint32_t b[3] = {w.b[0], w.b[1], w.b[2]};
        const uint32_t bkey = morton3(b[0], b[1], b[2]);
        int cell0[3];  // global cell of the block's first cell
#pragma unroll
        for (int q = 0; q < 3; ++q) cell0[q] = G.origin[q] + 4 * (int)b[q];
        Item wn;
        bool have_next;
        {
            if constexpr (HAS_G2P) {
#pragma unroll
                // staged as (v, m): the gather then reads each node's v with one
                // 12-byte LDS load from a 16-byte aligned address
                for (int t = 0; t < 2; ++t)
                    vstage[so[t]] =
                        sin[t] ? make_float4(sv[t].y, sv[t].z, sv[t].w, sv[t].x) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            if constexpr (HAS_P2G) {
                if (cs == start && tid == 0) blk_touch = 0u;
            }
            __syncthreads();
            // Wave priorities by phase: the particle loop 2, the accumulate 1,
            // the flush 0.  Two work-groups share each SIMD, and a wave in an
            // earlier phase of its pass then issues first (its loads start
            // sooner; the flush's atomics need no answer): 1.8-2 % on the
            // fused transfer against equal priorities.
            __builtin_amdgcn_s_setprio(2);
            uint32_t touch = 0;  // touched_blocks of this thread's particles
#pragma unroll 1
            for (int kk = 0; kk < XF_PER_THREAD; ++kk) {
                const uint32_t i = cs + (tid ^ flip) + (uint32_t)(kk * XF_THREADS);
                bool skip = false;
                uint32_t early_rank = 0u;  // G2P2G: the core rank taken before the stress
                if constexpr (LEARN) {
                    if (__ballot(i < ce) == 0ull) break;
                    skip = i >= ce;
                } else {
                    if (i >= ce) break;
                }
                if constexpr (LEARN) {
                    load_particle<PERMUTE>(a.Pin, a.order, min(i, ce - 1u), cur);
                } else {
                    cur = nxt;
                    if (kk + 1 < XF_PER_THREAD && i + XF_THREADS < ce)
                        load_particle<PERMUTE>(a.Pin, a.order, i + XF_THREADS, nxt);
                }
                const uint32_t src = cur.src;
                const ParticleArrays &Pin = a.Pin;
                const ParticleArrays &Pout = a.Pout;
                const uint32_t flags0 = cur.flags;
                uint32_t flags = flags0;
                if (!skip && (flags & F_TOMB)) {
                    // Culled since the last sort: keeps its slot, takes no part.
                    if (PERMUTE) {
                        for (int f = 0; f < NFIELDS; ++f) Pout.at(i)[f * TILE] = Pin.at(src)[f * TILE];
                    }
                    if (a.keys_out) a.keys_out[i] = G.tomb_key | a.mover_bit;
                    if constexpr (LEARN) {
                        skip = true;
                    } else {
                        continue;
                    }
                }
                float xp[3], F[9];
#pragma unroll
                for (int q = 0; q < 3; ++q) xp[q] = cur.x[q];
#pragma unroll
                for (int k = 0; k < 9; ++k) F[k] = cur.F[k];
                const float m = cur.m, vol = cur.vol, p0 = cur.p0, p1 = cur.p1;
                const uint32_t id = cur.id;
                // Collider bits: a particle near a collider gathers from and
                // scatters to its twin nodes as well (collide.hpp).
                uint32_t cbits = 0;
                if constexpr (COLL) cbits = a.X.collider_bits[id];
                float v[3], C[9];
                float J = 1.f;
                float R[9], U[9], sv[3] = {1.f, 1.f, 1.f};
                float hard = 1.0f;
                bool inside = true;
                uint32_t failed = 0;  // status bits of a failure detected in this transfer
                float A[9];
                float jp_new = 1.f;                  // snow: Jp after the advance
                float xnet[MLP_IN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                float E_net = 0.f;
                bool want_net = false;               // the network gives this particle's stress
                float tau_l[LEARN ? 9 : 1];          // the learned kernel's stress of the new F

                if constexpr (HAS_G2P) {
                  if (!skip) {
                    // collect_velocity: v = sum w v_i, C = 4/h^2 sum w v_i (x_i - x_p)^T.
                    float f[3], w[3][3];
                    int c[3];
                    bool in_region = true;
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const float xn = div_by(xp[q], h, G.inv_h);
                        const float fl = floorf(xn - 0.5f);
                        f[q] = xn - fl;
                        c[q] = (int)fl - cell0[q];
                        in_region = in_region && c[q] >= -1 && c[q] <= 4;
                        bspline_weights(f[q], w[q]);
                    }
                    if (COLL && (cbits >> 16)) {
                        // the melded velocities of this particle's twins
                        const int g0[3] = {c[0] + 4 * (int)b[0], c[1] + 4 * (int)b[1], c[2] + 4 * (int)b[2]};
                        const bool ok = node_in_grid(G, g0[0], g0[1], g0[2]) &&
                                        node_in_grid(G, g0[0] + 2, g0[1] + 2, g0[2] + 2);
                        if (!ok) atomicOr(&ds->status, MPM_STATUS_OUT_OF_GRID | (KID_G2P << 16));
                        const TwinEntry *tab = a.coll.twins.tab[(S1.epoch + 1u) & 1u];
                        gather_stencil(
                            w, f, apic,
                            [&](int i3, int j3, int k3) {
                                return ok ? twin_velocity(tab, a.coll.twins.mask,
                                                          node_index(G, (uint32_t)(g0[0] + i3), (uint32_t)(g0[1] + j3),
                                                                     (uint32_t)(g0[2] + k3)),
                                                          cbits, ds)
                                          : make_float4(0.f, 0.f, 0.f, 0.f);
                            },
                            v, C);
                    } else if (in_region) {
                        const float4 *base = vstage + (c[0] + 1) + VS_Y * (c[1] + 1) + VS_Z * (c[2] + 1);
                        gather_stencil(
                            w, f, apic,
                            [&](int i3, int j3, int k3) {
                                const float4 q = base[i3 + VS_Y * j3 + VS_Z * k3];
                                return make_float4(q.w, q.x, q.y, q.z);
                            },
                            v, C);
                    } else {
                        const int g0[3] = {c[0] + 4 * (int)b[0], c[1] + 4 * (int)b[1], c[2] + 4 * (int)b[2]};
                        const bool ok = node_in_grid(G, g0[0], g0[1], g0[2]) &&
                                        node_in_grid(G, g0[0] + 2, g0[1] + 2, g0[2] + 2);
                        if (!ok) atomicOr(&ds->status, MPM_STATUS_OUT_OF_GRID | (KID_G2P << 16));
                        gather_stencil(
                            w, f, apic,
                            [&](int i3, int j3, int k3) {
                                return ok ? a.B.vgrid[node_index(G, (uint32_t)(g0[0] + i3), (uint32_t)(g0[1] + j3),
                                                               (uint32_t)(g0[2] + k3))]
                                          : make_float4(0.f, 0.f, 0.f, 0.f);
                            },
                            v, C);
                    }

                    if constexpr (LIMITS) {
                        // limit_time_step_by_velocity / _by_deformation (limit_time_step.rs:199-222)
                        lim_v = fmaxf(lim_v, sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
#pragma unroll
                        for (int k = 0; k < 9; ++k) lim_c = fmaxf(lim_c, fabsf(C[k]));
                        lim_any = 1u;
                        continue;
                    }
                    // advance_particles: x += dt v, F += (C F) dt, plasticity, energy.
                    float CF[9];
                    matmul3(C, F, CF);
#pragma unroll
                    for (int k = 0; k < 9; ++k) F[k] += CF[k] * dt;
#pragma unroll
                    for (int q = 0; q < 3; ++q) xp[q] += v[q] * dt;
                    float energy = 0.f;
                    J = det3(F);
                    if ((SET & M_FLUID) && (flags & F_FLUID)) {
                        // advance_particles.rs:76-86: singular values <- (s1 s2 s3)^(1/3)
                        // (nalgebra's, >= 0), i.e. F <- |J|^(1/3) R with R the orthogonal
                        // polar factor of F (a reflection when det F < 0).
                        if (!(J > 0.f) || !polar3(F, R)) {
                            float V[9];
                            if (svd3_call(F, U, sv, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
                            if (sv[2] < 0.f) {
#pragma unroll
                                for (int k = 6; k < 9; ++k) V[k] = -V[k];  // U diag(1, 1, -1) V^T
                            }
                            matmul3_bt(U, V, R);
                            J = sv[0] * sv[1] * sv[2];
                        }
                        const float g = cbrtf(fabsf(J));
#pragma unroll
                        for (int k = 0; k < 9; ++k) F[k] = g * R[k];
                        J = det3(F);
                        if (LAST) energy = energy_inviscid(p0, (int)p1, J);
                    } else if ((SET & M_SNOW) && (flags & F_SNOW)) {
                        float V[9];
                        if (svd3_call(F, U, sv, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
                        float jp = a.X.jp[id];
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            const float cl = fminf(fmaxf(sv[q], S1.snow_lo), S1.snow_hi);
                            jp *= sv[q] / cl;
                            sv[q] = cl;
                        }
                        a.X.jp[id] = jp;
                        jp_new = jp;
                        recompose(U, sv, V, F);
                        J = sv[0] * sv[1] * sv[2];
                        if constexpr (LAST) {
                            // the factors of the new F, for the closure that follows
                            if (a.svd_out) {
#pragma unroll
                                for (int k = 0; k < 9; ++k) a.svd_out[(uint64_t)k * a.tau_stride + i] = U[k];
#pragma unroll
                                for (int q = 0; q < 3; ++q) a.svd_out[(uint64_t)(9 + q) * a.tau_stride + i] = sv[q];
                            }
                        }
                        hard = expf(S1.snow_hardening * (1.0f - jp));
                        if (LAST) energy = energy_fixed_corotated(p0 * hard, p1 * hard, sv);
                    } else {
                        if ((SET & M_SAND) && (flags & F_SAND)) {
                            // advance_particles.rs:50-71: Drucker-Prager return map in log strain.
                            float V[9];
                            if (svd3_call(F, U, sv, V)) atomicOr(&ds->status, MPM_STATUS_SVD_FP64);
                            const float alpha = a.X.sand_alpha[id];
                            // nalgebra's singular values are >= 0: the log strain of |s|, the
                            // sign of det F kept in the recomposition (a reflection for det F < 0)
                            const float sg = sv[2] < 0.f ? -1.f : 1.f;
                            const float e[3] = {logf(fabsf(sv[0])), logf(fabsf(sv[1])), logf(fabsf(sv[2]))};
                            const float etr = e[0] + e[1] + e[2];
                            const float eh[3] = {e[0] - etr / 3.f, e[1] - etr / 3.f, e[2] - etr / 3.f};
                            const float ehn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]);
                            if (etr < 0.f && ehn > 0.f) {
                                const float dg = ehn + (3.f * p1 + 2.f * p0) / 2.f / p0 * etr * alpha;
                                if (dg > 0.f) {
                                    float d[3];
#pragma unroll
                                    for (int q = 0; q < 3; ++q) d[q] = expf(e[q] - dg / ehn * eh[q]);
                                    d[2] *= sg;
                                    recompose(U, d, V, F);
                                }
                            } else {
                                const float one[3] = {1.f, 1.f, sg};
                                recompose(U, one, V, F);
                            }
                            J = det3(F);
                        }
                        if ((SET & M_FCR) && (flags & F_FCR)) {
                            if (!polar3(F, R)) {
                                float V[9];
                                svd3_call(F, U, sv, V);
                                matmul3_bt(U, V, R);
                            }
                            if (LAST) energy = energy_fixed_corotated_polar(p0, p1, F, R, J);
                        } else {
                            if (!(J > 0.f)) {
                                // try_elastic_energy_neo_hookean fails: EnergyError, particle FAILED
                                // (advance_particles.rs:73-74).
                                failed = MPM_STATUS_ENERGY_ERROR;
                            } else if (LAST) {
                                energy = energy_neo_hookean(p0, p1, F, J);
                            }
                        }
                    }

                    // cull_particles: strict inequalities.
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        // batched scenes: within the particle's own copy of the domain
                        const float xl = (SET & M_BATCH) ? scene_local(xp[q], q, S1) : xp[q];
                        inside = inside && xl > S1.domain_min[q] && xl < S1.domain_max[q];
                    }
                    if (!inside) flags |= F_TOMB;
                    // The next substep's stress needs det F >= INVERSE_EPS
                    // (prepare_tmp.wesl:74-78): a particle below it is FAILED
                    // and takes no part in any later scatter of the call.
                    if (inside && !(J >= INVERSE_EPS)) failed |= MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED;
                    if (failed) {
                        flags |= F_FAILED;
                        report_failure(ds, failed | (KID_G2P << 16), pidx);
                    }

                    if (a.keys_out) {
                        const uint32_t key = inside ? particle_block_key(xp, G, ds) : G.tomb_key;
                        a.keys_out[i] = key | (key != bkey ? a.mover_bit : 0u);
                    }
#pragma unroll
                    for (int q = 0; q < 3; ++q) Pout.x(q, i) = xp[q];
#pragma unroll
                    for (int k = 0; k < 9; ++k) Pout.F(k, i) = F[k];
                    if (LAST) {
#pragma unroll
                        for (int q = 0; q < 3; ++q) Pout.v(q, i) = v[q];
#pragma unroll
                        for (int k = 0; k < 9; ++k) Pout.A(k, i) = C[k];
                        // a failed energy evaluation leaves the stored energy
                        if (!(failed & MPM_STATUS_ENERGY_ERROR)) a.X.energy[id] = energy;
                    }
                    if (PERMUTE || flags != flags0) Pout.flags(i) = flags;
                    if (PERMUTE) {
                        Pout.m(i) = m;
                        Pout.vol(i) = vol;
                        Pout.p0(i) = p0;
                        Pout.p1(i) = p1;
                        Pout.id(i) = id;
                    }
                    if (LAST || !inside || failed) {
                        if constexpr (LEARN) {
                            skip = true;
                        } else {
                            continue;
                        }
                    }
                    if constexpr (!LEARN) {
                        if constexpr (MODE == XM_G2P2G && LIGHT) {
                            // The scatter's core rank, taken now so that the LDS
                            // atomic's latency hides behind the stress (the cell is
                            // the payload's: nothing moves the particle before it).
                            int ce[3];
                            bool core_e = true;
#pragma unroll
                            for (int q = 0; q < 3; ++q) {
                                const float xn = div_by(xp[q], h, G.inv_h);
                                ce[q] = (int)floorf(xn - 0.5f) - cell0[q];
                                core_e = core_e && ce[q] >= 0 && ce[q] <= 3;
                            }
                            if (core_e)
                                early_rank = atomicAdd(&core_count[(uint32_t)(ce[0] | (ce[1] << 2) | (ce[2] << 4))], 1u);
                        }
                        // Stress of the new F for the next substep's scatter.
                        float tau[9];
                        kirchhoff<SET>(flags, p0, p1, hard, F, J, R, U, sv, tau);
                        const bool visc = flags & F_VISC;
                        float vd = 0.f, vb = 0.f;
                        if ((SET & M_VISC) && visc) {
                            vd = a.X.visc_dyn[id];
                            vb = a.X.visc_bulk[id];
                        }
                        fused_affine<SET>(m, vol, scaling, C, tau, J, visc, vd, vb, A);
                    } else if (!skip) {
                        // the network's inputs; inside the training hull it gives the
                        // stress, else the analytic model does (before the network runs,
                        // so that F need not live through it)
                        if (flags & F_SNOW) {
                            closure_features(sv, p0, p1, jp_new, xnet, E_net);
                            want_net = true;
#pragma unroll
                            for (int q = 0; q < MLP_IN; ++q)
                                want_net = want_net && xnet[q] >= a.learn.hull_lo[q] && xnet[q] <= a.learn.hull_hi[q];
                        }
                        if (!want_net) kirchhoff<SET>(flags, p0, p1, hard, F, J, R, U, sv, tau_l);
                    }
                  }
                  if constexpr (LEARN) {
                    // The network on the wave's particles (every lane takes part), then
                    // the stress: learned inside the hull, analytic outside it, both
                    // for the particles the judge samples (closure.hip, k_closure).
                    float ynet[3];
                    closure_outputs(*a.learn.net, xnet, want_net, ynet);
                    if (!skip) {
                        float *tau = tau_l;
                        const bool snow = (flags & F_SNOW) != 0;
                        if (want_net) {
                            closure_tau(U, sv, p0, p1, hard, E_net, ynet, tau);
                            ++n_eval;
                        } else if (snow) {
                            ++n_fallback;
                        }
                        if (snow && a.learn.judge_stride && id % a.learn.judge_stride == a.learn.judge_phase) {
                            float ref[9];
                            tau_fixed_corotated_svd(p0 * hard, p1 * hard, U, sv, ref);
#pragma unroll
                            for (int k = 0; k < 9; ++k) {
                                err2 += (double)(tau[k] - ref[k]) * (double)(tau[k] - ref[k]);
                                ref2 += (double)ref[k] * (double)ref[k];
                            }
                            ++n_judged;
                        }
                        const bool visc = flags & F_VISC;
                        float vd = 0.f, vb = 0.f;
                        if ((SET & M_VISC) && visc) {
                            vd = a.X.visc_dyn[id];
                            vb = a.X.visc_bulk[id];
                        }
                        fused_affine<SET>(m, vol, scaling, C, tau, J, visc, vd, vb, A);
                    }
                  }
                } else {
                    // First substep of a call: the state holds v and C.
#pragma unroll
                    for (int q = 0; q < 3; ++q) v[q] = Pin.v(q, src);
#pragma unroll
                    for (int k = 0; k < 9; ++k) C[k] = Pin.A(k, src);
                    const bool ok = affine_from_state<SET>(flags, F, C, m, vol, p0, p1, id, a.X, scaling,
                                                           S1.snow_hardening, ds, a.tau_ext, a.tau_stride, i, A);
                    if (PERMUTE) {
                        for (int f = 0; f < NFIELDS; ++f) Pout.at(i)[f * TILE] = Pin.at(src)[f * TILE];
                    }
                    if (!ok) {
                        flags |= F_FAILED;
                        Pout.flags(i) = flags;
                        report_failure(ds, MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED | (KID_P2G << 16), pidx);
                        continue;
                    }
                }

                if constexpr (HAS_P2G) if (!skip) {
                    // external_force of the substep that starts, then the payload
                    //   (f, m | q | h A columns): contribution at stencil node (i,j,k)
                    //   w_ijk (m v + A d) = w_ijk (q + i c0 + j c1 + k c2).
                    if constexpr (COLL) {
                        // collide (collide.rs:21-206) runs after the sort and before
                        // the external force of the substep that starts
                        const uint32_t b0 = cbits;
                        collide_particle(a.coll.mesh, G, dt, xp, v, cbits);
                        if (cbits != b0) a.X.collider_bits[id] = cbits;
                    }
                    external_velocity<GOALS>(S1, a.X, id, xp, v);
                    float pl[16];
                    int c[3];
                    bool in_region = true;
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const float xn = div_by(xp[q], h, G.inv_h);
                        const float fl = floorf(xn - 0.5f);
                        pl[q] = xn - fl;
                        c[q] = (int)fl - cell0[q];
                        in_region = in_region && c[q] >= -1 && c[q] <= 4;
                    }
                    if (in_region) touch |= touched_blocks(c);
                    pl[3] = m;
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        pl[4 + r] = m * v[r] - h * (A[r] * pl[0] + A[3 + r] * pl[1] + A[6 + r] * pl[2]);
                        pl[7 + r] = h * A[r];
                        pl[10 + r] = h * A[3 + r];
                        pl[13 + r] = h * A[6 + r];
                    }
                    // Core cell: rank-major slot; shell cell or a rank past the
                    // core slots: overflow entry; overflow full or outside the
                    // region: straight to the global grid.
                    float4 *dst = nullptr;
                    uint32_t stride = 0;
                    if (in_region) {
                        const bool core = c[0] >= 0 && c[0] <= 3 && c[1] >= 0 && c[1] <= 3 && c[2] >= 0 && c[2] <= 3;
                        if (core) {
                            const uint32_t l = (uint32_t)(c[0] | (c[1] << 2) | (c[2] << 4));
                            const uint32_t r =
                                (MODE == XM_G2P2G && LIGHT) ? early_rank : atomicAdd(&core_count[l], 1u);
                            if (r < (uint32_t)XF_RANKS) {
                                dst = payc + r * 64u + l;
                                stride = XF_RANKS * 64;
                            }
                        }
                        if (!dst) {
                            const uint32_t e = atomicAdd(&nx, 1u);
                            if (e < (uint32_t)XF_OVF) {
                                dst = payx + e;
                                stride = XF_OVF;
                                xcell[e] = (uint8_t)((c[0] + 1) + 6 * (c[1] + 1) + 36 * (c[2] + 1));
                            }
                        }
                    }
                    if (dst) {
                        dst[0] = make_float4(pl[0], pl[1], pl[2], pl[3]);
                        dst[stride] = make_float4(pl[4], pl[5], pl[6], pl[7]);
                        dst[2 * stride] = make_float4(pl[8], pl[9], pl[10], pl[11]);
                        dst[3 * stride] = make_float4(pl[12], pl[13], pl[14], pl[15]);
                    } else {
                        const int gc[3] = {c[0] + 4 * (int)b[0], c[1] + 4 * (int)b[1], c[2] + 4 * (int)b[2]};
                        scatter_global(a.B, G, S1.epoch, ds, gc, pl);
                    }
                    if constexpr (COLL) {
                        if (cbits >> 16) {
                            const int gc[3] = {c[0] + 4 * (int)b[0], c[1] + 4 * (int)b[1], c[2] + 4 * (int)b[2]};
                            twin_scatter(a.coll.twins.tab[S1.epoch & 1u], a.coll.twins.mask, G, ds, gc, pl, cbits);
                        }
                    }
                }
            }

            if constexpr (HAS_P2G) {
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) touch |= (uint32_t)__shfl_xor((int)touch, o, 64);
                if (lane == 0) atomicOr(&blk_touch, touch);
            }

            // The next pass: the block's next pass_cap particles, or the next block.
            if (ce < end) {
                wn = w;
                wn.cs = ce;
                have_next = true;
            } else {
                have_next = get_block(w.j + 1, wn);
            }
            if (have_next) issue(wn, (npass + 1u) & 1u ? 255u : 0u);

            if constexpr (HAS_P2G) {
                __syncthreads();
                __builtin_amdgcn_s_setprio(1);
                for (uint32_t s = tid; s < (uint32_t)ARENA4; s += XF_THREADS) shared4[s] = make_float4(0.f, 0.f, 0.f, 0.f);
                __syncthreads();
                if (wave < 3u) {
                    // The block's own 64 cells, one stencil plane per wave.
                    const uint32_t cnt = min(core_count[lane], (uint32_t)XF_RANKS);

                    accumulate_plane<true>(
                        (int)wave, cnt, (int)(lane & 3u) + 1, (int)((lane >> 2) & 3u) + 1, (int)(lane >> 4) + 1,
                        [&](uint32_t r, float4 &a0, float4 &a1, float4 &a2, float4 &a3) {
                            const float4 *q = payc + r * 64u + lane;
                            a0 = q[0];
                            a1 = q[XF_RANKS * 64];
                            a2 = q[2 * XF_RANKS * 64];
                            a3 = q[3 * XF_RANKS * 64];
                        },
                        planes + wave * PSLOT);
                } else {
                    // Overflow entries (shell cells, ranks past the core slots):
                    // bucketed by cell; lane = occupied cell, 64 at a time.
                    const uint32_t ne = min(nx, (uint32_t)XF_OVF);
                    if (ne > 0) {
                        for (uint32_t q = lane; q < (uint32_t)NCELL; q += 64u) xcount[q] = 0u;
                        __builtin_amdgcn_wave_barrier();
                        for (uint32_t e = lane; e < ne; e += 64u)
                            xrank[e] = (uint8_t)atomicAdd(&xcount[xcell[e]], 1u);
                        __builtin_amdgcn_wave_barrier();
                        // Exclusive scans of the counts and of the occupied cells, 4 cells per lane.
                        uint32_t cnt4[4], sum = 0, occ = 0;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const uint32_t q = 4u * lane + (uint32_t)t;
                            cnt4[t] = q < (uint32_t)NCELL ? xcount[q] : 0u;
                            sum += cnt4[t];
                            occ += cnt4[t] ? 1u : 0u;
                        }
                        uint32_t incl = sum, iocc = occ;
#pragma unroll
                        for (int o = 1; o < 64; o <<= 1) {
                            const uint32_t y = __shfl_up(incl, o, 64), yo = __shfl_up(iocc, o, 64);
                            if (lane >= (uint32_t)o) {
                                incl += y;
                                iocc += yo;
                            }
                        }
                        const uint32_t nocc = __shfl(iocc, 63, 64);
                        uint32_t run = incl - sum, orun = iocc - occ;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const uint32_t q = 4u * lane + (uint32_t)t;
                            if (q < (uint32_t)NCELL) xstart[q] = run;
                            if (cnt4[t]) xocc[orun++] = (uint8_t)q;
                            run += cnt4[t];
                        }
                        __builtin_amdgcn_wave_barrier();
                        for (uint32_t e = lane; e < ne; e += 64u) xperm[xstart[xcell[e]] + xrank[e]] = (uint8_t)e;
                        __builtin_amdgcn_wave_barrier();
#pragma unroll 1
                        for (uint32_t o0 = 0; o0 < nocc; o0 += 64u) {
                            const uint32_t o = o0 + lane;
                            const uint32_t q = o < nocc ? xocc[o] : 0u;
                            const uint32_t cnt = o < nocc ? xcount[q] : 0u;
                            const uint32_t st = cnt ? xstart[q] : 0u;
                            const int cx = (int)(q % 6u), cy = (int)((q / 6u) % 6u), cz = (int)(q / 36u);
#pragma unroll 1
                            for (int i = 0; i < 3; ++i)
                                accumulate_plane<false>(
                                    i, cnt, cx, cy, cz,
                                    [&](uint32_t r, float4 &a0, float4 &a1, float4 &a2, float4 &a3) {
                                        const float4 *px = payx + xperm[st + r];
                                        a0 = px[0];
                                        a1 = px[XF_OVF];
                                        a2 = px[2 * XF_OVF];
                                        a3 = px[3 * XF_OVF];
                                    },
                                    shell);
                        }
                    }
                }
                // Last pass of the block: list the blocks around it that its
                // particles reach as candidates for the grid update and flag
                // them (plain stores: several entries may list one block, and
                // the grid update keeps the one whose exchange finds the flag).
                if (ce >= end && tid < (uint32_t)NCAND) {
                    const int dx = (int)(tid % 3u) - 1, dy = (int)((tid / 3u) % 3u) - 1, dz = (int)(tid / 9u) - 1;
                    const int bx = (int)b[0] + dx, by = (int)b[1] + dy, bz = (int)b[2] + dz;
                    uint32_t T = NO_BLOCK;
                    if (((blk_touch >> tid) & 1u) && bx >= 0 && by >= 0 && bz >= 0 && bx < G.nb[0] && by < G.nb[1] &&
                        bz < G.nb[2]) {
                        T = block_linear(G, (uint32_t)bx, (uint32_t)by, (uint32_t)bz);
                        a.B.blk_flag[T] = 1u;
                    }
                    a.B.cand[NCAND * bi + tid] = T;
                }
                __syncthreads();
                __builtin_amdgcn_s_setprio(0);
                // Flush the arenas to the grid with global float atomics, one
                // component per lane: the eight lanes of a group add the four
                // components of two x-neighbouring nodes that share a 32-byte
                // sector, so that each atomic request to memory (the float
                // atomics are carried out beyond the L2) carries whole nodes.
                // Zero components are skipped.  The region's x runs in classes,
                // the pairs (1, 2), (3, 4), (5, 6) and (0, 7), and each class
                // reads the plane arenas that can reach it only (the planes of
                // waves 0-2 cover x in [1 + w, 4 + w]): 22 LDS reads per thread
                // instead of 32, all in flight together.  Thread t takes
                // component t & 3 of region x = class pair[(t >> 2) & 1] at
                // (y, z) index 8 (4 h + wave) + ((t >> 3) & 7), h = 0, 1.
                static_assert(4 * 64 == XF_THREADS, "four waves of eight (y, z) positions");
                if constexpr (LIGHT) {
                    const uint32_t comp = tid & 3u, xsel = (tid >> 2) & 1u, pos = (tid >> 3) & 7u;
                    const float *shellf = reinterpret_cast<const float *>(shell) + comp;
                    const float *planef = reinterpret_cast<const float *>(planes) + comp;
                    float *gridf = reinterpret_cast<float *>(a.B.grid) + comp;
                    // region slot (rx, ry, rz); pmask: the planes that may reach its class
                    auto flush_slot = [&](int rx, int ry, int rz, uint32_t pmask) {
                        float g = shellf[4 * arena_slot(rx, ry, rz)];
#pragma unroll
                        for (int w = 0; w < 3; ++w) {
                            if (!((pmask >> w) & 1u)) continue;
                            const int u = rx - 1 - w;
                            const bool reach = u >= 0 && u <= 3;
                            const float o = planef[4 * (w * PSLOT + plane_slot(reach ? u : 0, ry, rz))];
                            g += reach ? o : 0.f;
                        }
                        if (!(g != 0.f)) return;
                        const int x = 4 * (int)b[0] + rx - 1, y = 4 * (int)b[1] + ry - 1, z = 4 * (int)b[2] + rz - 1;
                        if (!node_in_grid(G, x, y, z)) {
                            atomicOr(&ds->status, MPM_STATUS_OUT_OF_GRID | (KID_P2G << 16));
                            return;
                        }
                        unsafeAtomicAdd(gridf + 4u * node_index(G, (uint32_t)x, (uint32_t)y, (uint32_t)z), g);
                    };
#pragma unroll
                    for (int hh = 0; hh < 2; ++hh) {
                        const int idx = (hh * 4 + (int)wave) * 8 + (int)pos;
                        const int ry = idx & 7, rz = idx >> 3;
                        flush_slot(3 + (int)xsel, ry, rz, 7u);
                        flush_slot(1 + (int)xsel, ry, rz, 3u);
                        flush_slot(5 + (int)xsel, ry, rz, 6u);
                        flush_slot(xsel ? 7 : 0, ry, rz, 0u);
                    }
                } else {
                    // The kernels with the SVD, mesh colliders or the network (whose
                    // registers leave less room): thread t takes component t & 3 of
                    // region slot 64 k + (t >> 2) in each of the 8 region sub-blocks
                    // k, four sub-blocks at a time, every plane read and dropped
                    // where it does not reach.
                    const uint32_t lo = tid >> 2, comp = tid & 3u;
                    const int lx = (int)(lo & 3u), ly = (int)((lo >> 2) & 3u), lz = (int)(lo >> 4);
                    const float *shellf = reinterpret_cast<const float *>(shell) + comp;
                    const float *planef = reinterpret_cast<const float *>(planes) + comp;
                    float *gridf = reinterpret_cast<float *>(a.B.grid) + comp;
#pragma unroll 4
                    for (int k = 0; k < 8; ++k) {
                        const int rx = lx | ((k & 1) << 2), ry = ly | ((k & 2) << 1), rz = lz | (k & 4);
                        float g = shellf[4 * (64 * k + (int)lo)];
#pragma unroll
                        for (int w = 0; w < 3; ++w) {
                            const int u = rx - 1 - w;
                            const bool reach = u >= 0 && u <= 3;
                            const float o = planef[4 * (w * PSLOT + plane_slot(reach ? u : 0, ry, rz))];
                            g += reach ? o : 0.f;
                        }
                        if (!(g != 0.f)) continue;
                        const int x = 4 * (int)b[0] + rx - 1, y = 4 * (int)b[1] + ry - 1, z = 4 * (int)b[2] + rz - 1;
                        if (!node_in_grid(G, x, y, z)) {
                            atomicOr(&ds->status, MPM_STATUS_OUT_OF_GRID | (KID_P2G << 16));
                            continue;
                        }
                        unsafeAtomicAdd(gridf + 4u * node_index(G, (uint32_t)x, (uint32_t)y, (uint32_t)z), g);
                    }
                }
                if (tid < 64u) core_count[tid] = 0u;
                if (tid == 0) nx = 0u;
                __syncthreads();
            } else {
                __syncthreads();  // the next pass restages the velocities
            }
        }

        w = wn;
        have = have_next;
        ++npass;
    }
    if constexpr (LEARN) {
        // the closure's statistics of this work-group's particles, into its slot
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            n_eval += (uint32_t)__shfl_xor((int)n_eval, o, 64);
            n_fallback += (uint32_t)__shfl_xor((int)n_fallback, o, 64);
            n_judged += (uint32_t)__shfl_xor((int)n_judged, o, 64);
            err2 += __shfl_xor(err2, o, 64);
            ref2 += __shfl_xor(ref2, o, 64);
        }
        if (lane == 0 && (n_eval | n_fallback | n_judged)) {
            ClosureStats &S = a.learn.stats[blockIdx.x % CLOSURE_STATS_SLOTS];
            atomicAdd(&S.evaluated, (unsigned long long)n_eval);
            atomicAdd(&S.fallbacks, (unsigned long long)n_fallback);
            atomicAdd(&S.judged, (unsigned long long)n_judged);
            unsafeAtomicAdd(&S.err2, err2);
            unsafeAtomicAdd(&S.ref2, ref2);
        }
    }
    if constexpr (LIMITS) {
        // non-negative floats order as their bit patterns
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lim_v = fmaxf(lim_v, __shfl_xor(lim_v, o, 64));
            lim_c = fmaxf(lim_c, __shfl_xor(lim_c, o, 64));
            lim_any |= (uint32_t)__shfl_xor((int)lim_any, o, 64);
        }
        if (lane == 0) {
            atomicMax(&ds->limit_bits[0], __float_as_uint(lim_v));
            atomicMax(&ds->limit_bits[1], __float_as_uint(lim_c));
            if (lim_any) atomicMax(&ds->limit_bits[2], 1u);
        }
    }
}

// ------------------------------------------------------------ grid update
// Grids of up to this many blocks, and at most DENSE_FLAG_SCAN_RATIO times
// as many as the candidate slots of the launch, are updated from their block
// flags (4 bytes per block, read once); other ones from the candidate slots.
constexpr uint32_t DENSE_FLAG_SCAN_BLOCKS = 1u << 22;
constexpr uint64_t DENSE_FLAG_SCAN_RATIO = 4;
// A prime above every block count: i -> (i * FLAG_SCAN_STRIDE) mod n is a
// permutation of [0, n).
constexpr uint64_t FLAG_SCAN_STRIDE = 4294967291ull;

// One wave per touched block T (lane = node): each flagged block is processed
// once: v = p / m, and the accumulation grid is cleared for the next scatter.
// The flagged blocks are found in one of two ways:
//   * grids of up to DENSE_FLAG_SCAN_BLOCKS blocks, and up to
//     DENSE_FLAG_SCAN_RATIO times the candidate slots: the flags themselves,
//     32 per wave step, at the permuted indices (32 q + lane) *
//     FLAG_SCAN_STRIDE mod n, so that every wave meets about the same number
//     of flagged blocks however the particles cluster;
//   * otherwise (large grids, small scenes): the candidate slots written by
//     the transfer kernel (NCAND
//     per block-list entry, then the overflow list), 64 at a time; a block
//     listed by several entries is processed by the one whose exchange finds
//     its flag set.
__global__ __launch_bounds__(256) void k_grid_update(GridBuffers B, GridGeom G, const SubstepParams *params,
                                                     uint32_t *__restrict__ partial, DeviceScalars *ds) {
    if (ds->halt) return;
    const SubstepParams S = load_invariant(params + ds->grid_param);
    const uint32_t *ovf = B.ovf[S.epoch & 1u];
    // The next transfer reads the next substep's entry.
    if (blockIdx.x == 0 && threadIdx.x == 0) ds->param_index = S.index + 1u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gw = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint32_t nw = gridDim.x * (blockDim.x / 64);
    const uint32_t nlist = ds->num_blocks * (uint32_t)NCAND;
    const uint32_t novf = min(ds->ovf_count[S.epoch & 1u], G.num_blocks_total);
    if (gw == 0 && lane == 0) ds->ovf_count[(S.epoch + 1u) & 1u] = 0u;  // for the next transfer
    const uint32_t ncand = nlist + novf;
    const uint32_t x = lane & 3u, y = (lane >> 2) & 3u, z = lane >> 4;
    uint32_t active = 0;
    // The blocks that lane l found flagged (mine = its index, else NO_BLOCK),
    // processed two at a time.
    auto run = [&](uint32_t mine) {
        uint64_t todo = __ballot(mine != NO_BLOCK);
        // One block's nodes from the node values g (lane = node).
        auto update = [&](uint32_t T, const float4 g) {
            const int bx = (int)(T % (uint32_t)G.nb[0]), by = (int)((T / (uint32_t)G.nb[0]) % (uint32_t)G.nb[1]),
                      bz = (int)(T / ((uint32_t)G.nb[0] * (uint32_t)G.nb[1]));
            const uint32_t node = T * 64u + lane;
            const float m = g.x;
            float v[3] = {0.f, 0.f, 0.f};
            if (m > 0.f) {
                // meld_grid without collider twins (rust/crates/cpu/src/phase/meld_grid.rs:61-66).
                v[0] = g.y / m;
                v[1] = g.z / m;
                v[2] = g.w / m;
                grid_options(S, G, 4 * bx + (int)x, 4 * by + (int)y, 4 * bz + (int)z, v);
                ++active;
            }
            B.vgrid[node] = make_float4(m, v[0], v[1], v[2]);
            B.grid[node] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lane == 0) B.blk_epoch[T] = S.epoch;
        };
        // Two blocks at a time: both loads are in flight before either is used
        // (four, with four copies of the update, measured slower).
        while (todo) {
            const int l0 = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1ull;
            const uint32_t T0 = __shfl(mine, l0, 64);
            uint32_t T1 = NO_BLOCK;
            if (todo) {
                const int l1 = __ffsll((unsigned long long)todo) - 1;
                todo &= todo - 1ull;
                T1 = __shfl(mine, l1, 64);
            }
            const float4 g0 = B.grid[T0 * 64u + lane];
            const float4 g1 = T1 != NO_BLOCK ? B.grid[T1 * 64u + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
            update(T0, g0);
            if (T1 != NO_BLOCK) update(T1, g1);
        }
    };
    const uint32_t n = G.num_blocks_total;
    if (n <= DENSE_FLAG_SCAN_BLOCKS && (uint64_t)n <= DENSE_FLAG_SCAN_RATIO * ((uint64_t)ds->num_blocks * NCAND + novf)) {
        // 32 flags per wave step (half the lanes): twice the waves share
        // the flagged blocks, fewer of them per wave
        for (uint32_t q = gw; q * 32u < n; q += nw) {
            const uint32_t i = q * 32u + lane;
            uint32_t mine = NO_BLOCK;
            if (lane < 32u && i < n) {
                const uint32_t T = (uint32_t)(((uint64_t)i * FLAG_SCAN_STRIDE) % n);
                if (B.blk_flag[T]) {
                    B.blk_flag[T] = 0u;
                    mine = T;
                }
            }
            run(mine);
        }
    } else {
        for (uint32_t chunk = gw; chunk * 64u < ncand; chunk += nw) {
            const uint32_t c = chunk * 64u + lane;
            uint32_t mine = c < nlist ? B.cand[c] : (c < ncand ? ovf[c - nlist] : NO_BLOCK);
            if (mine != NO_BLOCK && atomicExch(&B.blk_flag[mine], 0u) == 0u) mine = NO_BLOCK;
            run(mine);
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

// Particles that are not in any block (tombstoned, sorted to the end): when
// the particle set is being permuted (Pout differs from Pin) they are copied
// to their sorted slot unchanged; their keys are the tombstone key.
__global__ __launch_bounds__(256) void k_copy_dead(ParticleArrays Pin, ParticleArrays Pout,
                                                   const uint32_t *__restrict__ order, uint64_t n,
                                                   uint32_t *__restrict__ keys_out, uint32_t tomb, DeviceScalars *ds) {
    if (ds->halt) return;
    const uint64_t live = ds->live_count;
    const bool copy = Pin.base != Pout.base;
    for (uint64_t i = live + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (copy) {
            const uint32_t p = order[i];
            for (int f = 0; f < NFIELDS; ++f) Pout.at(i)[f * TILE] = Pin.at(p)[f * TILE];
        }
        if (keys_out) keys_out[i] = tomb;
    }
}

// The live particles of a stored state whose det F is below INVERSE_EPS are
// flagged FAILED (the check of prepare_tmp.wesl:74-78 on a state that a call
// stops at before its first substep).
__global__ __launch_bounds__(256) void k_flag_inverted(ParticleArrays P, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t fl = P.flags(i);
        if (fl & F_TOMB) continue;
        float F[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) F[k] = P.F(k, i);
        if (!(det3(F) >= INVERSE_EPS)) P.flags(i) = fl | F_FAILED;
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
        for (int q = 0; q < 3; ++q) {
            x[3ull * o + q] = P.x(q, i);
            v[3ull * o + q] = P.v(q, i);
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
    for (int q = 0; q < 3; ++q) so[3 * i + q] = s[q];
}

int groups_per_cu(const void *kernel, int threads) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return per_cu;
}

// Work-groups of a persistent launch: the engine's persistent_blocks, or all
// resident at once (the occupancy of the kernel times the CU count); a
// multiple of 8 for the XCD mapping.  Each work-group takes a fixed share of
// the block list, so any count is correct.
int resident_groups(int per_cu, const TransferLaunch &L) {
    if (L.persistent_blocks > 0) return std::max(8, L.persistent_blocks & ~7);
    return std::max(8, (per_cu * L.cu_count) & ~7);
}

template <int MODE, bool PERMUTE, bool GOALS, uint32_t SET>
hipError_t launch_one(const TransferLaunch &L, const XferArgs &a, hipStream_t stream) {
    static const int per_cu =
        groups_per_cu(reinterpret_cast<const void *>(&k_transfer<MODE, PERMUTE, GOALS, SET>), XF_THREADS);
    const int groups = resident_groups(per_cu, L);
    hipLaunchKernelGGL((k_transfer<MODE, PERMUTE, GOALS, SET>), dim3(groups), dim3(XF_THREADS), 0, stream, a);
    return hipGetLastError();
}

template <int MODE, uint32_t SET>
hipError_t dispatch_flags(const TransferLaunch &L, bool permute, bool goals, const XferArgs &a, hipStream_t stream) {
    if constexpr (MODE == XM_LIMITS) {
        if (permute) return hipErrorInvalidValue;
        return launch_one<MODE, false, false, SET>(L, a, stream);
    } else if constexpr (MODE == XM_G2P) {
        return permute ? launch_one<MODE, true, false, SET>(L, a, stream) : launch_one<MODE, false, false, SET>(L, a, stream);
    } else {
        if (permute)
            return goals ? launch_one<MODE, true, true, SET>(L, a, stream) : launch_one<MODE, true, false, SET>(L, a, stream);
        return goals ? launch_one<MODE, false, true, SET>(L, a, stream) : launch_one<MODE, false, false, SET>(L, a, stream);
    }
}

template <int MODE>
hipError_t dispatch_set(const TransferLaunch &L, bool permute, bool goals, bool learn, const XferArgs &a,
                        hipStream_t stream) {
    if constexpr (MODE == XM_G2P2G) {
        // the learned snow stress inside the fused kernel (no colliders: the learned loop
        // runs its closure pass between gather and scatter then)
        if (learn && !L.colliders) {
            if (L.batched) return dispatch_flags<MODE, SET_ALL | M_BATCH | M_LEARN>(L, permute, goals, a, stream);
            return dispatch_flags<MODE, SET_ALL | M_LEARN>(L, permute, goals, a, stream);
        }
    }
    if (learn) return hipErrorInvalidValue;
    if (L.colliders) {
        if (L.model_set == SET_POLAR) return dispatch_flags<MODE, SET_POLAR | M_COLL>(L, permute, goals, a, stream);
        return dispatch_flags<MODE, SET_ALL | M_COLL>(L, permute, goals, a, stream);
    }
    if (L.batched) {
        if (L.model_set == SET_POLAR) return dispatch_flags<MODE, SET_POLAR | M_BATCH>(L, permute, goals, a, stream);
        return dispatch_flags<MODE, SET_ALL | M_BATCH>(L, permute, goals, a, stream);
    }
    if (L.model_set == SET_POLAR) return dispatch_flags<MODE, SET_POLAR>(L, permute, goals, a, stream);
    return dispatch_flags<MODE, SET_ALL>(L, permute, goals, a, stream);
}

}  // namespace

hipError_t launch_transfer(const TransferLaunch &L, const TransferCall &t, hipStream_t stream) {
    XferArgs a;
    a.Pin = t.Pin;
    a.Pout = t.permute ? t.Pout : t.Pin;
    a.X = t.X;
    a.order = t.order;
    a.blk_start = t.blk_start;
    a.blk_key = t.blk_key;
    a.G = t.G;

    a.params = t.params;
    a.B = t.B;
    a.keys_out = t.keys_out;
    a.mover_bit = t.mover_bit;
    a.ds = t.ds;
    a.tau_ext = t.tau_ext;
    a.tau_stride = t.tau_stride;
    a.svd_out = t.svd_out;
    a.pass_cap = (uint32_t)XF_THREADS * (uint32_t)std::min(std::max(L.particles_per_lane, 1), XF_PER_THREAD);
    if (L.colliders) {
        if (!t.coll) return hipErrorInvalidValue;
        a.coll = *t.coll;
    }
    if (t.learn) a.learn = *t.learn;
    const bool goals = t.goals;
    switch (t.kind) {
        case TransferKind::P2G: return dispatch_set<XM_P2G>(L, t.permute, goals, false, a, stream);
        case TransferKind::G2P2G: return dispatch_set<XM_G2P2G>(L, t.permute, goals, t.learn != nullptr, a, stream);
        case TransferKind::G2P: return dispatch_set<XM_G2P>(L, t.permute, false, false, a, stream);
        case TransferKind::LIMITS: return dispatch_set<XM_LIMITS>(L, t.permute, false, false, a, stream);
    }
    return hipErrorInvalidValue;
}

// The grid update runs on a fixed number of waves (32 per CU), so that its
// per-wave partial counts have a fixed size, in work-groups of
// L.grid_update_threads (a tuning knob: 1, 2 or 4 waves).
uint32_t grid_update_waves(const TransferLaunch &L) { return (uint32_t)std::max(32, 32 * L.cu_count); }
uint32_t grid_update_groups(const TransferLaunch &L) {
    return grid_update_waves(L) / (uint32_t)(L.grid_update_threads / 64);
}

hipError_t launch_grid_update(const TransferLaunch &L, const GridBuffers &B, const GridGeom &G,
                              const SubstepParams *params, uint32_t *partial, DeviceScalars *ds, hipStream_t stream) {
    hipLaunchKernelGGL(k_grid_update, dim3(grid_update_groups(L)), dim3(L.grid_update_threads), 0, stream, B, G, params,
                       partial, ds);
    return hipGetLastError();
}

hipError_t launch_sum_partials(const TransferLaunch &L, const uint32_t *partial, DeviceScalars *ds,
                               hipStream_t stream) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, stream, partial, grid_update_waves(L), ds);
    return hipGetLastError();
}

hipError_t launch_copy_dead(const ParticleArrays &Pin, const ParticleArrays &Pout, const uint32_t *order, uint64_t n,
                            uint32_t *keys_out, uint32_t tomb, DeviceScalars *ds, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_copy_dead, dim3(grid), dim3(256), 0, stream, Pin, Pout, order, n, keys_out, tomb, ds);
    return hipGetLastError();
}

hipError_t launch_flag_inverted(const ParticleArrays &P, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_flag_inverted, dim3(grid), dim3(256), 0, stream, P, n);
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
