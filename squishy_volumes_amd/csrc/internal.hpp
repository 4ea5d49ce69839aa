// internal.hpp - device data layout and kernel launchers shared by the
// engine's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpm_engine.h"
#include "device_math.hpp"

namespace svamd {

constexpr uint32_t TOMB_KEY_FLAG = 0xffffffffu;

// One set of permuted particle arrays in tiles of 32 particles (AoSoA): tile
// t holds NFIELDS rows of 32 floats, so field f of particle i lives at
//   base + t * NFIELDS * 32 + f * 32 + (i & 31),   t = i >> 5.
// One vector address per particle reaches every field through the load's
// immediate offset (f * 128 bytes < 4 KiB), and a wave of 64 consecutive
// particles touches two full 128-byte segments per field.  F and A are nine
// component fields each, component k = 3*col + row.  A holds the fused
// affine momentum m C - dt 4/h^2 V0 (tau + J sigma_visc) between substeps and
// the APIC matrix C at the end of a step_to call.  p0/p1 are mu/lambda for
// solids and the bulk modulus/exponent for fluids; id is the original
// particle index (the reference's sort_map).
enum : int { FX = 0, FV = 3, FF = 6, FA = 15, FM = 24, FVOL = 25, FP0 = 26, FP1 = 27, FFLAGS = 28, FID = 29, NFIELDS = 30 };
constexpr int TILE = 32;

struct ParticleArrays {
    float *base;
    __host__ __device__ __forceinline__ float *at(uint64_t i) const {
        return base + (i >> 5) * (uint64_t)(NFIELDS * TILE) + (i & 31u);
    }
    __host__ __device__ __forceinline__ float &x(int a, uint64_t i) const { return at(i)[(FX + a) * TILE]; }
    __host__ __device__ __forceinline__ float &v(int a, uint64_t i) const { return at(i)[(FV + a) * TILE]; }
    __host__ __device__ __forceinline__ float &F(int k, uint64_t i) const { return at(i)[(FF + k) * TILE]; }
    __host__ __device__ __forceinline__ float &A(int k, uint64_t i) const { return at(i)[(FA + k) * TILE]; }
    __host__ __device__ __forceinline__ float &m(uint64_t i) const { return at(i)[FM * TILE]; }
    __host__ __device__ __forceinline__ float &vol(uint64_t i) const { return at(i)[FVOL * TILE]; }
    __host__ __device__ __forceinline__ float &p0(uint64_t i) const { return at(i)[FP0 * TILE]; }
    __host__ __device__ __forceinline__ float &p1(uint64_t i) const { return at(i)[FP1 * TILE]; }
    __host__ __device__ __forceinline__ uint32_t &flags(uint64_t i) const {
        return *reinterpret_cast<uint32_t *>(&at(i)[FFLAGS * TILE]);
    }
    __host__ __device__ __forceinline__ uint32_t &id(uint64_t i) const {
        return *reinterpret_cast<uint32_t *>(&at(i)[FID * TILE]);
    }
};

// Arrays indexed by original particle index: never permuted.
struct AuxArrays {
    float *visc_dyn;
    float *visc_bulk;
    float *sand_alpha;
    float *jp;
    float *energy;
    float *goal_a[3];
    float *goal_b[3];
    uint32_t *goal_flags_a;
    uint32_t *goal_flags_b;
    uint32_t *collider_bits;   // per particle (the reference keeps them with the position)
};

// ---- mesh colliders (collide.hip, collide.hpp)
// The mesh on the device: topology, the per-call frames and their
// interpolation for the current substep, and the candidate triangles of each
// grid block (every triangle within the forget distance of a point of the
// block, in ascending order).
struct MeshDev {
    uint32_t ntri = 0, nvert = 0, ncolliders = 0;
    const uint32_t *tri = nullptr;       // 3 vertex indices per triangle
    const uint32_t *tri_coll = nullptr;  // collider of each triangle
    const uint32_t *opp = nullptr;       // 3 opposite triangles per triangle (ab, bc, ca)
    const uint32_t *vtx_off = nullptr;   // triangles at each vertex: vtx_tri[vtx_off[v] .. vtx_off[v + 1])
    const uint32_t *vtx_tri = nullptr;
    const float *pos_a = nullptr, *pos_b = nullptr;    // 3 per vertex, frames a and b
    const float *fric_a = nullptr, *fric_b = nullptr;  // per triangle
    const float *damp_a = nullptr, *damp_b = nullptr;
    float4 *vvel = nullptr;              // vertex velocities of the call
    float4 *vpos = nullptr;              // interpolated for the substep
    float4 *vnorm = nullptr;
    float4 *tnorm = nullptr;
    float *tfric = nullptr, *tdamp = nullptr;
    const uint32_t *blk_off = nullptr;   // num_blocks_total + 1
    const uint32_t *blk_tri = nullptr;
    float forget = 0.f, accept = 0.f;    // 2.2 h and 2 h (InputConsts::forget_distance / accept_distance)
};

// Twin nodes: the nodes that particles near a collider scatter to under their
// collider bits, beside the plain grid that holds every particle's
// contribution.  A hash table by node index; an entry holds up to TWIN_SLOTS
// distinct bit patterns.  After the meld, mp holds (m, v) of the twin.
constexpr int TWIN_SLOTS = 4;
constexpr uint32_t TWIN_EMPTY = 0xffffffffu;
constexpr int TWIN_PROBES = 64;
struct __align__(16) TwinEntry {
    uint32_t node;
    uint32_t bits[TWIN_SLOTS];
    uint32_t pad[3];
    float4 mp[TWIN_SLOTS];
};
struct TwinTables {
    TwinEntry *tab[2] = {nullptr, nullptr};  // by parity of the substep's epoch
    uint32_t mask = 0;                        // capacity - 1 (a power of two)
};

// Dense block grid.  Local node (x, y, z) = global node - origin lives in
// block (x>>2, y>>2, z>>2), linear block index bx + nb0*(by + nb1*bz), and at
// slot (x&3) + 4*(y&3) + 16*(z&3) of the block's 64 nodes.
struct GridGeom {
    int origin[3];
    int nb[3];
    float h;
    float inv_h;         // 1/h rounded to nearest: x / h is div_by(x, h, inv_h)
    uint32_t axis_bits;  // Morton bits per axis; 2^axis_bits >= nb[a]
    uint32_t tomb_key;   // all-ones in 3*axis_bits bits: never a valid block
    uint32_t num_blocks_total;
};

// Device-side counters, one instance in device memory.
struct DeviceScalars {
    uint32_t num_blocks;    // occupied blocks this substep
    uint32_t live_count;    // particles with a valid key (sorted to the front)
    uint32_t status;        // MPM_STATUS_* bits | kernel id << 16
    uint32_t active_nodes;  // nodes with m > 0 after the last grid update (summed on request)
    uint32_t ovf_count[2];  // overflow-list lengths, by parity of the grid's epoch
    uint32_t param_index;   // entry of the substep parameter table the next transfer reads
                            // (the grid update sets it to its own entry + 1)
    uint32_t grid_param;    // entry the next grid update reads (the transfer sets it)
    uint32_t fail_index;    // smallest parameter-table entry (transfer index) of a transfer that
                            // flagged a particle FAILED in the call (all ones: none)
    uint32_t limit_bits[3]; // the LIMITS transfer's largest |v| and |C_ij| (float bits), and 1 when
                            // it saw a live particle
    uint32_t halt;          // adaptive steps on the device: nonzero once the call's loop has
                            // ended, so that the kernels of the substeps launched past that
                            // point return at once (zero otherwise)
};

// A particle whose deformation gradient has det F below this before its
// stress evaluation is flagged FAILED (INVERSE_EPS, rust/crates/util/src/consts.rs:12;
// rust/crates/gpu/src/shaders/prepare_tmp.wesl:74-78).
constexpr float INVERSE_EPS = 1e-7f;

// Candidate blocks listed per block-list entry for the grid update: the 27
// blocks around the entry's block that its region reaches.
constexpr int NCAND = 27;

// Block key of a position: Morton code of the 4^3 block holding the base
// cell floor(x/h - 1/2) (position_to_shift_quadratic,
// rust/crates/cpu/src/kernels.rs:46-49), or the tombstone key (and an
// OUT_OF_GRID status) when the block is outside the allocated grid.
__device__ __forceinline__ uint32_t particle_block_key(const float x[3], const GridGeom &G, DeviceScalars *ds) {
    uint32_t b[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float xn = div_by(x[a], G.h, G.inv_h) - 0.5f;
        ok = ok && isfinite(xn);
        const float fl = floorf(xn);
        ok = ok && fl > -1.0e9f && fl < 1.0e9f;
        const int c = (int)fl - G.origin[a];
        ok = ok && c >= 0 && (c >> 2) <= G.nb[a] - 2;
        b[a] = (uint32_t)(c >> 2);
    }
    if (!ok) {
        atomicOr(&ds->status, MPM_STATUS_OUT_OF_GRID | (1u << 16));
        return G.tomb_key;
    }
    return morton3(b[0], b[1], b[2]);
}

struct SubstepParams {
    float dt;          // the step of the scatter and the grid update
    float dt_force;    // the step of the external force and collide of this substep, and of the
                       // advance that a gather of this entry completes (equal to dt unless the
                       // step adapts within the substep)
    float h;
    float gravity[3];  // interpolated gravity of the substep
    float factor_b;    // frame factor (goal interpolation)
    int has_goals;
    int has_b;
    uint32_t options;  // MPM_OPT_*
    float plane_z;
    float wall_min[3];
    float wall_max[3];
    float domain_min[3];
    float domain_max[3];
    float snow_lo, snow_hi, snow_hardening;
    float sphere_center[3], sphere_radius;
    // Batched scenes: copy k of the domain spans [scene_origin + k period,
    // scene_origin + (k + 1) period) along each axis, k = 0 .. scene_last.
    float scene_period[3], scene_inv_period[3], scene_last[3], scene_origin[3];
    uint32_t epoch;    // substep counter, stamps grid blocks
    uint32_t index;    // this entry's index in the substep parameter table
};

// Coordinate a of a position within its own scene of a batch (the domain
// repeated with period scene_period[a]); unchanged for one scene.
__device__ __forceinline__ float scene_local(float x, int a, const SubstepParams &S) {
    if (!(S.scene_last[a] > 0.f)) return x;
    const float k = fminf(fmaxf(floorf((x - S.scene_origin[a]) * S.scene_inv_period[a]), 0.f), S.scene_last[a]);
    return x - k * S.scene_period[a];
}

// Grid-side extensions without a reference counterpart, applied to the
// velocity v of a node with mass at local node coordinates (lx, ly, lz): grid
// gravity, the sticky plane, box walls and the sphere (separating slip).
__device__ __forceinline__ void grid_options(const SubstepParams &S, const GridGeom &G, int lx, int ly, int lz,
                                             float v[3]) {
    if (S.options & MPM_OPT_GRID_GRAVITY) {
#pragma unroll
        for (int q = 0; q < 3; ++q) v[q] += S.gravity[q] * S.dt;
    }
    if (S.options & (MPM_OPT_STICKY_PLANE | MPM_OPT_BOX_WALLS | MPM_OPT_SPHERE)) {
        const float xi[3] = {scene_local((float)(lx + G.origin[0]) * S.h, 0, S),
                             scene_local((float)(ly + G.origin[1]) * S.h, 1, S),
                             scene_local((float)(lz + G.origin[2]) * S.h, 2, S)};
        if ((S.options & MPM_OPT_STICKY_PLANE) && xi[2] < S.plane_z) v[0] = v[1] = v[2] = 0.f;
        if (S.options & MPM_OPT_BOX_WALLS) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (xi[q] < S.wall_min[q] && v[q] < 0.f) v[q] = 0.f;
                if (xi[q] > S.wall_max[q] && v[q] > 0.f) v[q] = 0.f;
            }
        }
        if (S.options & MPM_OPT_SPHERE) {
            const float d[3] = {xi[0] - S.sphere_center[0], xi[1] - S.sphere_center[1], xi[2] - S.sphere_center[2]};
            const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 < S.sphere_radius * S.sphere_radius && r2 > 0.f) {
                const float vn = (v[0] * d[0] + v[1] * d[1] + v[2] * d[2]) / r2;
                if (vn < 0.f) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) v[q] -= vn * d[q];
                }
            }
        }
    }
}

enum KernelId : uint32_t {
    KID_KEYS = 1,
    KID_SORT = 2,
    KID_BLOCKS = 3,
    KID_P2G = 4,
    KID_GRID = 5,
    KID_G2P = 6,
    KID_COPY = 7
};

// ---- sort.hip
size_t radix_sort_temp_bytes(uint64_t n);
hipError_t radix_sort(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out, uint32_t *vals_out,
                      uint32_t *keys_tmp, uint32_t *vals_tmp, uint64_t n, int key_bits, void *temp,
                      hipStream_t stream);
size_t scan_temp_bytes(uint64_t n);
hipError_t exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t *total, uint64_t n, void *temp,
                          hipStream_t stream);
hipError_t launch_keys(const ParticleArrays &P, const GridGeom &G, uint64_t n, uint32_t *keys, DeviceScalars *ds,
                       hipStream_t stream);
size_t block_list_temp_bytes(uint64_t n);
hipError_t launch_block_list(const uint32_t *keys, uint64_t n, uint32_t tomb_key, uint32_t *blk_start,
                             uint32_t *blk_key, uint32_t capacity, DeviceScalars *ds, void *temp, hipStream_t stream);
// The re-sort of keys that a transfer wrote, for key ranges of up to
// 2^BUCKET_SORT_MAX_BITS keys: a counting sort over every key.  The transfer
// marks with BUCKET_MOVER a particle whose key differs from that of the block
// it was processed in; the others (stayers) come in key order, since the
// slots are grouped by block in key order.  Stayers keep their input order
// within a key (they take their slots from a prefix count); movers follow
// their key's stayers in the order of the atomics that place them.  It writes
// order[] (sorted slot -> input slot), the occupied-block list and the live
// count, as radix_sort and launch_block_list together do.
constexpr int BUCKET_SORT_MAX_BITS = 22;
constexpr uint32_t BUCKET_MOVER = 0x80000000u;
struct BucketBuffers {
    uint32_t *counts = nullptr, *stay = nullptr;       // per key: particles, stayers (zero between sorts)
    uint32_t *cursor = nullptr, *stay_base = nullptr;  // per key: the movers' next slot, the stayers' base
    uint32_t *tiles = nullptr;                         // bucket_tile_sums(key_bits, n) entries
    uint32_t *span_stay = nullptr;                     // per span of input slots: bucket_span_count(n)
};
uint64_t bucket_span_count(uint64_t n);
size_t bucket_tile_sums(int key_bits, uint64_t n);
hipError_t bucket_sort(const uint32_t *keys, uint64_t n, int key_bits, const BucketBuffers &B, uint32_t *order,
                       uint32_t *blk_start, uint32_t *blk_key, uint32_t capacity, DeviceScalars *ds,
                       hipStream_t stream);
hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t stream);
hipError_t launch_reset_scalars(DeviceScalars *ds, hipStream_t stream);

// ---- transfer.hip
struct TransferLaunch {
    int persistent_blocks;  // work-groups of the block-loop kernels; 0: all resident ones
    uint32_t model_set;     // SET_POLAR or SET_ALL
    bool batched;           // batched scenes (mpm_consts.scene_repeat): the M_BATCH kernels
    int cu_count;
    int particles_per_lane; // particles per thread per pass of a block, 1..3 (pass = 256 x this)
    int grid_update_threads = 256;  // work-group size of the grid update: 64, 128 or 256
    bool colliders = false; // a collider mesh is set: the M_COLL kernels
};

// The three modes of the transfer kernel: the first scatter of a call (from
// the stored v, C, F), gather + advance + next scatter, and the last gather.
enum class TransferKind { P2G, G2P2G, G2P, LIMITS };

// Grid-side buffers shared by the transfer kernel and the grid update.
struct GridBuffers {
    float4 *grid;               // (m, p): the scatter adds into it with float atomics
    float4 *vgrid;              // (m, v) after the grid update
    uint32_t *blk_flag;         // set for blocks that received mass; cleared by the grid update
    uint32_t *blk_epoch;        // epoch of the block's last grid update
    uint32_t *cand;             // NCAND candidate blocks per list entry (a block may be listed
                                // by several entries)
    uint32_t *ovf[2];           // overflow lists of claimed blocks, by parity of the epoch
};

// Collider state a transfer needs (mesh set: the M_COLL kernels).
struct ColliderCall {
    MeshDev mesh;
    TwinTables twins;
};

struct LearnedCall;
struct StagedNet;

struct TransferCall {
    TransferKind kind;
    bool permute;               // slot i takes Pin[order[i]] and is written to Pout; else in place in Pin
    ParticleArrays Pin, Pout;
    AuxArrays X;
    const uint32_t *order;
    const uint32_t *blk_start;
    const uint32_t *blk_key;
    GridGeom G;
    const SubstepParams *params; // the engine's table, read at entry ds->param_index: the
                                 // substep of the scatter (the gather's dt, domain and
                                 // plasticity constants are the same every substep)
    bool goals;                 // that substep has goal positions
    GridBuffers B;
    uint32_t *keys_out;         // null: keys not needed
    uint32_t mover_bit = 0;     // BUCKET_MOVER: keys_out marks particles that left their block
    DeviceScalars *ds;
    const float *tau_ext = nullptr;  // P2G: snow stress from the learned closure (rows of tau_stride)
    uint64_t tau_stride = 0;
    float *svd_out = nullptr;        // G2P: U and s of each snow particle's new F (12 rows of tau_stride)
    const ColliderCall *coll = nullptr;  // mesh colliders, or null
    const LearnedCall *learn = nullptr;  // G2P2G: the learned snow stress evaluated in the kernel
};
hipError_t launch_transfer(const TransferLaunch &L, const TransferCall &t, hipStream_t stream);
uint32_t grid_update_waves(const TransferLaunch &L);
hipError_t launch_grid_update(const TransferLaunch &L, const GridBuffers &B, const GridGeom &G,
                              const SubstepParams *params, uint32_t *partial, DeviceScalars *ds, hipStream_t stream);
hipError_t launch_sum_partials(const TransferLaunch &L, const uint32_t *partial, DeviceScalars *ds,
                               hipStream_t stream);
hipError_t launch_copy_dead(const ParticleArrays &Pin, const ParticleArrays &Pout, const uint32_t *order,
                            uint64_t n, uint32_t *keys_out, uint32_t tomb, DeviceScalars *ds, hipStream_t stream);
hipError_t launch_flag_inverted(const ParticleArrays &P, uint64_t n, hipStream_t stream);

// ---- collide.hip
// The mesh for the substep of parameter entry ds->param_index.
hipError_t launch_mesh_interp(const MeshDev &m, const SubstepParams *params, DeviceScalars *ds, hipStream_t stream);
// The twins of the substep just scattered melded against the plain grid's
// totals (before the grid update clears them); the other parity's table
// cleared for the next scatter.
hipError_t launch_twin_meld(const TwinTables &t, const float4 *grid, const GridGeom &G, const SubstepParams *params,
                            DeviceScalars *ds, hipStream_t stream);
hipError_t launch_twin_clear(const TwinTables &t, hipStream_t stream);
hipError_t launch_unpermute(const ParticleArrays &P, uint64_t n, float *x, float *v, float *F, float *C,
                            uint32_t *flags, hipStream_t stream);
hipError_t launch_stress(const uint32_t *flags, const mpm_particle_parameters *params, const float *F,
                         const float *jp, float snow_hardening, float *tau, float *energy, uint64_t n,
                         hipStream_t stream);
hipError_t launch_svd(const float *F, float *U, float *s, float *V, uint64_t n, hipStream_t stream);

// ---- closure.hip: the learned stress model, a 6-64-64-3 perceptron.  Its
// FP32 parameters are one padded array: W1 [64][16] (input columns 6..15
// zero), b1 [64], W2 [64][64], b2 [64], W3 [16][64] (output rows 3..15
// zero), b3 [16]; W[out][in], so layer k computes z = W h + b.
constexpr int MLP_IN = 6, MLP_H = 64, MLP_OUT = 3, MLP_KIN = 16;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_KIN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + 16 * MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 16;
hipError_t launch_mlp_forward(const float *P, const float *x, float *y, uint64_t n, hipStream_t stream);
hipError_t launch_mlp_grad(const float *P, const float *x, const float *t, uint64_t n, float scale, float *grad,
                           float *loss, hipStream_t stream);
hipError_t launch_mlp_adam(float *P, float *G, float *M, float *V, float lr, float b1, float b2, float eps, int step,
                           hipStream_t stream);

// Network inputs of a snow particle: (s_q - 1) / SIGMA for its singular
// values s_0 >= s_1 >= s_2, log10(E / 1e5), (nu - 0.3) / 0.1 and
// (Jp - 1) / JP; outputs: principal Kirchhoff stresses in units of
// E exp(xi (1 - Jp)) STRESS (the hardened Young's modulus).
// The network's output is the principal stress minus its linear-elastic
// part, in units of E h CLOSURE_STRESS_SCALE m (h the hardening factor,
// m = max(|s - 1| / 0.025, 1e-6) the strain magnitude): that remainder is
// second order in the strain, so the output is first order in it, and
// 0.025^2 makes it order one at the snow model's compression limit.  The
// network's absolute error then scales with the strain, as the stress does.
constexpr float CLOSURE_SIGMA_SCALE = 0.025f, CLOSURE_JP_SCALE = 0.1f, CLOSURE_STRESS_SCALE = 0.025f * 0.025f;
__host__ __device__ __forceinline__ float closure_strain_scale(const float s[3]) {
    const float e0 = s[0] - 1.f, e1 = s[1] - 1.f, e2 = s[2] - 1.f;
    return fmaxf(sqrtf(e0 * e0 + e1 * e1 + e2 * e2) * (1.0f / CLOSURE_SIGMA_SCALE), 1e-6f);
}
struct ClosureStats {
    unsigned long long evaluated, fallbacks, judged, pad;
    double err2, ref2;  // sum |tau - tau_analytic|^2 and sum |tau_analytic|^2 over the kept particles
};
// The closure kernel runs at most this many work-groups, each summing into
// its own statistics slot.
constexpr uint32_t CLOSURE_STATS_SLOTS = 4096;
struct ClosureLaunch {
    ParticleArrays P;
    const uint32_t *order;  // slot i reads P at order[i], or null
    AuxArrays X;
    const float *params;
    float *tau;             // row k of slot i at tau[k * stride + i]
    uint64_t stride, n;
    float hull_lo[MLP_IN], hull_hi[MLP_IN];
    float hardening;
    uint32_t judge_stride, judge_phase;
    ClosureStats *stats;    // CLOSURE_STATS_SLOTS slots
    const float *svd = nullptr;  // U (rows 0-8) and s (rows 9-11) of each snow particle's F by
                                 // P slot (stride rows), written by the gather; null: the
                                 // closure takes the SVD itself
    const DeviceScalars *ds = nullptr;  // its halt flag (adaptive steps on the device), or null
};
hipError_t launch_closure(const ClosureLaunch &c, hipStream_t stream);
// The network and its judge for the fused gather-scatter (M_LEARN kernels).
struct LearnedCall {
    const StagedNet *net;   // the parameters packed as the lanes read them (launch_pack_net)
    float hull_lo[MLP_IN], hull_hi[MLP_IN];
    uint32_t judge_stride, judge_phase;
    ClosureStats *stats;    // CLOSURE_STATS_SLOTS slots, by work-group
};
hipError_t launch_pack_net(const float *params, StagedNet *out, hipStream_t stream);
uint32_t step_limit_groups(uint64_t n);
// AdaptiveTimeStepState (rust/crates/cpu/src/adaptive_time_step_state.rs:13-70)
// kept on the device for a call with adaptive steps, with the call's time:
// three one-thread kernels per substep update it from the limit reductions
// and write the steps into the two parameter-table entries of the substep
// (0: scatter and grid update, 1: gather and advance).
enum : uint32_t { ADAPT_RUNNING = 0, ADAPT_DONE = 1, ADAPT_FAILED = 2, ADAPT_ZERO_STEP = 3, ADAPT_WRONG_FRAME = 4,
                  ADAPT_TWIN_OVERFLOW = 5 };
struct DeviceAdaptive {
    double time, target;      // the call's time (advanced on the device) and target
    double fps;               // frames per second
    uint64_t frame;           // frame a of the call's input
    float gravity_a[3], gravity_b[3];
    float h;
    float max_time_step, by_velocity, by_deformation, by_sound, by_isolated;
    float prior[11];          // the last pushes, oldest at head
    uint32_t nprior, head;
    uint32_t epoch0;          // the grid epoch at the call's start
    uint32_t substeps;        // substeps advanced by the call
    uint32_t stop;            // ADAPT_*: why the loop ended (RUNNING while it runs)
    float last_advance;       // the step of the last advance
    float dt_force;           // the allowed step at the start of the current substep
};
hipError_t launch_adapt_begin(DeviceAdaptive *A, SubstepParams *table, DeviceScalars *ds, hipStream_t stream);
hipError_t launch_adapt_force(DeviceAdaptive *A, SubstepParams *table, DeviceScalars *ds, const float4 *limits,
                              hipStream_t stream);
hipError_t launch_adapt_integrate(DeviceAdaptive *A, SubstepParams *table, DeviceScalars *ds, hipStream_t stream);

// The live particles' bounding box and largest speed into out[0..7]
// ({lo xyz, hi xyz, max |v|, 1 if any live particle}); part holds
// 2 step_limit_groups(n) float4.
hipError_t launch_particle_bounds(const ParticleArrays &P, uint64_t n, float4 *part, float *out, hipStream_t stream);
// The time-step limits of the stored state into out: {by_sound, by_isolated,
// max |v|, by_deformation}; motion = false skips v and C (max |v| 0,
// by_deformation +inf).
hipError_t launch_step_limits(const ParticleArrays &P, uint64_t n, float h, bool motion, float4 *part, float4 *out,
                              const DeviceScalars *ds, hipStream_t stream);
hipError_t launch_closure_pairs(const ParticleArrays &P, const AuxArrays &X, uint64_t n, float hardening,
                                uint32_t stride, uint32_t phase, float *x, float *t, uint64_t capacity,
                                unsigned long long *count, hipStream_t stream);

}  // namespace svamd
