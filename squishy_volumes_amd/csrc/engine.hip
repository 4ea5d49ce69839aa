// engine.hip - host side of the MLS-MPM engine and its C ABI.
//
// mpm_create / mpm_step_to / mpm_read_particles stand where GpuState::
// from_io_state, GpuState::produce_next_state and the state download of
// rust/crates/gpu/src/gpu_state.rs:38-675 stand.  The substep loop, the time
// accumulation and the frame-factor check follow CpuState::produce_next_state
// (rust/crates/cpu/src/cpu_state.rs:146-197) and FrameInput::frame_factor
// (rust/crates/xpu/src/frame_input.rs:266-278).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <deque>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "device_math.hpp"
#include "closure_net.hpp"
#include "internal.hpp"

using namespace svamd;

namespace {

enum Stage { ST_KEYS = 0, ST_SORT, ST_BLOCKS, ST_P2G, ST_GRID, ST_G2P2G, ST_G2P, ST_COPY, ST_CLOSURE, ST_COUNT };
const char *kStageNames[ST_COUNT] = {"keys",        "sort",       "block_list", "p2g",    "grid_update",
                                     "g2p2g",       "g2p",        "copy_dead",  "closure"};

thread_local std::string g_last_error;

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            g_last_error = std::string(#expr) + ": " + hipGetErrorString(_e);                      \
            return MPM_ERR_HIP;                                                                    \
        }                                                                                          \
    } while (0)

struct EventRecord {
    int stage;
    hipEvent_t a, b;
};

}  // namespace

// A domain whose dense grid would take more than this (two float4 node
// grids) gets a grid window fitted to its particles instead.
constexpr uint64_t WINDOW_FROM_BYTES = 4ull << 30;

// AdaptiveTimeStepState (rust/crates/cpu/src/adaptive_time_step_state.rs:13-70):
// the current limits (a missing limit is f32::MAX, as unwrap_or gives it) and
// the history of up to 11 pushed limits; the allowed step is the smallest of
// them and max_time_step.  It lives as long as the engine, as the reference's
// lives as long as its CpuState.
struct AdaptiveStep {
    float max_time_step = FLT_MAX;
    float by_velocity = FLT_MAX, by_deformation = FLT_MAX, by_isolated = FLT_MAX, by_sound = FLT_MAX;
    std::deque<float> prior;
    float without_prior() const {
        return std::min({max_time_step, by_velocity, by_deformation, by_sound, by_isolated});
    }
    float allowed() const {
        float a = without_prior();
        for (float p : prior) a = std::min(a, p);
        return a;
    }
    void push_current_limit() {
        if (prior.size() > 10) prior.pop_front();  // TIME_STEP_HISTORY_LENGTH
        prior.push_back(without_prior());
    }
};

struct mpm_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    mpm_consts consts{};
    uint64_t n = 0;
    GridGeom geom{};
    int key_bits = 0;
    TransferLaunch launch{};

    // Three particle sets.  A call never writes the set that held its
    // starting state (`start`): its permuting transfers alternate between the
    // other two, so that the call can be run again from its start (a
    // particle that fails stops the call at the failing substep).
    ParticleArrays P[3]{};
    int cur = 0;
    int start = 0;
    AuxArrays X{};
    float *jp_start = nullptr;    // Jp at the start of the call (scenes with snow)

    uint32_t *keys = nullptr, *keys_sorted = nullptr, *order = nullptr;
    void *sort_temp = nullptr;
    // The counting sort's buffers, for 2^bucket_bits keys (0: none held).
    BucketBuffers bucket{};
    int bucket_bits = 0;
    uint32_t *blk_start = nullptr, *blk_key = nullptr;
    uint32_t blk_capacity = 0;
    void *blk_temp = nullptr;
    float4 *grid = nullptr, *vgrid = nullptr;
    uint32_t *blk_flag = nullptr, *blk_epoch = nullptr;
    uint32_t *cand = nullptr;     // NCAND candidate touched blocks per occupied block
    uint32_t *ovf[2] = {nullptr, nullptr};  // overflow lists of claimed blocks, by epoch parity
    uint32_t *partial = nullptr;  // per-wave active-node counts of the grid update
    bool keys_valid = false;      // e->keys holds the keys of the current particle order
    bool pending_permute = false; // a sort ran since the last transfer
    int resort_interval = 8;      // substeps between sorts
    DeviceScalars *ds = nullptr;
    DeviceScalars *ds_host = nullptr;  // pinned

    // Substep parameters of a call, one entry per substep, uploaded before
    // its first launch; transfers read entry ds->param_index, which each
    // grid update advances.  Kernel arguments are therefore the same in every
    // sort cycle, and a cycle captured as a HIP graph replays unchanged.
    SubstepParams *param_table = nullptr;  // device
    SubstepParams *param_host = nullptr;   // pinned staging
    uint32_t param_capacity = 0;
    bool use_graph = true;
    hipGraphExec_t cycle_graph[3][3][2] = {};  // [set read][set written][goals]
    int graph_interval = 0;                // resort interval the graphs were captured with

    double time = 0;
    uint32_t epoch = 0;
    uint32_t grid_epoch = 0;   // epoch stamped by the last grid update
    bool grid_valid = false;   // a grid update has run
    bool stepped = false;
    bool goals_allocated = false;

    std::vector<mpm_particle_parameters> parameters;
    std::vector<float> initial_positions;
    std::vector<uint32_t> collider_bits;
    std::vector<uint8_t> dead_at_create;

    // Learned closure (mpm_set_closure): the snow stress of each substep's
    // scatter comes from the network.
    bool closure_on = false;
    float *closure_params = nullptr;  // MLP_NPARAM
    float *closure_tau = nullptr;     // 9 rows of n
    float *closure_svd = nullptr;     // 12 rows of n: U, s of snow particles from the last gather
    bool svd_fresh = false;           // closure_svd holds the current F's factors
    ClosureStats *closure_stats = nullptr;
    float hull_lo[MLP_IN] = {0}, hull_hi[MLP_IN] = {0};
    uint32_t judge_stride = 0, judge_phase = 0;
    StagedNet *closure_net = nullptr;  // the parameters packed for the fused kernel
    bool learned_fused = false;         // mpm_tuning.learned_fused
    unsigned long long *pair_count = nullptr;  // mpm_closure_pairs
    float4 *limit_part = nullptr, *limit_out = nullptr;  // mpm_time_step_limits
    DeviceAdaptive *adapt_dev = nullptr;    // adaptive steps: the device state
    DeviceAdaptive *adapt_host = nullptr;   // and its pinned host copy
    // mpm_read_particles staging in original order (allocated on first use)
    float *rb_x = nullptr, *rb_v = nullptr, *rb_F = nullptr, *rb_C = nullptr;
    uint32_t *rb_flags = nullptr;
    int64_t cancel_after = -1;  // test hook: calls behave as cancelled before this substep

    // Grid window (mpm_tuning.grid_window): the block grid covers the base
    // cells win_lo..win_hi of the domain's dom_lo..dom_hi, refitted to the
    // particles at the start of a call, instead of the whole domain.
    bool windowed = false;
    long long dom_lo[3] = {0, 0, 0}, dom_hi[3] = {0, 0, 0};
    long long win_lo[3] = {0, 0, 0}, win_hi[3] = {0, 0, 0};
    uint32_t grid_block_cap = 0;     // blocks the grid buffers hold
    uint32_t blk_off_capacity = 0;   // entries of the collider block offsets
    float4 *bounds_part = nullptr;
    float *bounds_out = nullptr;

    // Adaptive time steps (mpm_set_adaptive_time_steps): two limits per substep.
    bool adaptive = false;
    AdaptiveStep ats;

    // Mesh colliders (mpm_set_collider_mesh).
    bool mesh_set = false;
    ColliderCall coll{};                   // device view: mesh and twin tables
    std::vector<uint32_t> mesh_tri;        // host copy of the triangles
    std::vector<float> mesh_pos_a, mesh_pos_b;  // the call's frames (host), for the candidate lists
    std::vector<void *> mesh_allocs;       // freed when the mesh changes
    uint32_t blk_tri_capacity = 0;
    uint32_t twin_capacity = 0;

    bool profiling = false;
    uint32_t profile_mask = 0;  // stages timed while profiling (bit = Stage)
    std::vector<EventRecord> records;
    std::vector<hipEvent_t> event_pool;
    double stage_ms[ST_COUNT] = {0};
    uint64_t stage_launches[ST_COUNT] = {0};

    std::vector<void *> allocations;

    ~mpm_engine() {
        if (device >= 0) (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        free_mesh();
        for (void *p : allocations) (void)hipFree(p);
        if (param_table) (void)hipFree(param_table);
        if (param_host) (void)hipHostFree(param_host);
        if (ds_host) (void)hipHostFree(ds_host);
        if (adapt_host) (void)hipHostFree(adapt_host);
        for (auto &r : records) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void drop_graphs() {
        for (auto &plane : cycle_graph)
            for (auto &row : plane)
                for (auto &g : row)
                    if (g) {
                        (void)hipGraphExecDestroy(g);
                        g = nullptr;
                    }
    }

    void free_mesh() {
        for (void *p : mesh_allocs) (void)hipFree(p);
        mesh_allocs.clear();
        coll = ColliderCall{};
        blk_tri_capacity = 0;
        twin_capacity = 0;
        blk_off_capacity = 0;
    }

    // Device memory of the collider mesh (freed with the mesh).
    template <class T>
    hipError_t mesh_alloc(T *&ptr, size_t count) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256));
        if (e == hipSuccess) mesh_allocs.push_back(p);
        ptr = static_cast<T *>(p);
        return e;
    }
    template <class T>
    void mesh_free(T *ptr) {
        auto it = std::find(mesh_allocs.begin(), mesh_allocs.end(), (void *)ptr);
        if (it != mesh_allocs.end()) {
            (void)hipFree(*it);
            mesh_allocs.erase(it);
        }
    }

    template <class T>
    hipError_t alloc(T *&ptr, size_t count) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256));
        if (e == hipSuccess) allocations.push_back(p);
        ptr = static_cast<T *>(p);
        return e;
    }

    template <class T>
    void release(T *&ptr) {
        if (!ptr) return;
        auto it = std::find(allocations.begin(), allocations.end(), (void *)ptr);
        if (it != allocations.end()) {
            (void)hipFree(*it);
            allocations.erase(it);
        }
        ptr = nullptr;
    }

    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        // Timing only, read after the stream has synchronized: no system-scope
        // fence (its cache writeback would slow the kernels it measures).
        hipEvent_t e = nullptr;
        (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
        return e;
    }

    void begin(int stage, EventRecord &r) {
        r.a = r.b = nullptr;
        if (!profiling || !((profile_mask >> stage) & 1u)) return;
        r.stage = stage;
        r.a = take_event();
        r.b = take_event();
        (void)hipEventRecord(r.a, stream);
    }
    void end(EventRecord &r) {
        if (!r.a) return;
        (void)hipEventRecord(r.b, stream);
        records.push_back(r);
    }
    void harvest() {
        for (auto &r : records) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                stage_ms[r.stage] += ms;
                stage_launches[r.stage] += 1;
            }
            event_pool.push_back(r.a);
            event_pool.push_back(r.b);
        }
        records.clear();
    }
};

namespace {

bool finite3(const float *v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// The cells of the domain along each axis (base cells floor(x / h - 0.5) of
// the positions inside it; with batched scenes, of every copy).
void domain_cells(const mpm_consts &c, long long lo[3], long long hi[3]) {
    const double h = c.grid_node_size;
    for (int a = 0; a < 3; ++a) {
        double extent_max = c.domain_max[a];
        if (c.scene_repeat[a] > 1) extent_max += (double)c.scene_period[a] * (c.scene_repeat[a] - 1);
        lo[a] = (long long)std::floor((double)c.domain_min[a] / h - 0.5);
        hi[a] = (long long)std::floor(extent_max / h - 0.5);
    }
}

// The block grid whose particles' base cells may lie in lo..hi (two cells of
// stencil around them, whole 4^3 blocks): at most 1023 blocks per axis (30-bit
// Morton keys) and 96 GiB for the two node grids.
int geometry_for(const long long lo[3], const long long hi[3], double h, GridGeom &g, int &key_bits) {
    int nbmax = 0;
    for (int a = 0; a < 3; ++a) {
        if (!(hi[a] - lo[a] < 4 * 1023ll)) return MPM_ERR_GRID_TOO_LARGE;
        const long long origin = 4 * (long long)std::floor(((double)lo[a] - 2.0) / 4.0);
        const long long cmax = hi[a] + 2 - origin;
        const long long nb = (cmax >> 2) + 2;
        if (origin < -(1ll << 30) || origin > (1ll << 30) || nb > 1023) return MPM_ERR_GRID_TOO_LARGE;
        g.origin[a] = (int)origin;
        g.nb[a] = (int)nb;
        nbmax = std::max<int>(nbmax, (int)nb);
    }
    uint32_t k = 1;
    while ((1u << k) < (uint32_t)nbmax) ++k;
    if (k > 10) return MPM_ERR_GRID_TOO_LARGE;
    g.axis_bits = k;
    key_bits = (int)(3 * k);
    g.tomb_key = (key_bits == 32) ? 0xffffffffu : ((1u << key_bits) - 1u);
    g.h = (float)h;
    g.inv_h = 1.0f / g.h;
    uint64_t total = (uint64_t)g.nb[0] * g.nb[1] * g.nb[2];
    // Two float4 node grids: accumulation and velocity.
    if (total * 64ull * 16ull * 2ull > (96ull << 30)) return MPM_ERR_GRID_TOO_LARGE;
    g.num_blocks_total = (uint32_t)total;
    return MPM_OK;
}

int compute_geometry(const mpm_consts &c, GridGeom &g, int &key_bits) {
    long long lo[3], hi[3];
    domain_cells(c, lo, hi);
    return geometry_for(lo, hi, c.grid_node_size, g, key_bits);
}

// Whole 32-particle tiles, rounded to an even count so a wave of 64 never
// reads past the allocation.
size_t particle_tiles(uint64_t n) { return (size_t)((n + 63) / 64) * 2; }

hipError_t alloc_particles(mpm_engine *e, ParticleArrays &P) {
    return e->alloc(P.base, particle_tiles(e->n) * NFIELDS * TILE);
}

SubstepParams make_params(const mpm_engine *e, const float gravity[3], float factor_b, bool goals, bool has_b) {
    SubstepParams S{};
    const mpm_consts &c = e->consts;
    S.dt = c.time_step;
    S.dt_force = c.time_step;
    S.h = c.grid_node_size;
    for (int a = 0; a < 3; ++a) {
        S.gravity[a] = gravity[a];
        S.wall_min[a] = c.wall_min[a];
        S.wall_max[a] = c.wall_max[a];
        S.domain_min[a] = c.domain_min[a];
        S.domain_max[a] = c.domain_max[a];
        S.sphere_center[a] = c.sphere_center[a];
        const bool rep = c.scene_repeat[a] > 1;
        S.scene_period[a] = rep ? c.scene_period[a] : 0.f;
        S.scene_inv_period[a] = rep ? 1.0f / c.scene_period[a] : 0.f;
        S.scene_last[a] = rep ? (float)(c.scene_repeat[a] - 1) : 0.f;
        // copies meet in the middle of the gap between them
        S.scene_origin[a] = rep ? c.domain_min[a] - 0.5f * (c.scene_period[a] - (c.domain_max[a] - c.domain_min[a])) : 0.f;
    }
    S.sphere_radius = c.sphere_radius;
    S.factor_b = factor_b;
    S.has_goals = goals ? 1 : 0;
    S.has_b = has_b ? 1 : 0;
    S.options = c.options;
    S.plane_z = c.plane_z;
    S.snow_lo = 1.0f - c.snow_compression;
    S.snow_hi = 1.0f + c.snow_stretch;
    S.snow_hardening = c.snow_hardening;
    S.epoch = e->epoch;
    return S;
}

// FrameInput::frame_factor and the gravity interpolation of
// rust/crates/cpu/src/phase/interpolate_input.rs:24-28, evaluated in the same
// f64 -> f32 order as the reference.
int frame_params(const mpm_engine *e, const mpm_frame_input *in, double time, float gravity[3], float &factor_b) {
    double frame_time = time * (double)e->consts.frames_per_second;
    uint64_t frame_low = (uint64_t)std::floor(frame_time);
    uint64_t frame = in ? in->frame : 0;
    if (frame_low != frame) return MPM_ERR_WRONG_FRAME;
    factor_b = (float)std::fmod(frame_time, 1.0);
    float factor_a = 1.0f - factor_b;
    const float *ga = in ? in->gravity_a : nullptr;
    const float *gb = in ? (in->has_b ? in->gravity_b : in->gravity_a) : nullptr;
    for (int a = 0; a < 3; ++a) gravity[a] = ga ? factor_a * ga[a] + factor_b * gb[a] : 0.0f;
    return MPM_OK;
}

int upload_goals(mpm_engine *e, const mpm_frame_input *in, bool &goals) {
    goals = in && in->particle_flags_a && in->goal_positions_a && (!in->has_b || (in->particle_flags_b && in->goal_positions_b));
    if (!goals) return MPM_OK;
    const size_t n = e->n;
    if (!e->goals_allocated) {
        for (int a = 0; a < 3; ++a) {
            HIP_TRY(e->alloc(e->X.goal_a[a], n));
            HIP_TRY(e->alloc(e->X.goal_b[a], n));
        }
        HIP_TRY(e->alloc(e->X.goal_flags_a, n));
        HIP_TRY(e->alloc(e->X.goal_flags_b, n));
        e->goals_allocated = true;
        e->drop_graphs();  // captured with the old (absent) goal arrays
    }
    std::vector<float> tmp(n);
    for (int which = 0; which < 2; ++which) {
        const float *src = which == 0 ? in->goal_positions_a : (in->has_b ? in->goal_positions_b : in->goal_positions_a);
        for (int a = 0; a < 3; ++a) {
            for (size_t i = 0; i < n; ++i) tmp[i] = src[3 * i + a];
            HIP_TRY(hipMemcpy(which == 0 ? e->X.goal_a[a] : e->X.goal_b[a], tmp.data(), n * 4, hipMemcpyHostToDevice));
        }
    }
    HIP_TRY(hipMemcpy(e->X.goal_flags_a, in->particle_flags_a, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(e->X.goal_flags_b, in->has_b ? in->particle_flags_b : in->particle_flags_a, n * 4,
                      hipMemcpyHostToDevice));
    return MPM_OK;
}

GridBuffers grid_buffers(const mpm_engine *e) {
    GridBuffers B;
    B.grid = e->grid;
    B.vgrid = e->vgrid;
    B.blk_flag = e->blk_flag;
    B.blk_epoch = e->blk_epoch;
    B.cand = e->cand;
    B.ovf[0] = e->ovf[0];
    B.ovf[1] = e->ovf[1];
    return B;
}

// The parameters of the call's substeps, params[t] for substep t (the
// engine's epoch + t), to the device table, and the table index to 0.
int upload_params(mpm_engine *e, std::vector<SubstepParams> &params) {
    if (params.empty()) return MPM_OK;
    if (params.size() > e->param_capacity) {
        e->drop_graphs();  // they hold the old table's address
        if (e->param_table) (void)hipFree(e->param_table);
        if (e->param_host) (void)hipHostFree(e->param_host);
        e->param_table = nullptr;
        e->param_host = nullptr;
        e->param_capacity = 0;
        const uint32_t cap = (uint32_t)std::max<size_t>(params.size(), 1024);
        HIP_TRY(hipMalloc((void **)&e->param_table, cap * sizeof(SubstepParams)));
        HIP_TRY(hipHostMalloc((void **)&e->param_host, cap * sizeof(SubstepParams), hipHostMallocDefault));
        e->param_capacity = cap;
    }
    for (size_t t = 0; t < params.size(); ++t) {
        params[t].index = (uint32_t)t;
        params[t].epoch = e->epoch + (uint32_t)t;
    }
    // The previous call has drained the stream: the staging buffer is free.
    std::memcpy(e->param_host, params.data(), params.size() * sizeof(SubstepParams));
    HIP_TRY(hipMemcpyAsync(e->param_table, e->param_host, params.size() * sizeof(SubstepParams),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(&e->ds->param_index, 0, 2 * sizeof(uint32_t), e->stream));
    return MPM_OK;
}

// Sort, block list: everything before a transfer that permutes.
int run_sort(mpm_engine *e) {
    EventRecord r{};
    ParticleArrays &P = e->P[e->cur];
    HIP_TRY(launch_reset_scalars(e->ds, e->stream));
    // keys that a transfer wrote (stayers in key order) take the counting sort
    const bool counting = e->keys_valid && e->key_bits <= e->bucket_bits;
    if (!e->keys_valid) {
        e->begin(ST_KEYS, r);
        HIP_TRY(launch_keys(P, e->geom, e->n, e->keys, e->ds, e->stream));
        e->end(r);
    }
    if (counting) {
        // the counting sort writes the block list too
        e->begin(ST_SORT, r);
        HIP_TRY(bucket_sort(e->keys, e->n, e->key_bits, e->bucket, e->order, e->blk_start, e->blk_key,
                            e->blk_capacity, e->ds, e->stream));
        e->end(r);
    } else {
        e->begin(ST_SORT, r);
        HIP_TRY(radix_sort(e->keys, nullptr, e->keys_sorted, e->order, nullptr, nullptr, e->n, e->key_bits,
                           e->sort_temp, e->stream));
        e->end(r);
        e->begin(ST_BLOCKS, r);
        HIP_TRY(launch_block_list(e->keys_sorted, e->n, e->geom.tomb_key, e->blk_start, e->blk_key, e->blk_capacity,
                                  e->ds, e->blk_temp, e->stream));
        e->end(r);
    }
    e->keys_valid = false;
    e->pending_permute = true;
    return MPM_OK;
}

// The counting sort's buffers for the geometry's key range, when it has at
// most 2^BUCKET_SORT_MAX_BITS keys (larger ranges take the radix sort).  The
// per-key counts start at zero; every sort leaves them so.
int ensure_buckets(mpm_engine *e) {
    if (e->key_bits > BUCKET_SORT_MAX_BITS || e->key_bits <= e->bucket_bits) return MPM_OK;
    BucketBuffers &B = e->bucket;
    e->release(B.counts);
    e->release(B.stay);
    e->release(B.cursor);
    e->release(B.stay_base);
    e->release(B.tiles);
    e->bucket_bits = 0;
    const size_t nkeys = (size_t)1 << e->key_bits;
    if (e->alloc(B.counts, nkeys) || e->alloc(B.stay, nkeys) || e->alloc(B.cursor, nkeys) ||
        e->alloc(B.stay_base, nkeys) || e->alloc(B.tiles, bucket_tile_sums(e->key_bits, e->n)))
        return MPM_ERR_OUT_OF_MEMORY;
    if (!B.span_stay && e->alloc(B.span_stay, bucket_span_count(e->n))) return MPM_ERR_OUT_OF_MEMORY;
    HIP_TRY(hipMemsetAsync(B.counts, 0, nkeys * sizeof(uint32_t), e->stream));
    HIP_TRY(hipMemsetAsync(B.stay, 0, nkeys * sizeof(uint32_t), e->stream));
    e->bucket_bits = e->key_bits;
    return MPM_OK;
}

// The set a permuting transfer writes: neither the one it reads nor the one
// holding the call's starting state.
int out_set(const mpm_engine *e) {
    for (int b = 0; b < 3; ++b)
        if (b != e->cur && b != e->start) return b;
    return (e->cur + 1) % 3;
}

// The grid buffers for `blocks` blocks (freed first when held): the two node
// grids, the block flags and epochs, the overflow lists, and the block list
// and candidate slots for min(n, blocks) occupied blocks.
int alloc_grid(mpm_engine *e, uint32_t blocks) {
    e->release(e->grid);
    e->release(e->vgrid);
    e->release(e->blk_flag);
    e->release(e->blk_epoch);
    e->release(e->ovf[0]);
    e->release(e->ovf[1]);
    e->release(e->blk_start);
    e->release(e->blk_key);
    e->release(e->cand);
    e->grid_block_cap = 0;
    const uint64_t max_blocks = std::max<uint64_t>(std::min<uint64_t>(e->n, blocks), 1);
    const uint64_t nodes = (uint64_t)blocks * 64ull;
    if (e->alloc(e->grid, nodes) || e->alloc(e->vgrid, nodes) || e->alloc(e->blk_flag, blocks) ||
        e->alloc(e->blk_epoch, blocks) || e->alloc(e->ovf[0], blocks) || e->alloc(e->ovf[1], blocks) ||
        e->alloc(e->blk_start, max_blocks) || e->alloc(e->blk_key, max_blocks) || e->alloc(e->cand, NCAND * max_blocks))
        return MPM_ERR_OUT_OF_MEMORY;
    e->grid_block_cap = blocks;
    e->blk_capacity = (uint32_t)max_blocks;
    return MPM_OK;
}

// An empty grid of the current geometry (as after creation).
int clear_grid(mpm_engine *e) {
    const size_t nblk = e->geom.num_blocks_total;
    HIP_TRY(hipMemsetAsync(e->grid, 0, nblk * 64 * sizeof(float4), e->stream));
    HIP_TRY(hipMemsetAsync(e->vgrid, 0, nblk * 64 * sizeof(float4), e->stream));
    HIP_TRY(hipMemsetAsync(e->blk_flag, 0, nblk * 4, e->stream));
    HIP_TRY(hipMemsetAsync(e->blk_epoch, 0xff, nblk * 4, e->stream));
    return MPM_OK;
}

bool window_is_domain(const mpm_engine *e) {
    for (int a = 0; a < 3; ++a)
        if (e->win_lo[a] != e->dom_lo[a] || e->win_hi[a] != e->dom_hi[a]) return false;
    return true;
}

// The grid window over base cells lo..hi: a new geometry, buffers grown when
// it needs more blocks than they hold, an empty grid, and everything built
// on the old geometry dropped (captured graphs, block keys, collider block
// lists and twin tables).
int set_window(mpm_engine *e, const long long lo[3], const long long hi[3]) {
    GridGeom g{};
    int kb = 0;
    int rc = geometry_for(lo, hi, e->consts.grid_node_size, g, kb);
    if (rc != MPM_OK) return rc;
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->drop_graphs();
    if (g.num_blocks_total > e->grid_block_cap) {
        const uint64_t want = std::min<uint64_t>((uint64_t)g.num_blocks_total * 5 / 4, 1023ull * 1023 * 1023);
        if ((rc = alloc_grid(e, (uint32_t)std::max<uint64_t>(want, g.num_blocks_total))) != MPM_OK) return rc;
    }
    e->geom = g;
    e->key_bits = kb;
    for (int a = 0; a < 3; ++a) {
        e->win_lo[a] = lo[a];
        e->win_hi[a] = hi[a];
    }
    if ((rc = ensure_buckets(e)) != MPM_OK) return rc;
    if ((rc = clear_grid(e)) != MPM_OK) return rc;
    e->keys_valid = false;
    e->grid_valid = false;
    if (e->mesh_set && g.num_blocks_total + 1 > e->blk_off_capacity) {
        MeshDev &M = e->coll.mesh;
        uint32_t *p = nullptr;
        if (M.blk_off) e->mesh_free(M.blk_off);
        if (e->mesh_alloc(p, (size_t)g.num_blocks_total + 1) != hipSuccess) return MPM_ERR_OUT_OF_MEMORY;
        M.blk_off = p;
        e->blk_off_capacity = g.num_blocks_total + 1;
    }
    if (e->launch.colliders) HIP_TRY(launch_twin_clear(e->coll.twins, e->stream));
    return MPM_OK;
}

// The live particles' bounding box and largest speed (out[0..7], see
// launch_particle_bounds).  Synchronizes the stream.
int particle_bounds(mpm_engine *e, float out[8]) {
    if (!e->bounds_part) {
        if (e->alloc(e->bounds_part, 2 * (size_t)step_limit_groups(e->n)) != hipSuccess ||
            e->alloc(e->bounds_out, 8) != hipSuccess)
            return MPM_ERR_OUT_OF_MEMORY;
    }
    HIP_TRY(launch_particle_bounds(e->P[e->cur], e->n, e->bounds_part, e->bounds_out, e->stream));
    HIP_TRY(hipMemcpyAsync(out, e->bounds_out, 8 * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return MPM_OK;
}

// The window a call to `target` needs: the particles' base cells grown by
// how far they can travel in the call at their largest speed under the
// frame's gravity (times `scale`, plus two blocks), within the domain.  The
// window is kept when it holds that box and is not 64 times larger; else it
// is refitted with a quarter of the box's extent as slack on each side.
int fit_window(mpm_engine *e, double target, const mpm_frame_input *in, double scale) {
    float b[8];
    int rc = particle_bounds(e, b);
    if (rc != MPM_OK) return rc;
    if (!(b[7] > 0.f)) return MPM_OK;  // no live particle
    const double h = e->consts.grid_node_size;
    const double T = std::max(0.0, target - e->time);
    double g = 0.0;
    if (in) {
        for (const float *gv : {in->gravity_a, in->gravity_b})
            g = std::max(g, std::sqrt((double)gv[0] * gv[0] + (double)gv[1] * gv[1] + (double)gv[2] * gv[2]));
    }
    const double disp = (std::isfinite(b[6]) ? (double)b[6] : 0.0) * T + 0.5 * g * T * T;
    const double mcells = std::min(scale * (1.25 * disp / h + 8.0), 1e12);
    const long long m = (long long)std::ceil(mcells);
    long long need_lo[3], need_hi[3];
    double vol_need = 1.0, vol_win = 1.0;
    bool covered = true;
    for (int a = 0; a < 3; ++a) {
        const double cl = std::floor((double)b[a] / h - 0.5), ch = std::floor((double)b[3 + a] / h - 0.5);
        const double lo = std::max((double)e->dom_lo[a], std::min(cl, (double)e->dom_hi[a]) - (double)m);
        const double hi = std::min((double)e->dom_hi[a], std::max(ch, (double)e->dom_lo[a]) + (double)m);
        need_lo[a] = (long long)lo;
        need_hi[a] = std::max((long long)hi, need_lo[a]);
        covered = covered && need_lo[a] >= e->win_lo[a] && need_hi[a] <= e->win_hi[a];
        vol_need *= (double)(need_hi[a] - need_lo[a] + 1);
        vol_win *= (double)(e->win_hi[a] - e->win_lo[a] + 1);
    }
    if (covered && vol_win <= 64.0 * vol_need) return MPM_OK;
    long long lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const long long slack = std::max<long long>(8, (need_hi[a] - need_lo[a]) / 4);
        lo[a] = std::max(e->dom_lo[a], need_lo[a] - slack);
        hi[a] = std::min(e->dom_hi[a], need_hi[a] + slack);
    }
    return set_window(e, lo, hi);
}

// One transfer launch.  After a sort the transfer reads slot i from
// order[i] and writes another particle set; otherwise it works in place.
// want_keys: write the block keys of the new positions for the next sort.
int run_transfer(mpm_engine *e, TransferKind kind, bool goals, bool want_keys) {
    EventRecord r{};
    TransferCall t{};
    t.kind = kind;
    t.permute = e->pending_permute;
    const int out = out_set(e);
    t.Pin = e->P[e->cur];
    t.Pout = e->P[out];
    t.X = e->X;
    t.order = e->order;
    t.blk_start = e->blk_start;
    t.blk_key = e->blk_key;
    t.G = e->geom;
    t.params = e->param_table;
    t.goals = goals;
    t.B = grid_buffers(e);
    t.keys_out = want_keys ? e->keys : nullptr;
    t.mover_bit = e->key_bits <= e->bucket_bits ? BUCKET_MOVER : 0u;
    t.ds = e->ds;
    if (kind == TransferKind::P2G && e->closure_on) {
        t.tau_ext = e->closure_tau;
        t.tau_stride = e->n;
    }
    // In the learned loop the gather hands the closure its SVD factors.
    const bool svd_out = kind == TransferKind::G2P && e->closure_on && !t.permute && e->closure_svd;
    if (svd_out) {
        t.svd_out = e->closure_svd;
        t.tau_stride = e->n;
    }
    // The learned snow stress inside the fused gather-scatter.
    LearnedCall lc{};
    const bool learn = kind == TransferKind::G2P2G && e->closure_on;
    if (learn) {
        if (!e->learned_fused || e->launch.colliders) return MPM_ERR_INVALID_ARGUMENT;
        lc.net = e->closure_net;
        for (int q = 0; q < MLP_IN; ++q) {
            lc.hull_lo[q] = e->hull_lo[q];
            lc.hull_hi[q] = e->hull_hi[q];
        }
        lc.judge_stride = e->judge_stride;
        lc.judge_phase = e->judge_phase;
        lc.stats = e->closure_stats;
        t.learn = &lc;
    }
    const int stage = kind == TransferKind::P2G ? ST_P2G : (kind == TransferKind::G2P2G ? ST_G2P2G : ST_G2P);
    if (kind == TransferKind::LIMITS && t.permute) return MPM_ERR_INVALID_ARGUMENT;
    if (e->launch.colliders) {
        // the mesh of the substep whose scatter this transfer runs
        if (kind != TransferKind::G2P) HIP_TRY(launch_mesh_interp(e->coll.mesh, e->param_table, e->ds, e->stream));
        t.coll = &e->coll;
    }
    e->begin(stage, r);
    HIP_TRY(launch_transfer(e->launch, t, e->stream));
    e->end(r);
    if (learn && e->judge_stride) e->judge_phase = (e->judge_phase + 1) % e->judge_stride;
    if (kind == TransferKind::LIMITS) return MPM_OK;  // reads the state, writes only the reduction
    if (t.permute || want_keys) {
        // Slots past the live particles: copied when permuting, tombstone keys.
        e->begin(ST_COPY, r);
        HIP_TRY(launch_copy_dead(t.Pin, t.permute ? t.Pout : t.Pin, e->order, e->n, t.keys_out, e->geom.tomb_key,
                                 e->ds, e->stream));
        e->end(r);
    }
    if (t.permute) e->cur = out;
    e->pending_permute = false;
    e->keys_valid = want_keys;
    e->svd_fresh = svd_out;
    return MPM_OK;
}

// Grid update of the grid the last transfer scattered (epoch e->epoch).
int run_grid(mpm_engine *e) {
    EventRecord r{};
    e->begin(ST_GRID, r);
    if (e->launch.colliders)
        HIP_TRY(launch_twin_meld(e->coll.twins, e->grid, e->geom, e->param_table, e->ds, e->stream));
    HIP_TRY(launch_grid_update(e->launch, grid_buffers(e), e->geom, e->param_table, e->partial, e->ds, e->stream));
    e->end(r);
    e->grid_epoch = e->epoch;
    e->grid_valid = true;
    e->epoch += 1;
    return MPM_OK;
}

// The learned stress of the snow particles for the scatter that follows
// (which permutes when a sort is pending, so slot i reads order[i] then).
int run_closure(mpm_engine *e) {
    EventRecord r{};
    ClosureLaunch c{};
    c.P = e->P[e->cur];
    c.order = e->pending_permute ? e->order : nullptr;
    c.X = e->X;
    c.params = e->closure_params;
    c.tau = e->closure_tau;
    c.stride = e->n;
    c.n = e->n;
    for (int q = 0; q < MLP_IN; ++q) {
        c.hull_lo[q] = e->hull_lo[q];
        c.hull_hi[q] = e->hull_hi[q];
    }
    c.hardening = e->consts.snow_hardening;
    c.judge_stride = e->judge_stride;
    c.judge_phase = e->judge_phase;
    c.stats = e->closure_stats;
    // the factors of the last gather, by the same slots (a sort since then
    // only adds the order that c.order applies)
    c.svd = e->svd_fresh ? e->closure_svd : nullptr;
    c.ds = e->ds;
    e->begin(ST_CLOSURE, r);
    HIP_TRY(launch_closure(c, e->stream));
    e->end(r);
    if (e->judge_stride) e->judge_phase = (e->judge_phase + 1) % e->judge_stride;
    return MPM_OK;
}

// One sort cycle of the plain loop (resort_interval iterations, launched by
// `iterations(u)` for u = 0 .. K - 1) as a HIP graph: captured from the same
// launches on the first use for this particle buffer and goal setting, then
// replayed.  The launches read their substep from the parameter table, so
// the replay needs no argument update; the host-side state advances as the
// captured launches advanced it.
template <class Iter>
int run_cycle_graph(mpm_engine *e, bool goals, Iter iterations) {
    const int K = std::max(1, e->resort_interval);
    if (e->graph_interval != K) {
        e->drop_graphs();
        e->graph_interval = K;
    }
    // The state after the cycle: one permuting transfer, K grid updates, the
    // next cycle's keys written.
    const int cur_after = out_set(e);
    const uint32_t epoch_after = e->epoch + (uint32_t)K;
    hipGraphExec_t &exec = e->cycle_graph[e->cur][cur_after][goals ? 1 : 0];
    if (!exec) {
        HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
        int rc = MPM_OK;
        for (int u = 0; u < K && rc == MPM_OK; ++u) rc = iterations((size_t)u);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
        if (rc != MPM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIP_TRY(ce);
        const hipError_t ie = hipGraphInstantiateWithFlags(&exec, graph, 0);
        (void)hipGraphDestroy(graph);
        HIP_TRY(ie);
        if (e->cur != cur_after || e->epoch != epoch_after || !e->keys_valid || e->pending_permute) {
            g_last_error = "captured sort cycle left an unexpected engine state";
            return MPM_ERR_HIP;
        }
    } else {
        e->cur = cur_after;
        e->keys_valid = true;
        e->pending_permute = false;
        e->grid_epoch = epoch_after - 1u;
        e->epoch = epoch_after;
        e->grid_valid = true;
    }
    HIP_TRY(hipGraphLaunch(exec, e->stream));
    return MPM_OK;
}

int fetch_scalars(mpm_engine *e) {
    if (e->grid_valid) HIP_TRY(launch_sum_partials(e->launch, e->partial, e->ds, e->stream));
    HIP_TRY(hipMemcpyAsync(e->ds_host, e->ds, sizeof(DeviceScalars), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return MPM_OK;
}

// Where a call started: the particle set holding its starting state (which
// the call does not write), the host-side bookkeeping and the time.
struct CallStart {
    int set;
    uint32_t epoch, judge_phase;
    double time;
    bool stepped;
};

CallStart begin_call(mpm_engine *e) {
    e->start = e->cur;
    return CallStart{e->cur, e->epoch, e->judge_phase, e->time, e->stepped};
}

// Back to the start of the call: its particle set and Jp, and an empty grid
// (the accumulation grid, block flags and overflow lists as after creation).
int restore_call_start(mpm_engine *e, const CallStart &c) {
    e->cur = c.set;
    e->start = c.set;
    e->keys_valid = false;
    e->pending_permute = false;
    e->svd_fresh = false;
    e->epoch = c.epoch;
    e->judge_phase = c.judge_phase;
    e->time = c.time;
    e->stepped = c.stepped;
    e->grid_valid = false;
    if (e->jp_start && e->n > 0)
        HIP_TRY(hipMemcpyAsync(e->X.jp, e->jp_start, e->n * 4, hipMemcpyDeviceToDevice, e->stream));
    const size_t nblk = e->geom.num_blocks_total;
    HIP_TRY(hipMemsetAsync(e->grid, 0, nblk * 64 * sizeof(float4), e->stream));
    HIP_TRY(hipMemsetAsync(e->blk_flag, 0, nblk * 4, e->stream));
    HIP_TRY(hipMemsetAsync(e->blk_epoch, 0xff, nblk * 4, e->stream));
    HIP_TRY(hipMemsetAsync(e->ds->ovf_count, 0, sizeof(e->ds->ovf_count), e->stream));
    HIP_TRY(hipMemsetAsync(&e->ds->status, 0, 4, e->stream));
    HIP_TRY(hipMemsetAsync(&e->ds->fail_index, 0xff, 4, e->stream));
    if (e->launch.colliders) HIP_TRY(launch_twin_clear(e->coll.twins, e->stream));
    return MPM_OK;
}

struct LoopReport {
    uint64_t done = 0;            // substeps advanced
    uint64_t transfers = 0;       // index of the last transfer (a gather writing the state) + 1
    uint64_t graph_launches = 0;
    bool canceled = false;
};

// Substeps 0 .. k - 1 of the uploaded parameter table (times[t] is the start
// of substep t).  A set cancel flag is seen before a substep's work: before
// the first one the call does nothing; later, the substep whose grid update
// has run is finished with a gather that writes the state, so that x, v, F
// and C stay one consistent state.
int run_loop(mpm_engine *e, const std::vector<double> &times, size_t k, bool goals, const volatile int32_t *cancel,
             bool cancellable, LoopReport &rep) {
    int rc;
    // the cancel flag, or the test hook's substep (mpm_debug_cancel_after)
    auto canceled = [&](size_t t) {
        return cancellable && ((cancel && *cancel) || (e->cancel_after >= 0 && t >= (uint64_t)e->cancel_after));
    };
    const size_t K = (size_t)std::max(1, e->resort_interval);
    auto sorts_before = [&](size_t t) { return t == 0 || (t < k && t % K == 0); };
    auto advanced = [&](size_t t) {
        e->time = times[t - 1] + (double)e->consts.time_step;
        ++rep.done;
        rep.transfers = t + 1;
    };
    // With the learned closure in its own pass (the default), substep t is a
    // gather (of substep t - 1, it writes the state), the sort when due, the
    // closure, a scatter from the state and the grid update.  With
    // mpm_tuning.learned_fused = 1 (and no colliders) the learned loop is the
    // plain one: the closure gives the first scatter's snow stress, and the
    // fused kernels evaluate the network after each advance.
    const bool separate = e->closure_on && (!e->learned_fused || e->launch.colliders);
    for (size_t t = 0; separate && k > 0 && t <= k; ++t) {
        if (t > 0) {
            if ((rc = run_transfer(e, TransferKind::G2P, false, t == k || sorts_before(t))) != MPM_OK) return rc;
            advanced(t);
        }
        if (t == k) break;
        if (canceled(t)) {
            rep.canceled = true;
            break;
        }
        if (sorts_before(t) && (rc = run_sort(e)) != MPM_OK) return rc;
        if ((rc = run_closure(e)) != MPM_OK) return rc;
        if ((rc = run_transfer(e, TransferKind::P2G, goals, false)) != MPM_OK) return rc;
        if ((rc = run_grid(e)) != MPM_OK) return rc;
    }
    // Plain loop.  Iteration t: the sort when due, transfer t, grid update t.
    auto iteration = [&](size_t t) -> int {
        int r;
        if (sorts_before(t) && (r = run_sort(e)) != MPM_OK) return r;
        const bool want_keys = t == k || sorts_before(t + 1);
        const TransferKind kind = t == 0 ? TransferKind::P2G : (t == k ? TransferKind::G2P : TransferKind::G2P2G);
        if (kind == TransferKind::P2G && e->closure_on && (r = run_closure(e)) != MPM_OK) return r;
        if ((r = run_transfer(e, kind, goals, want_keys)) != MPM_OK) return r;
        if (t < k && (r = run_grid(e)) != MPM_OK) return r;
        return MPM_OK;
    };
    // A sort cycle t .. t + K - 1 that starts after a transfer which wrote
    // the keys and ends with one that writes the next cycle's keys has the
    // same launches every time: it runs as one HIP graph launch.
    // (not with the learned closure: its judge samples other particles every substep)
    const bool graphs = e->use_graph && !e->profiling && !e->closure_on;
    for (size_t t = 0; !separate && k > 0 && t <= k;) {
        if (canceled(t)) {
            if (t > 0) {
                if ((rc = run_transfer(e, TransferKind::G2P, false, true)) != MPM_OK) return rc;
                advanced(t);
            }
            rep.canceled = true;
            break;
        }
        const bool cycle = graphs && t > 0 && t % K == 0 && t + K < k && e->keys_valid && !e->pending_permute;
        const size_t steps = cycle ? K : 1;
        if (cycle) {
            if ((rc = run_cycle_graph(e, goals, [&](size_t u) { return iteration(t + u); })) != MPM_OK) return rc;
            ++rep.graph_launches;
        } else if ((rc = iteration(t)) != MPM_OK) {
            return rc;
        }
        for (size_t u = t; u < t + steps; ++u)
            if (u > 0) advanced(u);
        t += steps;
    }
    return MPM_OK;
}

// Twin tables of `cap` entries each (a power of two), cleared.
int set_twin_capacity(mpm_engine *e, uint32_t cap) {
    for (int k = 0; k < 2; ++k)
        if (e->coll.twins.tab[k]) e->mesh_free(e->coll.twins.tab[k]);
    e->coll.twins = TwinTables{};
    e->twin_capacity = 0;
    e->drop_graphs();  // they hold the old tables' addresses
    for (int k = 0; k < 2; ++k)
        if (e->mesh_alloc(e->coll.twins.tab[k], cap) != hipSuccess) return MPM_ERR_OUT_OF_MEMORY;
    e->coll.twins.mask = cap - 1u;
    e->twin_capacity = cap;
    HIP_TRY(launch_twin_clear(e->coll.twins, e->stream));
    return MPM_OK;
}

// The call's collider frames: vertex positions, frictions and dampings of
// frames a and b (b = a without it), the vertex velocities (b - a) * fps or 0
// (linear_vertex_velocities, rust/crates/xpu/src/frame_input.rs:333-347), and
// the candidate triangles of each grid block: every triangle whose bounding
// box over both frames, grown by the forget distance (update_bvh,
// frame_input.rs:349-389), reaches the block, in ascending order.
int upload_mesh_frames(mpm_engine *e, const mpm_frame_input *in) {
    if (!e->mesh_set) return MPM_OK;
    MeshDev &M = e->coll.mesh;
    if (!in || !in->vertex_positions_a || !in->triangle_frictions_a || !in->triangle_dampings_a)
        return MPM_ERR_INVALID_ARGUMENT;
    const bool b = in->has_b && in->vertex_positions_b && in->triangle_frictions_b && in->triangle_dampings_b;
    const size_t nv = M.nvert, nt = M.ntri;
    e->mesh_pos_a.assign(in->vertex_positions_a, in->vertex_positions_a + 3 * nv);
    const float *pb = b ? in->vertex_positions_b : in->vertex_positions_a;
    e->mesh_pos_b.assign(pb, pb + 3 * nv);
    HIP_TRY(hipMemcpyAsync((void *)M.pos_a, e->mesh_pos_a.data(), nv * 12, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync((void *)M.pos_b, e->mesh_pos_b.data(), nv * 12, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync((void *)M.fric_a, in->triangle_frictions_a, nt * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync((void *)M.fric_b, b ? in->triangle_frictions_b : in->triangle_frictions_a, nt * 4,
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync((void *)M.damp_a, in->triangle_dampings_a, nt * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync((void *)M.damp_b, b ? in->triangle_dampings_b : in->triangle_dampings_a, nt * 4,
                           hipMemcpyHostToDevice, e->stream));
    std::vector<float4> vel(nv);
    const float fps = (float)e->consts.frames_per_second;
    for (size_t v = 0; v < nv; ++v) {
        float q[3];
        for (int a = 0; a < 3; ++a) q[a] = b ? (e->mesh_pos_b[3 * v + a] - e->mesh_pos_a[3 * v + a]) * fps : 0.f;
        vel[v] = make_float4(q[0], q[1], q[2], 0.f);
    }
    HIP_TRY(hipMemcpyAsync(M.vvel, vel.data(), nv * 16, hipMemcpyHostToDevice, e->stream));
    // Candidate lists per block of the dense grid.
    const GridGeom &G = e->geom;
    const double h = G.h, margin = (double)M.forget + 0.01 * h;
    const size_t nblk = G.num_blocks_total;
    std::vector<uint32_t> off(nblk + 1, 0);
    std::vector<int> range(6 * nt);
    for (size_t t = 0; t < nt; ++t) {
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = 1e300;
            hi[a] = -1e300;
        }
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = e->mesh_tri[3 * t + k];
            for (const std::vector<float> *P : {&e->mesh_pos_a, &e->mesh_pos_b})
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], (double)(*P)[3 * v + a]);
                    hi[a] = std::max(hi[a], (double)(*P)[3 * v + a]);
                }
        }
        bool empty = false;
        for (int a = 0; a < 3; ++a) {
            const double c0 = std::floor((lo[a] - margin) / h) - G.origin[a], c1 = std::floor((hi[a] + margin) / h) - G.origin[a];
            const double b0 = std::max(0.0, std::floor(c0 / 4.0)), b1 = std::min((double)G.nb[a] - 1, std::floor(c1 / 4.0));
            if (!(b0 <= b1)) empty = true;
            range[6 * t + 2 * a] = (int)b0;
            range[6 * t + 2 * a + 1] = (int)b1;
        }
        if (empty) range[6 * t] = 1, range[6 * t + 1] = 0;
    }
    auto for_blocks = [&](size_t t, auto fn) {
        const int *r = &range[6 * t];
        for (int z = r[4]; z <= r[5]; ++z)
            for (int y = r[2]; y <= r[3]; ++y)
                for (int x = r[0]; x <= r[1]; ++x) fn((size_t)x + (size_t)G.nb[0] * ((size_t)y + (size_t)G.nb[1] * z));
    };
    for (size_t t = 0; t < nt; ++t) for_blocks(t, [&](size_t B) { ++off[B + 1]; });
    for (size_t B = 0; B < nblk; ++B) off[B + 1] += off[B];
    const uint32_t total = off[nblk];
    std::vector<uint32_t> list(std::max<uint32_t>(total, 1));
    {
        std::vector<uint32_t> fill(off.begin(), off.end() - 1);
        for (size_t t = 0; t < nt; ++t) for_blocks(t, [&](size_t B) { list[fill[B]++] = (uint32_t)t; });
    }
    if (total > e->blk_tri_capacity) {
        if (M.blk_tri) e->mesh_free(M.blk_tri);
        uint32_t *p = nullptr;
        const uint32_t cap = std::max<uint32_t>(total + total / 2, 1024);
        if (e->mesh_alloc(p, cap) != hipSuccess) return MPM_ERR_OUT_OF_MEMORY;
        M.blk_tri = p;
        e->blk_tri_capacity = cap;
        e->drop_graphs();
    }
    HIP_TRY(hipMemcpyAsync((void *)M.blk_off, off.data(), (nblk + 1) * 4, hipMemcpyHostToDevice, e->stream));
    if (total > 0)
        HIP_TRY(hipMemcpyAsync((void *)M.blk_tri, list.data(), (size_t)total * 4, hipMemcpyHostToDevice, e->stream));
    // the host vectors are freed when this returns: the copies complete first
    HIP_TRY(hipStreamSynchronize(e->stream));
    return MPM_OK;
}

// run_adaptive's result for an overflow of the collider twins' table: the
// call runs again from its start with a larger one.
constexpr int RERUN_LARGER_TWINS = 1000;

// The substep loop with adaptive time steps, following the phase order of
// CpuState::run_phase (rust/crates/cpu/src/phase/mod.rs:27-75) with
// adaptive_time_steps: per substep the external force and collide take the
// allowed step, the sound-speed and isolated-particle limits of the state
// (limit_time_step_before_force) set the step of the scatter, and the
// velocity and deformation limits of the gathered v and C
// (limit_time_step_before_integrate) the step of the advance and of the time
// (limit_time_step.rs:25-223).  The limits are device reductions and the
// AdaptiveTimeStepState lives on the device with the call's time: three
// one-thread kernels per substep push the limits and write the steps, the
// frame factor and the gravity into parameter entries 0 (the scatter) and 1
// (the gather).  The host launches the substeps a sort cycle at a time
// without waiting; the substeps past the end of the loop return at once
// (DeviceScalars::halt), and the host reads the device state between cycles
// (for the cancel flag and the sort it then launches).  A particle that
// fails stops the call after the substep in which it failed, with the state
// its gather wrote; `stop` receives the failure's status bits.
int run_adaptive(mpm_engine *e, double target_time, const mpm_frame_input *input, bool goals,
                 const volatile int32_t *cancel, LoopReport &rep, uint32_t &stop) {
    int rc;
    stop = 0;
    if (!(e->time < target_time)) return MPM_OK;
    const size_t K = (size_t)std::max(1, e->resort_interval);
    const uint32_t fail_bits = MPM_STATUS_ENERGY_ERROR | MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED;
    if (!e->limit_part &&
        (e->alloc(e->limit_part, step_limit_groups(e->n)) != hipSuccess || e->alloc(e->limit_out, 1) != hipSuccess))
        return MPM_ERR_OUT_OF_MEMORY;
    if (!e->adapt_dev) {
        if (e->alloc(e->adapt_dev, 1) != hipSuccess) return MPM_ERR_OUT_OF_MEMORY;
        HIP_TRY(hipHostMalloc((void **)&e->adapt_host, sizeof(DeviceAdaptive), hipHostMallocDefault));
    }
    // The device state from the engine's AdaptiveTimeStepState.
    DeviceAdaptive &A = *e->adapt_host;
    std::memset(&A, 0, sizeof(A));
    A.time = e->time;
    A.target = target_time;
    A.fps = (double)e->consts.frames_per_second;
    A.frame = input ? input->frame : 0;
    for (int a = 0; a < 3; ++a) {
        A.gravity_a[a] = input ? input->gravity_a[a] : 0.f;
        A.gravity_b[a] = input ? (input->has_b ? input->gravity_b[a] : input->gravity_a[a]) : 0.f;
    }
    A.h = e->consts.grid_node_size;
    A.max_time_step = e->ats.max_time_step;
    A.by_velocity = e->ats.by_velocity;
    A.by_deformation = e->ats.by_deformation;
    A.by_sound = e->ats.by_sound;
    A.by_isolated = e->ats.by_isolated;
    for (float v : e->ats.prior) A.prior[A.nprior++] = v;
    const uint32_t epoch0 = e->epoch;
    const uint32_t grid_epoch0 = e->grid_epoch;
    const bool grid_valid0 = e->grid_valid;
    A.epoch0 = epoch0;
    A.stop = ADAPT_RUNNING;
    HIP_TRY(hipMemcpyAsync(e->adapt_dev, &A, sizeof(A), hipMemcpyHostToDevice, e->stream));
    // Template entries: the fields that stay for the call; the device writes the rest.
    float g0[3] = {0.f, 0.f, 0.f};
    const SubstepParams tmpl = make_params(e, g0, 0.f, goals, input && input->has_b);
    std::vector<SubstepParams> two{tmpl, tmpl};
    if ((rc = upload_params(e, two)) != MPM_OK) return rc;
    HIP_TRY(hipMemsetAsync(&e->ds->halt, 0, 4, e->stream));
    // The host's view of the device state after a synchronization.
    auto read_state = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(e->adapt_host, e->adapt_dev, sizeof(DeviceAdaptive), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(e->ds_host, e->ds, sizeof(DeviceScalars), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        return MPM_OK;
    };
    // One substep's launches; want_keys: its gather writes the keys of the next sort.
    auto substep = [&](bool want_keys) -> int {
        int r;
        HIP_TRY(launch_adapt_begin(e->adapt_dev, e->param_table, e->ds, e->stream));
        HIP_TRY(launch_step_limits(e->P[e->cur], e->n, e->consts.grid_node_size, false, e->limit_part, e->limit_out,
                                   e->ds, e->stream));
        HIP_TRY(launch_adapt_force(e->adapt_dev, e->param_table, e->ds, e->limit_out, e->stream));
        if (e->closure_on && (r = run_closure(e)) != MPM_OK) return r;
        if ((r = run_transfer(e, TransferKind::P2G, goals, false)) != MPM_OK) return r;
        if ((r = run_grid(e)) != MPM_OK) return r;
        if ((r = run_transfer(e, TransferKind::LIMITS, false, false)) != MPM_OK) return r;
        HIP_TRY(launch_adapt_integrate(e->adapt_dev, e->param_table, e->ds, e->stream));
        return run_transfer(e, TransferKind::G2P, false, want_keys);
    };
    size_t t = 0;
    uint32_t done_before = 0;
    int cur_before = e->cur;
    bool sorted = false;
    for (;;) {
        if (t > 0) {
            if ((rc = read_state()) != MPM_OK) return rc;
            // The first substep of a cycle runs unless the call ends with an
            // error; then its permuting transfer did not run either.
            if (sorted && A.substeps == done_before && A.stop != ADAPT_RUNNING) {
                e->cur = cur_before;
                e->pending_permute = false;
            }
            if (A.stop != ADAPT_RUNNING || (e->ds_host->status & fail_bits)) break;
        }
        if (A.substeps >= MPM_MAX_SUBSTEPS_PER_CALL) return MPM_ERR_INVALID_ARGUMENT;
        if ((cancel && *cancel) || (e->cancel_after >= 0 && (int64_t)A.substeps >= e->cancel_after)) {
            rep.canceled = true;
            break;
        }
        // A cycle of substeps up to the next sort (or to the test hook's cancel).
        size_t chunk = K - (t % K);
        if (e->cancel_after >= 0) chunk = std::min<size_t>(chunk, (size_t)(e->cancel_after - (int64_t)A.substeps));
        cur_before = e->cur;
        done_before = A.substeps;
        sorted = t % K == 0;
        if (sorted && (rc = run_sort(e)) != MPM_OK) return rc;
        for (size_t u = 0; u < chunk; ++u)
            if ((rc = substep(u + 1 == chunk || (t + u + 1) % K == 0)) != MPM_OK) return rc;
        t += chunk;
    }
    if ((rc = read_state()) != MPM_OK) return rc;
    HIP_TRY(hipMemsetAsync(&e->ds->halt, 0, 4, e->stream));
    // The engine's state from the device's.
    e->time = A.time;
    e->epoch = epoch0 + A.substeps;
    e->keys_valid = false;
    e->pending_permute = false;
    e->svd_fresh = false;  // the last gather launched may not have run
    e->ats.by_velocity = A.by_velocity;
    e->ats.by_deformation = A.by_deformation;
    e->ats.by_sound = A.by_sound;
    e->ats.by_isolated = A.by_isolated;
    e->ats.prior.clear();
    for (uint32_t k = 0; k < A.nprior; ++k) e->ats.prior.push_back(A.prior[(A.head + k) % 11u]);
    rep.done = A.substeps;
    if (A.substeps > 0) {
        e->grid_valid = true;
        e->grid_epoch = epoch0 + A.substeps - 1u;
    } else {
        e->grid_valid = grid_valid0;
        e->grid_epoch = grid_epoch0;
    }
    if (A.stop == ADAPT_ZERO_STEP) return MPM_ERR_ZERO_TIME_STEP;
    if (A.stop == ADAPT_WRONG_FRAME) return MPM_ERR_WRONG_FRAME;
    if (A.stop == ADAPT_TWIN_OVERFLOW) return RERUN_LARGER_TWINS;
    if (e->ds_host->status & fail_bits) {
        stop = e->ds_host->status;
        // a failure in the last gather before a cycle's end (the device stops
        // at the next substep's start and adjusts the time itself)
        if (A.stop != ADAPT_FAILED && (stop & MPM_STATUS_ENERGY_ERROR) && A.substeps > 0)
            e->time -= (double)A.last_advance;
    }
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_set_adaptive_time_steps(mpm_engine *e, int32_t adaptive, float max_time_step) {
    if (!e || !(max_time_step > 0.f) || (adaptive && e->launch.batched)) return MPM_ERR_INVALID_ARGUMENT;
    e->adaptive = adaptive != 0;
    e->ats.max_time_step = std::min(max_time_step, FLT_MAX);
    return MPM_OK;
}

int mpm_set_collider_mesh(mpm_engine *e, const mpm_collider_mesh *mesh) {
    if (!e) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->drop_graphs();
    e->free_mesh();
    e->mesh_set = false;
    e->launch.colliders = false;
    e->mesh_tri.clear();
    if (!mesh || mesh->num_triangles == 0) return MPM_OK;
    if (e->launch.batched || !mesh->triangles || !mesh->triangle_colliders) return MPM_ERR_INVALID_ARGUMENT;
    const uint32_t nv = mesh->num_vertices, nt = mesh->num_triangles;
    uint32_t ncol = 0;
    for (uint32_t t = 0; t < nt; ++t) {
        if (mesh->triangle_colliders[t] >= 16) return MPM_ERR_INVALID_ARGUMENT;
        ncol = std::max(ncol, mesh->triangle_colliders[t] + 1);
        for (int k = 0; k < 3; ++k)
            if (mesh->triangles[3 * t + k] >= nv) return MPM_ERR_INVALID_ARGUMENT;
    }
    // Topology: the triangle across each edge (ab, bc, ca) and the triangles at
    // each vertex in ascending order (Topology, rust/crates/mesh_util/src/mesh.rs).
    std::vector<uint32_t> tri(mesh->triangles, mesh->triangles + 3 * (size_t)nt);
    std::vector<std::pair<uint64_t, uint32_t>> edges;
    edges.reserve(3 * (size_t)nt);
    auto ekey = [](uint32_t a, uint32_t b) { return a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a; };
    for (uint32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) edges.push_back({ekey(tri[3 * t + k], tri[3 * t + (k + 1) % 3]), t});
    std::sort(edges.begin(), edges.end());
    std::vector<uint32_t> opp(3 * (size_t)nt, 0xffffffffu);
    for (uint32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            const uint64_t key = ekey(tri[3 * t + k], tri[3 * t + (k + 1) % 3]);
            for (auto it = std::lower_bound(edges.begin(), edges.end(), std::make_pair(key, 0u));
                 it != edges.end() && it->first == key; ++it)
                if (it->second != t) {
                    opp[3 * t + k] = it->second;
                    break;
                }
        }
    std::vector<uint32_t> voff(nv + 1, 0), vtri(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) ++voff[tri[3 * t + k] + 1];
    for (uint32_t v = 0; v < nv; ++v) voff[v + 1] += voff[v];
    {
        std::vector<uint32_t> fill(voff.begin(), voff.end() - 1);
        for (uint32_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) vtri[fill[tri[3 * t + k]]++] = t;
    }
    MeshDev &M = e->coll.mesh;
    M.nvert = nv;
    M.ntri = nt;
    M.ncolliders = ncol;
    M.forget = e->consts.grid_node_size * 2.2f;  // InputConsts::forget_distance
    M.accept = e->consts.grid_node_size * 2.0f;  // InputConsts::accept_distance
    uint32_t *d_tri, *d_coll, *d_opp, *d_voff, *d_vtri, *d_boff;
    float *pa, *pb, *fa, *fb, *da, *db;
    const size_t nblk = e->geom.num_blocks_total;
    if (e->mesh_alloc(d_tri, 3 * (size_t)nt) || e->mesh_alloc(d_coll, nt) || e->mesh_alloc(d_opp, 3 * (size_t)nt) ||
        e->mesh_alloc(d_voff, nv + 1) || e->mesh_alloc(d_vtri, 3 * (size_t)nt) || e->mesh_alloc(pa, 3 * (size_t)nv) ||
        e->mesh_alloc(pb, 3 * (size_t)nv) || e->mesh_alloc(fa, nt) || e->mesh_alloc(fb, nt) || e->mesh_alloc(da, nt) ||
        e->mesh_alloc(db, nt) || e->mesh_alloc(M.vvel, nv) || e->mesh_alloc(M.vpos, nv) || e->mesh_alloc(M.vnorm, nv) ||
        e->mesh_alloc(M.tnorm, nt) || e->mesh_alloc(M.tfric, nt) || e->mesh_alloc(M.tdamp, nt) ||
        e->mesh_alloc(d_boff, nblk + 1)) {
        e->free_mesh();
        return MPM_ERR_OUT_OF_MEMORY;
    }
    M.tri = d_tri;
    M.tri_coll = d_coll;
    M.opp = d_opp;
    M.vtx_off = d_voff;
    M.vtx_tri = d_vtri;
    M.pos_a = pa;
    M.pos_b = pb;
    M.fric_a = fa;
    M.fric_b = fb;
    M.damp_a = da;
    M.damp_b = db;
    M.blk_off = d_boff;
    e->blk_off_capacity = (uint32_t)(nblk + 1);
    auto up = [&](void *dst, const void *src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); };
    HIP_TRY(up(d_tri, tri.data(), tri.size() * 4));
    HIP_TRY(up(d_coll, mesh->triangle_colliders, (size_t)nt * 4));
    HIP_TRY(up(d_opp, opp.data(), opp.size() * 4));
    HIP_TRY(up(d_voff, voff.data(), voff.size() * 4));
    HIP_TRY(up(d_vtri, vtri.data(), vtri.size() * 4));
    e->mesh_tri = std::move(tri);
    // twin tables sized for a shell of near-collider nodes; grown on overflow
    uint32_t cap = 1u << 14;
    while (cap < std::min<uint64_t>(e->n / 2, 1ull << 26)) cap <<= 1;
    int rc = set_twin_capacity(e, cap);
    if (rc != MPM_OK) {
        e->free_mesh();
        return rc;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->mesh_set = true;
    e->launch.colliders = true;
    return MPM_OK;
}

const char *mpm_error_string(int result) {
    switch (result) {
        case MPM_OK: return "ok";
        case MPM_ERR_ENERGY: return "a particle's deformation gradient lost its positive determinant (energy error)";
        case MPM_ERR_CLOSE_TO_INVERTED: return "a particle is too close to inversion";
        case MPM_ERR_INVALID_ARGUMENT: return "invalid argument";
        case MPM_ERR_HIP: return g_last_error.empty() ? "HIP error" : g_last_error.c_str();
        case MPM_ERR_NO_DEVICE: return "no HIP device";
        case MPM_ERR_OUT_OF_MEMORY: return "out of device memory";
        case MPM_ERR_GRID_TOO_LARGE: return "the domain's dense block grid is too large";
        case MPM_ERR_CANCELED: return "the computation was canceled";
        case MPM_ERR_WRONG_FRAME: return "the frame input does not contain the current time";
        case MPM_ERR_ZERO_TIME_STEP: return "the time step is zero";
        case MPM_ERR_OUT_OF_GRID: return "a particle left the allocated grid";
        case MPM_ERR_CAPACITY: return "capacity exceeded";
        default: return "unknown error";
    }
}

int mpm_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int mpm_create(const mpm_consts *consts, const mpm_particles_view *pv, const mpm_tuning *tuning,
               mpm_engine **out_engine) {
    if (!consts || !pv || !out_engine) return MPM_ERR_INVALID_ARGUMENT;
    *out_engine = nullptr;
    if (!(consts->grid_node_size > 0.f) || !(consts->time_step >= 0.f) || consts->frames_per_second == 0)
        return MPM_ERR_INVALID_ARGUMENT;
    if (!finite3(consts->domain_min) || !finite3(consts->domain_max)) return MPM_ERR_INVALID_ARGUMENT;
    for (int a = 0; a < 3; ++a) {
        if (!(consts->domain_min[a] < consts->domain_max[a])) return MPM_ERR_INVALID_ARGUMENT;
        // copies of a batch are at least 4 nodes apart: no stencil reaches across
        if (consts->scene_repeat[a] > 1 &&
            !(consts->scene_period[a] >= consts->domain_max[a] - consts->domain_min[a] + 4.f * consts->grid_node_size))
            return MPM_ERR_INVALID_ARGUMENT;
    }
    if (pv->count > 0 && (!pv->flags || !pv->parameters || !pv->positions || !pv->position_gradients ||
                          !pv->velocities || !pv->velocity_gradients))
        return MPM_ERR_INVALID_ARGUMENT;
    if (pv->count >= (1ull << 31)) return MPM_ERR_CAPACITY;
    int ndev = mpm_device_count();
    if (ndev <= 0) return MPM_ERR_NO_DEVICE;
    if (consts->device < 0 || consts->device >= ndev) return MPM_ERR_INVALID_ARGUMENT;

    mpm_engine *e = new (std::nothrow) mpm_engine();
    if (!e) return MPM_ERR_OUT_OF_MEMORY;
    e->device = consts->device;
    e->consts = *consts;
    e->n = pv->count;
    e->time = pv->time;
    auto fail = [&](int code) {
        delete e;
        return code;
    };
    // The grid covers the whole domain, or (mpm_tuning.grid_window, or a
    // domain too large for one dense grid) a window fitted to the particles.
    domain_cells(*consts, e->dom_lo, e->dom_hi);
    int rc = geometry_for(e->dom_lo, e->dom_hi, consts->grid_node_size, e->geom, e->key_bits);
    const int window_mode = tuning ? tuning->grid_window : 0;
    e->windowed = window_mode > 0 ||
                  (window_mode == 0 && (rc != MPM_OK || (uint64_t)e->geom.num_blocks_total * 64ull * 32ull > WINDOW_FROM_BYTES));
    if (!e->windowed && rc != MPM_OK) return fail(rc);
    if (e->windowed) {
        // the first window: the live particles' cells, two blocks around them
        // and slack, as fit_window would choose it
        double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
        const double h = consts->grid_node_size;
        for (uint64_t i = 0; i < pv->count; ++i) {
            if (pv->flags[i] & MPM_TOMBSTONED) continue;
            for (int a = 0; a < 3; ++a) {
                const double c = std::floor((double)pv->positions[3 * i + a] / h - 0.5);
                if (!std::isfinite(c)) continue;
                lo[a] = std::min(lo[a], c);
                hi[a] = std::max(hi[a], c);
            }
        }
        long long wl[3], wh[3];
        for (int a = 0; a < 3; ++a) {
            if (!(lo[a] <= hi[a])) lo[a] = hi[a] = (double)e->dom_lo[a];
            const double need_lo = std::max((double)e->dom_lo[a], std::min(lo[a], (double)e->dom_hi[a]) - 8.0);
            const double need_hi = std::min((double)e->dom_hi[a], std::max(hi[a], (double)e->dom_lo[a]) + 8.0);
            const long long slack = std::max<long long>(8, (long long)(need_hi - need_lo) / 4);
            wl[a] = std::max(e->dom_lo[a], (long long)need_lo - slack);
            wh[a] = std::min(e->dom_hi[a], std::max((long long)need_hi, (long long)need_lo) + slack);
            e->win_lo[a] = wl[a];
            e->win_hi[a] = wh[a];
        }
        if ((rc = geometry_for(wl, wh, consts->grid_node_size, e->geom, e->key_bits)) != MPM_OK) return fail(rc);
    } else {
        for (int a = 0; a < 3; ++a) {
            e->win_lo[a] = e->dom_lo[a];
            e->win_hi[a] = e->dom_hi[a];
        }
    }
    if (hipSetDevice(e->device) != hipSuccess) return fail(MPM_ERR_HIP);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return fail(MPM_ERR_HIP);

    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, e->device) != hipSuccess) return fail(MPM_ERR_HIP);
    e->launch.persistent_blocks = 0;
    e->launch.cu_count = prop.multiProcessorCount;
    e->launch.particles_per_lane = 3;
    if (tuning) {
        if (tuning->persistent_blocks > 0) e->launch.persistent_blocks = tuning->persistent_blocks;
        if (tuning->resort_interval > 0) e->resort_interval = tuning->resort_interval;
        if (tuning->graph < 0) e->use_graph = false;
        if (tuning->particles_per_lane < 0 || tuning->particles_per_lane > 3) return fail(MPM_ERR_INVALID_ARGUMENT);
        if (tuning->particles_per_lane > 0) e->launch.particles_per_lane = tuning->particles_per_lane;
        if (tuning->learned_fused > 0) e->learned_fused = true;
        const int gut = tuning->grid_update_threads;
        if (gut != 0 && gut != 64 && gut != 128 && gut != 256) return fail(MPM_ERR_INVALID_ARGUMENT);
        if (gut) e->launch.grid_update_threads = gut;
        if (tuning->grid_window < -1 || tuning->grid_window > 1) return fail(MPM_ERR_INVALID_ARGUMENT);
    }
    // Kernels are specialised on the constitutive models present: snow and
    // sand need the full SVD, every other model a polar decomposition at most.
    e->launch.batched = consts->scene_repeat[0] > 1 || consts->scene_repeat[1] > 1 || consts->scene_repeat[2] > 1;
    e->launch.model_set = SET_POLAR;
    for (uint64_t i = 0; i < pv->count; ++i)
        if (pv->flags[i] & (MPM_SNOW | MPM_USE_SAND_ALPHA)) {
            e->launch.model_set = SET_ALL;
            break;
        }

    const size_t n = e->n;
    if (alloc_particles(e, e->P[0]) != hipSuccess || alloc_particles(e, e->P[1]) != hipSuccess ||
        alloc_particles(e, e->P[2]) != hipSuccess)
        return fail(MPM_ERR_OUT_OF_MEMORY);
    if (e->launch.model_set == SET_ALL && e->alloc(e->jp_start, n) != hipSuccess) return fail(MPM_ERR_OUT_OF_MEMORY);
    if (e->alloc(e->X.visc_dyn, n) || e->alloc(e->X.visc_bulk, n) || e->alloc(e->X.sand_alpha, n) ||
        e->alloc(e->X.jp, n) || e->alloc(e->X.energy, n) || e->alloc(e->X.collider_bits, n))
        return fail(MPM_ERR_OUT_OF_MEMORY);
    char *st = nullptr, *bt = nullptr;
    if (e->alloc(e->keys, n) || e->alloc(e->keys_sorted, n) || e->alloc(e->order, n) ||
        e->alloc(st, radix_sort_temp_bytes(n)) || e->alloc(bt, block_list_temp_bytes(n)) || e->alloc(e->ds, 1) ||
        e->alloc(e->partial, grid_update_waves(e->launch)))
        return fail(MPM_ERR_OUT_OF_MEMORY);
    if ((rc = alloc_grid(e, e->geom.num_blocks_total)) != MPM_OK) return fail(rc);
    e->sort_temp = st;
    e->blk_temp = bt;
    if ((rc = ensure_buckets(e)) != MPM_OK) return fail(rc);

    if (hipHostMalloc((void **)&e->ds_host, sizeof(DeviceScalars), hipHostMallocDefault) != hipSuccess)
        return fail(MPM_ERR_OUT_OF_MEMORY);
    if (clear_grid(e) != MPM_OK || hipMemsetAsync(e->ds, 0, sizeof(DeviceScalars), e->stream) != hipSuccess)
        return fail(MPM_ERR_HIP);

    // Host AoS -> device SoA.
    e->parameters.assign(pv->parameters, pv->parameters + n);
    e->initial_positions.assign(3 * n, 0.f);
    if (pv->initial_positions) std::copy(pv->initial_positions, pv->initial_positions + 3 * n, e->initial_positions.begin());
    e->collider_bits.assign(n, 0u);
    if (pv->collider_bits) std::copy(pv->collider_bits, pv->collider_bits + n, e->collider_bits.begin());
    e->dead_at_create.assign(n, 0);
    std::vector<float> col(n);
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    // Particle fields are staged in the device's tiled layout and uploaded
    // in one copy.
    std::vector<float> tiles(particle_tiles(n) * NFIELDS * TILE, 0.f);
    auto put = [&](int f, size_t i, float value) { tiles[(i >> 5) * (NFIELDS * TILE) + f * TILE + (i & 31)] = value; };
    auto putu = [&](int f, size_t i, uint32_t value) {
        std::memcpy(&tiles[(i >> 5) * (NFIELDS * TILE) + f * TILE + (i & 31)], &value, 4);
    };
    const mpm_particle_parameters *pp = pv->parameters;
    for (size_t i = 0; i < n; ++i) {
        const bool fluid = (pv->flags[i] & MPM_IS_FLUID) != 0;
        for (int a = 0; a < 3; ++a) {
            put(FX + a, i, pv->positions[3 * i + a]);
            put(FV + a, i, pv->velocities[3 * i + a]);
        }
        for (int k = 0; k < 9; ++k) {
            put(FF + k, i, pv->position_gradients[9 * i + k]);
            put(FA + k, i, pv->velocity_gradients[9 * i + k]);
        }
        put(FM, i, pp[i].mass);
        put(FVOL, i, pp[i].initial_volume);
        put(FP0, i, fluid ? pp[i].bulk_modulus : pp[i].mu);
        put(FP1, i, fluid ? (float)pp[i].exponent : pp[i].lambda);
        putu(FFLAGS, i, pv->flags[i]);
        putu(FID, i, (uint32_t)i);
        e->dead_at_create[i] = (pv->flags[i] & MPM_TOMBSTONED) ? 1 : 0;
    }
    ParticleArrays &P = e->P[0];
    bool ok = up(P.base, tiles.data(), tiles.size() * 4);
    std::vector<float>().swap(tiles);
    auto upf = [&](float *dst, auto getter) {
        for (size_t i = 0; i < n; ++i) col[i] = getter(i);
        return up(dst, col.data(), n * 4);
    };
    ok = ok && upf(e->X.visc_dyn, [&](size_t i) { return pp[i].viscosity_dynamic; });
    ok = ok && upf(e->X.visc_bulk, [&](size_t i) { return pp[i].viscosity_bulk; });
    ok = ok && upf(e->X.sand_alpha, [&](size_t i) { return pp[i].sand_alpha; });
    ok = ok && upf(e->X.jp, [&](size_t i) { return pv->plastic_jp ? pv->plastic_jp[i] : 1.0f; });
    ok = ok && upf(e->X.energy, [&](size_t i) { return pv->elastic_energies ? pv->elastic_energies[i] : 0.0f; });
    ok = ok && up(e->X.collider_bits, e->collider_bits.data(), n * 4);
    if (!ok) return fail(MPM_ERR_HIP);
    if (hipStreamSynchronize(e->stream) != hipSuccess) return fail(MPM_ERR_HIP);
    *out_engine = e;
    return MPM_OK;
}

void mpm_destroy(mpm_engine *engine) { delete engine; }

double mpm_time(const mpm_engine *engine) { return engine ? engine->time : 0.0; }

void *mpm_stream(mpm_engine *engine) { return engine ? (void *)engine->stream : nullptr; }

}  // extern "C"

namespace {

// One call on the current grid (see mpm_step_to).
int step_call(mpm_engine *e, double target_time, const mpm_frame_input *input, const volatile int32_t *cancel,
              mpm_status *status) {
    if (!e) return MPM_ERR_INVALID_ARGUMENT;
    if (status) std::memset(status, 0, sizeof(*status));
    HIP_TRY(hipSetDevice(e->device));
    if (e->adaptive) {
        bool goals = false;
        int rc = upload_goals(e, input, goals);
        if (rc != MPM_OK) return rc;
        if ((rc = upload_mesh_frames(e, input)) != MPM_OK) return rc;
        {
            float g[3], fb;
            if (e->time < target_time && (rc = frame_params(e, input, e->time, g, fb)) != MPM_OK) return rc;
        }
        const CallStart cs = begin_call(e);
        const AdaptiveStep ats0 = e->ats;
        HIP_TRY(hipMemsetAsync(&e->ds->status, 0, 4, e->stream));
        HIP_TRY(hipMemsetAsync(&e->ds->fail_index, 0xff, 4, e->stream));
        if (e->jp_start && e->n > 0)
            HIP_TRY(hipMemcpyAsync(e->jp_start, e->X.jp, e->n * 4, hipMemcpyDeviceToDevice, e->stream));
        LoopReport rep;
        uint32_t stop = 0;
        rc = run_adaptive(e, target_time, input, goals, cancel, rep, stop);
        // The table of collider twins overflowed: the call runs again from its
        // start with a larger table (gpu_state.rs:562-599).
        for (int redo = 0; rc == RERUN_LARGER_TWINS; ++redo) {
            if (redo >= 6 || e->twin_capacity >= (1u << 28)) return MPM_ERR_CAPACITY;
            if ((rc = set_twin_capacity(e, e->twin_capacity * 4)) != MPM_OK) return rc;
            if ((rc = restore_call_start(e, cs)) != MPM_OK) return rc;
            e->ats = ats0;
            rep = LoopReport{};
            rc = run_adaptive(e, target_time, input, goals, cancel, rep, stop);
        }
        if (rc != MPM_OK) return rc;
        if (rep.done > 0) e->stepped = true;
        if ((rc = fetch_scalars(e)) != MPM_OK) return rc;
        if (e->profiling) e->harvest();
        const uint32_t st = e->ds_host->status | stop;
        int result = rep.canceled ? MPM_ERR_CANCELED : MPM_OK;
        if (st & MPM_STATUS_INDIRECT_LIMIT_EXCEEDED) result = MPM_ERR_CAPACITY;
        else if (st & MPM_STATUS_ENERGY_ERROR) result = MPM_ERR_ENERGY;
        else if (st & MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED) result = MPM_ERR_CLOSE_TO_INVERTED;
        else if (result == MPM_OK && (st & MPM_STATUS_OUT_OF_GRID)) result = MPM_ERR_OUT_OF_GRID;
        if (status) {
            status->device_status = st;
            status->result = result;
            status->substeps = rep.done;
            status->time = e->time;
            status->active_blocks = e->ds_host->num_blocks;
            status->active_nodes = e->ds_host->active_nodes;
            status->svd_fp64_fallbacks = (st & MPM_STATUS_SVD_FP64) ? 1 : 0;
            status->grid_blocks = e->geom.num_blocks_total;
        }
        return result;
    }
    if (e->consts.time_step == 0.0f) return MPM_ERR_ZERO_TIME_STEP;

    // The substep times of the call, accumulated as the reference does
    // (f32 time step added to an f64 time, cpu_state.rs:187-190).
    std::vector<double> times;
    for (double t = e->time; t < target_time; t += (double)e->consts.time_step) {
        if (times.size() >= MPM_MAX_SUBSTEPS_PER_CALL) return MPM_ERR_INVALID_ARGUMENT;
        times.push_back(t);
    }
    std::vector<SubstepParams> params;
    params.reserve(times.size() + 1);
    bool goals = false;
    int rc = upload_goals(e, input, goals);
    if (rc != MPM_OK) return rc;
    if ((rc = upload_mesh_frames(e, input)) != MPM_OK) return rc;
    for (double t : times) {
        float g[3], fb;
        rc = frame_params(e, input, t, g, fb);
        if (rc != MPM_OK) return rc;
        params.push_back(make_params(e, g, fb, goals, input && input->has_b));
    }
    // Entry k (the last gather's) repeats entry k - 1.
    if (!params.empty()) params.push_back(params.back());

    // The call's starting point, kept so that the call can be run again up to
    // a failing substep: the particle set (never written by the call), Jp,
    // the epoch and the time.
    const CallStart cs = begin_call(e);
    HIP_TRY(hipMemsetAsync(&e->ds->status, 0, 4, e->stream));
    HIP_TRY(hipMemsetAsync(&e->ds->fail_index, 0xff, 4, e->stream));
    if (e->jp_start && e->n > 0)
        HIP_TRY(hipMemcpyAsync(e->jp_start, e->X.jp, e->n * 4, hipMemcpyDeviceToDevice, e->stream));
    if ((rc = upload_params(e, params)) != MPM_OK) return rc;

    LoopReport rep;
    if ((rc = run_loop(e, times, times.size(), goals, cancel, true, rep)) != MPM_OK) return rc;
    if ((rc = fetch_scalars(e)) != MPM_OK) return rc;
    // The table of collider twin nodes overflowed: the call runs again from
    // its start with a larger table, as the reference redoes a frame whose
    // hash table was too small (rust/crates/gpu/src/gpu_state.rs:562-599).
    for (int redo = 0; e->launch.colliders && (e->ds_host->status & MPM_STATUS_TABLE_TRIES_EXCEEDED); ++redo) {
        if (redo >= 6 || e->twin_capacity >= (1u << 28)) return MPM_ERR_CAPACITY;
        if ((rc = set_twin_capacity(e, e->twin_capacity * 4)) != MPM_OK) return rc;
        if ((rc = restore_call_start(e, cs)) != MPM_OK) return rc;
        if ((rc = upload_params(e, params)) != MPM_OK) return rc;
        rep = LoopReport{};
        if ((rc = run_loop(e, times, times.size(), goals, cancel, true, rep)) != MPM_OK) return rc;
        if ((rc = fetch_scalars(e)) != MPM_OK) return rc;
    }
    uint32_t st = e->ds_host->status;
    uint64_t done = rep.done;

    // A particle flagged FAILED stops the call at the substep whose advance
    // failed (CpuState::produce_next_state returns the state stored there,
    // cpu_state.rs:178-183; the wgpu backend ends the frame's work at the
    // first error, prepare_tmp.wesl:74-78).  Transfer f gathered and advanced
    // substep f - 1.  Unless it was the call's last gather, which wrote the
    // whole state, the call runs again from its start to a last gather at f.
    const uint32_t f = e->ds_host->fail_index;
    if (f != 0xffffffffu && rep.transfers > 0 && f < rep.transfers - 1) {
        if ((rc = restore_call_start(e, cs)) != MPM_OK) return rc;
        uint32_t stop_bits = 0;
        if (f == 0) {
            // The stored state itself is too close to inversion: it stays,
            // with the particles flagged.
            HIP_TRY(launch_flag_inverted(e->P[e->cur], e->n, e->stream));
            stop_bits = MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED | (KID_P2G << 16);
            done = 0;
        } else {
            std::vector<SubstepParams> redo(params.begin(), params.begin() + f);
            redo.push_back(redo.back());
            if ((rc = upload_params(e, redo)) != MPM_OK) return rc;
            LoopReport rep2;
            if ((rc = run_loop(e, times, f, goals, nullptr, false, rep2)) != MPM_OK) return rc;
            done = rep2.done;
        }
        if ((rc = fetch_scalars(e)) != MPM_OK) return rc;
        // The failure bits of the run that stopped there (an FP32 rerun
        // reproduces them unless the particle sat on the threshold).
        const uint32_t fail_bits = MPM_STATUS_ENERGY_ERROR | MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED;
        const uint32_t st2 = e->ds_host->status | stop_bits;
        st = (st2 & fail_bits) ? st2 : (st2 | (st & fail_bits));
    }
    // An energy error leaves the time at the start of the failing substep:
    // the reference returns before adding the step (cpu_state.rs:178-190).
    if (f != 0xffffffffu && f >= 1 && (st & MPM_STATUS_ENERGY_ERROR) && done > 0) e->time = times[done - 1];
    if (done > 0) e->stepped = true;
    if (e->profiling) e->harvest();
    int result = rep.canceled ? MPM_ERR_CANCELED : MPM_OK;
    if (st & MPM_STATUS_INDIRECT_LIMIT_EXCEEDED) result = MPM_ERR_CAPACITY;
    else if (st & MPM_STATUS_ENERGY_ERROR) result = MPM_ERR_ENERGY;
    else if (st & MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED) result = MPM_ERR_CLOSE_TO_INVERTED;
    else if (result == MPM_OK && (st & MPM_STATUS_OUT_OF_GRID)) result = MPM_ERR_OUT_OF_GRID;
    if (status) {
        status->device_status = st;
        status->result = result;
        status->substeps = done;
        status->time = e->time;
        status->active_blocks = e->ds_host->num_blocks;
        status->active_nodes = e->ds_host->active_nodes;
        status->svd_fp64_fallbacks = (st & MPM_STATUS_SVD_FP64) ? 1 : 0;
        status->graph_launches = rep.graph_launches;
        status->grid_blocks = e->geom.num_blocks_total;
    }
    return result;
}

}  // namespace

extern "C" {

// With a grid window, the window is fitted to the particles first; a call
// in which a particle left it (MPM_STATUS_OUT_OF_GRID while the window is
// not yet the whole domain) runs again from its start with a window grown
// by a factor of four each time.
int mpm_step_to(mpm_engine *e, double target_time, const mpm_frame_input *input, const volatile int32_t *cancel,
                mpm_status *status) {
    if (!e) return MPM_ERR_INVALID_ARGUMENT;
    if (!e->windowed) return step_call(e, target_time, input, cancel, status);
    if (status) std::memset(status, 0, sizeof(*status));
    HIP_TRY(hipSetDevice(e->device));
    int rc = fit_window(e, target_time, input, 1.0);
    if (rc != MPM_OK) return rc;
    const CallStart cs = begin_call(e);
    const AdaptiveStep ats0 = e->ats;
    if (e->jp_start && e->n > 0)
        HIP_TRY(hipMemcpyAsync(e->jp_start, e->X.jp, e->n * 4, hipMemcpyDeviceToDevice, e->stream));
    for (int attempt = 1;; ++attempt) {
        mpm_status st{};
        rc = step_call(e, target_time, input, cancel, &st);
        const bool short_grid = (st.device_status & MPM_STATUS_OUT_OF_GRID) && !window_is_domain(e);
        if (!short_grid || attempt >= 8 || (rc < 0 && rc != MPM_ERR_OUT_OF_GRID && rc != MPM_ERR_CANCELED)) {
            st.grid_blocks = e->geom.num_blocks_total;
            st.window_redos = (uint64_t)(attempt - 1);
            if (status) *status = st;
            return rc;
        }
        if ((rc = restore_call_start(e, cs)) != MPM_OK) return rc;
        e->ats = ats0;
        if ((rc = fit_window(e, target_time, input, std::pow(4.0, attempt))) != MPM_OK) return rc;
    }
}

int mpm_read_particles(mpm_engine *e, mpm_particles_view *out) {
    if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
    if (out->count != e->n) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(e->device));
    const size_t n = e->n;
    out->time = e->time;
    if (n == 0) return MPM_OK;
    // Staging buffers of the engine, allocated once (a failed allocation
    // leaves nothing half-made: the engine frees what it holds).
    if (!e->rb_flags) {
        if (e->alloc(e->rb_x, n * 3) != hipSuccess || e->alloc(e->rb_v, n * 3) != hipSuccess ||
            e->alloc(e->rb_F, n * 9) != hipSuccess || e->alloc(e->rb_C, n * 9) != hipSuccess ||
            e->alloc(e->rb_flags, n) != hipSuccess) {
            e->rb_flags = nullptr;
            return MPM_ERR_OUT_OF_MEMORY;
        }
    }
    float *x = e->rb_x, *v = e->rb_v, *F = e->rb_F, *C = e->rb_C;
    uint32_t *fl = e->rb_flags;
    int rc = MPM_OK;
    if (launch_unpermute(e->P[e->cur], n, x, v, F, C, fl, e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess) {
        rc = MPM_ERR_HIP;
    } else {
        auto down = [&](void *dst, const void *src, size_t bytes) {
            if (dst && rc == MPM_OK && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = MPM_ERR_HIP;
        };
        down(out->positions, x, n * 12);
        down(out->velocities, v, n * 12);
        down(out->position_gradients, F, n * 36);
        down(out->velocity_gradients, C, n * 36);
        down(out->flags, fl, n * 4);
        down(out->elastic_energies, e->X.energy, n * 4);
        down(out->plastic_jp, e->X.jp, n * 4);
    }
    if (rc != MPM_OK) return rc;
    if (out->parameters) std::copy(e->parameters.begin(), e->parameters.end(), out->parameters);
    if (out->initial_positions) std::copy(e->initial_positions.begin(), e->initial_positions.end(), out->initial_positions);
    if (out->collider_bits) {
        if (e->mesh_set) {
            HIP_TRY(hipMemcpy(out->collider_bits, e->X.collider_bits, n * 4, hipMemcpyDeviceToHost));
        } else {
            // Without mesh colliders every live particle's collider bits are set
            // to 0 in each substep (rust/crates/cpu/src/phase/collide.rs:53-58).
            for (size_t i = 0; i < n; ++i)
                out->collider_bits[i] = (e->stepped && !e->dead_at_create[i]) ? 0u : e->collider_bits[i];
        }
    }
    return MPM_OK;
}

static int read_grid_impl(mpm_engine *e, mpm_grid_view *out, bool momentum) {
    if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(e->device));
    const uint64_t nblk = e->geom.num_blocks_total;
    std::vector<float4> g(nblk * 64);
    std::vector<uint32_t> ep(nblk);
    HIP_TRY(hipStreamSynchronize(e->stream));
    HIP_TRY(hipMemcpy(g.data(), momentum ? e->grid : e->vgrid, nblk * 64 * 16, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ep.data(), e->blk_epoch, nblk * 4, hipMemcpyDeviceToHost));
    uint64_t count = 0;
    const int *nb = e->geom.nb;
    for (uint64_t B = 0; B < nblk; ++B) {
        if (!momentum && (!e->grid_valid || ep[B] != e->grid_epoch)) continue;
        const int bx = (int)(B % nb[0]), by = (int)((B / nb[0]) % nb[1]), bz = (int)(B / ((uint64_t)nb[0] * nb[1]));
        for (int s = 0; s < 64; ++s) {
            const float4 q = g[B * 64 + s];
            if (!(q.x != 0.f)) continue;
            if (count < out->capacity) {
                const int id[3] = {e->geom.origin[0] + 4 * bx + (s & 3), e->geom.origin[1] + 4 * by + ((s >> 2) & 3),
                                   e->geom.origin[2] + 4 * bz + (s >> 4)};
                if (out->node_ids)
                    for (int a = 0; a < 3; ++a) out->node_ids[3 * count + a] = id[a];
                if (out->collider_bits) out->collider_bits[count] = 0u;
                if (out->masses) out->masses[count] = q.x;
                if (out->velocities) {
                    out->velocities[3 * count] = q.y;
                    out->velocities[3 * count + 1] = q.z;
                    out->velocities[3 * count + 2] = q.w;
                }
            }
            ++count;
        }
    }
    out->count = count;
    return count > out->capacity ? MPM_ERR_CAPACITY : MPM_OK;
}

int mpm_read_grid(mpm_engine *e, mpm_grid_view *out) { return read_grid_impl(e, out, false); }
int mpm_read_grid_momentum(mpm_engine *e, mpm_grid_view *out) { return read_grid_impl(e, out, true); }

int mpm_run_stage(mpm_engine *e, int32_t stage, const mpm_frame_input *input) {
    if (!e) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(e->device));
    float g[3], fb;
    int rc = frame_params(e, input, e->time, g, fb);
    if (rc != MPM_OK) return rc;
    bool goals = false;
    if ((rc = upload_goals(e, input, goals)) != MPM_OK) return rc;
    if ((rc = upload_mesh_frames(e, input)) != MPM_OK) return rc;
    std::vector<SubstepParams> one{make_params(e, g, fb, goals, input && input->has_b)};
    if ((rc = upload_params(e, one)) != MPM_OK) return rc;
    switch (stage) {
        case 0: rc = run_sort(e); break;
        case 1: rc = run_transfer(e, TransferKind::P2G, goals, false); break;
        case 2: rc = run_grid(e); break;
        case 3:
            rc = run_transfer(e, TransferKind::G2P, false, true);
            if (rc == MPM_OK) {
                e->time += (double)e->consts.time_step;
                e->stepped = true;
            }
            break;
        default: return MPM_ERR_INVALID_ARGUMENT;
    }
    if (rc != MPM_OK) return rc;
    return fetch_scalars(e);
}

int mpm_set_profiling(mpm_engine *e, int32_t enabled) {
    if (!e) return MPM_ERR_INVALID_ARGUMENT;
    e->profiling = enabled != 0;
    e->profile_mask = enabled == 1 ? ((1u << ST_COUNT) - 1u) : (uint32_t)enabled;
    return MPM_OK;
}

int mpm_kernel_times(mpm_engine *e, int32_t reset, int32_t capacity, const char **names, double *total_ms,
                     uint64_t *launches, int32_t *count) {
    if (!e || capacity < 0) return MPM_ERR_INVALID_ARGUMENT;
    int c = std::min<int>(capacity, ST_COUNT);
    for (int i = 0; i < c; ++i) {
        if (names) names[i] = kStageNames[i];
        if (total_ms) total_ms[i] = e->stage_ms[i];
        if (launches) launches[i] = e->stage_launches[i];
    }
    if (count) *count = c;
    if (reset) {
        for (int i = 0; i < ST_COUNT; ++i) {
            e->stage_ms[i] = 0;
            e->stage_launches[i] = 0;
        }
    }
    return MPM_OK;
}

// ---- per-kernel entry points
size_t mpm_dev_radix_sort_temp_bytes(uint64_t n) { return radix_sort_temp_bytes(n); }

int mpm_dev_radix_sort(const uint32_t *keys_in, const uint32_t *values_in, uint32_t *keys_out, uint32_t *values_out,
                       uint64_t n, int32_t key_bits, void *temp, void *stream) {
    if (n > 0 && (!keys_in || !keys_out || !values_out || !temp)) return MPM_ERR_INVALID_ARGUMENT;
    if (key_bits < 0 || key_bits > 32 || n >= (1ull << 32)) return MPM_ERR_INVALID_ARGUMENT;
    if (keys_in == keys_out || (values_in && values_in == values_out)) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(radix_sort(keys_in, values_in, keys_out, values_out, nullptr, nullptr, n, key_bits, temp,
                       static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

size_t mpm_dev_scan_temp_bytes(uint64_t n) { return scan_temp_bytes(n); }

int mpm_dev_exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t *total, uint64_t n, void *temp, void *stream) {
    if (!temp || (n > 0 && (!in || !out))) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(exclusive_scan(in, out, total, n, temp, static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

int mpm_dev_copy(const void *src, void *dst, uint64_t bytes, void *stream) {
    if (bytes % 16 != 0 || (bytes > 0 && (!src || !dst))) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(launch_copy(src, dst, bytes, static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

int mpm_dev_stress(const uint32_t *flags, const mpm_particle_parameters *params, const float *F, const float *jp,
                   float snow_hardening, float *tau, float *energy, uint64_t n, void *stream) {
    if (n > 0 && (!flags || !params || !F || !tau)) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(launch_stress(flags, params, F, jp, snow_hardening, tau, energy, n, static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

int mpm_dev_svd(const float *F, float *U, float *s, float *V, uint64_t n, void *stream) {
    if (n > 0 && (!F || !U || !s || !V)) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(launch_svd(F, U, s, V, n, static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

int32_t mpm_mlp_param_count(void) { return MLP_NPARAM; }

int mpm_set_closure(mpm_engine *e, const mpm_closure *c) {
    if (!e) return MPM_ERR_INVALID_ARGUMENT;
    if (!c) {
        e->closure_on = false;
        return MPM_OK;
    }
    if (!c->params || c->judge_stride < 0) return MPM_ERR_INVALID_ARGUMENT;
    for (int q = 0; q < MLP_IN; ++q)
        if (!(c->hull_lo[q] <= c->hull_hi[q])) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(e->device));
    if (!e->closure_params) {
        if (e->alloc(e->closure_params, MLP_NPARAM) != hipSuccess ||
            e->alloc(e->closure_tau, 9 * std::max<uint64_t>(e->n, 1)) != hipSuccess ||
            e->alloc(e->closure_svd, 12 * std::max<uint64_t>(e->n, 1)) != hipSuccess ||
            e->alloc(e->closure_stats, CLOSURE_STATS_SLOTS) != hipSuccess)
            return MPM_ERR_OUT_OF_MEMORY;
        HIP_TRY(hipMemsetAsync(e->closure_stats, 0, CLOSURE_STATS_SLOTS * sizeof(ClosureStats), e->stream));
    }
    // Host or device memory (the parameters of a network trained on the GPU).
    HIP_TRY(hipMemcpyAsync(e->closure_params, c->params, MLP_NPARAM * sizeof(float), hipMemcpyDefault, e->stream));
    if (!e->closure_net && e->alloc(e->closure_net, 1) != hipSuccess) return MPM_ERR_OUT_OF_MEMORY;
    HIP_TRY(launch_pack_net(e->closure_params, e->closure_net, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (int q = 0; q < MLP_IN; ++q) {
        e->hull_lo[q] = c->hull_lo[q];
        e->hull_hi[q] = c->hull_hi[q];
    }
    e->judge_stride = (uint32_t)c->judge_stride;
    e->judge_phase = 0;
    e->closure_on = true;
    return MPM_OK;
}

int mpm_closure_pairs(mpm_engine *e, uint32_t stride, uint32_t phase, float *x, float *t, uint64_t capacity,
                      uint64_t *count) {
    if (!e || !count || stride == 0 || phase >= stride || (capacity > 0 && (!x || !t)))
        return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(e->device));
    if (!e->pair_count && e->alloc(e->pair_count, 1) != hipSuccess) return MPM_ERR_OUT_OF_MEMORY;
    HIP_TRY(hipMemsetAsync(e->pair_count, 0, sizeof(unsigned long long), e->stream));
    HIP_TRY(launch_closure_pairs(e->P[e->cur], e->X, e->n, e->consts.snow_hardening, stride, phase, x, t, capacity,
                                 e->pair_count, e->stream));
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&c, e->pair_count, sizeof(c), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *count = (uint64_t)c;
    return c > capacity ? MPM_ERR_CAPACITY : MPM_OK;
}

int mpm_read_time_step_limits(mpm_engine *e, mpm_time_step_limits *out) {
    if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipSetDevice(e->device));
    if (!e->limit_part &&
        (e->alloc(e->limit_part, step_limit_groups(e->n)) != hipSuccess || e->alloc(e->limit_out, 1) != hipSuccess))
        return MPM_ERR_OUT_OF_MEMORY;
    HIP_TRY(launch_step_limits(e->P[e->cur], e->n, e->consts.grid_node_size, true, e->limit_part, e->limit_out, e->ds,
                               e->stream));
    float4 r;
    HIP_TRY(hipMemcpyAsync(&r, e->limit_out, sizeof(r), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    out->by_sound = r.x;
    out->by_isolated = r.y;
    out->by_velocity = r.z > 0.f ? 0.5f * e->consts.grid_node_size / r.z : INFINITY;
    out->by_deformation = r.w;
    return MPM_OK;
}

int mpm_debug_cancel_after(mpm_engine *e, int64_t substep) {
    if (!e) return MPM_ERR_INVALID_ARGUMENT;
    e->cancel_after = substep < 0 ? -1 : substep;
    return MPM_OK;
}

int mpm_set_time_step(mpm_engine *e, float time_step) {
    if (!e || !(time_step >= 0.f) || !std::isfinite(time_step)) return MPM_ERR_INVALID_ARGUMENT;
    e->consts.time_step = time_step;
    return MPM_OK;
}

int mpm_read_closure_stats(mpm_engine *e, int32_t reset, mpm_closure_stats *out) {
    if (!e || !out) return MPM_ERR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    if (!e->closure_stats) return MPM_OK;
    HIP_TRY(hipSetDevice(e->device));
    std::vector<ClosureStats> h(CLOSURE_STATS_SLOTS);
    HIP_TRY(hipMemcpyAsync(h.data(), e->closure_stats, h.size() * sizeof(ClosureStats), hipMemcpyDeviceToHost,
                           e->stream));
    if (reset) HIP_TRY(hipMemsetAsync(e->closure_stats, 0, h.size() * sizeof(ClosureStats), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (const ClosureStats &s : h) {
        out->evaluated += s.evaluated;
        out->fallbacks += s.fallbacks;
        out->judged += s.judged;
        out->err2 += s.err2;
        out->ref2 += s.ref2;
    }
    return MPM_OK;
}

int mpm_mlp_forward(const float *params, const float *x, float *y, uint64_t n, void *stream) {
    if (!params || (n > 0 && (!x || !y))) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(launch_mlp_forward(params, x, y, n, static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

int mpm_mlp_grad(const float *params, const float *x, const float *targets, uint64_t n, float scale, float *grad,
                 float *loss, void *stream) {
    if (!params || !grad || !loss || (n > 0 && (!x || !targets))) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(launch_mlp_grad(params, x, targets, n, scale, grad, loss, static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

int mpm_mlp_adam(float *params, float *grad, float *m, float *v, float lr, float beta1, float beta2, float eps,
                 int32_t step, void *stream) {
    if (!params || !grad || !m || !v || step < 1) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(launch_mlp_adam(params, grad, m, v, lr, beta1, beta2, eps, step, static_cast<hipStream_t>(stream)));
    return MPM_OK;
}

}  // extern "C"
