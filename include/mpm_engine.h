/*
 * mpm_engine.h - C ABI of the MI355X MLS-MPM substep engine.
 *
 * The engine replaces the wgpu backend of squishy_volumes behind the backend
 * contract that the compute thread dispatches to
 * (rust/crates/core/src/compute_thread.rs:100-163):
 *
 *   GpuState::from_io_state     rust/crates/gpu/src/gpu_state.rs:38-190
 *   GpuState::produce_next_state rust/crates/gpu/src/gpu_state.rs:331-675
 *   CpuState::to_io_state        rust/crates/cpu/src/cpu_state.rs:71-136
 *
 * Every pointer in the engine API (mpm_create .. mpm_read_grid) is a HOST
 * pointer; the engine owns all device memory.  The mpm_dev_* entry points
 * further below run one kernel stage on DEVICE pointers and exist for the
 * per-kernel parity tests and the per-kernel benchmark.
 *
 * Threading: one engine is used from one thread at a time (the reference calls
 * its backend only from the simulation's compute thread,
 * rust/crates/core/src/compute_thread.rs:78).  No C++ exception crosses the
 * ABI; every entry point returns an mpm_result code.
 */
#ifndef SQUISHY_VOLUMES_AMD_MPM_ENGINE_H
#define SQUISHY_VOLUMES_AMD_MPM_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- particle flags ------------------------------------------------------
 * Bits 0-6 have the values of ParticleFlags
 * (rust/crates/file_frame/src/particles.rs:23-33).  Bits 8-9 select the two
 * constitutive models that the reference does not have (fixed-corotated and
 * snow plasticity); a solid without them is Neo-Hookean as in the reference. */
enum {
    MPM_IS_SOLID = 1u << 0,
    MPM_IS_FLUID = 1u << 1,
    MPM_USE_VISCOSITY = 1u << 2,
    MPM_USE_SAND_ALPHA = 1u << 3,
    MPM_HAS_GOAL = 1u << 4,
    MPM_TOMBSTONED = 1u << 5,
    MPM_FAILED = 1u << 6,
    MPM_FIXED_COROTATED = 1u << 8,
    MPM_SNOW = 1u << 9
};

/* ---- device status word --------------------------------------------------
 * Low 16 bits: error bits, high 16 bits: id of the reporting kernel, as in
 * GpuStatus (rust/crates/gpu/src/status.rs:18-77).  Bits 1-8 keep the
 * reference values; 16.. are errors this engine adds. */
enum {
    MPM_STATUS_TABLE_TRIES_EXCEEDED = 1u,      /* the table of collider twin nodes overflowed (the call
                                                  is redone with a larger one, as the reference redoes a
                                                  frame, rust/crates/gpu/src/gpu_state.rs:562-599) */
    MPM_STATUS_TABLE_ENTRY_MISSING = 2u,       /* a gather found no twin node of its collider bits */
    MPM_STATUS_INDIRECT_LIMIT_EXCEEDED = 4u,   /* a block held more particles than a work list can */
    MPM_STATUS_PARTICLE_CLOSE_TO_INVERTED = 8u,/* det F < 1e-7 before the stress evaluation */
    MPM_STATUS_ENERGY_ERROR = 16u,             /* det F <= 0 after the F update (CPU EnergyError) */
    MPM_STATUS_OUT_OF_GRID = 32u,              /* a live particle's stencil left the allocated grid */
    MPM_STATUS_SVD_FP64 = 64u                  /* information: a near-degenerate F took the FP64 SVD */
};

/* ---- result codes -------------------------------------------------------- */
typedef enum mpm_result {
    MPM_OK = 0,
    /* Inner errors: the state is valid and readable, the error is reported after
     * it (the inner Result of produce_next_state,
     * rust/core/src/compute_thread.rs:165-169). */
    MPM_ERR_ENERGY = 1,            /* a particle's det F went <= 0; it is flagged FAILED */
    MPM_ERR_CLOSE_TO_INVERTED = 2, /* GpuShaderError::ParticleCloseToInverted: det F < 1e-7 */
    /* Fatal errors (the outer Err). */
    MPM_ERR_INVALID_ARGUMENT = -1,
    MPM_ERR_HIP = -2,
    MPM_ERR_NO_DEVICE = -3,
    MPM_ERR_OUT_OF_MEMORY = -4,
    MPM_ERR_GRID_TOO_LARGE = -5,
    MPM_ERR_CANCELED = -6,         /* Error::Canceled (rust/crates/cpu/src/errors.rs:11) */
    MPM_ERR_WRONG_FRAME = -7,      /* FrameInputError::WrongFrameLoaded (xpu/src/frame_input.rs:13) */
    MPM_ERR_ZERO_TIME_STEP = -8,   /* Error::ZeroTimeStep */
    MPM_ERR_OUT_OF_GRID = -9,
    MPM_ERR_CAPACITY = -10
} mpm_result;

/* ---- engine options (mpm_consts.options) ---------------------------------
 * Grid-side forces and boundary conditions have no reference counterpart: the
 * reference adds gravity to particle velocities before P2G
 * (rust/crates/cpu/src/phase/external_force.rs:44-46) and has no grid BC. */
enum {
    MPM_OPT_GRID_GRAVITY = 1u << 0,  /* add dt*g on the grid instead of on particles */
    MPM_OPT_STICKY_PLANE = 1u << 1,  /* nodes with z < plane_z get v = 0 */
    MPM_OPT_BOX_WALLS = 1u << 2,     /* separating slip on the faces of [wall_min, wall_max] */
    MPM_OPT_SPHERE = 1u << 3         /* separating slip on a sphere: nodes inside lose their
                                        velocity component into it */
};

/* Per-simulation constants: InputConsts (rust/crates/file_input/src/header.rs:9-69)
 * after scaling, plus the fixed time step of ComputeSettings
 * (rust/crates/core/src/simulation.rs:242-250). */
typedef struct mpm_consts {
    float grid_node_size;        /* scaled_grid_node_size() */
    float time_step;             /* fixed dt (see mpm_set_adaptive_time_steps) */
    float domain_min[3];         /* scaled_domain_min(); particles are culled outside */
    float domain_max[3];
    uint32_t frames_per_second;
    uint32_t options;            /* MPM_OPT_* */
    float plane_z;               /* MPM_OPT_STICKY_PLANE */
    float wall_min[3];           /* MPM_OPT_BOX_WALLS */
    float wall_max[3];
    /* snow plasticity (MPM_SNOW particles): singular values are clamped to
     * [1 - snow_compression, 1 + snow_stretch]; Lame parameters scale with
     * exp(snow_hardening * (1 - Jp)). */
    float snow_compression;
    float snow_stretch;
    float snow_hardening;
    int32_t device;              /* HIP device ordinal; GpuState's adapter name */
    float sphere_center[3];      /* MPM_OPT_SPHERE */
    float sphere_radius;
    /* Batched scenes (no reference counterpart): with scene_repeat[a] > 1 the
     * domain [domain_min, domain_max] is one scene's, repeated scene_repeat[a]
     * times along axis a with period scene_period[a] (> the domain's width).
     * The grid covers all copies; the cull, walls, plane and sphere act on
     * each particle's and node's position within its own copy, copies
     * meeting in the middle of the gap between them (scene_period[a] must be
     * at least the domain's width + 4 grid_node_size, else
     * MPM_ERR_INVALID_ARGUMENT).  0 or 1 means one copy. */
    float scene_period[3];
    uint32_t scene_repeat[3];
} mpm_consts;

/* Per-particle parameters, the flattened form of ParticleParameters
 * (rust/crates/file_frame/src/particles.rs:61-80) as the reference's device
 * struct ParticleParametersDevice (rust/crates/gpu/src/particle_parameters.rs:15-24). */
typedef struct mpm_particle_parameters {
    float mass;
    float initial_volume;
    float viscosity_dynamic;     /* used iff MPM_USE_VISCOSITY */
    float viscosity_bulk;
    float mu;                    /* solids */
    float lambda;
    float sand_alpha;            /* used iff MPM_USE_SAND_ALPHA */
    float bulk_modulus;          /* fluids */
    int32_t exponent;
} mpm_particle_parameters;

/* A view of Particles (rust/crates/file_frame/src/particles.rs:92-109).  All
 * arrays hold `count` entries in the caller's (original) particle order.
 * Vectors are xyz triples, matrices are column-major 3x3 ([[f32;3];3] with the
 * outer index the column), as in IoState.  plastic_jp is an extension (snow
 * plasticity); NULL means 1 for every particle on input and "not wanted" on
 * output. */
typedef struct mpm_particles_view {
    uint64_t count;
    double time;                         /* IoState::time */
    uint32_t *flags;
    mpm_particle_parameters *parameters;
    float *elastic_energies;
    uint32_t *collider_bits;
    float *positions;                    /* count*3 */
    float *position_gradients;           /* count*9 */
    float *velocities;                   /* count*3 */
    float *velocity_gradients;           /* count*9 */
    float *initial_positions;            /* count*3 */
    float *plastic_jp;                   /* count, optional */
} mpm_particles_view;

/* Per-frame inputs: the parts of FrameInput (rust/crates/xpu/src/frame_input.rs:33-278)
 * the substep reads.  `a` is the frame that contains the current time, `b`
 * the next one (or absent at the end of the input).  The collider fields are
 * the frames' InputFrame data of the mesh set with mpm_set_collider_mesh
 * (vertex positions, and the per-triangle friction and damping); they are
 * required while a mesh is set and ignored otherwise.  The mesh is
 * interpolated per substep with the frame factor, as the reference's
 * interpolate_input does (rust/crates/cpu/src/phase/interpolate_input.rs:18-106),
 * and its vertex velocities are (b - a) * frames_per_second, or 0 without b
 * (linear_vertex_velocities, rust/crates/xpu/src/frame_input.rs:333-347). */
typedef struct mpm_frame_input {
    uint64_t frame;                      /* index of frame a */
    float gravity_a[3];
    float gravity_b[3];
    int32_t has_b;
    const uint32_t *particle_flags_a;    /* count, original order; NULL: no goals */
    const uint32_t *particle_flags_b;
    const float *goal_positions_a;       /* count*3 */
    const float *goal_positions_b;
    const float *vertex_positions_a;     /* num_vertices*3 of the collider mesh */
    const float *vertex_positions_b;
    const float *triangle_frictions_a;   /* num_triangles */
    const float *triangle_frictions_b;
    const float *triangle_dampings_a;    /* num_triangles */
    const float *triangle_dampings_b;
} mpm_frame_input;

/* The collider mesh's topology (Topology, rust/crates/mesh_util/src/mesh.rs;
 * set up once per simulation like the reference's FrameInput): triangles as
 * vertex index triples, and the collider (0..15) each triangle belongs to.
 * The engine derives the triangles across each edge (triangle_opposites) and
 * the triangles at each vertex itself.  Collider bits per particle hold, for
 * collider c, bit 16 + c "near" and bit c "side" (collider_bits.rs:9-38). */
typedef struct mpm_collider_mesh {
    uint32_t num_vertices;
    uint32_t num_triangles;
    const uint32_t *triangles;           /* num_triangles*3 */
    const uint32_t *triangle_colliders;  /* num_triangles, each < 16 */
} mpm_collider_mesh;

/* Report of one mpm_step_to call. */
typedef struct mpm_status {
    uint32_t device_status;              /* the device status word (bits above) */
    int32_t result;                      /* the mpm_result of the call */
    uint64_t substeps;                   /* substeps run by this call */
    double time;                         /* simulation time after the call */
    uint64_t active_blocks;              /* occupied 4^3 blocks in the last substep */
    uint64_t active_nodes;               /* grid nodes with mass > 0 in the last substep */
    uint64_t svd_fp64_fallbacks;         /* 1 when an SVD was recomputed in FP64 during the call */
    uint64_t graph_launches;             /* sort cycles run as one HIP graph launch by the call */
    uint64_t grid_blocks;                /* 4^3 blocks of the block grid (the window's, with
                                            mpm_tuning.grid_window) */
    uint64_t window_redos;               /* times the call ran again with a larger grid window */
} mpm_status;

/* Grid nodes, as GridNodes (rust/crates/file_frame/src/grid_nodes.rs:9-15):
 * the active nodes (mass > 0) after the last grid update. */
typedef struct mpm_grid_view {
    uint64_t capacity;                   /* entries the arrays can take */
    uint64_t count;                      /* entries written */
    int32_t *node_ids;                   /* count*3 */
    uint32_t *collider_bits;
    float *masses;
    float *velocities;                   /* count*3 */
} mpm_grid_view;

typedef struct mpm_engine mpm_engine;

/* Engine tuning (the autotuner's knobs).  Zero fields take defaults. */
typedef struct mpm_tuning {
    int32_t resort_interval;             /* substeps between particle sorts (default 8); between
                                            sorts a particle keeps its slot and may drift one cell
                                            out of its block; farther drift takes a slower path */
    int32_t persistent_blocks;           /* work-groups of the transfer kernel (default: all resident) */
    int32_t graph;                       /* -1: launch every kernel of the substep loop; otherwise
                                            (default) each whole sort cycle inside a call is one
                                            launch of a captured HIP graph (not while profiling,
                                            not with the learned closure) */
    int32_t particles_per_lane;          /* particles each thread of the transfer kernel takes per
                                            pass of a block, 1..3 (default 3: passes of 768) */
    int32_t grid_update_threads;         /* work-group size of the grid update kernel: 64, 128 or
                                            256 threads (0: 256); its wave count stays 32 per CU */
    int32_t learned_fused;               /* 1: the learned closure evaluated after each advance
                                            inside the fused gather-scatter (not with colliders);
                                            otherwise (default) in its own pass between the gather
                                            and a scatter from the state, which measures faster */
    int32_t grid_window;                 /* 1: the block grid is a window fitted to the particles
                                            at the start of every call (grown and the call run
                                            again when a particle leaves it), so that its memory
                                            follows the particles, not the domain; -1: the grid
                                            covers the whole domain; 0 (default): a window when
                                            the whole domain would take more than 4 GiB of grid
                                            or more than 1023 blocks along an axis */
} mpm_tuning;

/* ---- engine API (host pointers) ------------------------------------------ */

/* GpuState::from_io_state (rust/crates/gpu/src/gpu_state.rs:38-190).  Copies
 * the particles to the device.  The block grid covers the domain, or a
 * window fitted to the particles (mpm_tuning.grid_window); fails with
 * MPM_ERR_GRID_TOO_LARGE when the grid it needs has more than 1023 blocks
 * along an axis or more than 96 GiB of nodes (a window: at mpm_step_to). */
int mpm_create(const mpm_consts *consts, const mpm_particles_view *particles,
               const mpm_tuning *tuning, mpm_engine **out_engine);

/* GpuState::produce_next_state (rust/crates/gpu/src/gpu_state.rs:337-675):
 * runs substeps of the fixed time step while time < target_time.  `cancel`
 * (may be NULL) is polled between substeps like Harness::check
 * (rust/crates/xpu/src/harness.rs:65-76), and between sort cycles where they
 * run as HIP graphs: set before the first substep, the call does nothing;
 * later, the substep whose grid update has run is finished (its gather
 * writes x, v, F, C) and the call returns MPM_ERR_CANCELED with that
 * consistent state.
 * A particle that fails stops the call at the failing substep
 * (CpuState::produce_next_state, rust/crates/cpu/src/cpu_state.rs:178-183):
 *   - det F <= 0 after a Neo-Hookean (or sand) particle's advance
 *     (try_elastic_energy_neo_hookean, advance_particles.rs:73-74):
 *     MPM_ERR_ENERGY; the state holds x, F after that advance and v, C of
 *     that substep, the time stays at the substep's start;
 *   - det F < 1e-7 after any live particle's advance, i.e. before its next
 *     stress evaluation (prepare_tmp.wesl:74-78): MPM_ERR_CLOSE_TO_INVERTED;
 *     the state and the time are at the end of that substep (the wgpu
 *     backend reports the frame's target time instead);
 *   - det F < 1e-7 in the stored state: MPM_ERR_CLOSE_TO_INVERTED before any
 *     substep, the state unchanged.
 * The particle is flagged MPM_FAILED and takes no part in any scatter.  The
 * engine finds the failing substep by running the call again from its start
 * (it keeps the starting state for that), so such a call costs up to twice
 * its substeps.
 * A call runs at most MPM_MAX_SUBSTEPS_PER_CALL substeps: a longer span, or a
 * step too small to advance the f64 time, returns MPM_ERR_INVALID_ARGUMENT
 * before any work. */
#define MPM_MAX_SUBSTEPS_PER_CALL (1u << 20)
int mpm_step_to(mpm_engine *engine, double target_time, const mpm_frame_input *input,
                const volatile int32_t *cancel, mpm_status *status);

/* CpuState::to_io_state (rust/crates/cpu/src/cpu_state.rs:71-136): writes the
 * particle state in original order into caller-owned arrays (count must match;
 * any array may be NULL to skip it). */
int mpm_read_particles(mpm_engine *engine, mpm_particles_view *out);

/* The grid half of to_io_state (store_grid = true). */
int mpm_read_grid(mpm_engine *engine, mpm_grid_view *out);

/* Mesh colliders (the collide phase, rust/crates/cpu/src/phase/collide.rs:21-206,
 * collider bits and the grid's twin nodes, meld_grid.rs:16-69): from the next
 * call on, every substep sets each live particle's collider bits from the
 * closest triangle of each collider within 2.2 h, corrects the velocity of a
 * particle that crossed to the other side (friction, damping, push-back), and
 * particles whose bits are incompatible exchange no momentum through the
 * grid.  NULL removes the mesh.  Not with batched scenes
 * (MPM_ERR_INVALID_ARGUMENT). */
int mpm_set_collider_mesh(mpm_engine *engine, const mpm_collider_mesh *mesh);

/* Drop of GpuState. */
void mpm_destroy(mpm_engine *engine);

/* Text of an mpm_result (the Display of rust/crates/gpu/src/errors.rs and
 * rust/crates/cpu/src/errors.rs). */
const char *mpm_error_string(int result);

/* Simulation time of the engine (CpuState::time, rust/crates/cpu/src/cpu_state.rs). */
double mpm_time(const mpm_engine *engine);

/* Per-kernel timing (in place of the profiling_output_file of
 * GpuState::from_io_state, rust/crates/gpu/src/gpu_state.rs:38-190) of the
 * substeps run since the last call with reset=1:
 * names[i] / total_ms[i] / launches[i] for i < *count (at most `capacity`).
 * Timing uses HIP events on the engine's stream and is enabled by
 * mpm_set_profiling: enabled = 0 off, 1 every stage, otherwise a bit mask
 * with bit i for the i-th name (each timed launch adds two event records to
 * the stream, which cost a little time of their own). */
int mpm_set_profiling(mpm_engine *engine, int32_t enabled);
int mpm_kernel_times(mpm_engine *engine, int32_t reset, int32_t capacity, const char **names,
                     double *total_ms, uint64_t *launches, int32_t *count);

/* The engine's HIP stream (hipStream_t) so that callers can order work (no
 * counterpart in the reference). */
void *mpm_stream(mpm_engine *engine);

/* Number of HIP devices (GpuContext::available_gpus, rust/crates/gpu/src/context.rs). */
int mpm_device_count(void);

/* ---- per-kernel entry points (device pointers, stream may be NULL) -------- */

/* Stable LSD radix sort of `n` u32 keys by their low `key_bits` bits, carrying
 * u32 values.  values_in == NULL sorts the identity permutation.  The result
 * equals a stable CPU sort by key (sort_on_cpu, rust/crates/gpu/src/util.rs:65-69;
 * RadixSort::record_all_rounds, rust/crates/gpu/src/radix_sort/mod.rs:209-238).
 * The temp buffer takes mpm_dev_radix_sort_temp_bytes(n) bytes. */
size_t mpm_dev_radix_sort_temp_bytes(uint64_t n);
int mpm_dev_radix_sort(const uint32_t *keys_in, const uint32_t *values_in, uint32_t *keys_out,
                       uint32_t *values_out, uint64_t n, int32_t key_bits, void *temp,
                       void *stream);

/* The same sort without a temp buffer, for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Its working
 * memory is a device array of 34 MiB that each GPU holds from the library's
 * load to the process's end; inputs of more than
 * mpm_dev_radix_sort_noalloc_max_keys() keys (2^22) return MPM_ERR_CAPACITY.
 * Calls on one GPU share that array: calls that overlap in time on one GPU
 * (other streams, other threads) race and must be serialised by the caller.
 * The call allocates nothing. */
uint64_t mpm_dev_radix_sort_noalloc_max_keys(void);
int mpm_dev_radix_sort_noalloc(const uint32_t *keys_in, const uint32_t *values_in, uint32_t *keys_out,
                               uint32_t *values_out, uint64_t n, int32_t key_bits, void *stream);

/* Exclusive prefix sum of u32 (prefix_sum_on_cpu, rust/crates/gpu/src/util.rs:54-63);
 * *total receives the sum (device pointer). */
size_t mpm_dev_scan_temp_bytes(uint64_t n);
int mpm_dev_exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t *total, uint64_t n,
                           void *temp, void *stream);

/* Copy kernel of known size (float4 loads/stores), used to calibrate the HBM
 * byte counters of the profiler. */
int mpm_dev_copy(const void *src, void *dst, uint64_t bytes, void *stream);

/* Per-particle constitutive kernel for the parity tests: for `n` particles
 * (SoA device arrays, F column-major 9 per particle) writes the Kirchhoff
 * stress tau = P F^T (9 per particle, column-major), the elastic energy density
 * and, for SVD-based models, the signed singular values (3 per particle). */
int mpm_dev_stress(const uint32_t *flags, const mpm_particle_parameters *params,
                   const float *F, const float *jp, float snow_hardening, float *tau,
                   float *energy, uint64_t n, void *stream);

/* 3x3 signed SVD F = U diag(s) V^T with U, V rotations (n matrices, column-major). */
int mpm_dev_svd(const float *F, float *U, float *s, float *V, uint64_t n, void *stream);

/* ---- Learned constitutive closure (no counterpart in the reference, which
 * has no learned stress model; it stands in for the analytic snow stress).
 * A 6-64-64-3 perceptron with ReLU hidden layers, evaluated with BF16
 * matrix-core products and FP32 accumulation.  Its FP32 parameters are one
 * device array of mpm_mlp_param_count() floats, padded:
 *   W1 [64][16] (columns 6..15 zero), b1 [64], W2 [64][64], b2 [64],
 *   W3 [16][64] (rows 3..15 zero), b3 [16];  layer k: z = W h + b.
 * Samples are row-major: x [n][6], y and targets [n][3].  Weights, inputs,
 * hidden activations and back-propagated errors are rounded to BF16 (nearest
 * even) where they enter a product; biases, sums and outputs are FP32. */
int32_t mpm_mlp_param_count(void);
int mpm_mlp_forward(const float *params, const float *x, float *y, uint64_t n, void *stream);
/* Adds sum((y - t)^2) over the n samples to *loss and scale times its
 * gradient to grad (scale = 1/N gives the gradient of the mean over N). */
int mpm_mlp_grad(const float *params, const float *x, const float *targets, uint64_t n, float scale,
                 float *grad, float *loss, void *stream);
/* One Adam step (bias-corrected, step >= 1) on params; clears grad. */
int mpm_mlp_adam(float *params, float *grad, float *m, float *v, float lr, float beta1, float beta2,
                 float eps, int32_t step, void *stream);

/* The learned closure in the substep: with it set, the stress of every snow
 * particle in each substep's scatter comes from the network, evaluated on
 * the particle's sorted singular values s (SVD of F), E, nu and Jp.  Network
 * inputs: (s_q - 1) / 0.025, log10(E / 1e5), (nu - 0.3) / 0.1, (Jp - 1) / 0.1;
 * outputs: the principal Kirchhoff stresses minus their linear-elastic part
 * h (2 mu (s - 1) + lambda tr(s - 1)), in units of 0.025^2 E h m, with
 * h = exp(xi (1 - Jp)) (xi = snow_hardening) and m = max(|s - 1| / 0.025,
 * 1e-6) (|.| the Euclidean norm); the closure adds the linear part
 * (FP32) and composes tau = U diag(linear + out) U^T.  A particle whose
 * inputs leave [hull_lo, hull_hi] (the training set's range) takes the
 * analytic stress; every judge_stride-th particle (by id, rotating each
 * substep) takes both and the difference goes into the statistics.  A
 * substep then runs as a gather writing the state, the sort when due, the
 * closure, a scatter from the state and the grid update.  No counterpart in
 * the reference.  params is a host or device pointer (copied on the engine's
 * stream; device memory must be complete when the call is made); NULL turns
 * it off. */
typedef struct mpm_closure {
    const float *params;             /* mpm_mlp_param_count() floats */
    float hull_lo[6], hull_hi[6];
    int32_t judge_stride;            /* 0: no judge */
} mpm_closure;
typedef struct mpm_closure_stats {
    uint64_t evaluated;              /* network stresses used */
    uint64_t fallbacks;              /* analytic stresses outside the hull */
    uint64_t judged;                 /* particles checked against the analytic stress */
    double err2, ref2;               /* sum |tau - tau_analytic|^2, sum |tau_analytic|^2 (judged) */
} mpm_closure_stats;
int mpm_set_closure(mpm_engine *engine, const mpm_closure *closure);
/* Statistics since the last reset. */
int mpm_read_closure_stats(mpm_engine *engine, int32_t reset, mpm_closure_stats *out);
/* Training pairs from the engine's current state, for fine-tuning the
 * closure online: every live snow particle whose id % stride == phase gives
 * one row of network inputs x (6 floats) and targets t (3 floats: the
 * analytic principal Kirchhoff stresses 2 mu h (s - 1) s + lambda h (J - 1) J,
 * h = exp(xi (1 - Jp)), minus their linear part, in the output units
 * above).  x and t are device arrays of capacity rows; *count receives the number of rows (in an
 * unspecified order); MPM_ERR_CAPACITY when it exceeds the capacity (the
 * first capacity rows are written).  Synchronizes the engine's stream.  No
 * counterpart in the reference. */
int mpm_closure_pairs(mpm_engine *engine, uint32_t stride, uint32_t phase, float *x, float *t, uint64_t capacity,
                      uint64_t *count);

/* One substep stage at a time on an engine, for the per-kernel parity tests
 * (as the reference tests its shaders one by one, in the test.rs of each
 * shader's directory under rust/crates/gpu/src):
 * stage 0 = sort, 1 = P2G from the stored state (stress from F and C),
 * 2 = grid update, 3 = G2P writing the state (x, v, F, C), as the last
 * substep of a call does.  Host-side grid/particle readback goes through
 * mpm_read_grid / mpm_read_particles. */
int mpm_run_stage(mpm_engine *engine, int32_t stage, const mpm_frame_input *input);

/* Test hook: the following mpm_step_to calls behave as if their cancel flag
 * were set before substep `substep` of the call (where sort cycles run as
 * HIP graphs, before the first cycle boundary at or after it); a negative
 * value turns it off.  Stands in for a Harness cancel arriving mid-call. */
int mpm_debug_cancel_after(mpm_engine *engine, int64_t substep);

/* Raw P2G result (mass, momentum before normalisation) of the last P2G stage
 * for the active nodes, same layout as mpm_read_grid (velocities hold momentum). */
int mpm_read_grid_momentum(mpm_engine *engine, mpm_grid_view *out);

/* ---- time-step limits ---------------------------------------------------- */

/* The adaptive time-step limits of the reference (limit_time_step.rs:25-223,
 * rust/crates/cpu/src/phase/) over the live particles of the engine's
 * current state, as device reductions: the sound-speed limit h / c (c from
 * the largest diagonal-space stiffness of the Neo-Hookean or inviscid stress
 * over J and the current density), the isolated-particle limit, the velocity
 * limit 0.5 h / max|v| and the deformation limit min 0.2 / max(|C_ij|, 1e-8).
 * A limit without particles is +inf (None in the reference).  The reference
 * takes these twice per substep; here a caller reads them between calls and
 * sets the next calls' step with mpm_set_time_step (per-call adaptivity).
 * Synchronizes the engine's stream. */
typedef struct mpm_time_step_limits {
    float by_velocity;
    float by_deformation;
    float by_sound;
    float by_isolated;
} mpm_time_step_limits;
int mpm_read_time_step_limits(mpm_engine *engine, mpm_time_step_limits *out);
/* The time step of the following mpm_step_to calls (>= 0; 0 makes them
 * return MPM_ERR_ZERO_TIME_STEP). */
int mpm_set_time_step(mpm_engine *engine, float time_step);

/* Per-substep adaptive time steps: CpuRunParameters { adaptive_time_steps,
 * max_time_step } of CpuState::produce_next_state (rust/crates/cpu/src/
 * cpu_state.rs:138-197), with the phases LimitTimeStepBeforeForce and
 * LimitTimeStepBeforeIntegrate (rust/crates/cpu/src/phase/mod.rs:27-75,
 * limit_time_step.rs:25-223).  With adaptive != 0 every following
 * mpm_step_to substep takes two limits as device reductions: the sound-speed
 * and isolated-particle limits of the state (after the grid is laid out,
 * before the scatter) and the velocity and deformation limits of the
 * gathered v and C (before the advance).  Each is pushed into the engine's
 * AdaptiveTimeStepState (adaptive_time_step_state.rs:13-70: a history of
 * the last 11 pushes, a missing limit counting as f32::MAX), whose allowed
 * step (their minimum and max_time_step) is the step of the phases that
 * follow: the collide and external force take the step allowed at the
 * substep's start, the scatter and the grid update the step after the first
 * push, the advance and the time the step after the second.  A step of 0
 * returns MPM_ERR_ZERO_TIME_STEP.  The state lives as long as the engine.
 * With adaptive == 0 the calls take consts.time_step (mpm_set_time_step).
 * max_time_step must be > 0 (FLT_MAX or +inf: no maximum).  Not with a
 * batched engine.  An adaptive substep synchronizes the stream twice and is
 * not captured in a graph. */
int mpm_set_adaptive_time_steps(mpm_engine *engine, int32_t adaptive, float max_time_step);

/* ---- frame files (host only, no device) --------------------------------- */

/* The reference's frame file of an IoState (IoState::write / IoState::read,
 * rust/crates/file_frame/src/io_state.rs:31-75): 32 magic bytes, 64 version
 * bytes ("0.3.4", the squishy_volumes_file_util version whose readers accept
 * the file, rust/crates/file_util/src/lib.rs:27-44), then time, Particles and
 * Option<GridNodes> as bincode 1.3's default functions encode them.  The
 * parameters' enum and options follow the flags (MPM_USE_VISCOSITY,
 * MPM_IS_FLUID, MPM_USE_SAND_ALPHA), as ParticleFlags::from does in the
 * reference; plastic_jp is not part of the format (decoded as 1).  grid NULL
 * writes None. */
uint64_t mpm_frame_file_size(const mpm_particles_view *particles, const mpm_grid_view *grid);
/* Encodes into out; *written receives the size (also on MPM_ERR_CAPACITY). */
int mpm_frame_encode(const mpm_particles_view *particles, const mpm_grid_view *grid, uint8_t *out,
                     uint64_t capacity, uint64_t *written);
/* Checks a frame file's magic, version and structure and returns its counts;
 * *grid_nodes is UINT64_MAX when the file holds no grid. */
int mpm_frame_decode_counts(const uint8_t *data, uint64_t size, uint64_t *particles, uint64_t *grid_nodes);
/* Decodes into views of the counts above (grid may be NULL to skip the grid;
 * MPM_ERR_CAPACITY with grid->count set when its capacity is short). */
int mpm_frame_decode(const uint8_t *data, uint64_t size, mpm_particles_view *particles, mpm_grid_view *grid);

#ifdef __cplusplus
}
#endif

#endif
