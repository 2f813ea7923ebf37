/*
 * efl.h - C ABI of the MI355X-native per-step Eulerian fluid pipeline.
 *
 * This is the drop-in boundary for the render-world half of bevy_eulerian_fluid:
 * the pass framework, the render-graph node and its 33 compute kernels
 * (src/plugin.rs, src/pipeline.rs, src/euler_fluid/render_node.rs:84-476 and
 * the WGSL shaders below src/euler_fluid).  A host (the Rust FFI crate, the C++
 * mirror in bevy_eulerian_fluid_amd/host, or Python via ctypes) keeps the
 * reference's main-world API (FluidPlugin, FluidSettings, FluidSource,
 * LocalForces = ForcesToFluid, ProjectionMethod, ReinitializeMethod,
 * FluidTextures) and calls these entry points instead of recording wgpu
 * compute passes.
 *
 * Conventions
 *  - Every entry point returns an int status: EFL_OK (0) or a negative EFL_E_*.
 *    Nothing unwinds or aborts across the ABI; efl_last_error() returns a
 *    message for the last failure on a context.
 *  - The library owns all device memory.  Callers own host arrays and pass
 *    them by pointer and length for the duration of one call only.
 *  - All GPU work of a context is issued in order on one HIP stream owned by
 *    the context; getters and efl_sync() synchronise.  efl_step() returns
 *    without waiting for the device for Gauss-Seidel, Jacobi and fixed-cycle
 *    multigrid projections; an MGPCG step waits for its solve's stopping state
 *    unless the context is asynchronous (efl_set_async), and learned and
 *    residual-target multigrid steps always do.
 *  - Grids are row-major [y][x], x fastest, y = 0 is the TOP row (world y is
 *    up).  u is (W+1) x H, v is W x (H+1), every other field is W x H.
 *    Velocities are in cells per second.
 *  - Structs are plain C, 4-byte scalars, no padding surprises (sizes are
 *    asserted in the implementation).
 */
#ifndef EFL_H
#define EFL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EFL_ABI_VERSION 1

/* Limits of the reference. */
#define EFL_MAX_FLUID_SOURCES 16  /* fluid_source_uniform.rs:19 MAX_FLUID_SOURCE */
#define EFL_MAX_SOLIDS 256        /* fluid_to_solid.rs:28 MAX_SOLIDS */
#define EFL_MAX_BATCH 64          /* instances advanced by one efl_step call */

/* Status codes. */
#define EFL_OK 0
#define EFL_E_INVALID (-1)        /* bad argument (null pointer, size, enum) */
#define EFL_E_HIP (-2)            /* a HIP runtime call failed */
#define EFL_E_OOM (-3)            /* device allocation failed */
#define EFL_E_STATE (-4)          /* call not valid in the current state */
#define EFL_E_RANGE (-5)          /* count beyond a documented maximum */
#define EFL_E_NOT_CONVERGED (-6)  /* residual target missed (reported, not fatal) */

typedef struct efl_ctx efl_ctx;
typedef struct efl_instance efl_instance;

/* ProjectionMethod (projection.rs:45-52). */
typedef enum {
  EFL_PROJECTION_JACOBI = 0,        /* 50 x (jacobi, jacobi_reverse), projection.rs:135-143 */
  EFL_PROJECTION_GAUSS_SEIDEL = 1,  /* SOR red-black GS, weight 1.9, gauss_seidel.rs:60-148 */
  EFL_PROJECTION_MULTIGRID = 2,     /* V-cycle(s) with RBGS smoother, multi_grid.rs:390-475 */
  EFL_PROJECTION_MGPCG = 3,         /* no counterpart: CG preconditioned by one symmetric V-cycle */
  EFL_PROJECTION_LEARNED = 4        /* no counterpart: flexible CG preconditioned by the learned
                                       network (efl_set_learned), MGPCG as the fallback */
} efl_projection_kind;

/* ReinitializeMethod (reinitialize_levelset.rs:36-46). */
typedef enum {
  EFL_REINIT_JUMP_FLOODING = 0,
  EFL_REINIT_FAST_ITERATIVE = 1
} efl_reinit_kind;

/* GaussSeidelConfig / MultiGridConfig (gauss_seidel.rs:33-42, multi_grid.rs:56-71)
 * plus the residual-target loop that the reference does not have (it runs
 * exactly one V-cycle per step).  max_cycles = 1 and rel_tol = 0 reproduce
 * the reference. */
typedef struct {
  int32_t kind;               /* efl_projection_kind */
  uint32_t gs_iterations;     /* GaussSeidelConfig.num_iterations (default 20) */
  uint32_t mg_pre_iterations; /* MultiGridConfig.pre_smooth_config (default 2) */
  uint32_t mg_post_iterations;/* MultiGridConfig.post_smooth_config (default 2) */
  uint32_t mg_coarsest_iterations; /* MultiGridConfig.coarsest_config (default 3) */
  uint32_t max_cycles;        /* V-cycles (MULTIGRID) or CG iterations (MGPCG) per step */
  float rel_tol;              /* stop when ||r||2/||b||2 <= rel_tol; 0 = run max_cycles */
  uint32_t check_every;       /* residual check period in cycles (0 = 1) */
} efl_projection;

/* FastIterativeMethodConfig (fast_iterative_method.rs:14-23). */
typedef struct {
  int32_t kind;               /* efl_reinit_kind */
  uint32_t fim_iterations;    /* default 10 */
} efl_reinit;

/* FluidSettings (settings.rs:110-114) + FluidPlugin::new(length_unit)
 * (euler_fluid.rs:43-52; dx = 1/length_unit, settings.rs:116-123) + the
 * entity's Transform (fluid_uniform.rs:108-121) + its required
 * ProjectionMethod / ReinitializeMethod components (settings.rs:100-109). */
typedef struct {
  uint32_t width;             /* FluidSettings.size.x */
  uint32_t height;            /* FluidSettings.size.y */
  float rho;
  float gravity[2];
  float length_unit;          /* > 0 */
  float transform[16];        /* column-major Mat4; all zero = identity */
  efl_projection projection;
  efl_reinit reinit;
} efl_settings;

/* FluidSourceData (fluid_source_uniform.rs:48-55): shape 0 = circle
 * (data[0] = radius), 1 = AABB (data = half_size); mode 0 = source,
 * 1 = sink.  center is the source Transform translation, in pixels with the
 * origin at the grid centre and y up. */
typedef struct {
  float center[2];
  float data[2];
  float velocity[2];
  uint32_t shape;
  uint32_t mode;
} efl_source;

/* ForceToFluid (apply_forces.rs:36-40): force in m/s^2, position in cells
 * (x right, y down, like the grid index). */
typedef struct {
  float force[2];
  float position[2];
} efl_force;

/* SolidObstacle (obstacle.rs:17-32, solid_obstacle.wgsl:29-41).  shape is
 * parry's ShapeType id: 0 circle, 1 cuboid, 2 capsule, 4 triangle.  Matrices
 * are column-major; inverse_transform must be the inverse of transform. */
typedef struct {
  uint32_t entity_id;
  uint32_t shape;
  float values[6];
  float transform[16];
  float inverse_transform[16];
  float linear_velocity[2];
  float angular_velocity;
  float reserved;
} efl_obstacle;

/* FluidToSolidForce (fluid_to_solid.rs:60-64), before the readback's
 * dt/dx scaling, in the std430 16-byte slot whose padding here carries the
 * torque about the centre of mass in the world frame (no counterpart; the
 * reference's torque mixes a world-frame arm with a grid-frame force). */
typedef struct {
  float force[2];
  float torque;
  float torque_world;
} efl_solid_force;

/* No counterpart (avian2d's rigid-body response): per-obstacle dynamics for
 * the library's own integrator.  density is in units of FluidSettings.rho,
 * area in cells^2, inertia is the second moment of area about the centre of
 * mass in cells^4.  dynamic = 0 keeps the obstacle prescribed. */
typedef struct {
  float density;
  float area;
  float inertia;
  uint32_t dynamic;
} efl_body;

/* Per-instance results of the last step. */
typedef struct {
  uint64_t steps;             /* completed steps since the last initialisation */
  uint32_t cycles;            /* V-cycles / CG iterations used by the last projection */
  int32_t converged;          /* 1 if rel_tol was met (or rel_tol == 0) */
  float rel_residual;         /* ||r||2 / ||b||2 after the last projection (if measured) */
  float max_norm_divergence;  /* dt * max|D| over fluid, non-solid cells (if measured) */
  float projection_ms;        /* device time of the projection stage (if timing on) */
  float step_ms;              /* device time of the whole step (if timing on) */
  uint32_t unconverged_solves;  /* projections with a residual target that ended without meeting it,
                                   since the instance was created (asynchronous ones counted on the device) */
  uint32_t reserved;
} efl_stats;

/* Fluid diagnostics (diagnostics/component.rs: FluidVolume,
 * FluidMinVelocityMagnitude, FluidMaxVelocityMagnitude). */
typedef struct {
  float volume;        /* cells with levelset_air < 0 (calculate_volume.wgsl) */
  float min_velocity;  /* min_velocity.wgsl: smallest cell-centred speed over fluid cells, cells/s */
  float max_velocity;  /* max_velocity.wgsl: largest cell-centred speed, cells/s */
  float reserved;
} efl_fluid_diagnostics;

/* Field identifiers for efl_get_field / efl_set_field / efl_field_info. */
typedef enum {
  EFL_FIELD_U0 = 0,           /* (W+1) x H f32; FluidTextures.u */
  EFL_FIELD_U1,               /* (W+1) x H f32 */
  EFL_FIELD_V0,               /* W x (H+1) f32; FluidTextures.v */
  EFL_FIELD_V1,               /* W x (H+1) f32 */
  EFL_FIELD_U_SOLID,          /* (W+1) x H f32; FluidTextures.u_solid */
  EFL_FIELD_V_SOLID,          /* W x (H+1) f32; FluidTextures.v_solid */
  EFL_FIELD_SOLID_ID,         /* W x H i32 */
  EFL_FIELD_U_VALID0,         /* (W+1) x H i32 */
  EFL_FIELD_U_VALID1,         /* (W+1) x H i32 */
  EFL_FIELD_V_VALID0,         /* W x (H+1) i32 */
  EFL_FIELD_V_VALID1,         /* W x (H+1) i32 */
  EFL_FIELD_DIV,              /* W x H f32 */
  EFL_FIELD_P0,               /* W x H f32 */
  EFL_FIELD_P1,               /* W x H f32 */
  EFL_FIELD_LS_AIR0,          /* W x H f32; FluidTextures.levelset_air */
  EFL_FIELD_LS_AIR1,          /* W x H f32 */
  EFL_FIELD_GRAD_LS,          /* W x H x 2 f32 */
  EFL_FIELD_LS_SOLID,         /* W x H f32; FluidTextures.levelset_solid */
  EFL_FIELD_AREA_FRACTION,    /* W x H x 4 f32 (i-1/2, i+1/2, j-1/2, j+1/2) */
  EFL_FIELD_FIM_LABELS0,      /* W x H u32 */
  EFL_FIELD_FIM_LABELS,       /* W x H u32 */
  EFL_FIELD_JFA_SEEDS0,       /* W x H x 2 f32 */
  EFL_FIELD_JFA_SEEDS1,       /* W x H x 2 f32 */
  EFL_FIELD_SOLID_BINS,       /* 4 x 256 i32 (force x, force y, torque, world torque) */
  EFL_FIELD_SOLID_FORCES,     /* 256 x 4 f32 (efl_solid_force) */
  EFL_FIELD_MG_R0,            /* W x H f32: multigrid residual, finest level */
  EFL_FIELD_COUNT
} efl_field;

/* Stages of EulerFluidNode::run (render_node.rs:193-476), each runnable on
 * its own for parity tests and profiling. */
typedef enum {
  EFL_STAGE_INIT_GRID_CENTER = 0, /* K1 initialize_grid_center */
  EFL_STAGE_INIT_GRID_EDGE,       /* K2 initialize_grid_edge */
  EFL_STAGE_SOURCE_ONESHOT,       /* K3 with the oneshot (init) uniform */
  EFL_STAGE_UPDATE_SOLID,         /* K4 */
  EFL_STAGE_AREA_FRACTION,        /* K5 */
  EFL_STAGE_ADVECT_VELOCITY,      /* K6, K7 */
  EFL_STAGE_APPLY_FORCES,         /* K8, K9 */
  EFL_STAGE_DIVERGENCE,           /* K10 */
  EFL_STAGE_PROJECTION,           /* K11-K17 per efl_projection */
  EFL_STAGE_SOLVE_VELOCITY,       /* K18, K19 */
  EFL_STAGE_EXTRAPOLATE_VELOCITY, /* K20-K23 */
  EFL_STAGE_ADVECT_LEVELSET,      /* K24 (cubic) */
  EFL_STAGE_SOURCE,               /* K3 with the per-step uniform */
  EFL_STAGE_REINITIALIZE,         /* K25-K27 or K28-K30 */
  EFL_STAGE_LEVELSET_GRADIENT,    /* K31 */
  EFL_STAGE_FLUID_TO_SOLID,       /* K32, K33 (bins are cleared first) */
  EFL_STAGE_GS_SWEEP,             /* one red + black sweep at weight 1.9 */
  EFL_STAGE_MG_VCYCLE,            /* one V-cycle (reference order) */
  EFL_STAGE_MG_RESIDUAL,          /* finest-level residual into EFL_FIELD_MG_R0 */
  EFL_STAGE_MG_SMOOTH,            /* two red-black sweeps at weight 1 on level 0 (pre-smoothing) */
  EFL_STAGE_GS_SWEEP_SPLIT,       /* GS_SWEEP dispatched one colour per kernel */
  EFL_STAGE_NN_PRECONDITION,      /* one application of the learned preconditioner to the CG
                                     residual (needs efl_set_learned and a learned projection) */
  EFL_STAGE_MG_BOTTOM,            /* the V-cycle's bottom (the levels one workgroup solves) */
  EFL_STAGE_CG_SMOOTH,            /* timing only (efl_time_stage): the level-0 launch of an MGPCG
                                     iteration that carries the x / r update (2 sweeps, residual,
                                     restriction); needs an MGPCG instance with a tiled level 1 */
  EFL_STAGE_COUNT
} efl_stage;

/* ---- context ---------------------------------------------------------- */

/* EulerFluidNode + pass pipelines for one device.  device = HIP ordinal. */
int efl_ctx_create(int device, efl_ctx** out);
int efl_ctx_destroy(efl_ctx* ctx);
const char* efl_last_error(const efl_ctx* ctx);
int efl_sync(efl_ctx* ctx);
/* Per-stage device timing (hipEvents) into efl_stats; off by default. */
int efl_set_timing(efl_ctx* ctx, int enabled);
/* No counterpart in the reference (it records wgpu command buffers each
 * frame).  When enabled, efl_step records a step's launches once as HIP
 * graphs (per instance list, dt and launch-relevant settings) and replays
 * them; results are identical to eager launches.  Off by default: at
 * 1024^2 eager launches measured about 2.5 % faster than replay (the
 * runtime leaves idle gaps inside replayed graphs).  Steps with timing,
 * diagnostics, a multigrid residual target or a learned projection always
 * run eagerly. */
int efl_set_graphs(efl_ctx* ctx, int enabled);
/* No counterpart in the reference.  By default (margin 0) an MGPCG or learned
 * step reads the solve's stopping state back to the host (one wait per queued
 * batch of iterations), so efl_step returns once the solve has finished on
 * the device and the iteration count is exactly what the tolerance needs.
 * With margin > 0 an MGPCG step never waits for its own solve: it queues the
 * largest of the last four known iteration counts plus `margin` (at most
 * max_cycles) and returns.  Iterations after the solve stopped return at once
 * on the device, so when the budget covers the solve the fields are the
 * synchronous ones bit for bit; a solve that needs more stops at the budget,
 * unconverged (efl_stats.unconverged_solves counts it).  Each instance keeps
 * at most two such solves in flight: a step waits for the statistics of the
 * step before last if the device has not finished it (frames in flight), so
 * the counts it budgets from are at most two steps old.  Statistics resolve
 * lazily (efl_get_stats waits for them, efl_sync resolves all).  Learned and
 * residual-target multigrid projections, diagnostics (efl_set_diagnostics)
 * and timing (efl_set_timing) still read back; the repair of sealed regions
 * after a failed solve is skipped. */
int efl_set_async(efl_ctx* ctx, uint32_t margin);
/* No counterpart in the reference: the number of times the step machinery of
 * this context has waited for the device (tests of efl_set_async): reads of a
 * synchronous solve's outcome, diagnostics, timing, and the two-in-flight
 * limit of asynchronous steps. */
int efl_host_waits(const efl_ctx* ctx, uint64_t* out);

/* construct_rigid_body_buffer_for_gpu (obstacle.rs:89-160): the obstacle
 * list is global to the context, as SolidObstaclesBuffer is global to the
 * app.  n <= EFL_MAX_SOLIDS. */
int efl_set_obstacles(efl_ctx* ctx, const efl_obstacle* obstacles, uint32_t n);

/* No counterpart (replaces avian2d for the fluid's bodies): when bodies are
 * set, each efl_step ends with a semi-implicit Euler step of the dynamic
 * obstacles, on the device, driven by gravity and the fluid->solid forces of
 * the instances stepped in that call.  With m = density * area and
 * J = density * inertia (cell units), the acceleration in cells/s^2 is
 *   a = (F / dx^2 + m_a a_prev + m g / dx) / (m + m_a),  m_a = 4 rho area,
 *   alpha = (T_world / dx^2 + J_a alpha_prev) / (J + J_a),  J_a = 2 rho J,
 * with F the force bins (y flipped to the world frame) and T_world the
 * world-frame torque; the added-mass terms keep the explicit coupling of
 * light bodies stable.  Bodies are kept inside the wall ring (inelastic).
 * n must equal the obstacle count; n = 0 turns integration off. */
int efl_set_bodies(efl_ctx* ctx, const efl_body* bodies, uint32_t n);
/* Current obstacle records of the context's list (after integration). */
int efl_get_obstacles(efl_ctx* ctx, efl_obstacle* out, uint32_t n);

/* ---- instances --------------------------------------------------------- */

/* watch_fluid_component (setup_components.rs:38-330): allocates the grid
 * buffer set, the multigrid pyramid (multi_grid.rs:85-131) and the
 * redistancing buffers (reinitialize_levelset.rs:107-160). */
int efl_instance_create(efl_ctx* ctx, const efl_settings* settings, efl_instance** out);
int efl_instance_destroy(efl_instance* inst);

/* No counterpart: give this instance its own obstacle list (and, with
 * bodies != NULL, dynamics), used instead of the context's.  This is what lets
 * instances of one batched efl_step carry different bodies.  n = 0 and
 * obstacles = NULL return the instance to the context's list. */
int efl_set_instance_obstacles(efl_instance* inst, const efl_obstacle* obstacles, const efl_body* bodies, uint32_t n);
/* Current obstacle records this instance sees (its own list or the context's). */
int efl_get_instance_obstacles(efl_instance* inst, efl_obstacle* out, uint32_t n);

/* Replace the projection / reinitialisation config (ProjectionMethod and
 * ReinitializeMethod are mutable components). */
int efl_set_projection(efl_instance* inst, const efl_projection* projection);
int efl_set_reinit(efl_instance* inst, const efl_reinit* reinit);
/* FluidSettings.gravity, rho and the fluid Transform may change per frame
 * (update_simulation_uniform, fluid_uniform.rs:108-121). */
int efl_set_uniforms(efl_instance* inst, float rho, const float gravity[2], const float transform[16]);

/* update_fluid_source_init_buffer / update_fluid_source_buffer
 * (fluid_source_uniform.rs:77-158): oneshot = 1 sets the sources applied at
 * initialisation (FluidSourceOneshot), oneshot = 0 the per-step sources.
 * n <= EFL_MAX_FLUID_SOURCES, else EFL_E_RANGE. */
int efl_set_sources(efl_instance* inst, const efl_source* sources, uint32_t n, int oneshot);

/* apply_forces_and_clear (apply_forces.rs:157-167): forces used by the next
 * step only, then cleared. */
int efl_set_forces(efl_instance* inst, const efl_force* forces, uint32_t n);

/* No counterpart (needed above 16 sources): a W x H level set that is
 * min-combined into levelset_air1 after the oneshot sources at the next
 * initialisation.  NULL clears it. */
int efl_init_levelset(efl_instance* inst, const float* levelset);

/* FluidStatus::Reset (fluid_status.rs:15-37): the next efl_step runs the
 * initialisation dispatch instead of a step. */
int efl_reset(efl_instance* inst);

/* The Uninitialized branch of EulerFluidNode::run (render_node.rs:219-277),
 * issued now. */
int efl_initialize(efl_instance* inst);

/* One fluid step for each of n instances (render_node.rs:205-217 loops over
 * fluid entities; here they are batched).  Uninitialised instances are
 * initialised instead, as the reference does.  dt = FluidTimeStep.  Which
 * projections make it wait for the device: see the conventions above and
 * efl_set_async. */
int efl_step(efl_ctx* ctx, efl_instance* const* insts, uint32_t n, float dt);

/* Run one stage of the step on one instance (parity tests, profiling). */
int efl_run_stage(efl_instance* inst, int stage, float dt);
/* Run a stage reps times back to back on the context stream, bracketed by
 * hipEvents; *ms = mean device time per repetition.  GS_SWEEP and MG_SMOOTH
 * time their tiled smoother launch alone: reps launches back to back after
 * one full run of the stage (stencil setup and the copy of x back to its
 * buffer stay outside the events); MG_BOTTOM likewise times the bottom
 * launch alone.  CG_SMOOTH starts an MGPCG solve of the current system from
 * a zero guess, runs the first half of its first iteration and then times
 * the level-0 launch that finishes it (as every iteration issues it) reps
 * times; the pressure is restored afterwards and the next step solves as
 * usual. */
int efl_time_stage(efl_instance* inst, int stage, float dt, uint32_t reps, float* ms);

/* ---- results ------------------------------------------------------------ */

/* Element count (floats or ints) and row width of a field. */
int efl_field_info(const efl_instance* inst, int field, size_t* count, uint32_t* width, uint32_t* height, uint32_t* channels);
/* Copy a field to / from host memory (count elements, 4 bytes each). */
int efl_get_field(efl_instance* inst, int field, void* host, size_t count);
int efl_set_field(efl_instance* inst, int field, const void* host, size_t count);
/* Device pointer of a field (valid until the instance is destroyed). */
int efl_field_device_ptr(efl_instance* inst, int field, void** dptr);
/* forces_to_solid readback (setup_components.rs:293-294): out[EFL_MAX_SOLIDS]. */
int efl_get_solid_forces(efl_instance* inst, efl_solid_force* out, uint32_t n);
/* When enabled, each step records ||r||/||b|| after the projection and
 * dt*max|D| after the velocity update (before extrapolation) into efl_stats;
 * each costs one reduction pass.  Off by default. */
int efl_set_diagnostics(efl_instance* inst, int enabled);
/* Measure ||r||/||b|| and dt*max|D| of the current fields (forces a sync). */
int efl_measure(efl_instance* inst, float dt, float* rel_residual, float* max_norm_divergence);
int efl_get_stats(efl_instance* inst, efl_stats* out);
/* The reference's diagnostics passes (diagnostics/plugin.rs:20-104; its node
 * runs them on u0 / v0 / levelset_air before each frame's step) on the
 * current fields: the same 16 x 16 blocks over the floor(W/16) x floor(H/16)
 * cover and the same second stage, where each of 512 lanes adds its block
 * values l, l + 512, ... before the lanes are folded, so the speeds are a
 * max / min of sums when there are more than 512 blocks and min_velocity is 0
 * when there are fewer.  Synchronous (one device round trip). */
int efl_get_fluid_diagnostics(efl_instance* inst, efl_fluid_diagnostics* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same diagnostics as efl_get_fluid_diagnostics, over
 * caller-owned device fields levelset_air (width x height), u ((width + 1) x
 * height) and v (width x (height + 1)), written as {volume, min_velocity,
 * max_velocity} to the caller's device buffer out (3 floats).  Enqueued on
 * ctx's stream and asynchronous: read out after efl_sync.  Allocates nothing:
 * the per-block partials go to a scratch array the library holds on each GPU
 * (EFL_NOALLOC_MAX_BLOCKS 16 x 16 blocks, 768 KiB; grids up to 4096 x 4096),
 * larger grids return EFL_E_RANGE.  Calls on one GPU share that array, so
 * calls that overlap on one GPU (two contexts' streams) race. */
#define EFL_NOALLOC_MAX_BLOCKS 65536
int efl_fluid_diagnostics_noalloc(efl_ctx* ctx, const float* levelset_air, const float* u, const float* v,
                                  uint32_t width, uint32_t height, float* out);

/* ---- autotuner (no counterpart in the reference) -----------------------
 *
 * efl_autotune times this instance's pressure solve on its own current
 * system for each candidate and keeps the fastest: multigrid sweep counts
 * (pre, post, coarsest) from a fixed list, the size of the bottom solve
 * (levels of at most 1024 or 4096 cells in one workgroup), the tile height of
 * the smoother launches (64 x 16 or 64 x 8 cells per workgroup, or by level
 * size: 64 x 8 below 2^20 cells, the default; level 0 of an MGPCG
 * preconditioner stays at 64 x 16) and, for a learned projection, the
 * convolution strip width of each network level.  It runs
 * the first stages of a step (solids to divergence) to form a fresh system,
 * restores the initial guess before every trial, and leaves velocities,
 * level set and pressure as they were; every candidate meets the same
 * rel_tol, so the answer's accuracy does not change.  A candidate replaces the
 * current choice only if it is at least 3 % faster.  Only projections with a
 * residual target are tuned (MGPCG, MULTIGRID with rel_tol > 0, LEARNED);
 * others return EFL_E_STATE.  With apply = 0 the settings are only reported. */
typedef struct {
  uint32_t pre, post, coarsest;   /* chosen sweep counts */
  uint32_t bottom_cells;          /* 1024 or 4096 */
  uint32_t nn_strips;             /* bit l: 64-pixel strips on network level l (else 16) */
  uint32_t tile_h;                /* smoother tile height: 16, 8 or 0 (by level size) */
  uint32_t trials;                /* timed solves */
  float best_ms;                  /* projection time of the chosen settings */
  float default_ms;               /* projection time of the settings on entry */
} efl_tuning;
int efl_autotune(efl_instance* inst, float dt, int apply, efl_tuning* result);
/* Apply settings that efl_autotune reported (a host can keep them and skip
 * the trials next time): pre, post, coarsest (1..8 each), bottom_cells (1024,
 * or 4096 where the device grants the bottom solve's LDS), tile_h (16, 8 or 0)
 * and, for an instance with a network, nn_strips; trials and the times are
 * ignored.  Waits for the instance's queued work first.  Values out of range
 * return EFL_E_INVALID and change nothing. */
int efl_apply_tuning(efl_instance* inst, const efl_tuning* tuning);

/* ---- learned preconditioner (no counterpart in the reference) -----------
 *
 * A 4-level convolutional U-Net (3x3 convolutions, 16/32/64/64 channels,
 * BF16 weights and activations, FP32 accumulation on MFMA) maps the CG
 * residual r (divided by sigma = rms(r)), the solid fraction
 * 1 - mean(area fractions) and clamp(phi_air / 8, -1, 1) to an
 * approximation of A^-1 r.  EFL_PROJECTION_LEARNED runs flexible CG with it
 * (beta = -(z . q_prev) / (p . q)_prev); when the relative residual stalls
 * (less than a factor 0.7 over 8 iterations) or max_cycles pass, the solve
 * continues as MGPCG from the current iterate.  Training runs on the device
 * with the library's own forward and backward kernels and Adam, on samples
 * (rhs, solution) that the plain solver produces.  Parameters are FP32
 * master weights, layer by layer: W[cout][cin][3][3] then bias[cout]. */
typedef struct efl_nn efl_nn;
size_t efl_nn_param_count(void);
double efl_nn_flops_per_pixel(void);   /* forward FLOPs per grid cell */
int efl_nn_create(efl_ctx* ctx, efl_nn** out);
int efl_nn_destroy(efl_nn* nn);
/* He-normal initialisation from a seed (SplitMix64). */
int efl_nn_init_params(efl_nn* nn, uint64_t seed);
int efl_nn_set_params(efl_nn* nn, const float* params, size_t n);
int efl_nn_get_params(efl_nn* nn, float* params, size_t n);
/* Training set of `capacity` samples of width x height (multiples of 8). */
int efl_nn_dataset_reserve(efl_nn* nn, uint32_t width, uint32_t height, uint32_t capacity);
/* Append a sample: rhs and target are W x H host arrays (the right-hand side
 * of a pressure solve and its solution, or a residual and the error it
 * leaves); the features come from the instance's current area fractions,
 * level set and pressure-unknown mask (set by its last projection), whose
 * size must match the dataset's. */
int efl_nn_add_sample(efl_nn* nn, efl_instance* inst, const float* rhs, const float* target);
int efl_nn_sample_count(const efl_nn* nn, uint32_t* count);
/* Copy sample `index` to the host: 5 planes of W x H floats (rhs / rms(rhs),
 * solid fraction, clamped level set, mask, target / rms(target): flexible CG
 * fixes every step length itself, so only the shape of A^-1 r is learned). */
int efl_nn_get_sample(const efl_nn* nn, uint32_t index, float* planes);
/* The loss efl_nn_train and efl_nn_eval minimise and report.
 * EFL_NN_LOSS_MSE (the default): mean over the batch's pixels of
 * mask (output - target)^2.  EFL_NN_LOSS_RESIDUAL: mean over the batch's
 * samples of 1 - (r . A y)^2 / (|r|^2 |A y|^2), y the output and A the
 * sample's own pressure operator (its stencil is stored with the sample):
 * the squared relative residual that the best multiple of y leaves, which is
 * what one flexible-CG step with the network achieves; targets are unused.
 * Other values return EFL_E_INVALID. */
enum { EFL_NN_LOSS_MSE = 0, EFL_NN_LOSS_RESIDUAL = 1 };
int efl_nn_set_loss(efl_nn* nn, int loss);
/* iterations Adam steps on random batches of `batch` samples (seeded);
 * loss_history (may be NULL) receives each step's loss. */
int efl_nn_train(efl_nn* nn, uint32_t iterations, uint32_t batch, float lr, uint64_t seed, float* loss_history);
/* Loss of samples [first, first + count) without updating; grads (may be
 * NULL, efl_nn_param_count() floats) receives d loss / d params, out (may
 * be NULL, count * W * H floats) the network outputs. */
int efl_nn_eval(efl_nn* nn, uint32_t first, uint32_t count, float* loss, float* grads, float* out);
/* Use nn for this instance's EFL_PROJECTION_LEARNED solves (NULL detaches). */
int efl_set_learned(efl_instance* inst, efl_nn* nn);
/* Iterations of the last learned solve: [0] flexible-CG iterations with the
 * network, [1] MGPCG iterations of the fallback (0 if it was not needed);
 * learned_rel (may be NULL): ||r|| / ||b|| when the network's part ended. */
int efl_nn_last_solve(efl_instance* inst, uint32_t out[2], float* learned_rel);

#ifdef __cplusplus
}
#endif

#endif /* EFL_H */
