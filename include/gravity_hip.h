/*
 * gravity_hip.h: C ABI of the MI355X particle step of Gravity.
 *
 * The drop-in boundary for the reference's OpenGL compute path: each entry
 * point names the Kotlin member of dev.gravity.GravityPipeline (or
 * GravityConfig) it replaces, as file:line in the reference
 * (K/ = src/main/kotlin/dev/gravity/). A JNI shim maps them one to one onto a
 * Kotlin facade with the same members.
 *
 * Conventions (K/GravityPipeline.kt:73-81 and the Kotlin require/check):
 *   - 0 is success; GS_EINVAL mirrors IllegalArgumentException (a bad
 *     argument), GS_ESTATE mirrors IllegalStateException (a GPU or state
 *     failure). gs_last_error() holds the message of the last failure on the
 *     calling thread.
 *   - Bodies are in the reference's units: position / viewport and
 *     velocity / viewport per axis; xyvv arrays are the reference's
 *     vec4(x, y, vx, vy) records, materials vec4(mass, friction, solid, group).
 *   - A handle owns every device buffer and one HIP stream; calls on one
 *     handle must not overlap (the reference's single GL thread).
 *   - Everything is asynchronous on the handle's stream except gs_download,
 *     gs_synchronize and the diagnostics, which block.
 */
#ifndef GRAVITY_HIP_H
#define GRAVITY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_OK 0
#define GS_EINVAL (-1)
#define GS_ESTATE (-2)

/* Flat copy of the GravityConfig fields the step reads
 * (K/GravityConfig.kt:610-860) plus the viewport and start state. */
typedef struct gs_config {
    /* viewport and start (K/GravityPipeline.kt:73-81, StartConfig) */
    int viewport_width;
    int viewport_height;
    int bounded;                        /* walls; false = torus */
    int mutual_gravity;
    int liquid;
    /* particles (ParticleConfig, K/GravityConfig.kt:610-626) */
    int capacity;
    int spatial_sort_interval;
    /* gravity (GravityFieldConfig, K/GravityConfig.kt:636-671) */
    int grid_width;
    int grid_height;
    float gravitational_constant;
    float softening_cells;
    float local_search_sigma_multiplier;
    int minimum_neighbor_radius_cells;
    int maximum_neighbor_radius_cells;
    int exact_neighbor_body_limit;
    int maximum_home_body_interactions;
    float minimum_distance_cell_fraction;
    float minimum_assignment_window_power;
    /* collisions (CollisionConfig, K/GravityConfig.kt:673-738) */
    float time_step;
    int maximum_substeps;
    float contact_diameter_pixels;
    int neighbor_cell_radius;
    float restitution_coefficient;
    float friction;
    float coincident_distance_squared_epsilon;
    float neighbor_skin_pixels;
    int sleep;
    float sleep_speed_pixels;
    /* liquid (LiquidConfig, K/GravityConfig.kt:838-855) */
    float kernel_radius_diameters;
    int liquid_iterations;
    float liquid_viscosity;
    /* uniform field in px/s^2 (GravityPipeline.fieldAccelerationX/Y, :108-109) */
    float field_acceleration_x;
    float field_acceleration_y;
    /* Extension without a reference counterpart: 0 replaces the contact
     * passes by plain drifts (the gravity-only benchmark configurations). */
    int collisions_enabled;
    /* adaptive collisions (CollisionConfig.adaptive and its fractions,
     * K/GravityConfig.kt:686-705): the device picks each step's substep
     * count (K/GravityPipeline.kt:1155-1211) */
    int adaptive_collisions;
    float allowed_displacement_fraction;
    float allowed_penetration_fraction;
    int dense_contact_threshold;
    float dense_body_fraction;
    int minimum_dense_bodies;
    /* the trajectory preview (TrajectoryConfig, K/GravityConfig.kt:739-758) */
    int trajectory_steps;
    float trajectory_maximum_time_step;
    float trajectory_target_segment_pixels;
    float trajectory_maximum_path_diagonals;
    float trajectory_minimum_dash_pixels;
    /* Coulomb friction of glued blocks against walls
     * (ProjectileConfig.blockFriction, K/GravityConfig.kt:773) */
    float block_friction;
} gs_config;

typedef struct gs_sim gs_sim;

/* GravityConfig() defaults (K/GravityConfig.kt:16-26, 610-855) on a
 * 3440 x 1440 torus with mutual gravity, as the reference's benchmarks run. */
void gs_config_defaults(gs_config *cfg);

/* GravityPipeline(config, viewportWidth, viewportHeight, profiler,
 * scenarioState, bounded) (K/GravityPipeline.kt:73-81, init :338-364).
 * scenario_xyvv (nullable) holds n_scenario normalised bodies; the rest of
 * the capacity is zero. Accelerations start at zero, as after init. */
int gs_create(const gs_config *cfg, const float *scenario_xyvv, int n_scenario, gs_sim **out);

/* GravityPipeline.close() (K/GravityPipeline.kt:1556). */
void gs_destroy(gs_sim *sim);

/* Property setters (K/GravityPipeline.kt:97-121): BOUNDED throws when the
 * FFT axis would exceed 2048 (supportsWalls, :1628). */
#define GS_BOUNDED 1
#define GS_MUTUAL_GRAVITY 2
#define GS_LIQUID 3
int gs_set_bool(gs_sim *sim, int key, int value);

/* fieldAccelerationX / fieldAccelerationY (K/GravityPipeline.kt:108-109). */
int gs_set_field(gs_sim *sim, float ax_px, float ay_px);

/* collisionSubsteps, coerced to 1..maximum (K/GravityPipeline.kt:147-151). */
int gs_set_substeps(gs_sim *sim, int substeps);

/* initializeAcceleration(activeBodies) (K/GravityPipeline.kt:366-370). */
int gs_initialize_acceleration(gs_sim *sim, int active);

/* step(activeBodies) (K/GravityPipeline.kt:445-494): one physics step. */
int gs_step(gs_sim *sim, int active);

/* Host upload of bodies first..first+count (uploadBodies,
 * K/GravityPipeline.kt:733-784; the test hook CollisionGpuAccess.upload,
 * T/CollisionGpuSupport.kt:51-66). materials may be NULL. */
int gs_upload(gs_sim *sim, int first, int count, const float *xyvv, const float *materials);

/* Host upload of accelerations (vec2 per body, normalised units). */
int gs_upload_accel(gs_sim *sim, int first, int count, const float *accel);

/* Blocking readback (the buffers the tests map, T/CollisionGpuSupport.kt:113-133). */
#define GS_BODIES_XYVV 0     /* float[4 * count] */
#define GS_PRESSURES 1       /* float[count] */
#define GS_ACCEL 2           /* float[2 * count] */
#define GS_MATERIALS 3       /* float[4 * count] */
#define GS_CONTACT_COUNTS 4  /* uint32[count] */
#define GS_SORT_DEST 5       /* uint32[count], new index of each body after the last sort */
#define GS_FORCE_MESH 6      /* float[2 * mesh nodes]: (Fx, Fy) per node the force gather reads, after a force build (under walls the padding rows and columns hold stale values) */
#define GS_CONTACTS 7        /* uint32[16 * count]: the 16 smallest contact indices, ascending, body-major */
int gs_download(gs_sim *sim, int what, int first, int count, void *dst);

/* stopBodies(activeBodies) (K/GravityPipeline.kt:612, S/stop_bodies.comp). */
int gs_stop_bodies(gs_sim *sim, int active);

/* wake() (K/GravityPipeline.kt:610). */
int gs_wake(gs_sim *sim);

/* projectileBodyCount(diameterPixels) (K/GravityPipeline.kt:180): the
 * bodies a projectile of that diameter holds, in *count. */
int gs_projectile_body_count(gs_sim *sim, float diameter_px, int *count);

/* spawnProjectile(firstBody, centerX, centerY, velocityX, velocityY,
 * diameterPixels) (K/GravityPipeline.kt:621-646): a hexagonally packed disc
 * of loose unit-mass bodies written at firstBody (pixels and px/s; under
 * walls the centre is kept a radius inside them). *spawned receives the
 * bodies written, 0 when they do not fit the capacity (fits, :720-722). */
int gs_spawn_projectile(gs_sim *sim, int first_body, float center_x, float center_y,
                        float velocity_x, float velocity_y, float diameter_px, int *spawned);

/* CircularProjectile.offsets / diameterForBodyCount (K/CircularProjectile.kt:
 * 10-60), host only (no handle, no GPU): the packing's centre offsets in px
 * (2 floats each; written when capacity allows, *count always) and the
 * smallest diameter on a 0.05 px grid holding `bodies` centres. */
int gs_projectile_offsets(float contact_diameter_px, float projectile_diameter_px, float *xy,
                          int capacity, int *count);
int gs_projectile_diameter_for_body_count(float contact_diameter_px, int bodies, float *diameter_px);

/* ClusterStatistics (K/GravityPipeline.kt:47-57) and measureCluster
 * (:793-905): a blocking readback of the active bodies and the reference's
 * physics-quality metric (circular mean centre, covariance area, core
 * areas, speeds, non-finite bodies), computed on the host in FP64. */
typedef struct gs_cluster_statistics {
    double center_x;
    double center_y;
    double covariance_area_pixels;
    double equivalent_diameter_pixels;
    double core90_area_pixels;
    double core99_area_pixels;
    double max_speed_pixels_per_second;
    double average_speed_pixels_per_second;
    int non_finite_bodies;
} gs_cluster_statistics;
int gs_measure_cluster(gs_sim *sim, int active, gs_cluster_statistics *out);

/* updateTrajectory(activeBodies, originX, originY, velocityX, velocityY)
 * (K/GravityPipeline.kt:514-598, S/trajectory.comp): the preview path of a
 * test particle from (origin_x, origin_y) px at (velocity_x, velocity_y) px/s
 * through the last force build and the current bodies of simulation 0, into
 * the handle's vertex buffer: trajectory_steps x 6 vertices
 * vec4(x, y, dash coordinate px, path fraction), unused ones (0, 0, 0, -1).
 * Asynchronous on the handle's stream. GS_EINVAL when the near-field limits
 * would put more than 2048 terms in one 64-cell chunk of the stencil. */
int gs_update_trajectory(gs_sim *sim, int active, float origin_x, float origin_y, float velocity_x,
                         float velocity_y);
/* bindTrajectoryForRender (:600-602) and trajectoryVertexCount (:242-243):
 * the device address of the vertex buffer and its vertex count (both
 * nullable), for a renderer that reads HIP memory or maps it through
 * HIP-GL interop. */
int gs_trajectory_view(gs_sim *sim, void **vertices, int *vertex_count);
/* Copies vertices [first, first + count) to host memory (4 floats each). Blocks. */
int gs_download_trajectory(gs_sim *sim, int first, int count, float *dst);

/* measureAdaptiveCollisions (K/GravityPipeline.kt:425-445,
 * AdaptiveCollisionStatistics :65-71): the adaptive state of simulation
 * `index` (0 for a plain handle) after the last step: the largest speed the
 * step started with, the deepest pair penetration, the substeps it chose, the
 * most contacts of one body and the bodies at or above the dense contact
 * count, the last three from the step's final pass. Blocks. */
typedef struct gs_adaptive_statistics {
    float maximum_speed_pixels_per_second;
    float maximum_penetration_pixels;
    int substeps;
    int maximum_contact_count;
    int dense_contact_bodies;
} gs_adaptive_statistics;
int gs_measure_adaptive_collisions(gs_sim *sim, int index, gs_adaptive_statistics *out);

/* GravityPipeline.supportsWalls(gravity) (K/GravityPipeline.kt:1628): 1 when
 * the doubled mesh fits the 2048-value FFT line, else 0. Host only. */
int gs_supports_walls(int grid_width, int grid_height);

/* bindBodiesForRender (K/GravityPipeline.kt:509-512, S/particles.vert:3-4):
 * the renderer reads the bodies (binding 0, vec4 x, y, vx, vy normalised)
 * and the pressures (binding 1, float). gs_render_view gives the device
 * addresses of the current buffers (valid until the next call that steps,
 * uploads or sorts; either output nullable). gs_render_copy copies `count`
 * bodies and pressures, asynchronously on the handle's stream after the work
 * already queued, into caller device memory (for example a GL buffer mapped
 * through hipGraphicsGLRegisterBuffer); either destination nullable. */
int gs_render_view(gs_sim *sim, void **bodies, void **pressures);
int gs_render_copy(gs_sim *sim, int count, void *bodies, void *pressures);

/* Glued blocks (K/RigidBlocks.kt): bodies that move as one rigid body. A
 * block's bodies are uploaded first (gs_upload, at centre + offset, their
 * materials naming the block as GravityPipeline.uploadBodies does); then
 * gs_rigid_glue is RigidBlocks.add (K/RigidBlocks.kt:280-301, called by
 * glueBlock, K/GravityPipeline.kt:692-717): bodies first_body ..
 * first_body + count - 1 at offsets_xy (2 floats each, px) from
 * (center_x, center_y), glued as block *group. The block's region of the
 * shape values holds its hull only: the distance field (GlueShape) is read
 * by the RIGID contact passes alone, which are not built, so gs_step,
 * gs_debug_resolve_collisions, gs_debug_resolve_liquid and the adjoint fail
 * with GS_ESTATE once a block is glued. Single-simulation handles only.
 * gs_debug_sort moves the members with their bodies (RigidBlocks.remap,
 * :391-399). */
int gs_rigid_glue(gs_sim *sim, int first_body, int count, const float *offsets_xy, float center_x,
                  float center_y, int *group);
/* spawnBlock(firstBody, minX, minY, maxX, maxY, density, velocityX,
 * velocityY, spin) without a mask (K/GravityPipeline.kt:647-717): the
 * rectangle filled with a hexagonal packing (RectangularBlock), uploaded as
 * uploadBodies does (mass = density, the block friction, solid, the block's
 * index; velocity v + spin x offset) and glued (gs_rigid_glue). *spawned
 * receives the bodies written: 0 when none fit the rectangle, the capacity
 * or the block slots. */
int gs_spawn_block(gs_sim *sim, int first_body, float min_x, float min_y, float max_x, float max_y,
                   float density, float velocity_x, float velocity_y, float spin, int *spawned);
/* RectangularBlock.offsets / bodyCount (K/RectangularBlock.kt:10-77), host
 * only: up to `capacity` offsets (2 floats each, px, row by row) into xy
 * (nullable), the full count into *count. */
int gs_block_offsets(float contact_diameter_px, float width_px, float height_px, float *xy, int capacity,
                     int *count);
/* RigidBlocks.project(bodies, materials, bounded, contacts)
 * (K/RigidBlocks.kt:309-357; S/rigid_reduce.comp, S/rigid_apply.comp) for
 * blocks nothing breaks: every block fitted to its members (momentum and
 * angular momentum kept) and its members snapped onto the fit; contacts != 0
 * after a contact pass (settles the contact memory and stops the blocks
 * against the pushes it asked for). Queued on the handle's stream. */
int gs_rigid_project(gs_sim *sim, int contacts);
/* RigidBlocks.size (K/RigidBlocks.kt:172-174): out[0] group indices in use,
 * out[1] tiles in use, out[2] member slots. */
int gs_rigid_size(gs_sim *sim, int *out);
/* The block buffers, as T/RigidBlockIntegrationTest.kt:100-108 reads them:
 * 32-bit words from `first`, `count` of them. Writable: the groups (but
 * for their tile and hull words, 12-15 of each) and the contact memory,
 * which the RIGID contact passes write. Blocks. */
#define GS_RIGID_GROUPS 0          /* 40 words per group (RigidGroup) */
#define GS_RIGID_CONTACT_MEMORY 1  /* int impacts, next impacts, states, idle passes, pair contacts: 66816 each */
#define GS_RIGID_MEMBERS 2         /* int per member slot, -1 for none */
#define GS_RIGID_REST 3            /* 2 floats per member slot (px) */
#define GS_RIGID_SHAPES 4          /* float: each block's hull corners from tiles.z */
#define GS_RIGID_TILE_GROUPS 5     /* int per tile */
int gs_rigid_read(gs_sim *sim, int what, int first, int count, void *dst);
int gs_rigid_write(gs_sim *sim, int what, int first, int count, const void *src);

/* Test hooks mirroring the reflection hooks of T/CollisionGpuSupport.kt:68-90:
 * resolveCollisions(count, kickDt, driftDt, writePressure, substep) (one
 * pass; substep < 0 is the final pass), sortBodies(count), buildForce +
 * applyKick of initializeAcceleration. */
int gs_debug_resolve_collisions(gs_sim *sim, int count, float kick_dt, float drift_dt,
                                int write_pressure, int substep);
int gs_debug_sort(gs_sim *sim, int count);
int gs_debug_resolve_liquid(gs_sim *sim, int count);

/* Readback of the force build's first half at the current positions of
 * `count` bodies (the density buffer of S/build_grid.comp:49-76 and the
 * spectrum of S/fft_init.comp + the forward transform): density receives
 * the fixed-point CIC deposit (uint32 per mesh node, row-major; rows the
 * walls leave empty are 0), spectrum its FP32 forward 2D DFT (2 floats per
 * mesh node), cells the physical cell table (2 uint32 per cell: bodies,
 * mass in 1/256). Each output is nullable. Blocks. */
int gs_debug_density(gs_sim *sim, int count, uint32_t *density, float *spectrum, uint32_t *cells);

/* gs_debug_density without any device allocation in the call. No
 * counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The fixed-point density is staged in a static device array of the
 * library, 2048 x 1024 mesh nodes (8 MiB), one per GPU, held until the
 * process ends: a larger mesh fails with GS_EINVAL, and so does a count
 * other than the handle's active body count (changing it would rebuild the
 * handle's tables). Every call on a GPU shares that GPU's array and clears
 * it first; calls that overlap on one GPU race. Blocks. */
int gs_debug_density_noalloc(gs_sim *sim, int count, uint32_t *density, float *spectrum,
                             uint32_t *cells);

/* validateFft(activeBodies) (K/GravityPipeline.kt:378-419): the density
 * spectrum against an FP64 transform of the same fixed-point deposit;
 * out[0] maximum absolute error, out[1] that over the largest reference
 * value, out[2] RMS error. Works on walls too (the reference checks the
 * periodic mesh only). Blocks. */
int gs_validate_fft(gs_sim *sim, int active, double *out);

/* Waits for the handle's stream. */
int gs_synchronize(gs_sim *sim);

/* Capacity and mesh of a handle: out[0] capacity, out[1..2] mesh w, h,
 * out[3..4] physical grid w, h, out[5] contact slots. */
int gs_query(gs_sim *sim, int *out);

/* Per-kernel timing of `steps` eager steps with HIP events on the handle's
 * stream (the reference's GpuProfiler stages, K/GpuTimer.kt:17-40). */
typedef struct gs_kernel_stat {
    char name[40];
    int launches;
    double total_ms;
} gs_kernel_stat;
int gs_profile_steps(gs_sim *sim, int steps, gs_kernel_stat *out, int capacity, int *written);

/* Steps through one captured HIP graph per step (default on); 0 launches
 * every kernel eagerly. */
int gs_set_graphs(gs_sim *sim, int enabled);

/* Launch shapes of the step. No reference counterpart: the reference's
 * workgroup sizes are fixed in its shaders. No choice changes a result.
 *   big_segment_limit    a P3M cell or sort tile of more bodies is ordered by
 *                        one workgroup's LDS sort instead of per-lane ranks
 *                        (0: never, the default; else at least 32)
 *   forward_item_bodies  bodies per forward-FFT row item, on average (the
 *                        deposit's load balance; 64 .. 2^24, default 4096)
 *   deposit_runs         1: sum a lane's bodies of one cell in registers
 *                        before the LDS atomics of the deposit (default 0)
 *   scan_direct_blocks   scans of more 1024-value blocks scan their block
 *                        totals in a launch of their own (default 1024) */
typedef struct gs_tuning {
    int big_segment_limit;
    int forward_item_bodies;
    int deposit_runs;
    int scan_direct_blocks;
} gs_tuning;
int gs_get_tuning(gs_sim *sim, gs_tuning *out);
/* GS_EINVAL for a value out of range. Takes effect from the next step. */
int gs_set_tuning(gs_sim *sim, const gs_tuning *tuning);
/* Autotuner: times `steps` graph-replayed steps of the handle's own
 * workload for each candidate (coordinate descent over the fields above,
 * starting from the current choice; a candidate must be 1 % faster) and
 * keeps the fastest. The state is copied on the device first and put back
 * before every trial and at the end, so the simulations continue exactly
 * as if untuned (the copies need about 80 bytes per body slot of device
 * memory, more with sleeping tiles, while it runs). chosen (nullable)
 * receives the choice; ms (nullable, 2 doubles) its ms per step and the
 * starting choice's. Blocks. */
int gs_autotune(gs_sim *sim, int steps, gs_tuning *chosen, double *ms);

/* Batched ensembles. No reference counterpart: the reference runs one
 * simulation per GravityPipeline. A batch handle holds `count` independent
 * simulations, each with its own configuration, viewport, body count and
 * P3M grid (they share time_step, maximum_substeps and
 * spatial_sort_interval), and steps all of them with one launch sequence
 * (one captured HIP graph): every kernel walks a table of the batch's body
 * blocks, mesh lines or tiles. Simulation i holds scenarios[i] (nullable)
 * of n_scenario[i] bodies. gs_* calls without an index act on simulation 0;
 * gs_step / gs_initialize_acceleration set simulation 0's active count and
 * advance the whole batch. */
int gs_batch_create(const gs_config *cfgs, int count, const float *const *scenarios,
                    const int *n_scenario, gs_sim **out);
/* Number of simulations in the handle (a plain handle holds 1). */
int gs_batch_size(gs_sim *sim);
/* Active bodies of simulation `index` from the next step on. */
int gs_batch_set_active(gs_sim *sim, int index, int active);
int gs_batch_upload(gs_sim *sim, int index, int first, int count, const float *xyvv,
                    const float *materials);
int gs_batch_download(gs_sim *sim, int index, int what, int first, int count, void *dst);
/* initializeAcceleration and step of every simulation, with their active counts. */
int gs_batch_initialize_acceleration(gs_sim *sim);
int gs_batch_step(gs_sim *sim);

/* Adjoint of the gravity step. No reference counterpart: the reference has
 * no derivatives. For a single simulation without contact passes
 * (collisions_enabled = 0), with mutual gravity, torus or walls (a wall's
 * clamp is differentiated as the identity).
 *
 * gs_adjoint_record: `steps` steps exactly like gs_step from the current
 * state (its acceleration initialised), keeping the state every
 * checkpoint_every steps. gs_adjoint_vjp: the vector-Jacobian product of
 * those steps by reversed leapfrog, recomputing one checkpoint interval at a
 * time: given dL/dx_T (float[2 * n], normalised units, the final body
 * order) it writes dL/dx_0 and dL/dv_0 (float[2 * n] each, nullable), with
 * respect to the normalised positions and velocities of the recording's
 * start, in its body order; the handle keeps the final state. The spatial
 * sorts of the recording are undone on the way back. gs_adjoint_order:
 * where[i] = the index that now holds the body that started at index i.
 * count must equal the recording's active body count (GS_EINVAL otherwise),
 * and every array holds 2 * count floats (where: count entries); a VJP after
 * the active count has changed fails with GS_ESTATE. The first recording of
 * a shape (capacity, steps, interval) allocates its checkpoint buffers; later
 * recordings of the same shape reuse them. */
int gs_adjoint_record(gs_sim *sim, int active, int steps, int checkpoint_every);
int gs_adjoint_vjp(gs_sim *sim, int count, const float *dL_dx, float *dL_dx0, float *dL_dv0);
int gs_adjoint_order(gs_sim *sim, int count, uint32_t *where);
/* Record + loss L = mean |minimum image((x_T - target) * viewport)|^2 (px^2,
 * targets float[2 * n] normalised, in the start order) + VJP: writes L and
 * dL/dv_0 (float[2 * n], normalised units, start order). */
int gs_vjp_initial_velocity(gs_sim *sim, int active, int steps, int checkpoint_every,
                            const float *targets, float *loss, float *dL_dv0);

/* Test hook of the adjoint: out = J_A(x)^T w, the transposed Jacobian of the
 * accelerations (force build + kick, normalised units) at the current
 * positions of `count` bodies, applied to w (float[2 * count] each). */
int gs_debug_acceleration_vjp(gs_sim *sim, int count, const float *w, float *out);

/* Message of the last failure on this thread, or "" (Kotlin exception text). */
const char *gs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
