/*
 * CPU oracle for the Gravity particle step: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP32 restatement of the reference's compute shaders and of the
 * Kotlin orchestration around them, following their float operation order.
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it, and only as the checker: the product path never calls it.
 *
 * Layout follows the reference buffers: bodies are vec4(x/W, y/H, vx/W, vy/H),
 * accelerations vec2 (normalised units), materials vec4(mass, friction,
 * solid, group), pressures float.
 *
 * Where the reference is nondeterministic (atomic linked-list order), the
 * oracle fixes a canonical order; the HIP path uses the same one:
 *   - spatial sort: stable by body index within a tile;
 *   - collision contacts: accumulated in ascending index of the other body;
 *   - P3M near-field cell lists: visited in descending body index (the order
 *     a sequential insertion into the reference's LIFO lists produces).
 */
#ifndef GRAVITY_ORACLE_H
#define GRAVITY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int width, height;        /* viewport, px */
    int bounded;              /* walls instead of the torus */
    int mutual_gravity;
    int liquid;
    int collide;              /* 0: drift-only passes (gravity-only configs) */
    int substeps;
    float dt;
    float diameter;           /* contact diameter, px */
    float restitution;
    float coincident_eps;     /* squared px */
    int grid_w, grid_h;       /* physical P3M grid */
    float gravity;            /* G */
    float softening_cells;
    float near_sigma;         /* localSearchSigmaMultiplier */
    int near_min, near_max;
    int exact_limit;          /* exactNeighborBodyLimit */
    int home_limit;           /* maximumHomeBodyInteractions */
    float min_dist_frac;
    float window_floor;       /* minimumAssignmentWindowPower */
    float field_x, field_y;   /* px/s^2 */
    int sort_interval;
    /* liquid */
    float kernel_radius_diameters;
    int liquid_iterations;
    float viscosity;
    /* sleeping tiles (CollisionConfig.sleep, sleepSpeedPixels) */
    int sleep;
    float sleep_speed;        /* px/s */
} orc_params;

/* ---- a: spatial sort ---------------------------------------------------- */
uint32_t orc_spread_bits(uint32_t v);
uint32_t orc_sort_key(float x, float y, int width_bits, int height_bits);
int orc_sort_axis_bits(int extent);
/* Stable counting sort by Morton tile key; dest[i] = new index of body i. */
void orc_spatial_sort(int n, int width, int height, float *bodies, float *accel,
                      float *materials, uint32_t *dest);

/* ---- c: collisions ------------------------------------------------------ */
int orc_collision_grid_cells(int extent, float diameter);
float orc_pair_distance_squared(uint32_t id, const float *pos, uint32_t other,
                                const float *other_pos, float vw, float vh,
                                int bounded);
/* Exact contact sets; lists[i*cap + k] ascending, counts[i] complete. */
void orc_contacts(const orc_params *p, int n, const float *bodies,
                  const float *materials, uint32_t *counts, uint32_t *lists,
                  int cap);
/* Drift of one pass, in place (collision_cache_drift.comp). */
void orc_drift(const orc_params *p, int n, float *bodies, const float *accel,
               float kick_dt, float drift_dt);
/* Contact solve of one pass: in -> out (collision_cache_solve.comp). */
void orc_solve(const orc_params *p, int n, const float *in, float *out,
               const float *materials, const uint32_t *counts,
               const uint32_t *lists, int cap, float *pressures,
               int write_pressure);
/* drift + contacts + solve; bodies updated in place. */
void orc_collide_pass(const orc_params *p, int n, float *bodies,
                      const float *accel, const float *materials,
                      float *pressures, float kick_dt, float drift_dt,
                      int write_pressure, uint32_t *counts_out);

/* ---- b: P3M gravity ----------------------------------------------------- */
int orc_mass_scale(float total_mass);
void orc_twiddles(int max_axis, float *tw /* 4*(max_axis-1) */);
void orc_twiddles_f64(int max_axis, double *tw);
/* In-place 2D complex FFT of values[2*w*h], rows then columns, radix-2 DIT. */
void orc_fft2d(float *values, int w, int h, const float *tw, int inverse);
void orc_fft2d_f64(double *values, int w, int h, const double *tw, int inverse);
void orc_kernel_init(int mw, int mh, int pw, int ph, float aspect, float G,
                     float total_mass, float sigma_cells, float *kx, float *ky);
void orc_convolve(int mw, int mh, float window_floor, const float *rho,
                  const float *kx, const float *ky, float *fx, float *fy);
/* CIC deposit (fixed point) + near-field cell tables. */
void orc_deposit(int n, const float *bodies, const float *materials, int mw,
                 int mh, int pw, int ph, int mass_scale, uint32_t *density,
                 int32_t *cell_count, uint32_t *cell_mass);
/* Near-field radii from the config (GravityPipeline.neighborRadiusX/Y). */
void orc_neighbor_radius(const orc_params *p, int *rx, int *ry);

/* A whole simulation (GravityPipeline), for step-level parity. */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_create(const orc_params *p, int n, const float *bodies,
                        const float *materials);
void orc_sim_destroy(orc_sim *s);
void orc_sim_set_bodies(orc_sim *s, const float *bodies);
void orc_sim_get(const orc_sim *s, float *bodies, float *accel,
                 float *pressures, float *materials);
void orc_sim_set_accel(orc_sim *s, const float *accel);
void orc_sim_build_force(orc_sim *s);
/* Force mesh after build_force: real parts of the inverse transforms. */
void orc_sim_forces(const orc_sim *s, float *fx, float *fy);
void orc_sim_density_spectrum(const orc_sim *s, float *spectrum);
void orc_sim_kick(orc_sim *s, int initial);
/* The trajectory preview (S/trajectory.comp, K/GravityPipeline.kt:514-598)
 * of a test particle from (origin_x, origin_y) px at (vx, vy) px/s through the
 * last force build and the current bodies: steps * 6 vertices vec4(x, y, dash
 * coordinate px, path fraction) into out, unused ones (0, 0, 0, -1). */
typedef struct orc_trajectory_params {
    int steps;
    float maximum_dt;            /* TrajectoryConfig.maximumTimeStep */
    float segment_px;            /* targetSegmentPixels */
    float path_diagonals;        /* maximumPathViewportDiagonals */
    float minimum_dash_px;       /* minimumDashPixels */
} orc_trajectory_params;
void orc_sim_trajectory(const orc_sim *s, float origin_x, float origin_y, float vx, float vy,
                        const orc_trajectory_params *t, float *out);
void orc_sim_initialize_acceleration(orc_sim *s);
void orc_sim_step(orc_sim *s);
/* GravityPipeline.wake(): every body awake for the next two steps. */
void orc_sim_wake(orc_sim *s);
/* Step stamp of the sleeping tiles, and the body-awake stamps of the last
 * step (a body was awake in it when its stamp equals the stamp - 1). */
uint32_t orc_sim_sleep_stamp(const orc_sim *s);
void orc_sim_body_awake(const orc_sim *s, uint32_t *awake);
/* Sort at the next step (as the reference's step counter would). */
long orc_sim_steps(const orc_sim *s);
void orc_sim_last_sort_dest(const orc_sim *s, uint32_t *dest);

/* ---- adjoint reference: FP64 gravity-only leapfrog ----------------------- */
/* `steps` steps from bodies with initial velocities offset by dv (nullable),
 * in FP64 with the CIC deposit unrounded and no spatial sort; returns
 * mean |minimum image((x_T - target) * viewport)|^2 (px^2). */
double orc_grav64_loss(const orc_params *p, int n, const float *bodies,
                       const float *mats, int steps, const float *targets,
                       const double *dv, double *final_xy);

/* ---- glued blocks: the rigid projection ---------------------------------- */
/* GlueShape.hullIndices (K/GlueShape.kt:158-190): corner indices of the
 * convex hull of the points, counter-clockwise from the lowest x; returns
 * their number (at most n). */
int orc_rigid_hull(const float *xs, const float *ys, int n, int *corners);
/* RigidBlocks.prepare and upload (K/RigidBlocks.kt:751-857) of a block of n
 * members at offsets_xy (px): rest_xy the offsets about their centroid,
 * centroid_xy that centroid, *inertia the sum of |r0|^2 (summed in double),
 * bounds the rest-frame extent (min x, max x, min y, max y). */
void orc_rigid_prepare(const float *offsets_xy, int n, float *rest_xy, float *centroid_xy,
                       double *inertia, float *bounds);
/* RigidBlocks.project (K/RigidBlocks.kt:309-357; S/rigid_reduce.comp,
 * S/rigid_apply.comp) for blocks nothing breaks and that are awake. groups:
 * 40 words per group (S/rigid_layout.glsl RigidGroup); memory: the five
 * contact planes of 66816 ints; members, rest (2 floats) and the tile groups
 * per member slot and tile; bodies vec4 (normalised), updated in place. */
typedef struct orc_rigid_params {
    float width, height;      /* viewport, px */
    int bounded;
    int contacts;             /* after a contact pass */
    int group_count;          /* group indices in use */
    int tile_count;           /* tiles in use */
    float substep_dt;         /* s */
    float body_radius;        /* px */
    float block_friction;
} orc_rigid_params;
void orc_rigid_project(const orc_rigid_params *p, void *groups, int32_t *memory, const float *shapes,
                       const int32_t *members, const float *rest, const int32_t *tile_groups,
                       float *bodies);

/* ---- d: liquid ---------------------------------------------------------- */
void orc_liquid_constants(float diameter, float radius_diameters, float *h,
                          float *poly6, float *spiky, float *rest_density,
                          float *relaxation);
void orc_sim_resolve_liquid(orc_sim *s);

#ifdef __cplusplus
}
#endif
#endif
