// extern "C" boundary (include/gravity_hip.h): argument checks with the
// reference's messages, exceptions mapped to return codes, a thread-local
// error string.
#include <cmath>
#include <cstring>
#include <string>

#include "engine.hpp"

struct gs_sim {
    gs::Engine* engine;
};

namespace {
thread_local std::string g_error;

int fail(int code, const char* msg) {
    g_error = msg;
    return code;
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        g_error.clear();
        return GS_OK;
    } catch (const gs::ArgumentError& e) {
        return fail(GS_EINVAL, e.what());
    } catch (const gs::StateError& e) {
        return fail(GS_ESTATE, e.what());
    } catch (const std::exception& e) {
        return fail(GS_ESTATE, e.what());
    }
}
}  // namespace

extern "C" {

void gs_config_defaults(gs_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->viewport_width = 3440;
    c->viewport_height = 1440;
    c->bounded = 0;
    c->mutual_gravity = 1;
    c->liquid = 0;
    c->capacity = 400000;
    c->spatial_sort_interval = 2;
    c->grid_width = 512;
    c->grid_height = 256;
    c->gravitational_constant = 0.035f;
    c->softening_cells = 0.85f;
    c->local_search_sigma_multiplier = 3.0f;
    c->minimum_neighbor_radius_cells = 1;
    c->maximum_neighbor_radius_cells = 6;
    c->exact_neighbor_body_limit = 2;
    c->maximum_home_body_interactions = 8;
    c->minimum_distance_cell_fraction = 0.05f;
    c->minimum_assignment_window_power = 0.08f;
    c->time_step = 1.0f / 120.0f;
    c->maximum_substeps = 16;
    c->contact_diameter_pixels = 1.12838f;
    c->neighbor_cell_radius = 1;
    c->restitution_coefficient = 0.0f;
    c->friction = 0.0f;
    c->coincident_distance_squared_epsilon = 1e-12f;
    c->neighbor_skin_pixels = 0.5f;
    c->sleep = 1;
    c->sleep_speed_pixels = 6.0f;
    c->kernel_radius_diameters = 2.0f;
    c->liquid_iterations = 12;
    c->liquid_viscosity = 0.1f;
    c->field_acceleration_x = 0.0f;
    c->field_acceleration_y = 0.0f;
    c->collisions_enabled = 1;
    c->adaptive_collisions = 0;
    c->allowed_displacement_fraction = 0.25f;
    c->allowed_penetration_fraction = 0.125f;
    c->dense_contact_threshold = 4;
    c->dense_body_fraction = 0.01f;
    c->minimum_dense_bodies = 256;
    c->trajectory_steps = 1024;
    c->trajectory_maximum_time_step = 1.0f;
    c->trajectory_target_segment_pixels = 6.0f;
    c->trajectory_maximum_path_diagonals = 1.0f;
    c->trajectory_minimum_dash_pixels = 3.0f;
    c->block_friction = 0.5f;
}

int gs_create(const gs_config* cfg, const float* state, int n, gs_sim** out) {
    if (!cfg || !out) return fail(GS_EINVAL, "null argument");
    if (n < 0) return fail(GS_EINVAL, "negative scenario size");
    if (n > 0 && !state) return fail(GS_EINVAL, "scenario data is null");
    *out = nullptr;
    return guarded([&] {
        std::vector<gs_config> cfgs{*cfg};
        std::vector<const float*> states{state};
        std::vector<int> counts{n};
        auto* e = new gs::Engine(cfgs, states, counts);
        *out = new gs_sim{e};
    });
}

int gs_batch_create(const gs_config* cfgs, int count, const float* const* states,
                    const int* n, gs_sim** out) {
    if (!cfgs || !out) return fail(GS_EINVAL, "null argument");
    if (count < 1) return fail(GS_EINVAL, "a batch holds at least one simulation");
    *out = nullptr;
    std::vector<gs_config> c(cfgs, cfgs + count);
    std::vector<const float*> st((size_t)count, nullptr);
    std::vector<int> k((size_t)count, 0);
    for (int i = 0; i < count; i++) {
        k[i] = n ? n[i] : 0;
        st[i] = states ? states[i] : nullptr;
        if (k[i] < 0) return fail(GS_EINVAL, "negative scenario size");
        if (k[i] > 0 && !st[i]) return fail(GS_EINVAL, "scenario data is null");
    }
    return guarded([&] {
        auto* e = new gs::Engine(c, st, k);
        *out = new gs_sim{e};
    });
}

void gs_destroy(gs_sim* sim) {
    if (!sim) return;
    delete sim->engine;
    delete sim;
}

#define SIM_OR_FAIL                                                    \
    if (!sim || !sim->engine) return fail(GS_EINVAL, "null handle");   \
    gs::Engine& e = *sim->engine;

int gs_set_bool(gs_sim* sim, int key, int value) {
    SIM_OR_FAIL
    return guarded([&] {
        switch (key) {
            case GS_BOUNDED: e.set_bounded(0, value != 0); break;
            case GS_MUTUAL_GRAVITY: e.set_mutual(0, value != 0); break;
            case GS_LIQUID: e.set_liquid(0, value != 0); break;
            default: throw gs::ArgumentError("unknown property");
        }
    });
}

int gs_set_field(gs_sim* sim, float ax, float ay) {
    SIM_OR_FAIL
    return guarded([&] { e.set_field(0, ax, ay); });
}

int gs_set_substeps(gs_sim* sim, int substeps) {
    SIM_OR_FAIL
    return guarded([&] { e.set_substeps(substeps); });
}

int gs_initialize_acceleration(gs_sim* sim, int active) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, active);
        e.initialize_acceleration();
    });
}

int gs_step(gs_sim* sim, int active) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, active);
        e.step();
    });
}

int gs_upload(gs_sim* sim, int first, int count, const float* xyvv, const float* mats) {
    SIM_OR_FAIL
    if (count > 0 && !xyvv) return fail(GS_EINVAL, "null bodies");
    return guarded([&] { e.upload(0, first, count, xyvv, mats); });
}

int gs_upload_accel(gs_sim* sim, int first, int count, const float* accel) {
    SIM_OR_FAIL
    if (count > 0 && !accel) return fail(GS_EINVAL, "null accelerations");
    return guarded([&] { e.upload_accel(0, first, count, accel); });
}

int gs_download(gs_sim* sim, int what, int first, int count, void* dst) {
    SIM_OR_FAIL
    if (count > 0 && !dst) return fail(GS_EINVAL, "null destination");
    return guarded([&] { e.download(0, what, first, count, dst); });
}

int gs_stop_bodies(gs_sim* sim, int active) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, active);
        e.stop_bodies(0);
    });
}

#define INDEX_OR_FAIL                                                          \
    if (index < 0 || index >= e.sims()) return fail(GS_EINVAL, "no such simulation");

int gs_batch_size(gs_sim* sim) {
    SIM_OR_FAIL
    return e.sims();
}

int gs_batch_set_active(gs_sim* sim, int index, int active) {
    SIM_OR_FAIL
    INDEX_OR_FAIL
    return guarded([&] { e.set_active(index, active); });
}

int gs_batch_upload(gs_sim* sim, int index, int first, int count, const float* xyvv,
                    const float* mats) {
    SIM_OR_FAIL
    INDEX_OR_FAIL
    if (count > 0 && !xyvv) return fail(GS_EINVAL, "null bodies");
    return guarded([&] { e.upload(index, first, count, xyvv, mats); });
}

int gs_batch_download(gs_sim* sim, int index, int what, int first, int count, void* dst) {
    SIM_OR_FAIL
    INDEX_OR_FAIL
    if (count > 0 && !dst) return fail(GS_EINVAL, "null destination");
    return guarded([&] { e.download(index, what, first, count, dst); });
}

int gs_batch_initialize_acceleration(gs_sim* sim) {
    SIM_OR_FAIL
    return guarded([&] { e.initialize_acceleration(); });
}

int gs_batch_step(gs_sim* sim) {
    SIM_OR_FAIL
    return guarded([&] { e.step(); });
}

int gs_adjoint_record(gs_sim* sim, int active, int steps, int checkpoint_every) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, active);
        e.adjoint_record(steps, checkpoint_every);
    });
}

int gs_adjoint_vjp(gs_sim* sim, int count, const float* dL_dx, float* dL_dx0, float* dL_dv0) {
    SIM_OR_FAIL
    if (!dL_dx) return fail(GS_EINVAL, "null gradient");
    return guarded([&] { e.adjoint_vjp(count, dL_dx, dL_dx0, dL_dv0); });
}

int gs_adjoint_order(gs_sim* sim, int count, uint32_t* where) {
    SIM_OR_FAIL
    if (!where) return fail(GS_EINVAL, "null destination");
    return guarded([&] { e.adjoint_order(count, where); });
}

int gs_vjp_initial_velocity(gs_sim* sim, int active, int steps, int checkpoint_every,
                            const float* targets, float* loss, float* dL_dv0) {
    SIM_OR_FAIL
    if (!targets || !dL_dv0) return fail(GS_EINVAL, "null argument");
    return guarded([&] {
        e.set_active(0, active);
        e.adjoint_record(steps, checkpoint_every);
        const gs::HostSim& h = e.sim(0);
        const size_t n = (size_t)h.n;
        std::vector<float> xyvv(4 * n), grad(2 * n, 0.0f);
        std::vector<uint32_t> where(n);
        e.download(0, GS_BODIES_XYVV, 0, (int)n, xyvv.data());
        e.adjoint_order((int)n, where.data());
        // L = mean |minimum image((x_T - target) * viewport)|^2, px^2
        const double W = h.cfg.viewport_width, H = h.cfg.viewport_height;
        double sum = 0.0;
        for (size_t i = 0; i < n; i++) {
            size_t f = where[i];
            double dx = ((double)xyvv[4 * f] - targets[2 * i]) * W;
            double dy = ((double)xyvv[4 * f + 1] - targets[2 * i + 1]) * H;
            if (!h.bounded) {
                dx -= std::floor(dx / W + 0.5) * W;
                dy -= std::floor(dy / H + 0.5) * H;
            }
            sum += dx * dx + dy * dy;
            grad[2 * f] = (float)(2.0 * dx * W / (double)n);
            grad[2 * f + 1] = (float)(2.0 * dy * H / (double)n);
        }
        if (loss) *loss = n ? (float)(sum / (double)n) : 0.0f;
        e.adjoint_vjp((int)n, grad.data(), nullptr, dL_dv0);
    });
}

int gs_debug_acceleration_vjp(gs_sim* sim, int count, const float* w, float* out) {
    SIM_OR_FAIL
    if (!w || !out) return fail(GS_EINVAL, "null argument");
    return guarded([&] {
        e.set_active(0, count);
        e.debug_accel_vjp(w, out);
    });
}

int gs_wake(gs_sim* sim) {
    SIM_OR_FAIL
    return guarded([&] { e.wake(0); });
}

int gs_debug_resolve_collisions(gs_sim* sim, int count, float kick, float drift, int pressure,
                                int substep) {
    SIM_OR_FAIL
    (void)substep;
    return guarded([&] {
        e.set_active(0, count);
        e.debug_collide(kick, drift, pressure != 0);
    });
}

int gs_debug_sort(gs_sim* sim, int count) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, count);
        e.debug_sort();
    });
}

int gs_debug_resolve_liquid(gs_sim* sim, int count) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, count);
        e.debug_liquid();
    });
}

int gs_debug_density(gs_sim* sim, int count, uint32_t* density, float* spectrum, uint32_t* cells) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, count);
        e.debug_density(density, spectrum, cells);
    });
}

int gs_debug_density_noalloc(gs_sim* sim, int count, uint32_t* density, float* spectrum,
                             uint32_t* cells) {
    SIM_OR_FAIL
    return guarded([&] {
        // Changing the active count would rebuild the handle's tables.
        if (count != e.sim(0).n) throw gs::ArgumentError("count must equal the active body count");
        e.debug_density_noalloc(density, spectrum, cells);
    });
}

int gs_validate_fft(gs_sim* sim, int active, double* out) {
    SIM_OR_FAIL
    if (!out) return fail(GS_EINVAL, "null output");
    return guarded([&] {
        e.set_active(0, active);
        e.validate_fft(out);
    });
}

int gs_projectile_body_count(gs_sim* sim, float diameter, int* count) {
    SIM_OR_FAIL
    if (!count) return fail(GS_EINVAL, "null output");
    return guarded([&] { *count = e.projectile_body_count(0, diameter); });
}

int gs_spawn_projectile(gs_sim* sim, int first, float cx, float cy, float vx, float vy, float diameter,
                        int* spawned) {
    SIM_OR_FAIL
    if (!spawned) return fail(GS_EINVAL, "null output");
    *spawned = 0;
    return guarded([&] { *spawned = e.spawn_projectile(0, first, cx, cy, vx, vy, diameter); });
}

int gs_projectile_offsets(float contact, float diameter, float* xy, int capacity, int* count) {
    if (!count) return fail(GS_EINVAL, "null output");
    return guarded([&] {
        std::vector<float> v;
        gs::projectile_offsets(contact, diameter, v);
        *count = (int)(v.size() / 2);
        if (xy && capacity >= *count) std::memcpy(xy, v.data(), v.size() * sizeof(float));
    });
}

int gs_projectile_diameter_for_body_count(float contact, int bodies, float* diameter) {
    if (!diameter) return fail(GS_EINVAL, "null output");
    return guarded([&] { *diameter = gs::projectile_diameter_for(contact, bodies); });
}

int gs_measure_cluster(gs_sim* sim, int active, gs_cluster_statistics* out) {
    SIM_OR_FAIL
    if (!out) return fail(GS_EINVAL, "null output");
    return guarded([&] {
        e.set_active(0, active);
        *out = e.measure_cluster(0);
    });
}

int gs_supports_walls(int grid_width, int grid_height) {
    return gs::supports_walls(grid_width, grid_height) ? 1 : 0;
}

int gs_render_view(gs_sim* sim, void** bodies, void** pressures) {
    SIM_OR_FAIL
    return guarded([&] { e.render_view(0, bodies, pressures); });
}

int gs_render_copy(gs_sim* sim, int count, void* bodies, void* pressures) {
    SIM_OR_FAIL
    return guarded([&] { e.render_copy(0, count, bodies, pressures); });
}

int gs_synchronize(gs_sim* sim) {
    SIM_OR_FAIL
    return guarded([&] { e.synchronize(); });
}

int gs_query(gs_sim* sim, int* out) {
    SIM_OR_FAIL
    if (!out) return fail(GS_EINVAL, "null output");
    const gs::HostSim& h = e.sim(0);
    out[0] = h.cap;
    out[1] = h.desc.mw;
    out[2] = h.desc.mh;
    out[3] = h.desc.pw;
    out[4] = h.desc.ph;
    out[5] = gs::CONTACT_SLOTS;
    return GS_OK;
}

int gs_profile_steps(gs_sim* sim, int steps, gs_kernel_stat* out, int capacity, int* written) {
    SIM_OR_FAIL
    if (steps < 0 || (capacity > 0 && !out)) return fail(GS_EINVAL, "bad profile arguments");
    return guarded([&] {
        auto stats = e.profile(steps);
        int k = 0;
        for (auto& s : stats) {
            if (k >= capacity) break;
            std::memset(&out[k], 0, sizeof(gs_kernel_stat));
            std::strncpy(out[k].name, std::get<0>(s).c_str(), sizeof(out[k].name) - 1);
            out[k].launches = std::get<1>(s);
            out[k].total_ms = std::get<2>(s);
            k++;
        }
        if (written) *written = k;
    });
}

int gs_rigid_glue(gs_sim* sim, int first_body, int count, const float* offsets_xy, float center_x,
                  float center_y, int* group) {
    SIM_OR_FAIL
    if (count > 0 && !offsets_xy) return fail(GS_EINVAL, "null argument");
    return guarded([&] {
        const int g = e.rigid_glue(first_body, count, offsets_xy, center_x, center_y);
        if (group) *group = g;
    });
}

int gs_spawn_block(gs_sim* sim, int first_body, float min_x, float min_y, float max_x, float max_y,
                   float density, float velocity_x, float velocity_y, float spin, int* spawned) {
    SIM_OR_FAIL
    return guarded([&] {
        const int k = e.spawn_block(first_body, min_x, min_y, max_x, max_y, density, velocity_x, velocity_y, spin);
        if (spawned) *spawned = k;
    });
}

int gs_block_offsets(float contact_diameter_px, float width_px, float height_px, float* xy, int capacity,
                     int* count) {
    return guarded([&] {
        std::vector<float> offsets;
        gs::block_offsets(contact_diameter_px, width_px, height_px, offsets);
        const int n = (int)(offsets.size() / 2);
        if (xy)
            for (int i = 0; i < 2 * std::min(n, std::max(capacity, 0)); i++) xy[i] = offsets[i];
        if (count) *count = n;
    });
}

int gs_rigid_project(gs_sim* sim, int contacts) {
    SIM_OR_FAIL
    return guarded([&] { e.rigid_project(contacts != 0); });
}

int gs_rigid_size(gs_sim* sim, int* out) {
    SIM_OR_FAIL
    if (!out) return fail(GS_EINVAL, "null argument");
    return guarded([&] { e.rigid_size(out); });
}

int gs_rigid_read(gs_sim* sim, int what, int first, int count, void* dst) {
    SIM_OR_FAIL
    if (count > 0 && !dst) return fail(GS_EINVAL, "null argument");
    return guarded([&] { e.rigid_access(what, first, count, dst, false); });
}

int gs_rigid_write(gs_sim* sim, int what, int first, int count, const void* src) {
    SIM_OR_FAIL
    if (count > 0 && !src) return fail(GS_EINVAL, "null argument");
    return guarded([&] { e.rigid_access(what, first, count, const_cast<void*>(src), true); });
}

int gs_set_graphs(gs_sim* sim, int enabled) {
    SIM_OR_FAIL
    e.set_graphs(enabled != 0);
    return GS_OK;
}

int gs_update_trajectory(gs_sim* sim, int active, float origin_x, float origin_y, float velocity_x,
                         float velocity_y) {
    SIM_OR_FAIL
    return guarded([&] {
        e.set_active(0, active);
        e.update_trajectory(origin_x, origin_y, velocity_x, velocity_y);
    });
}

int gs_trajectory_view(gs_sim* sim, void** vertices, int* vertex_count) {
    SIM_OR_FAIL
    return guarded([&] { e.trajectory_view(vertices, vertex_count); });
}

int gs_download_trajectory(gs_sim* sim, int first, int count, float* dst) {
    SIM_OR_FAIL
    if (count > 0 && !dst) return fail(GS_EINVAL, "null argument");
    return guarded([&] { e.download_trajectory(first, count, dst); });
}

int gs_measure_adaptive_collisions(gs_sim* sim, int index, gs_adaptive_statistics* out) {
    SIM_OR_FAIL
    if (!out) return fail(GS_EINVAL, "null argument");
    if (index < 0 || index >= e.sims()) return fail(GS_EINVAL, "simulation index out of range");
    return guarded([&] { *out = e.measure_adaptive(index); });
}

static void tuning_out(const gs::Tuning& t, gs_tuning* out) {
    out->big_segment_limit = t.bigLimit;
    out->forward_item_bodies = t.forwardItemBodies;
    out->deposit_runs = t.depositRuns;
    out->scan_direct_blocks = t.scanDirectBlocks;
}

int gs_get_tuning(gs_sim* sim, gs_tuning* out) {
    SIM_OR_FAIL
    if (!out) return fail(GS_EINVAL, "null argument");
    tuning_out(e.tuning(), out);
    return GS_OK;
}

int gs_set_tuning(gs_sim* sim, const gs_tuning* tuning) {
    SIM_OR_FAIL
    if (!tuning) return fail(GS_EINVAL, "null argument");
    return guarded([&] {
        gs::Tuning t;
        t.bigLimit = tuning->big_segment_limit;
        t.forwardItemBodies = tuning->forward_item_bodies;
        t.depositRuns = tuning->deposit_runs;
        t.scanDirectBlocks = tuning->scan_direct_blocks;
        e.set_tuning(t);
    });
}

int gs_autotune(gs_sim* sim, int steps, gs_tuning* chosen, double* ms) {
    SIM_OR_FAIL
    if (steps < 1) return fail(GS_EINVAL, "autotune needs at least one timed step");
    return guarded([&] {
        gs::Tuning t = e.autotune(steps, ms);
        if (chosen) tuning_out(t, chosen);
    });
}

const char* gs_last_error(void) { return g_error.c_str(); }

}  // extern "C"
