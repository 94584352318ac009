// Host-side members of GravityPipeline that the C ABI exports next to the
// step: projectile placement (K/CircularProjectile.kt:10-60,
// K/GravityPipeline.kt:180, 621-646, 720-779), the cluster diagnostic
// (K/GravityPipeline.kt:793-905), the walls capability check (:1628) and the
// render view of the state buffers (bindBodiesForRender, :509-512).
//
// Float expressions follow the Kotlin ones operation for operation (the
// library is built without contraction, as the JVM evaluates them).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"

namespace gs {

constexpr float MAXIMUM_UNIT_F = 0.99999994f;
constexpr int MAXIMUM_FFT_AXIS = 2048;

// CircularProjectile.offsets: hexagonally packed centres inside the disc a
// projectile of the given diameter leaves for body centres.
void projectile_offsets(float contactDiameter, float projectileDiameter, std::vector<float>& xy) {
    if (!(contactDiameter > 0.0f)) throw ArgumentError("Failed requirement.");
    if (!(projectileDiameter >= contactDiameter)) throw ArgumentError("Failed requirement.");
    xy.clear();
    const float bodyRadius = 0.5f * contactDiameter;
    const float availableRadius = 0.5f * projectileDiameter - bodyRadius;
    const float rowStep = contactDiameter * (float)(std::sqrt(3.0) * 0.5);
    int row = 0;
    float y = -availableRadius;
    while (y <= availableRadius) {
        const float rowOffset = (row & 1) == 0 ? 0.0f : 0.5f * contactDiameter;
        float x = -availableRadius + rowOffset;
        while (x <= availableRadius) {
            if (x * x + y * y <= availableRadius * availableRadius) {
                xy.push_back(x);
                xy.push_back(y);
            }
            x += contactDiameter;
        }
        y += rowStep;
        row++;
    }
}

// CircularProjectile.diameterForBodyCount: the smallest diameter on a
// 0.05 px grid whose packing holds `bodies` centres.
float projectile_diameter_for(float contactDiameter, int bodies) {
    if (!(contactDiameter > 0.0f) || bodies <= 0) throw ArgumentError("Failed requirement.");
    constexpr float STEP = 0.05f;
    const double d = (double)contactDiameter;
    const double bodyArea = d * d * std::sqrt(3.0) * 0.5;
    const double estimate = 2.0 * std::sqrt((double)bodies * bodyArea / M_PI) + d;
    float diameter = (float)std::max(d, estimate - 2.0 * d);
    std::vector<float> xy;
    auto count = [&](float D) {
        projectile_offsets(contactDiameter, D, xy);
        return (int)(xy.size() / 2);
    };
    while (diameter > contactDiameter && count(diameter) >= bodies)
        diameter = std::max(contactDiameter, diameter - contactDiameter);
    while (count(diameter) < bodies) diameter += STEP;
    return diameter;
}

// ToroidalSpace.wrapPixels
static float wrap_pixels(float value, int extent) {
    const float size = (float)extent;
    return value - std::floor(value / size) * size;
}

int Engine::spawn_projectile(int s, int firstBody, float centerX, float centerY, float velocityX,
                             float velocityY, float diameter) {
    HostSim& h = sims_.at(s);
    std::vector<float> offsets;
    projectile_offsets(h.cfg.contact_diameter_pixels, diameter, offsets);
    const int count = (int)(offsets.size() / 2);
    // fits (K/GravityPipeline.kt:720-722): a full capacity spawns nothing
    if (!(firstBody >= 0 && firstBody <= h.cap && count <= h.cap - firstBody)) return 0;
    float farthest = 0.0f;
    for (int k = 0; k < count; k++)
        farthest = std::max(farthest, (float)std::hypot((double)offsets[2 * k], (double)offsets[2 * k + 1]));
    const float radius = farthest + 0.5f * h.cfg.contact_diameter_pixels;
    const float W = (float)h.cfg.viewport_width, H = (float)h.cfg.viewport_height;
    auto coerce = [](float v, float lo, float hi) {
        if (lo > hi) throw ArgumentError("Cannot coerce value to an empty range");
        return v < lo ? lo : (v > hi ? hi : v);
    };
    const float cx = h.bounded ? coerce(centerX, radius, W - radius) : centerX;
    const float cy = h.bounded ? coerce(centerY, radius, H - radius) : centerY;
    std::vector<float> bodies((size_t)4 * count), mats((size_t)4 * count);
    for (int k = 0; k < count; k++) {
        const float px = cx + offsets[2 * k], py = cy + offsets[2 * k + 1];
        const float x = h.bounded ? coerce(px, 0.0f, W * MAXIMUM_UNIT_F) : wrap_pixels(px, h.cfg.viewport_width);
        const float y = h.bounded ? coerce(py, 0.0f, H * MAXIMUM_UNIT_F) : wrap_pixels(py, h.cfg.viewport_height);
        bodies[4 * k] = x / (float)h.cfg.viewport_width;
        bodies[4 * k + 1] = y / (float)h.cfg.viewport_height;
        bodies[4 * k + 2] = velocityX / (float)h.cfg.viewport_width;
        bodies[4 * k + 3] = velocityY / (float)h.cfg.viewport_height;
        // loose bodies: unit mass, the configured friction, not solid, no group
        mats[4 * k] = 1.0f;
        mats[4 * k + 1] = h.cfg.friction;
        mats[4 * k + 2] = 0.0f;
        mats[4 * k + 3] = 0.0f;
    }
    upload(s, firstBody, count, bodies.data(), mats.data());
    return count;
}

int Engine::projectile_body_count(int s, float diameter) const {
    std::vector<float> offsets;
    projectile_offsets(sims_.at(s).cfg.contact_diameter_pixels, diameter, offsets);
    return (int)(offsets.size() / 2);
}

gs_cluster_statistics Engine::measure_cluster(int s) {
    const HostSim& h = sims_.at(s);
    const int n = h.n;
    if (n < 1) throw ArgumentError("measureCluster needs at least one active body");
    std::vector<float> raw((size_t)4 * n);
    download(s, GS_BODIES_XYVV, 0, n, raw.data());
    const double W = (double)h.cfg.viewport_width, H = (double)h.cfg.viewport_height;
    const double TAU = 2.0 * M_PI;
    double cxCos = 0.0, cxSin = 0.0, cyCos = 0.0, cySin = 0.0, maxSpeed = 0.0, totalSpeed = 0.0;
    int nonFinite = 0;
    std::vector<double> xs((size_t)n), ys((size_t)n);
    for (int i = 0; i < n; i++) {
        const float rx = raw[4 * i], ry = raw[4 * i + 1], rvx = raw[4 * i + 2], rvy = raw[4 * i + 3];
        const bool finite = std::isfinite(rx) && std::isfinite(ry) && std::isfinite(rvx) && std::isfinite(rvy);
        if (!finite) nonFinite++;
        const double x = finite ? (double)rx : 0.0, y = finite ? (double)ry : 0.0;
        xs[i] = x;
        ys[i] = y;
        cxCos += std::cos(TAU * x);
        cxSin += std::sin(TAU * x);
        cyCos += std::cos(TAU * y);
        cySin += std::sin(TAU * y);
        const double speed = finite ? std::hypot((double)rvx * W, (double)rvy * H) : 0.0;
        maxSpeed = std::max(maxSpeed, speed);
        totalSpeed += speed;
    }
    auto wrapUnit = [](double v) { return v - std::floor(v); };
    auto minimumImage = [](double d) { return d - std::floor(d + 0.5); };
    const double ccx = wrapUnit(std::atan2(cxSin, cxCos) / TAU);
    const double ccy = wrapUnit(std::atan2(cySin, cyCos) / TAU);
    double meanX = 0.0, meanY = 0.0;
    for (int i = 0; i < n; i++) {
        xs[i] = (ccx + minimumImage(xs[i] - ccx)) * W;
        ys[i] = (ccy + minimumImage(ys[i] - ccy)) * H;
        meanX += xs[i];
        meanY += ys[i];
    }
    meanX /= n;
    meanY /= n;
    double xx = 0.0, xy = 0.0, yy = 0.0;
    for (int i = 0; i < n; i++) {
        const double x = xs[i] - meanX, y = ys[i] - meanY;
        xx += x * x;
        xy += x * y;
        yy += y * y;
    }
    xx /= n;
    xy /= n;
    yy /= n;
    const double area = 4.0 * M_PI * std::sqrt(std::max(xx * yy - xy * xy, 0.0));
    std::vector<double> sx = xs, sy = ys;
    std::sort(sx.begin(), sx.end());
    std::sort(sy.begin(), sy.end());
    const double medianX = sx[(size_t)n / 2], medianY = sy[(size_t)n / 2];
    std::vector<double> r2((size_t)n);
    for (int i = 0; i < n; i++) {
        const double x = xs[i] - medianX, y = ys[i] - medianY;
        r2[i] = x * x + y * y;
    }
    std::sort(r2.begin(), r2.end());
    auto coreArea = [&](double fraction) {
        int index = std::min(std::max((int)(fraction * n), 1), n) - 1;
        return M_PI * r2[(size_t)index] / fraction;
    };
    gs_cluster_statistics out{};
    out.center_x = (double)wrap_pixels((float)medianX, h.cfg.viewport_width);
    out.center_y = (double)wrap_pixels((float)medianY, h.cfg.viewport_height);
    out.covariance_area_pixels = area;
    out.equivalent_diameter_pixels = 2.0 * std::sqrt(area / M_PI);
    out.core90_area_pixels = coreArea(0.90);
    out.core99_area_pixels = coreArea(0.99);
    out.max_speed_pixels_per_second = maxSpeed;
    out.average_speed_pixels_per_second = totalSpeed / n;
    out.non_finite_bodies = nonFinite;
    return out;
}

gs_adaptive_statistics Engine::measure_adaptive(int s) {
    sims_.at(s);
    uint32_t st[ADAPT_WORDS];
    check(hipStreamSynchronize(st_), "adaptive state");
    check(hipMemcpy(st, adapt_ + (size_t)ADAPT_WORDS * s, sizeof(st), hipMemcpyDeviceToHost),
          "adaptive state");
    gs_adaptive_statistics out{};
    std::memcpy(&out.maximum_speed_pixels_per_second, &st[0], 4);
    std::memcpy(&out.maximum_penetration_pixels, &st[1], 4);
    out.substeps = (int)st[2];
    out.maximum_contact_count = (int)st[3];
    out.dense_contact_bodies = (int)st[4];
    return out;
}

// K/GravityPipeline.kt:514-598: the uniforms of S/trajectory.comp, the
// origin placed as placeX / placeY place it (:773-779).
void Engine::update_trajectory(float originX, float originY, float velocityX, float velocityY) {
    normalize();
    const HostSim& h = sims_[0];
    const SimDesc& d = h.desc;
    if (d.mutual) {
        const int cells = (2 * d.rx + 1) * (2 * d.ry + 1);
        const long worst = (long)d.homeLimit + (long)(std::min(cells, 64) - 1) * d.exactLimit;
        if (worst > TRAJ_TERMS)
            throw ArgumentError("the trajectory preview holds 2048 near-field terms per 64 stencil cells");
    }
    const float W = (float)h.cfg.viewport_width, H = (float)h.cfg.viewport_height;
    auto coerce = [](float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); };
    TrajectoryInputs in{};
    in.force = spec_;
    in.cellRec = cellRec_;
    in.cellBody = cellBody_;
    in.body = body_[bodyPar_];
    in.mats = mat_[auxPar_];
    in.steps = h.cfg.trajectory_steps;
    in.maximumDt = h.cfg.trajectory_maximum_time_step;
    in.segmentPx = h.cfg.trajectory_target_segment_pixels;
    in.maxPathPx = h.cfg.trajectory_maximum_path_diagonals * std::sqrt(W * W + H * H);
    in.minimumDashPx = h.cfg.trajectory_minimum_dash_pixels;
    in.originX = h.bounded ? coerce(originX, 0.0f, W * MAXIMUM_UNIT_F) : wrap_pixels(originX, h.cfg.viewport_width);
    in.originY = h.bounded ? coerce(originY, 0.0f, H * MAXIMUM_UNIT_F) : wrap_pixels(originY, h.cfg.viewport_height);
    in.velocityX = velocityX;
    in.velocityY = velocityY;
    launch_trajectory(st_, dSims_, in, trajectory_);
    check(hipGetLastError(), "trajectory");
}

void Engine::trajectory_view(void** vertices, int* count) {
    if (vertices) *vertices = trajectory_;
    if (count) *count = trajectoryVertices_;
}

void Engine::download_trajectory(int first, int count, float* dst) {
    if (first < 0 || count < 0 || first > trajectoryVertices_ || count > trajectoryVertices_ - first)
        throw ArgumentError("trajectory range exceeds its vertices");
    if (count == 0) return;
    check(hipStreamSynchronize(st_), "trajectory");
    check(hipMemcpy(dst, trajectory_ + first, (size_t)count * sizeof(float4), hipMemcpyDeviceToHost),
          "trajectory");
}

bool supports_walls(int gridWidth, int gridHeight) {
    return 2 * std::max(gridWidth, gridHeight) <= MAXIMUM_FFT_AXIS;
}

void Engine::render_view(int s, void** bodies, void** pressures) {
    const HostSim& h = sims_.at(s);
    if (bodies) *bodies = body_[bodyPar_] + h.base;
    if (pressures) *pressures = press_ + h.base;
}

void Engine::render_copy(int s, int count, void* bodies, void* pressures) {
    const HostSim& h = sims_.at(s);
    if (count < 0 || count > h.cap) throw ArgumentError("render count exceeds capacity");
    if (count == 0) return;
    if (bodies)
        check(hipMemcpyAsync(bodies, body_[bodyPar_] + h.base, (size_t)count * sizeof(float4),
                             hipMemcpyDeviceToDevice, st_), "render bodies");
    if (pressures)
        check(hipMemcpyAsync(pressures, press_ + h.base, (size_t)count * sizeof(float),
                             hipMemcpyDeviceToDevice, st_), "render pressures");
}

}  // namespace gs
