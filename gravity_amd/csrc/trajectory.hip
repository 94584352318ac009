// The trajectory preview (S/trajectory.comp, K/GravityPipeline.kt:514-598): a
// test particle launched from a point, integrated by kick-drift-kick with a
// step that keeps its segments near targetSegmentPixels, through the last
// force build (mesh force and near field of the physical cells) and the
// current bodies; each step emits up to three pieces (split where the path
// crosses a torus seam) as pairs of vertices vec4(x, y, dash coordinate px,
// path distance), the distances normalised by the path length at the end.
//
// The path is a sequential chain of steps, so one wave runs it: every lane
// carries the same particle state and performs the same scalar updates
// (identical operations give identical values), the lanes share each step's
// near field — lane i evaluates stencil cell i's terms (one per monopole or
// exact body) into LDS at offsets from a wave prefix sum — and lane 0 adds the
// terms in stencil order, as the shader's loop does, and hands the sum to the
// others. Lane step % 64 stores the step's vertices and normalises them.
//
// Float expressions follow the shader's operation for operation (the library
// is built without contraction; sqrt and the divisions round correctly, exp
// is the device library's FP32 exp).
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

constexpr int TRAJ_NT = 64;
constexpr int TRAJ_VERTICES_PER_STEP = 6;

// S/trajectory.comp:77-98: localAcceleration (the test particle is no body)
__device__ __forceinline__ float2 traj_local(const SimDesc& s, float sx, float sy, float tx,
                                             float ty, float mass) {
    float dx = sx - tx, dy = sy - ty;
    if (!s.bounded) {
        dx -= floorf(dx + 0.5f);
        dy -= floorf(dy + 0.5f);
    }
    const float px = dx * s.aspect, py = dy * 1.0f;
    const float r2 = px * px + py * py;
    const float lw = expf(-r2 / s.twoSigma2);
    const float scale = s.G * mass * lw / (s.totalMass * fmaxf(r2, s.minDist2));
    return make_float2(scale * px / s.aspect, scale * py / 1.0f);
}

__device__ __forceinline__ float traj_mix(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// S/trajectory.comp:100-122: interpolateForce
__device__ __forceinline__ float2 traj_mesh_force(const SimDesc& s, const float2* __restrict__ force,
                                                  float x, float y) {
    const float gx = x * (float)s.pw - 0.5f, gy = y * (float)s.ph - 0.5f;
    const float flx = floorf(gx), fly = floorf(gy);
    const int bx = (int)flx, by = (int)fly;
    const float fx = gx - flx, fy = gy - fly;
    const float2* F = force + s.meshBase;
    const int x0 = wrap_axis(bx, s.mw), x1 = wrap_axis(bx + 1, s.mw);
    const int y0 = wrap_axis(by, s.mh), y1 = wrap_axis(by + 1, s.mh);
    const float2 v00 = F[x0 + y0 * s.mw], v10 = F[x1 + y0 * s.mw];
    const float2 v01 = F[x0 + y1 * s.mw], v11 = F[x1 + y1 * s.mw];
    return make_float2(traj_mix(traj_mix(v00.x, v10.x, fx), traj_mix(v01.x, v11.x, fx), fy),
                       traj_mix(traj_mix(v00.y, v10.y, fx), traj_mix(v01.y, v11.y, fx), fy));
}

__device__ __forceinline__ int wave_exclusive_int(int v, int& total) {
    const int lane = (int)(threadIdx.x & 63u);
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(inc, (unsigned)o, 64);
        if (lane >= o) inc += t;
    }
    total = __shfl(inc, 63, 64);
    return inc - v;
}

// S/trajectory.comp:124-168: accelerationAt, the wave together.
__device__ float2 traj_acceleration(const SimDesc& s, const TrajectoryInputs& in, float x, float y,
                                    float2* terms) {
    if (!s.mutual) return make_float2(s.fieldX, s.fieldY);
    const float2 mf = traj_mesh_force(s, in.force, x, y);
    float ax = mf.x + s.fieldX, ay = mf.y + s.fieldY;
    const int lane = (int)threadIdx.x;
    const int hx = (int)floorf(x * (float)s.pw), hy = (int)floorf(y * (float)s.ph);
    const int DX = 2 * s.rx + 1, cells = DX * (2 * s.ry + 1);
    for (int base = 0; base < cells; base += TRAJ_NT) {
        const int i = base + lane;
        int nterm = 0;
        bool mono = false;
        uint4 rec = make_uint4(0u, 0u, 0u, 0u);
        int wx = 0, wy = 0;
        if (i < cells) {
            const int dy = i / DX - s.ry, dx = i % DX - s.rx;
            const int cx = hx + dx, cy = hy + dy;
            if (!s.bounded || (cx >= 0 && cy >= 0 && cx < s.pw && cy < s.ph)) {
                wx = wrap_axis(cx, s.pw);
                wy = wrap_axis(cy, s.ph);
                rec = in.cellRec[s.physBase + wx + wy * s.pw];
                const int count = (int)rec.y;
                const bool home = dx == 0 && dy == 0;
                if (!home && count > s.exactLimit) {
                    mono = true;
                    nterm = 1;
                } else {
                    nterm = min(count, home ? s.homeLimit : s.exactLimit);
                }
            }
        }
        int total;
        const int at = wave_exclusive_int(nterm, total);
        // (the host keeps a chunk's terms within TRAJ_TERMS; the bound here
        // only keeps a write inside the array)
        if (mono) {
            const float cxn = ((float)wx + 0.5f) / (float)s.pw;
            const float cyn = ((float)wy + 0.5f) / (float)s.ph;
            if (at < TRAJ_TERMS) terms[at] = traj_local(s, cxn, cyn, x, y, (float)rec.z / 256.0f);
        } else {
            for (int k = 0; k < nterm && at + k < TRAJ_TERMS; k++) {
                const uint32_t id = __float_as_uint(in.cellBody[rec.x + (uint32_t)k].w);
                const float4 o = in.body[id];
                terms[at + k] = traj_local(s, o.x, o.y, x, y, in.mats[id].x);
            }
        }
        __syncthreads();
        if (lane == 0) {
            for (int k = 0; k < min(total, TRAJ_TERMS); k++) {
                const float2 t = terms[k];
                ax += t.x;
                ay += t.y;
            }
        }
        __syncthreads();
    }
    return make_float2(__shfl(ax, 0, 64), __shfl(ay, 0, 64));
}

// S/trajectory.comp:170-186: integrationDt
__device__ __forceinline__ float traj_dt(const SimDesc& s, const TrajectoryInputs& in, float vx,
                                         float vy, float ax, float ay) {
    const float sx = vx * s.vw, sy = vy * s.vh, gx = ax * s.vw, gy = ay * s.vh;
    const float speed = sqrtf(sx * sx + sy * sy);
    const float accel = sqrtf(gx * gx + gy * gy);
    float r;
    if (accel > 1e-5f)
        r = (sqrtf(speed * speed + 2.0f * accel * in.segmentPx) - speed) / accel;
    else
        r = in.segmentPx / fmaxf(speed, 1e-5f);
    return clampf(r, s.dt, in.maximumDt);
}

// S/trajectory.comp:188-259: emitPeriodicStep, stored by the step's lane.
__device__ __forceinline__ void traj_emit(float4* __restrict__ out, int step, bool store, float sx,
                                          float sy, float dx, float dy, float startDistance,
                                          float stepDistance, float startDash, float stepDash) {
    float4* v = out + step * TRAJ_VERTICES_PER_STEP;
    if (store) {
        for (int k = 0; k < TRAJ_VERTICES_PER_STEP; k++) v[k] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    }
    const float ex = sx + dx, ey = sy + dy;
    float cuts[4];
    int cutCount = 2;
    cuts[0] = 0.0f;
    cuts[1] = 1.0f;
    if (ex < 0.0f) cuts[cutCount++] = -sx / dx;
    else if (ex >= 1.0f) cuts[cutCount++] = (1.0f - sx) / dx;
    if (ey < 0.0f) cuts[cutCount++] = -sy / dy;
    else if (ey >= 1.0f) cuts[cutCount++] = (1.0f - sy) / dy;
    for (int index = 1; index < cutCount; index++) {
        const float value = clampf(cuts[index], 0.0f, 1.0f);
        int insertion = index;
        while (insertion > 0 && cuts[insertion - 1] > value) {
            cuts[insertion] = cuts[insertion - 1];
            insertion--;
        }
        cuts[insertion] = value;
    }
    float unique[4];
    int uniqueCount = 0;
    for (int index = 0; index < cutCount; index++)
        if (uniqueCount == 0 || fabsf(cuts[index] - unique[uniqueCount - 1]) > 1e-6f)
            unique[uniqueCount++] = cuts[index];
    if (!store) return;
    for (int piece = 0; piece < uniqueCount - 1; piece++) {
        const float f1 = unique[piece], f2 = unique[piece + 1];
        const float mid = 0.5f * (f1 + f2);
        const float tx = floorf(sx + dx * mid), ty = floorf(sy + dy * mid);
        v[2 * piece] = make_float4(clampf(sx + dx * f1 - tx, 0.0f, 1.0f),
                                   clampf(sy + dy * f1 - ty, 0.0f, 1.0f),
                                   startDash + stepDash * f1, startDistance + stepDistance * f1);
        v[2 * piece + 1] = make_float4(clampf(sx + dx * f2 - tx, 0.0f, 1.0f),
                                       clampf(sy + dy * f2 - ty, 0.0f, 1.0f),
                                       startDash + stepDash * f2, startDistance + stepDistance * f2);
    }
}

// S/trajectory.comp:261-329
__global__ void __launch_bounds__(TRAJ_NT) k_trajectory(const SimDesc* __restrict__ sims,
                                                        TrajectoryInputs in,
                                                        float4* __restrict__ out) {
    __shared__ float2 terms[TRAJ_TERMS];
    const SimDesc& s = sims[0];
    const int lane = (int)threadIdx.x;
    float x = in.originX / s.vw, y = in.originY / s.vh;
    float vx = in.velocityX / s.vw, vy = in.velocityY / s.vh;
    float2 a = traj_acceleration(s, in, x, y, terms);
    float path = 0.0f, dash = 0.0f;
    bool finished = false;
    for (int step = 0; step < in.steps; step++) {
        const bool mine = step % TRAJ_NT == lane;
        if (finished) {
            if (mine) {
#pragma unroll
                for (int k = 0; k < TRAJ_VERTICES_PER_STEP; k++)
                    out[step * TRAJ_VERTICES_PER_STEP + k] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
            }
            continue;
        }
        const float dt = traj_dt(s, in, vx, vy, a.x, a.y);
        float hvx = vx + a.x * (0.5f * dt), hvy = vy + a.y * (0.5f * dt);
        float dx = hvx * dt, dy = hvy * dt;
        if (s.bounded) {
            // Walls act like body contacts: the normal motion stops, and
            // rebounds by the restitution coefficient (mix(v, -e v, hit)).
            const float nx = x + dx, ny = y + dy;
            const float cx = clampf(nx, 0.0f, MAXIMUM_UNIT), cy = clampf(ny, 0.0f, MAXIMUM_UNIT);
            const float hitX = nx != cx ? 1.0f : 0.0f, hitY = ny != cy ? 1.0f : 0.0f;
            hvx = hvx * (1.0f - hitX) + (-s.restitution * hvx) * hitX;
            hvy = hvy * (1.0f - hitY) + (-s.restitution * hvy) * hitY;
            dx = cx - x;
            dy = cy - y;
        }
        const float sdx = dx * s.vw, sdy = dy * s.vh;
        float stepDistance = sqrtf(sdx * sdx + sdy * sdy);
        const float remaining = in.maxPathPx - path;
        if (stepDistance > remaining) {
            const float k = remaining / stepDistance;
            dx *= k;
            dy *= k;
            stepDistance = remaining;
            finished = true;
        }
        const float speedMultiplier = 1.0f + sqrtf(hvx * hvx + hvy * hvy);
        const float dashLength = in.minimumDashPx * speedMultiplier;
        const float stepDash = stepDistance * in.minimumDashPx / dashLength;
        traj_emit(out, step, mine, x, y, dx, dy, path, stepDistance, dash, stepDash);
        path += stepDistance;
        dash += stepDash;
        if (s.bounded) {
            x = x + dx;
            y = y + dy;
        } else {
            x = x + dx - floorf(x + dx);
            y = y + dy - floorf(y + dy);
        }
        if (path >= in.maxPathPx - 1e-4f) {
            finished = true;
            continue;
        }
        a = traj_acceleration(s, in, x, y, terms);
        vx = hvx + a.x * (0.5f * dt);
        vy = hvy + a.y * (0.5f * dt);
    }
    // every lane reads back only the vertices it stored
    const float inverse = 1.0f / fmaxf(path, 1e-6f);
    for (int step = lane; step < in.steps; step += TRAJ_NT) {
        for (int k = 0; k < TRAJ_VERTICES_PER_STEP; k++) {
            float4 v = out[step * TRAJ_VERTICES_PER_STEP + k];
            if (v.w >= 0.0f) {
                v.w = clampf(v.w * inverse, 0.0f, 1.0f);
                out[step * TRAJ_VERTICES_PER_STEP + k] = v;
            }
        }
    }
}

void launch_trajectory(hipStream_t st, const SimDesc* sims, const TrajectoryInputs& in, float4* out) {
    launch(k_trajectory, dim3(1), dim3(TRAJ_NT), 0, st, sims, in, out);
}

}  // namespace gs
