// Hard-disc collision passes (S/collision_cache_{drift,count,solve}.comp,
// S/collision_material.glsl, S/collision_crowd.glsl), non-RIGID.
//
// One pass of the reference is drift -> broad phase -> count -> Jacobi solve.
// Here a pass is two launches:
//   count : clears the previous pass's grid cells, walks the 3x3 stencil of
//           this pass's grid, writes the contact count and the ascending list
//           of the CONTACT_SLOTS smallest contact indices;
//   solve : Jacobi correction in ascending contact order (a body with more
//           contacts continues past its slots by an ordered grid walk), walls,
//           integration, then, fused, the next pass's drift and its insertion
//           into the other grid buffer.
// The broad phase is a linked-cell grid of cells at least diameter + guard
// wide, rebuilt every pass; heads of a pass are cleared by the next pass's
// count kernel, so no clearing launch is needed. Two grid buffers alternate.
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

struct GridBuf {
    uint32_t* heads;   // per cell: first body + 1, 0 when empty
    float4* nodes;     // per body: x, y, next body + 1 (bits), unused
    int* cellOf;       // per body: the cell it was inserted in
};

__device__ __forceinline__ void grid_insert(const SimDesc& s, int g, float px, float py,
                                            const GridBuf& grid) {
    int cx = floor_index(px * (float)s.cgw, s.cgw);
    int cy = floor_index(py * (float)s.cgh, s.cgh);
    int c = s.cellBase + cx + cy * s.cgw;
    uint32_t old = atomicExch(&grid.heads[c], (uint32_t)(g + 1));
    grid.nodes[g] = make_float4(px, py, __uint_as_float(old), 0.0f);
    grid.cellOf[g] = c;
}

// S/collision_crowd.glsl:2-15: the lower index goes first; squares summed
// without contraction.
__device__ __forceinline__ float pair_distance_squared(int id, float px, float py, int other,
                                                       float ox, float oy, const SimDesc& s) {
    bool ownFirst = id < other;
    float ax = ownFirst ? px : ox, ay = ownFirst ? py : oy;
    float bx = ownFirst ? ox : px, by = ownFirst ? oy : py;
    float dx = ax * s.vw - bx * s.vw;
    float dy = ay * s.vh - by * s.vh;
    if (!s.bounded) {
        dx -= floorf(dx / s.vw + 0.5f) * s.vw;
        dy -= floorf(dy / s.vh + 0.5f) * s.vh;
    }
    return dx * dx + dy * dy;
}

// Stencil of one axis: 3 cells around c (wrapped on the torus, cut at walls),
// or every cell of a grid narrower than 3.
struct AxisRange { int first, count; };
__device__ __forceinline__ AxisRange axis_range(int c, int size) {
    if (size < 3) return {0, size};
    return {c - 1, 3};
}

template <typename F>
__device__ __forceinline__ void walk_stencil(const SimDesc& s, float px, float py,
                                             const GridBuf& grid, int guardLimit, F&& visit) {
    int cx = floor_index(px * (float)s.cgw, s.cgw);
    int cy = floor_index(py * (float)s.cgh, s.cgh);
    AxisRange ax = axis_range(cx, s.cgw), ay = axis_range(cy, s.cgh);
    int guard = 0;
    for (int j = 0; j < ay.count; j++) {
        int y = ay.first + j;
        if (s.bounded) { if (y < 0 || y >= s.cgh) continue; }
        else y = wrap_axis(y, s.cgh);
        for (int i = 0; i < ax.count; i++) {
            int x = ax.first + i;
            if (s.bounded) { if (x < 0 || x >= s.cgw) continue; }
            else x = wrap_axis(x, s.cgw);
            uint32_t e = grid.heads[s.cellBase + x + y * s.cgw];
            while (e != 0u && guard < guardLimit) {
                int other = (int)e - 1;
                float4 nd = grid.nodes[other];
                e = __float_as_uint(nd.z);
                guard++;
                visit(other, nd.x, nd.y);
            }
        }
    }
}

__device__ __forceinline__ bool is_hard(const SimDesc& s, const float4* mats, int a, int b) {
    if (!s.liquid) return true;
    return mats[a].z != 0.0f || mats[b].z != 0.0f;
}

// ---------------------------------------------------------------------------
// Drift (S/collision_cache_drift.comp:79-91) and insertion into the grid.
__global__ void __launch_bounds__(WG) k_drift_insert(const SimDesc* __restrict__ sims,
                                                     const BodyBlock* __restrict__ blocks,
                                                     float2* pos, float2* vel,
                                                     const float2* __restrict__ acc,
                                                     float kickDt, float driftDt, int insert,
                                                     GridBuf grid) {
    BodyBlock b = blocks[blockIdx.x];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float2 p = pos[g];
    if (kickDt != 0.0f || driftDt != 0.0f) {
        float2 v = vel[g];
        if (kickDt != 0.0f) {
            float2 a = acc[g];
            float h = 0.5f * kickDt;
            v.x += a.x * h;
            v.y += a.y * h;
            vel[g] = v;
        }
        p.x += v.x * driftDt;
        p.y += v.y * driftDt;
        if (s.bounded) {
            p.x = clampf(p.x, 0.0f, MAXIMUM_UNIT);
            p.y = clampf(p.y, 0.0f, MAXIMUM_UNIT);
        } else {
            p.x = fractf(p.x);
            p.y = fractf(p.y);
        }
        pos[g] = p;
    }
    if (insert) grid_insert(s, g, p.x, p.y, grid);
}

// ---------------------------------------------------------------------------
// Count (S/collision_cache_count.comp:61-95, 130-285 without RIGID): every
// hard pair with d^2 < D^2, strictly.
__global__ void __launch_bounds__(WG) k_count(const SimDesc* __restrict__ sims,
                                              const BodyBlock* __restrict__ blocks,
                                              GridBuf grid, GridBuf prev, int clearPrev,
                                              const float4* __restrict__ mats,
                                              uint32_t* __restrict__ counts,
                                              uint32_t* __restrict__ contacts, int planeStride,
                                              int guardLimit) {
    __shared__ uint32_t slots[CONTACT_SLOTS][WG];
    BodyBlock b = blocks[blockIdx.x];
    const SimDesc& s = sims[b.sim];
    int t = (int)threadIdx.x;
    int g = b.first + t;
    if (g - s.base >= s.n) return;
    if (clearPrev) prev.heads[prev.cellOf[g]] = 0u;
    float4 own = grid.nodes[g];
    float px = own.x, py = own.y;
    bool ownFluid = s.liquid && mats[g].z == 0.0f;
    uint32_t n = 0;
    walk_stencil(s, px, py, grid, guardLimit, [&](int other, float ox, float oy) {
        if (other == g) return;
        if (ownFluid && mats[other].z == 0.0f) return;
        float d2 = pair_distance_squared(g, px, py, other, ox, oy, s);
        if (!(d2 < s.diameter2)) return;
        // Keep the CONTACT_SLOTS smallest indices, ascending.
        uint32_t u = (uint32_t)other;
        int k = n < (uint32_t)CONTACT_SLOTS ? (int)n : CONTACT_SLOTS;
        if (k == CONTACT_SLOTS && u > slots[CONTACT_SLOTS - 1][t]) { n++; return; }
        int at = k == CONTACT_SLOTS ? CONTACT_SLOTS - 1 : k;
        while (at > 0 && slots[at - 1][t] > u) {
            slots[at][t] = slots[at - 1][t];
            at--;
        }
        slots[at][t] = u;
        n++;
    });
    counts[g] = n;
    int k = n < (uint32_t)CONTACT_SLOTS ? (int)n : CONTACT_SLOTS;
    for (int i = 0; i < k; i++) contacts[(size_t)i * planeStride + g] = slots[i][t];
}

// ---------------------------------------------------------------------------
// Solve (S/collision_cache_solve.comp:88-252, 274-478 without RIGID).

struct Accum {
    float pcx, pcy, vcx, vcy, pressure;
};

// S/collision_material.glsl:128-135
__device__ __forceinline__ void friction_correction(float rvx, float rvy, float nx, float ny,
                                                    float rvn, float normalChange,
                                                    float slidingShare, float friction,
                                                    float& ox, float& oy) {
    float tx = rvx - rvn * nx, ty = rvy - rvn * ny;
    float ts = sqrtf(tx * tx + ty * ty);
    if (ts <= 0.0f) { ox = 0.0f; oy = 0.0f; return; }
    float a = slidingShare * ts, c = friction * normalChange;
    float change = a < c ? a : c;
    float k = -change / ts;
    ox = tx * k;
    oy = ty * k;
}

// S/collision_cache_solve.comp:55-86
__device__ __forceinline__ void wall_contact(const SimDesc& s, float distance, float nx, float ny,
                                             float vx, float vy, float weight, float friction,
                                             int writePressure, Accum& acc) {
    float penetration = 0.5f * s.diameter - distance;
    if (penetration <= 0.0f) return;
    acc.pcx += nx * (penetration * weight);
    acc.pcy += ny * (penetration * weight);
    float ns = vx * nx + vy * ny;
    float e1 = 1.0f + s.restitution;
    if (ns < 0.0f) {
        float k = e1 * ns * weight;
        acc.vcx -= nx * k;
        acc.vcy -= ny * k;
    }
    if (friction > 0.0f) {
        float approach = -ns > 0.0f ? -ns : 0.0f;
        float fx, fy;
        friction_correction(vx, vy, nx, ny, ns, weight * (e1 * approach + penetration / s.subDt),
                            weight, friction, fx, fy);
        acc.vcx += fx;
        acc.vcy += fy;
    }
    if (writePressure) {
        float approach = -ns > 0.0f ? -ns : 0.0f;
        acc.pressure += weight * (penetration / s.subDt + approach);
    }
}

__device__ __forceinline__ void accumulate_contact(const SimDesc& s, int id, int other,
                                                   const float2* __restrict__ posIn,
                                                   const float2* __restrict__ velIn,
                                                   const float4* __restrict__ mats,
                                                   const uint32_t* __restrict__ counts,
                                                   const float2* __restrict__ fallback,
                                                   float ppx, float ppy, float vpx, float vpy,
                                                   float4 ownMat, uint32_t ownContacts,
                                                   int writePressure, Accum& acc) {
    float4 om = mats[other];
    if (s.liquid && !(ownMat.z != 0.0f || om.z != 0.0f)) return;
    float2 op = posIn[other];
    float dx = ppx - op.x * s.vw;
    float dy = ppy - op.y * s.vh;
    if (!s.bounded) {
        dx -= floorf(dx / s.vw + 0.5f) * s.vw;
        dy -= floorf(dy / s.vh + 0.5f) * s.vh;
    }
    float d2 = dx * dx + dy * dy;
    if (!(d2 < s.diameter2)) return;
    float eps = s.coincidentEps;
    float distance = sqrtf(d2 > eps ? d2 : eps);
    float nx, ny;
    if (d2 > eps) {
        nx = dx / distance;
        ny = dy / distance;
    } else {
        // S/collision_cache_solve.comp:105-112, cos/sin from a 65536-entry table.
        uint32_t lo = (uint32_t)min(id, other), hi = (uint32_t)max(id, other);
        uint32_t hash = lo * 1664525u + hi * 1013904223u;
        float2 c = fallback[hash & 65535u];
        nx = id < other ? c.x : -c.x;
        ny = id < other ? c.y : -c.y;
    }
    float penetration = s.diameter - distance;
    uint32_t oc = counts[other];
    uint32_t crowd = ownContacts > oc ? ownContacts : oc;
    float pairWeight = 1.0f / (float)(crowd > 1u ? crowd : 1u);
    float share = om.x / (ownMat.x + om.x);
    float2 ov = velIn[other];
    float ovx = ov.x * s.vw, ovy = ov.y * s.vh;
    float rvx = vpx - ovx, rvy = vpy - ovy;
    float rvn = rvx * nx + rvy * ny;
    float push = share * penetration * pairWeight;
    acc.pcx += nx * push;
    acc.pcy += ny * push;
    float e1 = 1.0f + s.restitution;
    if (rvn < 0.0f) {
        float k = share * e1 * rvn * pairWeight;
        acc.vcx -= nx * k;
        acc.vcy -= ny * k;
    }
    float friction = sqrtf(ownMat.y * om.y);
    if (friction > 0.0f) {
        float approach = -rvn > 0.0f ? -rvn : 0.0f;
        float fx, fy;
        friction_correction(rvx, rvy, nx, ny, rvn,
                            share * pairWeight * (e1 * approach + penetration / s.subDt),
                            share * pairWeight, friction, fx, fy);
        acc.vcx += fx;
        acc.vcy += fy;
    }
    if (writePressure) {
        float approach = -rvn > 0.0f ? -rvn : 0.0f;
        acc.pressure += 0.5f * pairWeight * (penetration / s.subDt + approach);
    }
}

__global__ void __launch_bounds__(WG) k_solve(const SimDesc* __restrict__ sims,
                                              const BodyBlock* __restrict__ blocks,
                                              const float2* __restrict__ posIn,
                                              const float2* __restrict__ velIn,
                                              float2* __restrict__ posOut,
                                              float2* __restrict__ velOut,
                                              const float4* __restrict__ mats,
                                              const uint32_t* __restrict__ counts,
                                              const uint32_t* __restrict__ contacts,
                                              int planeStride, GridBuf grid, int guardLimit,
                                              const float2* __restrict__ fallback,
                                              float* __restrict__ pressures, int writePressure,
                                              float nextDriftDt, int insertNext, GridBuf next) {
    BodyBlock b = blocks[blockIdx.x];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float2 p = posIn[g];
    float2 v = velIn[g];
    float4 ownMat = mats[g];
    float ppx = p.x * s.vw, ppy = p.y * s.vh;
    float vpx = v.x * s.vw, vpy = v.y * s.vh;
    Accum acc{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t own = counts[g];
    int listed = own < (uint32_t)CONTACT_SLOTS ? (int)own : CONTACT_SLOTS;
    uint32_t last = 0u;
    for (int i = 0; i < listed; i++) {
        last = contacts[(size_t)i * planeStride + g];
        accumulate_contact(s, g, (int)last, posIn, velIn, mats, counts, fallback, ppx, ppy, vpx,
                           vpy, ownMat, own, writePressure, acc);
    }
    if (own > (uint32_t)CONTACT_SLOTS) {
        // More contacts than slots: take the rest in ascending index, one
        // ordered walk of the grid per contact.
        bool ownFluid = s.liquid && ownMat.z == 0.0f;
        for (uint32_t rest = own - (uint32_t)CONTACT_SLOTS; rest > 0u; rest--) {
            uint32_t best = 0xffffffffu;
            walk_stencil(s, p.x, p.y, grid, guardLimit, [&](int other, float ox, float oy) {
                uint32_t u = (uint32_t)other;
                if (other == g || u <= last || u >= best) return;
                if (ownFluid && mats[other].z == 0.0f) return;
                float d2 = pair_distance_squared(g, p.x, p.y, other, ox, oy, s);
                if (d2 < s.diameter2) best = u;
            });
            if (best == 0xffffffffu) break;
            last = best;
            accumulate_contact(s, g, (int)best, posIn, velIn, mats, counts, fallback, ppx, ppy, vpx,
                               vpy, ownMat, own, writePressure, acc);
        }
    }
    if (s.bounded) {
        float weight = 1.0f / (float)(own + 1u);
        float fr = ownMat.y;
        wall_contact(s, ppx, 1.0f, 0.0f, vpx, vpy, weight, fr, writePressure, acc);
        wall_contact(s, s.vw - ppx, -1.0f, 0.0f, vpx, vpy, weight, fr, writePressure, acc);
        wall_contact(s, ppy, 0.0f, 1.0f, vpx, vpy, weight, fr, writePressure, acc);
        wall_contact(s, s.vh - ppy, 0.0f, -1.0f, vpx, vpy, weight, fr, writePressure, acc);
    }
    ppx += acc.pcx;
    ppy += acc.pcy;
    vpx += acc.vcx + acc.pcx / s.subDt;
    vpy += acc.vcy + acc.pcy / s.subDt;
    float2 op, ov;
    if (s.bounded) {
        op.x = clampf(ppx / s.vw, 0.0f, MAXIMUM_UNIT);
        op.y = clampf(ppy / s.vh, 0.0f, MAXIMUM_UNIT);
    } else {
        ppx = glsl_mod(ppx, s.vw);
        ppy = glsl_mod(ppy, s.vh);
        op.x = ppx / s.vw;
        op.y = ppy / s.vh;
    }
    ov.x = vpx / s.vw;
    ov.y = vpy / s.vh;
    if (writePressure && !(s.liquid && ownMat.z == 0.0f)) pressures[g] = acc.pressure;
    // The next pass's drift (S/collision_cache_drift.comp:83-91; only the
    // first pass of a step kicks, so there is none here).
    if (nextDriftDt != 0.0f) {
        op.x += ov.x * nextDriftDt;
        op.y += ov.y * nextDriftDt;
        if (s.bounded) {
            op.x = clampf(op.x, 0.0f, MAXIMUM_UNIT);
            op.y = clampf(op.y, 0.0f, MAXIMUM_UNIT);
        } else {
            op.x = fractf(op.x);
            op.y = fractf(op.y);
        }
    }
    posOut[g] = op;
    velOut[g] = ov;
    if (insertNext) grid_insert(s, g, op.x, op.y, next);
}

// ---------------------------------------------------------------------------
// Host launchers

static inline GridBuf grid_of(const CollideGrid& c) { return GridBuf{c.heads, c.nodes, c.cellOf}; }

void launch_drift_insert(hipStream_t st, const LaunchShape& ls, float2* pos, float2* vel,
                         const float2* acc, float kickDt, float driftDt, int insert,
                         const CollideGrid& grid) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_drift_insert, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       pos, vel, acc, kickDt, driftDt, insert, grid_of(grid));
}

void launch_count(hipStream_t st, const LaunchShape& ls, const CollideGrid& grid,
                  const CollideGrid& prev, int clearPrev, const float4* mats, uint32_t* counts,
                  uint32_t* contacts) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_count, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       grid_of(grid), grid_of(prev), clearPrev, mats, counts, contacts,
                       ls.planeStride, ls.guardLimit);
}

void launch_solve(hipStream_t st, const LaunchShape& ls, const float2* posIn, const float2* velIn,
                  float2* posOut, float2* velOut, const float4* mats, const uint32_t* counts,
                  const uint32_t* contacts, const CollideGrid& grid, const float2* fallback,
                  float* pressures, int writePressure, float nextDriftDt, int insertNext,
                  const CollideGrid& next) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_solve, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, posIn,
                       velIn, posOut, velOut, mats, counts, contacts, ls.planeStride,
                       grid_of(grid), ls.guardLimit, fallback, pressures, writePressure,
                       nextDriftDt, insertNext, grid_of(next));
}

}  // namespace gs
