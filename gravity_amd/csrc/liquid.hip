// Position-based liquid (Macklin & Mueller 2013) between the collision
// substeps and the gravity kick (S/liquid_{grid,lambda,apply,viscosity}.comp,
// S/liquid_kernel.glsl, K/LiquidSolver.kt:107-173).
//
// Each body searches the 3x3 liquid cells (cells >= the kernel radius)
// around the cell it had when the solve began, so i finds j exactly when j
// finds i. The reference keeps linked lists per cell and neighbour lists
// with a skin; here the bodies are counting-sorted by that start cell once
// per solve, so a cell is a contiguous, ascending run of body indices and
// no list can overflow. The sums run in the stencil's cell order (rows,
// then columns, -1..1) and ascending index within a cell, the canonical
// order the CPU oracle uses too (the reference's depends on atomics).
//
// Arithmetic follows the shaders' operation order; every kernel is
// compiled without contraction, so the results are bit-exact against the
// oracle.
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

__device__ __forceinline__ int liquid_cell(const SimDesc& s, float x, float y) {
    return s.liqCellBase + floor_index(x * (float)s.lgw, s.lgw) +
           floor_index(y * (float)s.lgh, s.lgh) * s.lgw;
}

// Every other body of the 3x3 start cells, in the canonical order, with the
// data `load` fetches for it. On a grid narrower than three cells a wrapped
// row or column met twice is taken once. Consecutive cells are contiguous
// runs of `items` (they are counting-sorted by cell), so a stencil row's
// cells, taken in the canonical column order -1, 0, 1, are at most two runs:
// one under walls or away from the torus's seam, two where the row wraps.
// The three rows' runs are taken as one sequence of six pieces, LIQ_BATCH
// entries at a time: a batch's list entries load together, then the
// neighbours' data together (each clamped to a valid entry), so a batch
// costs two memory round trips where an entry at a time cost two per
// neighbour; the visits keep the canonical order.
constexpr int LIQ_BATCH = 8;
template <typename Load, typename Visit>
__device__ __forceinline__ void liquid_neighbours(const SimDesc& s, int g, int home,
                                                  const uint32_t* __restrict__ prefix,
                                                  const uint32_t* __restrict__ items, Load&& load,
                                                  Visit&& visit) {
    const int local = home - s.liqCellBase;
    const int hx = local % s.lgw, hy = local / s.lgw, lgw = s.lgw;
    // The row's column runs [aLo, aHi] and [bLo, bHi] (empty when lo > hi).
    int aLo, aHi, bLo = 1, bHi = 0;
    if (s.bounded) {
        aLo = max(hx - 1, 0);
        aHi = min(hx + 1, lgw - 1);
    } else if (lgw >= 3 && hx == 0) {
        aLo = aHi = lgw - 1;
        bLo = 0;
        bHi = 1;
    } else if (lgw >= 3 && hx == lgw - 1) {
        aLo = lgw - 2;
        aHi = lgw - 1;
        bLo = bHi = 0;
    } else if (lgw >= 3) {
        aLo = hx - 1;
        aHi = hx + 1;
    } else if (lgw == 2 && hx == 0) {
        aLo = aHi = 1;
        bLo = bHi = 0;
    } else {
        aLo = 0;
        aHi = lgw - 1;
    }
    const bool bRun = bLo <= bHi;
    // Piece 2j + r is run r of stencil row j: entries [first, first + its
    // length) of `items`, entries [cum, next cum) of the sequence. A row
    // outside the walls, or met twice on a grid narrower than three rows,
    // reads the home cell's bounds and is empty.
    uint32_t first[6], cum[7];
    cum[0] = 0u;
    int ys[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        int y = hy + j - 1;
        bool ok;
        if (s.bounded) {
            ok = y >= 0 && y < s.lgh;
        } else {
            y = wrap_axis(y, s.lgh);
            ok = (j < 1 || y != ys[0]) && (j < 2 || y != ys[1]);
        }
        ys[j] = y;
        const int row = ok ? s.liqCellBase + y * lgw : home - hx;
        const uint32_t a0 = prefix[row + aLo], a1 = prefix[row + aHi + 1];
        const uint32_t b0 = prefix[row + (bRun ? bLo : aLo)], b1 = prefix[row + (bRun ? bHi + 1 : aLo)];
        first[2 * j] = a0;
        first[2 * j + 1] = b0;
        cum[2 * j + 1] = cum[2 * j] + (ok ? a1 - a0 : 0u);
        cum[2 * j + 2] = cum[2 * j + 1] + (ok && bRun ? b1 - b0 : 0u);
    }
    const uint32_t total = cum[6];
    using Data = decltype(load(0));
    for (uint32_t k0 = 0; k0 < total; k0 += LIQ_BATCH) {
        int id[LIQ_BATCH];
#pragma unroll
        for (int q = 0; q < LIQ_BATCH; q++) {
            const uint32_t e = min(k0 + (uint32_t)q, total - 1u);
            uint32_t pos = first[0] + e;
#pragma unroll
            for (int k = 1; k < 6; k++)
                if (e >= cum[k]) pos = first[k] + (e - cum[k]);
            id[q] = (int)items[pos];
        }
        Data d[LIQ_BATCH];
#pragma unroll
        for (int q = 0; q < LIQ_BATCH; q++) d[q] = load(id[q]);
#pragma unroll
        for (int q = 0; q < LIQ_BATCH; q++)
            if (k0 + (uint32_t)q < total && id[q] != g) visit(id[q], d[q]);
    }
}

// S/liquid_kernel.glsl pixelDelta: (first - second) * viewport, minimum
// image on the torus.
__device__ __forceinline__ void liquid_delta(const SimDesc& s, float ax, float ay, float bx,
                                             float by, float& dx, float& dy) {
    dx = (ax - bx) * s.vw;
    dy = (ay - by) * s.vh;
    // a difference under a quarter of the viewport is its own minimum image
    // (floor(d / w + 0.5) is 0 there), with no division
    if (!s.bounded && !(fabsf(dx) < 0.25f * s.vw && fabsf(dy) < 0.25f * s.vh)) {
        dx -= floorf(dx / s.vw + 0.5f) * s.vw;
        dy -= floorf(dy / s.vh + 0.5f) * s.vh;
    }
}

__device__ __forceinline__ float poly6(const SimDesc& s, float d2) {
    float x = s.lh2 - d2;
    return x > 0.0f ? s.lpoly6 * x * x * x : 0.0f;
}

// Gradient of the spiky kernel with respect to the first body of
// delta = first - second; coincident bodies take the collision solver's
// fixed pseudo-random direction per pair.
__device__ __forceinline__ void spiky(const SimDesc& s, float dx, float dy, float dist,
                                      uint32_t first, uint32_t second,
                                      const float2* __restrict__ fallback, float& gx, float& gy) {
    if (dist >= s.lh) { gx = 0.0f; gy = 0.0f; return; }
    float nx, ny;
    if (dist > 1e-4f) {
        nx = dx / dist;
        ny = dy / dist;
    } else {
        uint32_t lo = min(first, second), hi = max(first, second);
        float2 c = fallback[(lo * 1664525u + hi * 1013904223u) & 65535u];
        nx = first < second ? c.x : -c.x;
        ny = first < second ? c.y : -c.y;
    }
    float x = s.lh - dist;
    float m = s.lspiky * x * x;
    gx = nx * m;
    gy = ny * m;
}

// Virtual rows continuing the packing beyond a wall `distance` px away:
// density and its derivative with respect to the distance.
__device__ __forceinline__ void wall_density(const SimDesc& s, float distance, float& rho,
                                             float& deriv) {
    float rx = 0.0f, ry = 0.0f;
    float delta = distance - 0.5f * s.diameter + s.lrowStep;
    for (int row = 0; row < 8 && delta < s.lh; row++) {
        float dm = delta > 0.0f ? delta : 0.0f;
        float a2 = s.lh2 - dm * dm;
        float a = sqrtf(a2);
        float a5 = a2 * a2 * a;
        rx += a5 * a2;
        ry += -7.0f * dm * a5;
        delta += s.lrowStep;
    }
    rho = rx * s.lwallScale;
    deriv = ry * s.lwallScale;
}

// All four walls at pixel position (px, py): density, gradient.
__device__ __forceinline__ void walls_density(const SimDesc& s, float px, float py, float& rho,
                                              float& gx, float& gy) {
    rho = gx = gy = 0.0f;
    if (!s.bounded) return;
    float lo0, lo1, hi0, hi1, bo0, bo1, to0, to1;
    wall_density(s, px, lo0, lo1);
    wall_density(s, s.vw - px, hi0, hi1);
    wall_density(s, py, bo0, bo1);
    wall_density(s, s.vh - py, to0, to1);
    rho = lo0 + hi0 + bo0 + to0;
    gx = lo1 - hi1;
    gy = bo1 - to1;
}

// ---------------------------------------------------------------------------
// Start cells: bin -> scan -> scatter -> canon (ascending index per cell).

__global__ void __launch_bounds__(WG) k_liq_bin(const SimDesc* __restrict__ sims,
                                                const BodyBlock* __restrict__ blocks,
                                                const float4* __restrict__ body,
                                                uint32_t* liqCount, uint32_t* __restrict__ cellOf) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n || !s.liquid) return;
    float4 p = body[g];
    int c = liquid_cell(s, p.x, p.y);
    cellOf[g] = (uint32_t)c;
    atomicAdd(&liqCount[c], 1u);
}

__global__ void __launch_bounds__(WG) k_liq_scatter(const SimDesc* __restrict__ sims,
                                                    const BodyBlock* __restrict__ blocks,
                                                    const uint32_t* __restrict__ cellOf,
                                                    uint32_t* liqCount,
                                                    const uint32_t* __restrict__ prefix,
                                                    uint32_t* __restrict__ unordered) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n || !s.liquid) return;
    uint32_t c = cellOf[g];
    uint32_t rank = atomicSub(&liqCount[c], 1u) - 1u;
    unordered[prefix[c] + rank] = (uint32_t)g;
}

__global__ void __launch_bounds__(WG) k_liq_canon(int slots, const uint32_t* __restrict__ cellOf,
                                                  const uint32_t* __restrict__ prefix,
                                                  const uint32_t* __restrict__ unordered,
                                                  uint32_t* __restrict__ items) {
    int d = (int)(blockIdx.x * WG + threadIdx.x);
    if (d >= slots) return;
    uint32_t g = unordered[d];
    uint32_t c = cellOf[g];
    int s0 = (int)prefix[c], s1 = (int)prefix[c + 1];
    items[s0 + segment_rank<false>(unordered, s0, s1, g)] = g;
}

// ---------------------------------------------------------------------------
// S/liquid_lambda.comp:36-105: density constraint C = max(rho/rho0 - 1, 0)
// of a fluid body and its multiplier, stored times the body's mass; solid
// bodies add density around them but have no constraint.
__global__ void __launch_bounds__(WG) k_liq_lambda(const SimDesc* __restrict__ sims,
                                                   const BodyBlock* __restrict__ blocks,
                                                   const float4* __restrict__ body,
                                                   const float4* __restrict__ mats,
                                                   const uint32_t* __restrict__ cellOf,
                                                   const uint32_t* __restrict__ prefix,
                                                   const uint32_t* __restrict__ items,
                                                   const float2* __restrict__ fallback,
                                                   float* __restrict__ lambdas,
                                                   float* __restrict__ pressures,
                                                   int writePressure) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n || !s.liquid) return;
    float4 me = body[g];
    float4 m = mats[g];
    if (m.z != 0.0f) {
        lambdas[g] = 0.0f;
        return;
    }
    float density = 0.0f, ogx = 0.0f, ogy = 0.0f, others = 0.0f;
    liquid_neighbours(s, g, (int)cellOf[g], prefix, items,
                      [&](int j) {
                          const float4 o = body[j];
                          return make_float3(o.x, o.y, mats[j].z);
                      },
                      [&](int j, float3 o) {
        float dx, dy;
        liquid_delta(s, me.x, me.y, o.x, o.y, dx, dy);
        float d2 = dx * dx + dy * dy;
        if (d2 >= s.lh2) return;
        density += poly6(s, d2);
        float gx, gy;
        spiky(s, dx, dy, sqrtf(d2), (uint32_t)g, (uint32_t)j, fallback, gx, gy);
        gx /= s.lrest;
        gy /= s.lrest;
        ogx += gx;
        ogy += gy;
        if (o.z == 0.0f) others += gx * gx + gy * gy;
    });
    density += poly6(s, 0.0f);
    float wr, wgx, wgy;
    walls_density(s, me.x * s.vw, me.y * s.vh, wr, wgx, wgy);
    density += wr;
    ogx += wgx / s.lrest;
    ogy += wgy / s.lrest;
    float c = density / s.lrest - 1.0f;
    if (c < 0.0f) c = 0.0f;
    float lambda = -c / (ogx * ogx + ogy * ogy + others + s.lrelax);
    lambdas[g] = lambda * m.x;
    if (writePressure) pressures[g] = c * s.lpressureScale;
}

// S/liquid_apply.comp:38-92: Jacobi position correction, converted to
// velocity (dv = dx / dt); walls stop a body at its radius and take away the
// velocity into them. Bodies of simulations without liquid are copied.
__global__ void __launch_bounds__(WG) k_liq_apply(const SimDesc* __restrict__ sims,
                                                  const BodyBlock* __restrict__ blocks,
                                                  const float4* __restrict__ body,
                                                  float4* __restrict__ out,
                                                  const float4* __restrict__ mats,
                                                  const uint32_t* __restrict__ cellOf,
                                                  const uint32_t* __restrict__ prefix,
                                                  const uint32_t* __restrict__ items,
                                                  const float2* __restrict__ fallback,
                                                  const float* __restrict__ lambdas) {
    BodyBlock bb = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[bb.sim];
    int g = bb.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float4 b = body[g];
    if (!s.liquid) {
        out[g] = b;
        return;
    }
    float4 m = mats[g];
    float ownMass = m.x;
    float ownLambda = lambdas[g] / ownMass;
    float cx = 0.0f, cy = 0.0f;
    liquid_neighbours(s, g, (int)cellOf[g], prefix, items,
                      [&](int j) {
                          const float4 o = body[j];
                          return make_float3(o.x, o.y, lambdas[j]);
                      },
                      [&](int j, float3 o) {
        float dx, dy;
        liquid_delta(s, b.x, b.y, o.x, o.y, dx, dy);
        float d2 = dx * dx + dy * dy;
        if (d2 >= s.lh2) return;
        float coef = ownLambda + o.z / ownMass;
        if (coef == 0.0f) return;
        float gx, gy;
        spiky(s, dx, dy, sqrtf(d2), (uint32_t)g, (uint32_t)j, fallback, gx, gy);
        cx += coef * gx;
        cy += coef * gy;
    });
    if (m.z == 0.0f) {
        float wr, wgx, wgy;
        walls_density(s, b.x * s.vw, b.y * s.vh, wr, wgx, wgy);
        cx += ownLambda * wgx;
        cy += ownLambda * wgy;
    }
    cx /= s.lrest;
    cy /= s.lrest;
    float ppx = b.x * s.vw + cx, ppy = b.y * s.vh + cy;
    float vpx = b.z * s.vw + cx / s.dt, vpy = b.w * s.vh + cy / s.dt;
    if (s.bounded) {
        float r = s.lradius;
        float ix = clampf(ppx, r, s.vw - r), iy = clampf(ppy, r, s.vh - r);
        bool lowx = ppx < ix, lowy = ppy < iy, highx = ppx > ix, highy = ppy > iy;
        if ((lowx && vpx < 0.0f) || (highx && vpx > 0.0f)) vpx = 0.0f;
        if ((lowy && vpy < 0.0f) || (highy && vpy > 0.0f)) vpy = 0.0f;
        ppx = ix;
        ppy = iy;
    }
    float4 o;
    if (s.bounded) {
        o.x = clampf(ppx / s.vw, 0.0f, MAXIMUM_UNIT);
        o.y = clampf(ppy / s.vh, 0.0f, MAXIMUM_UNIT);
    } else {
        o.x = fractf(ppx / s.vw);
        o.y = fractf(ppy / s.vh);
    }
    o.z = vpx / s.vw;
    o.w = vpy / s.vh;
    out[g] = o;
}

struct BodySolid {
    float4 b;
    float solid;
};

// S/liquid_viscosity.comp:27-55: XSPH viscosity between fluid bodies.
__global__ void __launch_bounds__(WG) k_liq_viscosity(const SimDesc* __restrict__ sims,
                                                      const BodyBlock* __restrict__ blocks,
                                                      const float4* __restrict__ body,
                                                      float4* __restrict__ out,
                                                      const float4* __restrict__ mats,
                                                      const uint32_t* __restrict__ cellOf,
                                                      const uint32_t* __restrict__ prefix,
                                                      const uint32_t* __restrict__ items) {
    BodyBlock bb = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[bb.sim];
    int g = bb.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float4 b = body[g];
    if (!s.liquid || !s.lviscous || mats[g].z != 0.0f) {
        out[g] = b;
        return;
    }
    float vx = b.z * s.vw, vy = b.w * s.vh;
    float bx = 0.0f, by = 0.0f;
    liquid_neighbours(s, g, (int)cellOf[g], prefix, items,
                      [&](int j) { return BodySolid{body[j], mats[j].z}; },
                      [&](int j, BodySolid od) {
        if (od.solid != 0.0f) return;
        const float4 o = od.b;
        float dx, dy;
        liquid_delta(s, b.x, b.y, o.x, o.y, dx, dy);
        float d2 = dx * dx + dy * dy;
        if (d2 >= s.lh2) return;
        float w = poly6(s, d2);
        bx += (o.z * s.vw - vx) * w;
        by += (o.w * s.vh - vy) * w;
    });
    b.z = (vx + bx * s.lviscosity) / s.vw;
    b.w = (vy + by * s.lviscosity) / s.vh;
    out[g] = b;
}

// ---------------------------------------------------------------------------

void launch_liquid_bin(hipStream_t st, const LaunchShape& ls, const float4* body,
                       uint32_t* liqCount, uint32_t* cellOf) {
    if (ls.bodyBlocks == 0) return;
    launch(k_liq_bin, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, body,
                       liqCount, cellOf);
}

void launch_liquid_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* cellOf,
                           uint32_t* liqCount, const uint32_t* prefix, uint32_t* unordered) {
    if (ls.bodyBlocks == 0) return;
    launch(k_liq_scatter, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       cellOf, liqCount, prefix, unordered);
}

void launch_liquid_canon(hipStream_t st, int slots, const uint32_t* cellOf,
                         const uint32_t* prefix, const uint32_t* unordered, uint32_t* items) {
    if (slots <= 0) return;
    launch(k_liq_canon, dim3((slots + WG - 1) / WG), dim3(WG), 0, st, slots, cellOf,
                       prefix, unordered, items);
}

void launch_liquid_lambda(hipStream_t st, const LaunchShape& ls, const float4* body,
                          const float4* mats, const uint32_t* cellOf, const uint32_t* prefix,
                          const uint32_t* items, const float2* fallback, float* lambdas,
                          float* pressures, int writePressure) {
    if (ls.bodyBlocks == 0) return;
    launch(k_liq_lambda, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       body, mats, cellOf, prefix, items, fallback, lambdas, pressures,
                       writePressure);
}

void launch_liquid_apply(hipStream_t st, const LaunchShape& ls, const float4* body, float4* out,
                         const float4* mats, const uint32_t* cellOf, const uint32_t* prefix,
                         const uint32_t* items, const float2* fallback, const float* lambdas) {
    if (ls.bodyBlocks == 0) return;
    launch(k_liq_apply, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       body, out, mats, cellOf, prefix, items, fallback, lambdas);
}

void launch_liquid_viscosity(hipStream_t st, const LaunchShape& ls, const float4* body,
                             float4* out, const float4* mats, const uint32_t* cellOf,
                             const uint32_t* prefix, const uint32_t* items) {
    if (ls.bodyBlocks == 0) return;
    launch(k_liq_viscosity, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       body, out, mats, cellOf, prefix, items);
}

}  // namespace gs
