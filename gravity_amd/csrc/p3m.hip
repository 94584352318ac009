// P3M gravity, body side (S/build_grid.comp, S/kick.comp,
// K/GravityPipeline.kt:907-973, 1213-1261).
//
// The reference pushes every body onto a linked list of its physical cell
// with atomics; here the bodies are counting-sorted by physical cell into
// contiguous segments (bin -> scan -> scatter -> canon). Each segment is
// ordered by descending body index, the order a sequential insertion into
// the reference's LIFO lists gives, so the near field is deterministic.
// The segments serve three consumers: the forward row FFT, which gathers the
// fixed-point CIC deposit of a mesh row from three physical rows of bodies
// (no global atomics, no clearing); the cell monopoles; and the exact near
// field of the kick.
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

__device__ __forceinline__ int home_cell(const SimDesc& s, float px, float py) {
    int hx = (int)floorf(px * (float)s.pw), hy = (int)floorf(py * (float)s.ph);
    return s.physBase + wrap_axis(hx, s.pw) + wrap_axis(hy, s.ph) * s.pw;
}

// S/build_grid.comp:71-74: the home cell of every body.
__global__ void __launch_bounds__(WG) k_bin(const SimDesc* __restrict__ sims,
                                            const BodyBlock* __restrict__ blocks,
                                            const float4* __restrict__ body, uint32_t* cellCount,
                                            uint32_t* __restrict__ homeOf) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n || !s.mutual) return;
    float4 p = body[g];
    int c = home_cell(s, p.x, p.y);
    homeOf[g] = (uint32_t)c;
    atomicAdd(&cellCount[c], 1u);
}

__global__ void __launch_bounds__(WG) k_cell_scatter(const SimDesc* __restrict__ sims,
                                                     const BodyBlock* __restrict__ blocks,
                                                     const uint32_t* __restrict__ homeOf,
                                                     uint32_t* cellCount,
                                                     const uint32_t* __restrict__ cellPrefix,
                                                     uint32_t* __restrict__ slotBody) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n || !s.mutual) return;
    uint32_t c = homeOf[g];
    uint32_t rank = atomicSub(&cellCount[c], 1u) - 1u;
    slotBody[cellPrefix[c] + rank] = (uint32_t)g;
}

// Orders each cell by descending body index; the first body of a cell also
// sums the cell mass in 1/256 units (S/build_grid.comp:75), exactly.
__global__ void __launch_bounds__(WG) k_cell_canon(int slots, const uint32_t* __restrict__ homeOf,
                                                   const uint32_t* __restrict__ cellPrefix,
                                                   const uint32_t* __restrict__ slotBody,
                                                   const float4* __restrict__ body,
                                                   const float4* __restrict__ mats,
                                                   float4* __restrict__ cellBody,
                                                   uint32_t* __restrict__ cellMass) {
    int d = (int)(blockIdx.x * WG + threadIdx.x);
    if (d >= slots) return;
    uint32_t g = slotBody[d];
    uint32_t c = homeOf[g];
    int s0 = (int)cellPrefix[c], s1 = (int)cellPrefix[c + 1];
    int rank = 0;
    for (int e = s0; e < s1; e++) rank += slotBody[e] > g ? 1 : 0;
    float4 p = body[g];
    float m = mats[g].x;
    cellBody[s0 + rank] = make_float4(p.x, p.y, m, __uint_as_float(g));
    if (rank == 0) {
        uint32_t sum = 0u;
        for (int e = s0; e < s1; e++) sum += (uint32_t)rintf(mats[slotBody[e]].x * 256.0f);
        cellMass[c] = sum;
    }
}

__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// Cells of a stencil row whose tables are loaded together (the default
// 3-cell near-field radius in one batch; wider stencils take several).
constexpr int NEAR_BATCH = 7;

// Short-range term of one source (S/kick.comp:66-80) from its physical
// displacement (px = dx * aspect, py = dy) and its Gaussian weight.
__device__ __forceinline__ void near_term(const SimDesc& s, float px, float py, float weight,
                                          float mass, float& ax, float& ay) {
    float r2 = px * px + py * py;
    float q = s.gOverM * mass * weight * __builtin_amdgcn_rcpf(fmaxf(r2, s.minDist2));
    ax += q * px * s.invAspect;
    ay += q * py;
}

__device__ __forceinline__ float min_image(const SimDesc& s, float d) {
    return s.bounded ? d : d - floorf(d + 0.5f);
}

// S/kick.comp:105-156 (sleep handled by the caller's schedule).
//
// The mesh gather follows the reference's operations. The near field is the
// reference's sum in the reference's order, evaluated on the FP32-tolerance
// path: divisions by constants become reciprocal products, exp is the
// hardware exp2, and the Gaussian of a cell monopole is the product of its
// x and y factors, so a body takes 2(2r+1) exponentials for its monopoles
// instead of (2r+1)^2.
__global__ void __launch_bounds__(WG) k_kick(const SimDesc* __restrict__ sims,
                                             const BodyBlock* __restrict__ blocks,
                                             float4* __restrict__ body,
                                             float2* __restrict__ acc,
                                             const float2* __restrict__ force,
                                             const uint32_t* __restrict__ cellPrefix,
                                             const uint32_t* __restrict__ cellMass,
                                             const float4* __restrict__ cellBody, int initial) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float h = 0.5f * s.dt;
    float4 p = body[g];
    if (!s.mutual) {
        if (!initial) {
            p.z += s.fieldX * h;
            p.w += s.fieldY * h;
            body[g] = p;
        }
        acc[g] = make_float2(s.fieldX, s.fieldY);
        return;
    }
    // S/kick.comp:82-103: CIC gather of the mesh force (x = Fx, y = Fy).
    float gx = p.x * (float)s.pw - 0.5f, gy = p.y * (float)s.ph - 0.5f;
    float flx = floorf(gx), fly = floorf(gy);
    int bx = (int)flx, by = (int)fly;
    float fx = gx - flx, fy = gy - fly;
    int x0 = wrap_axis(bx, s.mw), x1 = wrap_axis(bx + 1, s.mw);
    int y0 = wrap_axis(by, s.mh), y1 = wrap_axis(by + 1, s.mh);
    const float2* F = force + s.meshBase;
    float2 v00 = F[x0 + y0 * s.mw], v10 = F[x1 + y0 * s.mw];
    float2 v01 = F[x0 + y1 * s.mw], v11 = F[x1 + y1 * s.mw];
    float ax = mixf(mixf(v00.x, v10.x, fx), mixf(v01.x, v11.x, fx), fy);
    float ay = mixf(mixf(v00.y, v10.y, fx), mixf(v01.y, v11.y, fx), fy);
    ax += s.fieldX;
    ay += s.fieldY;
    const float k2 = s.invTwoSigma2;
    int hx = (int)floorf(p.x * (float)s.pw), hy = (int)floorf(p.y * (float)s.ph);
    for (int dy = -s.ry; dy <= s.ry; dy++) {
        int cy = hy + dy;
        if (s.bounded && (cy < 0 || cy >= s.ph)) continue;
        int wy = wrap_axis(cy, s.ph);
        float py = min_image(s, ((float)wy + 0.5f) * s.invPh - p.y);
        float ey = __expf(-(py * py) * k2);
        const int rowBase = s.physBase + wy * s.pw;
        for (int dx0 = -s.rx; dx0 <= s.rx; dx0 += NEAR_BATCH) {
            uint32_t first[NEAR_BATCH], after[NEAR_BATCH], cmass[NEAR_BATCH];
#pragma unroll
            for (int k = 0; k < NEAR_BATCH; k++) {
                int dx = dx0 + k, cx = hx + dx;
                bool on = dx <= s.rx && (!s.bounded || (cx >= 0 && cx < s.pw));
                int c = rowBase + wrap_axis(cx, s.pw);
                first[k] = on ? cellPrefix[c] : 0u;
                after[k] = on ? cellPrefix[c + 1] : 0u;
                cmass[k] = on ? cellMass[c] : 0u;
            }
#pragma unroll
            for (int k = 0; k < NEAR_BATCH; k++) {
                int dx = dx0 + k, cx = hx + dx;
                if (dx > s.rx || (s.bounded && (cx < 0 || cx >= s.pw))) continue;
                int s0 = (int)first[k];
                int count = (int)(after[k] - first[k]);
                bool home = dx == 0 && dy == 0;
                if (!home && count > s.exactLimit) {
                    int wx = wrap_axis(cx, s.pw);
                    float px = min_image(s, ((float)wx + 0.5f) * s.invPw - p.x) * s.aspect;
                    float w = __expf(-(px * px) * k2) * ey;
                    near_term(s, px, py, w, (float)cmass[k] * (1.0f / 256.0f), ax, ay);
                } else {
                    int limit = home ? s.homeLimit : s.exactLimit;
                    int stop = s0 + (count < limit ? count : limit);
                    for (int t = s0; t < stop; t++) {
                        float4 o = cellBody[t];
                        if (__float_as_uint(o.w) == (uint32_t)g) continue;
                        float ox = min_image(s, o.x - p.x) * s.aspect;
                        float oy = min_image(s, o.y - p.y);
                        float w = __expf(-(ox * ox + oy * oy) * k2);
                        near_term(s, ox, oy, w, o.z, ax, ay);
                    }
                }
            }
        }
    }
    if (!initial) {
        p.z += ax * h;
        p.w += ay * h;
        body[g] = p;
    }
    acc[g] = make_float2(ax, ay);
}

void launch_bin(hipStream_t st, const LaunchShape& ls, const float4* body, uint32_t* cellCount,
                uint32_t* homeOf) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_bin, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, body,
                       cellCount, homeOf);
}

void launch_cell_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* homeOf,
                         uint32_t* cellCount, const uint32_t* cellPrefix, uint32_t* slotBody) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_cell_scatter, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       homeOf, cellCount, cellPrefix, slotBody);
}

void launch_cell_canon(hipStream_t st, int slots, const uint32_t* homeOf,
                       const uint32_t* cellPrefix, const uint32_t* slotBody, const float4* body,
                       const float4* mats, float4* cellBody, uint32_t* cellMass) {
    if (slots <= 0) return;
    hipLaunchKernelGGL(k_cell_canon, dim3((slots + WG - 1) / WG), dim3(WG), 0, st, slots, homeOf,
                       cellPrefix, slotBody, body, mats, cellBody, cellMass);
}

void launch_kick(hipStream_t st, const LaunchShape& ls, float4* body, float2* acc,
                 const float2* force, const uint32_t* cellPrefix, const uint32_t* cellMass,
                 const float4* cellBody, int initial) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_kick, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, body, acc,
                       force, cellPrefix, cellMass, cellBody, initial);
}

}  // namespace gs
