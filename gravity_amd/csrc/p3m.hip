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
                                            const float2* __restrict__ pos, uint32_t* cellCount,
                                            uint32_t* __restrict__ homeOf) {
    BodyBlock b = blocks[blockIdx.x];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n || !s.mutual) return;
    float2 p = pos[g];
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
    BodyBlock b = blocks[blockIdx.x];
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
                                                   const float2* __restrict__ pos,
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
    float2 p = pos[g];
    float m = mats[g].x;
    cellBody[s0 + rank] = make_float4(p.x, p.y, m, __uint_as_float(g));
    if (rank == 0) {
        uint32_t sum = 0u;
        for (int e = s0; e < s1; e++) sum += (uint32_t)rintf(mats[slotBody[e]].x * 256.0f);
        cellMass[c] = sum;
    }
}

// S/kick.comp:66-80
__device__ __forceinline__ void local_acceleration(const SimDesc& s, float sx, float sy, float tx,
                                                   float ty, float mass, float& ax, float& ay) {
    float dx = sx - tx, dy = sy - ty;
    if (!s.bounded) {
        dx -= floorf(dx + 0.5f);
        dy -= floorf(dy + 0.5f);
    }
    float px = dx * s.aspect, py = dy * 1.0f;
    float r2 = px * px + py * py;
    float lw = expf(-r2 / s.twoSigma2);
    float scale = s.G * mass * lw / (s.totalMass * (r2 > s.minDist2 ? r2 : s.minDist2));
    ax += scale * px / s.aspect;
    ay += scale * py / 1.0f;
}

__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// S/kick.comp:105-156 (sleep handled by the caller's schedule).
__global__ void __launch_bounds__(WG) k_kick(const SimDesc* __restrict__ sims,
                                             const BodyBlock* __restrict__ blocks,
                                             const float2* __restrict__ pos,
                                             float2* __restrict__ vel, float2* __restrict__ acc,
                                             const float2* __restrict__ force,
                                             const uint32_t* __restrict__ cellPrefix,
                                             const uint32_t* __restrict__ cellMass,
                                             const float4* __restrict__ cellBody, int initial) {
    BodyBlock b = blocks[blockIdx.x];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float h = 0.5f * s.dt;
    if (!s.mutual) {
        if (!initial) {
            float2 v = vel[g];
            v.x += s.fieldX * h;
            v.y += s.fieldY * h;
            vel[g] = v;
        }
        acc[g] = make_float2(s.fieldX, s.fieldY);
        return;
    }
    float2 p = pos[g];
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
    // Short-range complement over the (2rx+1) x (2ry+1) physical cells.
    int hx = (int)floorf(p.x * (float)s.pw), hy = (int)floorf(p.y * (float)s.ph);
    for (int dy = -s.ry; dy <= s.ry; dy++) {
        int cy = hy + dy;
        if (s.bounded && (cy < 0 || cy >= s.ph)) continue;
        int wy = wrap_axis(cy, s.ph);
        for (int dx = -s.rx; dx <= s.rx; dx++) {
            int cx = hx + dx;
            if (s.bounded && (cx < 0 || cx >= s.pw)) continue;
            int wx = wrap_axis(cx, s.pw);
            int c = s.physBase + wx + wy * s.pw;
            int s0 = (int)cellPrefix[c], s1 = (int)cellPrefix[c + 1];
            int count = s1 - s0;
            bool home = dx == 0 && dy == 0;
            if (!home && count > s.exactLimit) {
                float cxn = ((float)wx + 0.5f) / (float)s.pw;
                float cyn = ((float)wy + 0.5f) / (float)s.ph;
                float agg = (float)cellMass[c] / 256.0f;
                local_acceleration(s, cxn, cyn, p.x, p.y, agg, ax, ay);
            } else {
                int limit = home ? s.homeLimit : s.exactLimit;
                int end = s0 + (count < limit ? count : limit);
                for (int t = s0; t < end; t++) {
                    float4 o = cellBody[t];
                    if (__float_as_uint(o.w) != (uint32_t)g)
                        local_acceleration(s, o.x, o.y, p.x, p.y, o.z, ax, ay);
                }
            }
        }
    }
    if (!initial) {
        float2 v = vel[g];
        v.x += ax * h;
        v.y += ay * h;
        vel[g] = v;
    }
    acc[g] = make_float2(ax, ay);
}

void launch_bin(hipStream_t st, const LaunchShape& ls, const float2* pos, uint32_t* cellCount,
                uint32_t* homeOf) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_bin, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, pos,
                       cellCount, homeOf);
}

void launch_cell_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* homeOf,
                         uint32_t* cellCount, const uint32_t* cellPrefix, uint32_t* slotBody) {
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_cell_scatter, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       homeOf, cellCount, cellPrefix, slotBody);
}

void launch_cell_canon(hipStream_t st, int slots, const uint32_t* homeOf,
                       const uint32_t* cellPrefix, const uint32_t* slotBody, const float2* pos,
                       const float4* mats, float4* cellBody, uint32_t* cellMass) {
    if (slots <= 0) return;
    hipLaunchKernelGGL(k_cell_canon, dim3((slots + WG - 1) / WG), dim3(WG), 0, st, slots, homeOf,
                       cellPrefix, slotBody, pos, mats, cellBody, cellMass);
}

void launch_kick(hipStream_t st, const LaunchShape& ls, const float2* pos, float2* vel,
                 float2* acc, const float4* mats, const float2* force, const uint32_t* cellPrefix,
                 const uint32_t* cellMass, const float4* cellBody, int initial) {
    (void)mats;
    if (ls.bodyBlocks == 0) return;
    hipLaunchKernelGGL(k_kick, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, pos, vel,
                       acc, force, cellPrefix, cellMass, cellBody, initial);
}

}  // namespace gs
