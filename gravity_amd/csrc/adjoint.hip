// Adjoint of the gravity step: vector-Jacobian products of the P3M
// acceleration field, for the reversed leapfrog of Engine::adjoint_vjp.
//
// The forward leapfrog step n (drift-only passes, no collisions) is
//   vh = v_n + dt/2 A(x_n),  x_{n+1} = x_n + dt vh,  v_{n+1} = vh + dt/2 A(x_{n+1}),
// with A(x_i) = sum over mesh nodes of W(x_i - node) F(node) (CIC gather of
// F = K * rho, rho = sum_k m_k W(x_k - node)) plus the near field of the
// kick (S/kick.comp:66-80, 118-151). Going backwards, with w = vbar + dt xbar
// (the adjoint of vh), each step needs g = J_A(x_n)^T w once: the two
// half-kicks that read A(x_n) both see w. The terms of J_A^T w:
//   - the gather's own-position derivative: grad of W(x_i - node) against
//     w_i . F(node);
//   - the mesh: w spread with the CIC weights (transposed gather), the
//     transposed convolution K^T = -K (the force kernels are odd), i.e. the
//     forward FFT pipeline applied to wx - i wy, minus the real part of its
//     result is rhobar; each body then takes m_i grad W(x_i - node) against
//     rhobar (the deposit's derivative);
//   - the near field: the derivative of every term with respect to the
//     body's own position, and, for every body k that sees body i exactly,
//     the derivative of k's term with respect to i's position (gathered, so
//     no atomics). The near-field topology (cell membership, monopoles,
//     the home-cell cap) is frozen: it is piecewise constant in x.
// The fixed-point mass deposit is differentiated as the CIC deposit it
// rounds. All of it is FP32 with contraction (a tolerance path).
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

__device__ __forceinline__ float adj_min_image(const SimDesc& s, float d) {
    return s.bounded ? d : d - floorf(d + 0.5f);
}

// d(w . A)/d(px, py) for one near-field term A = C m phi(r2) (px/aspect, py),
// phi = exp(-r2 / 2 sigma^2) / max(r2, minDist2), px = dx * aspect, py = dy.
__device__ __forceinline__ void near_term_vjp(const SimDesc& s, float px, float py, float mass,
                                              float wx, float wy, float& dpx, float& dpy) {
#pragma clang fp contract(fast)
    const float k2 = s.invTwoSigma2;
    float r2 = px * px + py * py;
    float e = __expf(-r2 * k2);
    bool clamped = !(r2 > s.minDist2);
    float q = __builtin_amdgcn_rcpf(clamped ? s.minDist2 : r2);
    float phi = e * q;
    float dphi = -k2 * phi - (clamped ? 0.0f : phi * q);
    float sw = wx * px * s.invAspect + wy * py;
    float cm = s.gOverM * mass;
    dpx = cm * (2.0f * dphi * px * sw + phi * wx * s.invAspect);
    dpy = cm * (2.0f * dphi * py * sw + phi * wy);
}

__device__ __forceinline__ float2 adjoint_w(const float2* __restrict__ xbar,
                                            const float2* __restrict__ vbar, int i, float dt) {
    float2 a = xbar[i], b = vbar[i];
    return make_float2(b.x + dt * a.x, b.y + dt * a.y);
}

// The near-field terms of J_A^T w for any radius, a cell at a time: this
// body's own terms (S/kick.comp:118-151) against its position, and the terms
// of the bodies that see this body exactly against this body's position.
__device__ __forceinline__ void near_vjp_any(const SimDesc& s, const uint4* __restrict__ cellRec,
                                             const float4* __restrict__ cellBody,
                                             const float2* __restrict__ xbar,
                                             const float2* __restrict__ vbar, float dt, int g,
                                             float4 p, float mass, float2 w, int myCount,
                                             int myRank, float& gx, float& gy) {
#pragma clang fp contract(fast)
    const float cellMassUnit = 1.0f / 256.0f;
    int hx = (int)floorf(p.x * (float)s.pw), hy = (int)floorf(p.y * (float)s.ph);
    for (int dy = -s.ry; dy <= s.ry; dy++) {
        int cy = hy + dy;
        if (s.bounded && (cy < 0 || cy >= s.ph)) continue;
        int wy = wrap_axis(cy, s.ph);
        float cpy = adj_min_image(s, ((float)wy + 0.5f) * s.invPh - p.y);
        const uint4* row = cellRec + (s.physBase + wy * s.pw);
        for (int dx = -s.rx; dx <= s.rx; dx++) {
            int cx = hx + dx;
            if (s.bounded && (cx < 0 || cx >= s.pw)) continue;
            int wx = wrap_axis(cx, s.pw);
            uint4 rec = row[wx];
            int count = (int)rec.y;
            if (count == 0) continue;
            bool isHome = dx == 0 && dy == 0;
            float dpx, dpy;
            if (!isHome && count > s.exactLimit) {
                float cpx = adj_min_image(s, ((float)wx + 0.5f) * s.invPw - p.x) * s.aspect;
                near_term_vjp(s, cpx, cpy, (float)rec.z * cellMassUnit, w.x, w.y, dpx, dpy);
                gx -= s.aspect * dpx;
                gy -= dpy;
            } else {
                int limit = isHome ? s.homeLimit : s.exactLimit;
                int stop = (int)rec.x + (count < limit ? count : limit);
                for (int t = (int)rec.x; t < stop; t++) {
                    float4 o = cellBody[t];
                    if (__float_as_uint(o.w) == (uint32_t)g) continue;
                    float ox = adj_min_image(s, o.x - p.x) * s.aspect;
                    float oy = adj_min_image(s, o.y - p.y);
                    near_term_vjp(s, ox, oy, o.z, w.x, w.y, dpx, dpy);
                    gx -= s.aspect * dpx;
                    gy -= dpy;
                }
            }
            bool seen = isHome ? myRank < s.homeLimit : myCount <= s.exactLimit;
            if (!seen) continue;
            for (int t = (int)rec.x; t < (int)rec.x + count; t++) {
                float4 o = cellBody[t];
                int k = (int)__float_as_uint(o.w);
                if (k == g) continue;
                float2 wk = adjoint_w(xbar, vbar, k, dt);
                float ox = adj_min_image(s, p.x - o.x) * s.aspect;
                float oy = adj_min_image(s, p.y - o.y);
                near_term_vjp(s, ox, oy, mass, wk.x, wk.y, dpx, dpy);
                gx += s.aspect * dpx;
                gy += dpy;
            }
        }
    }
}

// The same terms for rx = ry = R, a stencil row at a time: the row's records
// load together, then the first body of every cell this body takes or is
// seen by (and the adjoint of each body that sees it), then the second ones,
// then any further ones, so a row costs a few memory round trips however
// many of its cells hold bodies taken one by one.
template <int R>
__device__ __forceinline__ void near_vjp_rows(const SimDesc& s, const uint4* __restrict__ cellRec,
                                              const float4* __restrict__ cellBody,
                                              const float2* __restrict__ xbar,
                                              const float2* __restrict__ vbar,
                                              const float2* __restrict__ wCell, float dt, int g,
                                              float4 p, float mass, float2 w, int myCount,
                                              int myRank, float& gx, float& gy) {
#pragma clang fp contract(fast)
    const float cellMassUnit = 1.0f / 256.0f;
    const int exactLimit = s.exactLimit, homeLimit = s.homeLimit;
    constexpr int D = 2 * R + 1;
    int hx = (int)floorf(p.x * (float)s.pw), hy = (int)floorf(p.y * (float)s.ph);
    uint32_t colOk = 0u;
    float cpx[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        int cx = hx + k - R;
        if (!s.bounded || (cx >= 0 && cx < s.pw)) colOk |= 1u << k;
        int wx = wrap_axis(cx, s.pw);
        cpx[k] = adj_min_image(s, ((float)wx + 0.5f) * s.invPw - p.x) * s.aspect;
    }
    // A body sees this one exactly if it shares its home cell and this body is
    // among the first homeLimit entries, or if this body's cell is exact.
    const bool seenOutside = myCount <= exactLimit;
    const bool seenHome = myRank < homeLimit;
#pragma unroll 1
    for (int j = 0; j < D; j++) {
        int cy = hy + j - R;
        if (s.bounded && (cy < 0 || cy >= s.ph)) continue;
        int wy = wrap_axis(cy, s.ph);
        float cpy = adj_min_image(s, ((float)wy + 0.5f) * s.invPh - p.y);
        const uint4* row = cellRec + (s.physBase + wy * s.pw);
        const bool homeRow = j == R;
        // Every wrapped column is a valid record: the loads are unconditional
        // (no load waits on a branch) and a column that does not take part
        // reads as an empty cell.
        uint4 rec[D];
#pragma unroll
        for (int k = 0; k < D; k++) rec[k] = row[wrap_axis(hx + k - R, s.pw)];
#pragma unroll
        for (int k = 0; k < D; k++)
            if (!((colOk >> k) & 1u)) rec[k].y = 0u;
        int take[D], need[D];
        uint32_t seenMask = 0u;
#pragma unroll
        for (int k = 0; k < D; k++) {
            int count = (int)rec[k].y;
            bool home = homeRow && k == R;
            take[k] = home ? min(count, homeLimit) : (count <= exactLimit ? count : 0);
            bool seen = count > 0 && (home ? seenHome : seenOutside);
            if (seen) seenMask |= 1u << k;
            need[k] = seen ? count : take[k];
        }
        float dpx, dpy;
#pragma unroll
        for (int k = 0; k < D; k++) {
            int count = (int)rec[k].y;
            bool home = homeRow && k == R;
            if (!home && count > exactLimit) {
                near_term_vjp(s, cpx[k], cpy, (float)rec[k].z * cellMassUnit, w.x, w.y, dpx, dpy);
                gx -= s.aspect * dpx;
                gy -= dpy;
            }
        }
        // Entries 0 and 1 of every cell, each round's loads together and
        // unconditional: an entry that is not needed reads slot 0 and its
        // adjoint this body's own.
#pragma unroll
        for (int t = 0; t < 2; t++) {
            float4 o[D];
#pragma unroll
            for (int k = 0; k < D; k++) o[k] = cellBody[need[k] > t ? rec[k].x + (uint32_t)t : 0u];
            // The entries' adjoint weights by slot, loaded with the entries.
            float2 wk[D];
#pragma unroll
            for (int k = 0; k < D; k++) wk[k] = wCell[need[k] > t ? rec[k].x + (uint32_t)t : 0u];
#pragma unroll
            for (int k = 0; k < D; k++) {
                if (need[k] <= t || __float_as_uint(o[k].w) == (uint32_t)g) continue;
                if (take[k] > t) {
                    float ox = adj_min_image(s, o[k].x - p.x) * s.aspect;
                    float oy = adj_min_image(s, o[k].y - p.y);
                    near_term_vjp(s, ox, oy, o[k].z, w.x, w.y, dpx, dpy);
                    gx -= s.aspect * dpx;
                    gy -= dpy;
                }
                if ((seenMask >> k) & 1u) {
                    float ox = adj_min_image(s, p.x - o[k].x) * s.aspect;
                    float oy = adj_min_image(s, p.y - o[k].y);
                    near_term_vjp(s, ox, oy, mass, wk[k].x, wk[k].y, dpx, dpy);
                    gx += s.aspect * dpx;
                    gy += dpy;
                }
            }
            // the next round's loads after this round's values are consumed (registers)
            asm volatile("" ::: "memory");
        }
        // Further entries: the home cell's, and every entry of a cell that sees
        // this body and holds more than two; FURTHER_BATCH at a time, the
        // loads clamped to the cell's list.
        constexpr int FURTHER_BATCH = 4;
#pragma unroll
        for (int k = 0; k < D; k++) {
            for (int t0 = 2; t0 < need[k]; t0 += FURTHER_BATCH) {
                float4 o[FURTHER_BATCH];
                float2 wo[FURTHER_BATCH];
#pragma unroll
                for (int q = 0; q < FURTHER_BATCH; q++) {
                    const uint32_t slot = rec[k].x + (uint32_t)min(t0 + q, need[k] - 1);
                    o[q] = cellBody[slot];
                    wo[q] = wCell[slot];
                }
#pragma unroll
                for (int q = 0; q < FURTHER_BATCH; q++) {
                    const int t = t0 + q;
                    if (t >= need[k] || (int)__float_as_uint(o[q].w) == g) continue;
                    if (take[k] > t) {
                        float ox = adj_min_image(s, o[q].x - p.x) * s.aspect;
                        float oy = adj_min_image(s, o[q].y - p.y);
                        near_term_vjp(s, ox, oy, o[q].z, w.x, w.y, dpx, dpy);
                        gx -= s.aspect * dpx;
                        gy -= dpy;
                    }
                    if ((seenMask >> k) & 1u) {
                        float ox = adj_min_image(s, p.x - o[q].x) * s.aspect;
                        float oy = adj_min_image(s, p.y - o[q].y);
                        near_term_vjp(s, ox, oy, mass, wo[q].x, wo[q].y, dpx, dpy);
                        gx += s.aspect * dpx;
                        gy += dpy;
                    }
                }
            }
        }
    }
}

// Per body: the gather's own-position term and the near-field terms of
// J_A^T w; also the value the transposed gather spreads, (wx, -wy).
// The adjoint weight w = vbar + dt * xbar of the body in every cell slot, so
// the near-field VJP loads a neighbour's weight with its entry (by slot)
// instead of through the entry's body index. Cell slots count the bodies of
// the simulations with mutual gravity only while body indices count every
// simulation's capacity: the two agree for the one simulation starting at
// body 0 that the engine's adjoint entry points require.
__global__ void __launch_bounds__(WG) k_adj_cell_weights(const SimDesc* __restrict__ sims,
                                                         const BodyBlock* __restrict__ blocks,
                                                         const float4* __restrict__ cellBody,
                                                         const float2* __restrict__ xbar,
                                                         const float2* __restrict__ vbar, float dt,
                                                         float2* __restrict__ wCell) {
#pragma clang fp contract(fast)
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int slot = b.first + (int)threadIdx.x;
    if (slot - s.base >= s.n || !s.mutual) return;
    // (the id is clamped to the simulation's bodies: every slot of a
    // simulation with mutual gravity holds one of them)
    const uint32_t id = min(__float_as_uint(cellBody[slot].w), (uint32_t)(s.base + s.n - 1));
    wCell[slot] = adjoint_w(xbar, vbar, (int)id, dt);
}

__global__ void __launch_bounds__(WG) k_adj_direct(const SimDesc* __restrict__ sims,
                                                   const BodyBlock* __restrict__ blocks,
                                                   const float4* __restrict__ body,
                                                   const float4* __restrict__ mats,
                                                   const float2* __restrict__ force,
                                                   const uint4* __restrict__ cellRec,
                                                   const float4* __restrict__ cellBody,
                                                   const float2* __restrict__ xbar,
                                                   const float2* __restrict__ vbar,
                                                   const float2* __restrict__ wCell, float dt,
                                                   float2* __restrict__ gbar,
                                                   float2* __restrict__ spread) {
#pragma clang fp contract(fast)
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float4 p = body[g];
    float mass = mats[g].x;
    float2 w = adjoint_w(xbar, vbar, g, dt);
    spread[g] = make_float2(w.x, -w.y);
    float gx = 0.0f, gy = 0.0f;
    // CIC gather: A = sum of bilinear weights times F; its gradient in x.
    if (s.mutual) {
        float fgx = p.x * (float)s.pw - 0.5f, fgy = p.y * (float)s.ph - 0.5f;
        float flx = floorf(fgx), fly = floorf(fgy);
        int bx = (int)flx, by = (int)fly;
        float fx = fgx - flx, fy = fgy - fly;
        int x0 = wrap_axis(bx, s.mw), x1 = wrap_axis(bx + 1, s.mw);
        int y0 = wrap_axis(by, s.mh), y1 = wrap_axis(by + 1, s.mh);
        const float2* F = force + s.meshBase;
        float2 v00 = F[x0 + y0 * s.mw], v10 = F[x1 + y0 * s.mw];
        float2 v01 = F[x0 + y1 * s.mw], v11 = F[x1 + y1 * s.mw];
        float d00 = w.x * v00.x + w.y * v00.y, d10 = w.x * v10.x + w.y * v10.y;
        float d01 = w.x * v01.x + w.y * v01.y, d11 = w.x * v11.x + w.y * v11.y;
        gx += (float)s.pw * ((d10 - d00) * (1.0f - fy) + (d11 - d01) * fy);
        gy += (float)s.ph * ((d01 - d00) * (1.0f - fx) + (d11 - d10) * fx);
    }
    if (s.mutual) {
        int hx = (int)floorf(p.x * (float)s.pw), hy = (int)floorf(p.y * (float)s.ph);
        int home = s.physBase + wrap_axis(hx, s.pw) + wrap_axis(hy, s.ph) * s.pw;
        uint4 own = cellRec[home];
        // The home cell holds this body, so its list is never empty; its
        // entries load RANK_BATCH at a time, clamped to the list.
        constexpr int RANK_BATCH = 8;
        int myCount = (int)own.y, myRank = 0;
        for (int t0 = 0; t0 < myCount; t0 += RANK_BATCH) {
            uint32_t id[RANK_BATCH];
#pragma unroll
            for (int q = 0; q < RANK_BATCH; q++)
                id[q] = __float_as_uint(cellBody[own.x + (uint32_t)min(t0 + q, myCount - 1)].w);
#pragma unroll
            for (int q = 0; q < RANK_BATCH; q++)
                if (t0 + q < myCount && id[q] == (uint32_t)g) myRank = t0 + q;
        }
        if (s.rx == 3 && s.ry == 3)
            near_vjp_rows<3>(s, cellRec, cellBody, xbar, vbar, wCell, dt, g, p, mass, w, myCount,
                             myRank, gx, gy);
        else
            near_vjp_any(s, cellRec, cellBody, xbar, vbar, dt, g, p, mass, w, myCount, myRank, gx,
                         gy);
    }
    gbar[g] = make_float2(gx, gy);
}

// The deposit's derivative against rhobar = -Re(result of the transposed
// convolution), then the step's update: xbar += coef * g, vbar = w.
__global__ void __launch_bounds__(WG) k_adj_finish(const SimDesc* __restrict__ sims,
                                                   const BodyBlock* __restrict__ blocks,
                                                   const float4* __restrict__ body,
                                                   const float4* __restrict__ mats,
                                                   const float2* __restrict__ result,
                                                   const float2* __restrict__ gbar,
                                                   float2* __restrict__ xbar,
                                                   float2* __restrict__ vbar, float dt,
                                                   float coef) {
#pragma clang fp contract(fast)
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    float4 p = body[g];
    float2 gsum = gbar[g];
    if (s.mutual) {
        float m = mats[g].x;
        float fgx = p.x * (float)s.pw - 0.5f, fgy = p.y * (float)s.ph - 0.5f;
        float flx = floorf(fgx), fly = floorf(fgy);
        int bx = (int)flx, by = (int)fly;
        float fx = fgx - flx, fy = fgy - fly;
        int x0 = wrap_axis(bx, s.mw), x1 = wrap_axis(bx + 1, s.mw);
        int y0 = wrap_axis(by, s.mh), y1 = wrap_axis(by + 1, s.mh);
        const float2* R = result + s.meshBase;
        float r00 = -R[x0 + y0 * s.mw].x, r10 = -R[x1 + y0 * s.mw].x;
        float r01 = -R[x0 + y1 * s.mw].x, r11 = -R[x1 + y1 * s.mw].x;
        gsum.x += m * (float)s.pw * ((r10 - r00) * (1.0f - fy) + (r11 - r01) * fy);
        gsum.y += m * (float)s.ph * ((r01 - r00) * (1.0f - fx) + (r11 - r10) * fx);
    }
    float2 w = adjoint_w(xbar, vbar, g, dt);
    float2 xb = xbar[g];
    xbar[g] = make_float2(xb.x + coef * gsum.x, xb.y + coef * gsum.y);
    vbar[g] = w;
}

// Undoes a spatial sort on the adjoint state: the value of old body i is
// the sorted value at dest[i].
__global__ void __launch_bounds__(WG) k_adj_unsort(const SimDesc* __restrict__ sims,
                                                   const BodyBlock* __restrict__ blocks,
                                                   const uint32_t* __restrict__ dest,
                                                   const float2* __restrict__ xin,
                                                   const float2* __restrict__ vin,
                                                   float2* __restrict__ xout,
                                                   float2* __restrict__ vout) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    uint32_t d = dest[g];
    xout[g] = xin[d];
    vout[g] = vin[d];
}

// The identity permutation of the recording's start: where[g] = g.
__global__ void __launch_bounds__(WG) k_adj_iota(const SimDesc* __restrict__ sims,
                                                 const BodyBlock* __restrict__ blocks,
                                                 uint32_t* __restrict__ where) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    where[g] = (uint32_t)g;
}

// Composes a sort into the running permutation: where[i] is the index now
// holding the body that started at index i.
__global__ void __launch_bounds__(WG) k_adj_compose(const SimDesc* __restrict__ sims,
                                                    const BodyBlock* __restrict__ blocks,
                                                    const uint32_t* __restrict__ dest,
                                                    uint32_t* __restrict__ where) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    where[g] = dest[where[g]];
}

void launch_adj_direct(hipStream_t st, const LaunchShape& ls, const float4* body,
                       const float4* mats, const float2* force, const uint4* cellRec,
                       const float4* cellBody, const float2* xbar, const float2* vbar, float dt,
                       float2* wCell, float2* gbar, float2* spread) {
    if (ls.bodyBlocks == 0) return;
    launch(k_adj_cell_weights, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, cellBody,
           xbar, vbar, dt, wCell);
    launch(k_adj_direct, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       body, mats, force, cellRec, cellBody, xbar, vbar, wCell, dt, gbar, spread);
}

void launch_adj_finish(hipStream_t st, const LaunchShape& ls, const float4* body,
                       const float4* mats, const float2* result, const float2* gbar, float2* xbar,
                       float2* vbar, float dt, float coef) {
    if (ls.bodyBlocks == 0) return;
    launch(k_adj_finish, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       body, mats, result, gbar, xbar, vbar, dt, coef);
}

void launch_adj_unsort(hipStream_t st, const LaunchShape& ls, const uint32_t* dest,
                       const float2* xin, const float2* vin, float2* xout, float2* vout) {
    if (ls.bodyBlocks == 0) return;
    launch(k_adj_unsort, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       dest, xin, vin, xout, vout);
}

void launch_adj_iota(hipStream_t st, const LaunchShape& ls, uint32_t* where) {
    if (ls.bodyBlocks == 0) return;
    launch(k_adj_iota, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, where);
}

void launch_adj_compose(hipStream_t st, const LaunchShape& ls, const uint32_t* dest,
                        uint32_t* where) {
    if (ls.bodyBlocks == 0) return;
    launch(k_adj_compose, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       dest, where);
}

}  // namespace gs
