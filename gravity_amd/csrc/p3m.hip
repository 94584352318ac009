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

// S/build_grid.comp:71-74: the home cell of every body; one atomic per run
// of equal cells in a wave.
__global__ void __launch_bounds__(WG) k_bin(const SimDesc* __restrict__ sims,
                                            const BodyBlock* __restrict__ blocks,
                                            const float4* __restrict__ body, uint32_t* cellCount,
                                            uint32_t* __restrict__ homeOf) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    const bool active = g - s.base < s.n && s.mutual;
    uint32_t c = 0xffffffffu;
    if (active) {
        float4 p = body[g];
        c = (uint32_t)home_cell(s, p.x, p.y);
        homeOf[g] = c;
    }
    const LaneRun run = lane_run(c);
    if (active && run.rank == 0) atomicAdd(&cellCount[c], (uint32_t)run.length);
}

__global__ void __launch_bounds__(WG) k_cell_scatter(const SimDesc* __restrict__ sims,
                                                     const BodyBlock* __restrict__ blocks,
                                                     const uint32_t* __restrict__ homeOf,
                                                     uint32_t* cellCount,
                                                     const uint32_t* __restrict__ cellPrefix,
                                                     uint32_t* __restrict__ slotBody,
                                                     uint32_t* __restrict__ slotCell) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    const bool active = g - s.base < s.n && s.mutual;
    const uint32_t c = active ? homeOf[g] : 0xffffffffu;
    const LaneRun run = lane_run(c);
    uint32_t top = 0u;
    if (active && run.rank == 0) top = atomicSub(&cellCount[c], (uint32_t)run.length);
    top = (uint32_t)__shfl((int)top, run.start, 64);
    if (!active) return;
    uint32_t rank = top - 1u - (uint32_t)run.rank;
    const uint32_t slot = cellPrefix[c] + rank;
    slotBody[slot] = (uint32_t)g;
    slotCell[slot] = c;
}

// Orders each cell by descending body index and sums the cell mass in 1/256
// units (S/build_grid.comp:75), exactly: the slots of a cell are contiguous,
// so a wave adds its runs of one cell in registers; a run that is a whole
// cell stores its sum, any other accumulates it with one atomic into the
// cell record and the near record (whose masses the cell scan zeroed). The
// bodies of a cell of at most NEAR_INLINE go into its near record as well.
// A cell of more than big.limit bodies is listed for k_cell_canon_big, which
// orders it.
__global__ void __launch_bounds__(WG) k_cell_canon(int slots, const uint32_t* __restrict__ slotCell,
                                                   const uint32_t* __restrict__ cellPrefix,
                                                   const uint32_t* __restrict__ slotBody,
                                                   const float4* __restrict__ body,
                                                   const float4* __restrict__ mats,
                                                   float4* __restrict__ cellBody,
                                                   uint4* __restrict__ cellRec,
                                                   uint4* __restrict__ cellNear, BigSegments big) {
    int d = (int)(blockIdx.x * WG + threadIdx.x);
    const bool active = d < slots;
    uint32_t c = 0xffffffffu, units = 0u;
    int s0 = 0, s1 = 0;
    if (active) {
        uint32_t g = slotBody[d];
        c = slotCell[d];
        // the body's own data load with its cell's bounds, ahead of the rank
        const float4 p = body[g];
        const float m = mats[g].x;
        s0 = (int)cellPrefix[c];
        s1 = (int)cellPrefix[c + 1];
        units = (uint32_t)rintf(m * 256.0f);
        if (is_big_segment(big, s1 - s0)) {
            if (d == s0) list_big_segment(big, s0, s1, 0);
        } else {
            const int rank = segment_rank<true>(slotBody, s0, s1, g);
            cellBody[s0 + rank] = make_float4(p.x, p.y, m, __uint_as_float(g));
            if (s1 - s0 <= NEAR_INLINE) {
                float* rec = reinterpret_cast<float*>(cellNear + 2 * (size_t)c);
                if (rank == 0) {
                    *reinterpret_cast<float2*>(rec + 2) = make_float2(p.x, p.y);
                    rec[7] = m;
                } else {
                    *reinterpret_cast<float2*>(rec + 4) = make_float2(p.x, p.y);
                    rec[6] = m;
                }
            }
        }
    }
    const LaneRun run = lane_run(c);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = (uint32_t)__shfl_up((int)units, (unsigned)o, 64);
        if (run.rank >= o) units += t;
    }
    if (active && run.rank == run.length - 1) {
        uint32_t* mass = &reinterpret_cast<uint32_t*>(cellRec)[4 * (size_t)c + 2];
        uint32_t* nearMass = &reinterpret_cast<uint32_t*>(cellNear)[8 * (size_t)c + 1];
        if (s0 + run.length == s1) {
            // the run is the whole cell: no other wave adds to it
            *mass = units;
            *nearMass = units;
        } else {
            atomicAdd(mass, units);
            atomicAdd(nearMass, units);
        }
    }
}

// The cells k_cell_canon listed as big, each ordered by descending body
// index by one workgroup.
__global__ void __launch_bounds__(BIG_NT) k_cell_canon_big(const uint32_t* __restrict__ slotBody,
                                                           const float4* __restrict__ body,
                                                           const float4* __restrict__ mats,
                                                           float4* __restrict__ cellBody,
                                                           BigSegments big) {
    order_big_segments<true>(big, slotBody, [&](const BigSegment& seg, int rank, uint32_t g) {
        const float4 p = body[g];
        cellBody[seg.s0 + (uint32_t)rank] = make_float4(p.x, p.y, mats[g].x, __uint_as_float(g));
    });
}

__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// The near field's constants, read once per body into registers: the
// per-cell helpers below take either this or the SimDesc itself (the same
// member names), and a loop over the stencil then reloads nothing. Walls or
// torus are part of the type, so a body term under walls carries no minimum
// image.
template <bool BOUNDED>
struct NearK {
    static constexpr bool bounded = BOUNDED;
    int pw, ph, exactLimit, homeLimit;
    float aspect, invAspect, gOverM, minDist2;
};
template <bool BOUNDED>
__device__ __forceinline__ NearK<BOUNDED> near_k(const SimDesc& s) {
    return NearK<BOUNDED>{s.pw, s.ph, s.exactLimit, s.homeLimit,
                          s.aspect, s.invAspect, s.gOverM, s.minDist2};
}

template <typename K>
__device__ __forceinline__ float min_image(const K& s, float d) {
    return s.bounded ? d : d - floorf(d + 0.5f);
}

// Short-range term of one source (S/kick.comp:66-80) from its physical
// displacement (px = dx * aspect, py = dy), its squared length r2, its
// Gaussian weight and its mass: G m w / (M max(r2, minDist2)) * (dx, dy).
template <typename K>
__device__ __forceinline__ void near_term(const K& s, float px, float py, float r2,
                                          float weight, float mass, float& ax, float& ay) {
#pragma clang fp contract(fast)
    float q = s.gOverM * mass * weight * __builtin_amdgcn_rcpf(fmaxf(r2, s.minDist2));
    ax += q * (px * s.invAspect);
    ay += q * py;
}

// The bodies of one cell taken one by one (a home cell, or a cell with at
// most exactNeighborBodyLimit bodies), skipping the body itself.
__device__ __forceinline__ void exact_cell(const SimDesc& s, const float4* __restrict__ cellBody,
                                           int first, int count, int g, float x, float y, float k2,
                                           float& ax, float& ay) {
#pragma clang fp contract(fast)
    for (int t = first; t < first + count; t++) {
        float4 o = cellBody[t];
        if (__float_as_uint(o.w) == (uint32_t)g) continue;
        float ox = min_image(s, o.x - x) * s.aspect;
        float oy = min_image(s, o.y - y);
        float r2 = ox * ox + oy * oy;
        near_term(s, ox, oy, r2, __expf(-r2 * k2), o.z, ax, ay);
    }
}

// S/kick.comp:82-103: CIC gather of the mesh force (x = Fx, y = Fy).
__device__ __forceinline__ float2 mesh_force(const SimDesc& s, const float2* __restrict__ force,
                                             float x, float y) {
    float gx = x * (float)s.pw - 0.5f, gy = y * (float)s.ph - 0.5f;
    float flx = floorf(gx), fly = floorf(gy);
    int bx = (int)flx, by = (int)fly;
    float fx = gx - flx, fy = gy - fly;
    int x0 = wrap_axis(bx, s.mw), x1 = wrap_axis(bx + 1, s.mw);
    int y0 = wrap_axis(by, s.mh), y1 = wrap_axis(by + 1, s.mh);
    const float2* F = force + s.meshBase;
    float2 v00 = F[x0 + y0 * s.mw], v10 = F[x1 + y0 * s.mw];
    float2 v01 = F[x0 + y1 * s.mw], v11 = F[x1 + y1 * s.mw];
    return make_float2(mixf(mixf(v00.x, v10.x, fx), mixf(v01.x, v11.x, fx), fy),
                       mixf(mixf(v00.y, v10.y, fx), mixf(v01.y, v11.y, fx), fy));
}

// The kick's near field (S/kick.comp:118-151) is the reference's sum over
// the (2rx+1) x (2ry+1)
// physical cells around the home cell: the home cell body by body (at most
// maximumHomeBodyInteractions list entries, the body itself skipped), cells
// with more than exactNeighborBodyLimit bodies as a monopole at the cell
// centre, the others body by body. The near field is evaluated on the
// FP32-tolerance path (contraction allowed, reciprocal products, hardware
// exp2), and a monopole's Gaussian is the product of its column and row
// factors; the mesh gather, the field and the half kick round as written.

// One source taken body by body (x, y its normalised position, z its mass).
template <typename K>
__device__ __forceinline__ void body_term(const K& s, float4 o, float x, float y, float k2,
                                          float& ax, float& ay) {
#pragma clang fp contract(fast)
    float ox = min_image(s, o.x - x) * s.aspect;
    float oy = min_image(s, o.y - y);
    float r2 = ox * ox + oy * oy;
    near_term(s, ox, oy, r2, __expf(-r2 * k2), o.z, ax, ay);
}

// Near field of rx = ry = R cells for exactNeighborBodyLimit <= NEAR_INLINE:
// the column offsets and Gaussian factors of a body are computed once; the
// home cell's bodies (its record hrec is loaded with the mesh force) load
// together; then each stencil row's 32-byte near records, which hold the
// count, the mass and, for a cell taken body by body, its bodies, from the
// workgroup's LDS stage of the records around its bodies or, for a block
// whose box does not fit the stage, from global memory (near_field_rows,
// the same code and the same order for both).

// The home cell: its first homeLimit list entries, the body itself skipped
// (under walls, none for a body on or beyond the far edge).
template <typename K>
__device__ __forceinline__ void near_home(const K& s, uint4 hrec,
                                          const float4* __restrict__ cellBody, int g, int hx,
                                          int hy, float x, float y, float k2, float& ax,
                                          float& ay) {
#pragma clang fp contract(fast)
    constexpr int HOME_BATCH = 8;
    const bool homeOk = !s.bounded || (hx >= 0 && hx < s.pw && hy >= 0 && hy < s.ph);
    const int take = homeOk ? min((int)hrec.y, s.homeLimit) : 0;
    // The home cell holds the body itself, so its list is never empty and
    // the loads, clamped to the list, are unconditional (no load waits on
    // a branch).
    const int last = max((int)hrec.y, 1) - 1;
    float4 hb[HOME_BATCH];
#pragma unroll
    for (int t = 0; t < HOME_BATCH; t++) hb[t] = cellBody[hrec.x + (uint32_t)min(t, last)];
#pragma unroll
    for (int t = 0; t < HOME_BATCH; t++)
        if (t < take && __float_as_uint(hb[t].w) != (uint32_t)g) body_term(s, hb[t], x, y, k2, ax, ay);
    for (int t = HOME_BATCH; t < take; t++) {
        float4 o = cellBody[hrec.x + (uint32_t)t];
        if (__float_as_uint(o.w) != (uint32_t)g) body_term(s, o, x, y, k2, ax, ay);
    }
}

// The column offsets, Gaussian column factors and the columns that take part.
template <int R>
__device__ __forceinline__ uint32_t near_columns(const SimDesc& s, int hx, float x, float k2,
                                                 float (&cpx)[2 * R + 1], float (&cex)[2 * R + 1]) {
#pragma clang fp contract(fast)
    uint32_t colOk = 0u;
#pragma unroll
    for (int k = 0; k < 2 * R + 1; k++) {
        int cx = hx + k - R;
        if (!s.bounded || (cx >= 0 && cx < s.pw)) colOk |= 1u << k;
        int wx = wrap_axis(cx, s.pw);
        float px = min_image(s, ((float)wx + 0.5f) * s.invPw - x) * s.aspect;
        cpx[k] = px;
        cex[k] = __expf(-(px * px) * k2);
    }
    return colOk;
}

// One stencil cell of a row from its near record (lo, hi): a monopole at the
// cell centre, or its one or two bodies.
template <typename K>
__device__ __forceinline__ void near_cell(const K& s, uint32_t count, uint4 lo, uint4 hi,
                                          float px, float ex, float py, float py2, float ey,
                                          float x, float y, float k2, float& ax, float& ay) {
#pragma clang fp contract(fast)
    const float cellMassUnit = 1.0f / 256.0f;
    if (count > (uint32_t)s.exactLimit) {
        near_term(s, px, py, px * px + py2, ex * ey, (float)lo.y * cellMassUnit, ax, ay);
    } else {
        if (count > 0u)
            body_term(s, make_float4(__uint_as_float(lo.z), __uint_as_float(lo.w),
                                     __uint_as_float(hi.w), 0.0f),
                      x, y, k2, ax, ay);
        if (count > 1u)
            body_term(s, make_float4(__uint_as_float(hi.x), __uint_as_float(hi.y),
                                     __uint_as_float(hi.z), 0.0f),
                      x, y, k2, ax, ay);
    }
}

// The row's y offset, its Gaussian factor, and whether it takes part.
__device__ __forceinline__ bool near_row(const SimDesc& s, int cy, float y, float k2, int& wy,
                                         float& py, float& py2, float& ey) {
#pragma clang fp contract(fast)
    if (s.bounded && (cy < 0 || cy >= s.ph)) return false;
    wy = wrap_axis(cy, s.ph);
    py = min_image(s, ((float)wy + 0.5f) * s.invPh - y);
    py2 = py * py;
    ey = __expf(-py2 * k2);
    return true;
}

// The rows from `Rows`, a record source: rec(cy, cx) gives the near record
// of the cell at unwrapped coordinates (cx, cy), from the workgroup's LDS
// stage (StagedRows) or from global memory (GlobalRows, for a block whose box
// does not fit the stage: the sparse outskirts of a disc, where about a third
// of a galaxy's blocks lie). Reading a row cell by cell would wait for every
// record in turn; here the records of a part of a row (both halves of each)
// load together, unconditionally (every address is a record of the stage or
// of the simulation; a record of a cell outside the walls, or a half a cell
// does not use, is read and left alone), and the cells are then added one
// by one in row order. Three cells a part keep the kernel within the
// registers of seven waves a SIMD, the most its LDS stage allows.
struct StagedRows {
    const uint4* stage;
    int x0, y0, bw;
    __device__ __forceinline__ const uint4* rec(const SimDesc&, int cy, int cx) const {
        return stage + 2 * ((cy - y0) * bw + (cx - x0));
    }
};
struct GlobalRows {
    const uint4* __restrict__ cellNear;
    __device__ __forceinline__ const uint4* rec(const SimDesc& s, int cy, int cx) const {
        return cellNear + 2 * (size_t)(s.physBase + wrap_axis(cy, s.ph) * s.pw + wrap_axis(cx, s.pw));
    }
};

template <int R, bool BOUNDED, typename Rows>
__device__ __forceinline__ void near_field_rows(const SimDesc& s, uint4 hrec, const Rows& rows,
                                                const float4* __restrict__ cellBody, int g,
                                                float x, float y, float& ax, float& ay) {
    const float k2 = s.invTwoSigma2;
    const NearK<BOUNDED> c = near_k<BOUNDED>(s);
    int hx = (int)floorf(x * (float)s.pw), hy = (int)floorf(y * (float)s.ph);
    constexpr int D = 2 * R + 1;
    constexpr int PART = 3;
    near_home(c, hrec, cellBody, g, hx, hy, x, y, k2, ax, ay);
    float cpx[D], cex[D];
    const uint32_t colOk = near_columns<R>(s, hx, x, k2, cpx, cex);
#pragma unroll 1
    for (int j = 0; j < D; j++) {
        int wy;
        float py, py2, ey;
        const int cy = hy + j - R;
        if (!near_row(s, cy, y, k2, wy, py, py2, ey)) continue;
        const uint32_t rowOk = j == R ? colOk & ~(1u << R) : colOk;
#pragma unroll
        for (int k0 = 0; k0 < D; k0 += PART) {
            uint4 lo[PART], hi[PART];
#pragma unroll
            for (int q = 0; q < PART; q++) {
                const uint4* rec = rows.rec(s, cy, hx + min(k0 + q, D - 1) - R);
                lo[q] = rec[0];
                hi[q] = rec[1];
            }
#pragma unroll
            for (int q = 0; q < PART; q++) {
                const int k = k0 + q;
                if (k < D) {
                    const uint32_t count = (rowOk >> k) & 1u ? lo[q].x : 0u;
                    near_cell(c, count, lo[q], hi[q], cpx[k], cex[k], py, py2, ey, x, y, k2, ax,
                              ay);
                }
            }
        }
    }
}

// Near field of any radius, a cell at a time.
__device__ __forceinline__ void near_field_any(const SimDesc& s, const uint4* __restrict__ cellRec,
                                               const float4* __restrict__ cellBody, int g,
                                               float x, float y, float& ax, float& ay) {
#pragma clang fp contract(fast)
    const float k2 = s.invTwoSigma2;
    const float cellMassUnit = 1.0f / 256.0f;
    int hx = (int)floorf(x * (float)s.pw), hy = (int)floorf(y * (float)s.ph);
    for (int dy = -s.ry; dy <= s.ry; dy++) {
        int cy = hy + dy;
        if (s.bounded && (cy < 0 || cy >= s.ph)) continue;
        int wy = wrap_axis(cy, s.ph);
        float py = min_image(s, ((float)wy + 0.5f) * s.invPh - y);
        float py2 = py * py;
        float ey = __expf(-py2 * k2);
        const uint4* row = cellRec + (s.physBase + wy * s.pw);
        for (int dx = -s.rx; dx <= s.rx; dx++) {
            int cx = hx + dx;
            if (s.bounded && (cx < 0 || cx >= s.pw)) continue;
            int wx = wrap_axis(cx, s.pw);
            uint4 rec = row[wx];
            int count = (int)rec.y;
            bool home = dx == 0 && dy == 0;
            if (!home && count > s.exactLimit) {
                float px = min_image(s, ((float)wx + 0.5f) * s.invPw - x) * s.aspect;
                near_term(s, px, py, px * px + py2, __expf(-(px * px) * k2) * ey,
                          (float)rec.z * cellMassUnit, ax, ay);
            } else if (count > 0) {
                int limit = home ? s.homeLimit : s.exactLimit;
                exact_cell(s, cellBody, (int)rec.x, count < limit ? count : limit, g, x, y, k2,
                           ax, ay);
            }
        }
    }
}

// S/kick.comp:105-156. A simulation with the default 3-cell near field (and
// exact cells of at most NEAR_INLINE bodies) takes near_field_rows<3>, any
// other near_field_any; the choice is per simulation, so a simulation kicks
// the same alone or in a batch.
//
// For the fixed near field the workgroup first stages in LDS the near records
// of the box of cells its bodies' stencils cover (the home cells' bounding box
// grown by the radius), one coalesced copy of up to KICK_STAGE records, and
// its lanes read their rows from there: the 256 bodies of a spatially sorted
// block share most of their 49 stencil cells, which each would otherwise
// fetch from L2 on its own. A block whose box is larger (a sparse region, a
// block across the torus seam, a non-finite position) reads global memory.
constexpr int KICK_STAGE = 640;   // near records a workgroup stages (20 KB of LDS)

// WALLS: the simulations this launch kicks (1 under walls, 0 on the torus);
// a block of a simulation of the other kind leaves at once (a batch that
// mixes both launches both).
template <int WALLS>
__global__ void __launch_bounds__(WG) k_kick(const SimDesc* __restrict__ sims,
                                             const BodyBlock* __restrict__ blocks,
                                             float4* __restrict__ body,
                                             float2* __restrict__ acc,
                                             const float2* __restrict__ force,
                                             const uint4* __restrict__ cellRec,
                                             const uint4* __restrict__ cellNear,
                                             const float4* __restrict__ cellBody, int initial,
                                             SleepBufs sleep) {
    __shared__ uint4 stage[2 * KICK_STAGE];
    __shared__ int waveBox[WG / 64][4];
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    if (s.bounded != WALLS) return;
    int g = b.first + (int)threadIdx.x;
    // Every lane stays to the staging barriers; a lane without a body to kick
    // leaves after them.
    bool live = g - s.base < s.n;
    // S/kick.comp:109-110: a sleeping body is not kicked.
    if (live && !initial && s.sleepOn) {
        const SleepState st = sleep.state[b.sim];
        if (sleep_on(s, st) && sleep.bodyAwake[g] != st.stamp) live = false;
    }
    float h = 0.5f * s.dt;
    float4 p = live ? body[g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (!s.mutual) {
        if (!live) return;
        if (!initial) {
            p.z += s.fieldX * h;
            p.w += s.fieldY * h;
            body[g] = p;
        }
        acc[g] = make_float2(s.fieldX, s.fieldY);
        return;
    }
    const bool fixed = s.rx == 3 && s.ry == 3 && s.exactLimit <= NEAR_INLINE;
    bool staged = false;
    int x0 = 0, y0 = 0, bw = 0;
    if (fixed) {
        constexpr int R = 3;
        // the home cells' bounding box over the workgroup's bodies
        const int hx = live ? (int)floorf(p.x * (float)s.pw) : 0x3fffffff;
        const int hy = live ? (int)floorf(p.y * (float)s.ph) : 0x3fffffff;
        int lox = hx, hix = live ? hx : -0x3fffffff, loy = hy, hiy = live ? hy : -0x3fffffff;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lox = min(lox, __shfl_xor(lox, o));
            hix = max(hix, __shfl_xor(hix, o));
            loy = min(loy, __shfl_xor(loy, o));
            hiy = max(hiy, __shfl_xor(hiy, o));
        }
        const int wave = (int)(threadIdx.x >> 6);
        if ((threadIdx.x & 63u) == 0u) {
            waveBox[wave][0] = lox;
            waveBox[wave][1] = hix;
            waveBox[wave][2] = loy;
            waveBox[wave][3] = hiy;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < WG / 64; w++) {
            lox = min(lox, waveBox[w][0]);
            hix = max(hix, waveBox[w][1]);
            loy = min(loy, waveBox[w][2]);
            hiy = max(hiy, waveBox[w][3]);
        }
        // Only boxes of cells inside the grid are staged (which also keeps
        // the sizes below from overflowing).
        if (lox >= 0 && hix < s.pw && loy >= 0 && hiy < s.ph && lox <= hix && loy <= hiy) {
            x0 = lox - R;
            y0 = loy - R;
            bw = hix - lox + 2 * R + 1;
            const int bh = hiy - loy + 2 * R + 1;
            staged = bw * bh <= KICK_STAGE;
            if (staged) {
                for (int i = (int)threadIdx.x; i < bw * bh; i += WG) {
                    const int cy = y0 + i / bw, cx = x0 + i % bw;
                    const size_t c = (size_t)(s.physBase + wrap_axis(cy, s.ph) * s.pw + wrap_axis(cx, s.pw));
                    stage[2 * i] = cellNear[2 * c];
                    stage[2 * i + 1] = cellNear[2 * c + 1];
                }
            }
        }
        __syncthreads();
    }
    if (!live) return;
    float ax, ay;
    if (fixed) {
        const uint4 hrec = cellRec[home_cell(s, p.x, p.y)];
        float2 mf = mesh_force(s, force, p.x, p.y);
        ax = mf.x + s.fieldX;
        ay = mf.y + s.fieldY;
        if (staged)
            near_field_rows<3, WALLS != 0>(s, hrec, StagedRows{stage, x0, y0, bw}, cellBody, g,
                                           p.x, p.y, ax, ay);
        else
            near_field_rows<3, WALLS != 0>(s, hrec, GlobalRows{cellNear}, cellBody, g, p.x, p.y,
                                           ax, ay);
    } else {
        float2 mf = mesh_force(s, force, p.x, p.y);
        ax = mf.x + s.fieldX;
        ay = mf.y + s.fieldY;
        near_field_any(s, cellRec, cellBody, g, p.x, p.y, ax, ay);
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
    launch(k_bin, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, body,
                       cellCount, homeOf);
}

void launch_cell_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* homeOf,
                         uint32_t* cellCount, const uint32_t* cellPrefix, uint32_t* slotBody,
                         uint32_t* slotCell) {
    if (ls.bodyBlocks == 0) return;
    launch(k_cell_scatter, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       homeOf, cellCount, cellPrefix, slotBody, slotCell);
}

void launch_cell_canon(hipStream_t st, int slots, const uint32_t* slotCell,
                       const uint32_t* cellPrefix, const uint32_t* slotBody, const float4* body,
                       const float4* mats, float4* cellBody, uint4* cellRec, uint4* cellNear,
                       const BigSegments& big) {
    if (slots <= 0) return;
    launch(k_cell_canon, dim3((slots + WG - 1) / WG), dim3(WG), 0, st, slots, slotCell,
                       cellPrefix, slotBody, body, mats, cellBody, cellRec, cellNear, big);
    if (big.limit != 0)
        launch(k_cell_canon_big, dim3(BIG_GRID), dim3(BIG_NT), 0, st, slotBody, body, mats,
               cellBody, big);
}

void launch_kick(hipStream_t st, const LaunchShape& ls, float4* body, float2* acc,
                 const float2* force, const uint4* cellRec, const uint4* cellNear,
                 const float4* cellBody, int initial) {
    if (ls.bodyBlocks == 0) return;
    if (ls.walls != WALLS_ALL)
        launch(k_kick<0>, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, body, acc,
               force, cellRec, cellNear, cellBody, initial, ls.sleep);
    if (ls.walls != WALLS_NONE)
        launch(k_kick<1>, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, body, acc,
               force, cellRec, cellNear, cellBody, initial, ls.sleep);
}

}  // namespace gs
