// Hard-disc collision passes (S/collision_cache_{drift,count,solve}.comp,
// S/collision_material.glsl, S/collision_crowd.glsl), non-RIGID.
//
// One pass of the reference is drift -> broad phase -> count -> Jacobi solve.
// Here a pass is two launches:
//   count : clears the previous pass's grid cells, walks the 3x3 stencil of
//           this pass's grid, writes the contact count and the ascending list
//           of the CONTACT_SLOTS smallest contact indices (kept sorted in
//           registers while walking; the indices they push out go to
//           overflow planes, which a crowded body then orders into the
//           listed planes, up to CONTACT_SORTED contacts);
//   solve : Jacobi correction in ascending contact order (a body with more
//           contacts than the contact array lists takes the rest six at a
//           time, in order, from the contact pool), walls,
//           integration, then, fused, the next pass's drift and its insertion
//           into the other grid buffer.
// The broad phase is a linked-cell grid of cells at least diameter + guard
// wide, rebuilt every pass; heads of a pass are cleared by the next pass's
// count kernel, so no clearing launch is needed. Two grid buffers alternate.
// Bodies are float4 (x, y, vx, vy): a neighbour is one 16-byte load. Grid
// nodes hold positions in pixels (x * width, y * height: the products the
// distance test forms), so a visit is two subtractions and a sum of squares.
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

struct GridBuf {
    uint32_t* heads;   // per cell: first body + 1, 0 when empty
    float4* nodes;     // per body: x, y in px, next body + 1 (bits), cx | cy << 16 (bits)
    int* cellOf;       // per body: the cell it was inserted in
};

// The broad-phase cell of a normalised position.
__device__ __forceinline__ int cell_x(const SimDesc& s, float x) {
    return floor_index(x * (float)s.cgw, s.cgw);
}
__device__ __forceinline__ int cell_y(const SimDesc& s, float y) {
    return floor_index(y * (float)s.cgh, s.cgh);
}

// A node keeps the body's cell as cx | cy << 16 (grids are at most 65535
// cells a side), so the count finds its stencil from the node alone.
__device__ __forceinline__ void grid_insert(const SimDesc& s, int g, float px, float py,
                                            const GridBuf& grid) {
    int cx = cell_x(s, px), cy = cell_y(s, py);
    int c = s.cellBase + cx + cy * s.cgw;
    uint32_t old = atomicExch(&grid.heads[c], (uint32_t)(g + 1));
    grid.nodes[g] = make_float4(px * s.vw, py * s.vh, __uint_as_float(old),
                                __uint_as_float((uint32_t)cx | ((uint32_t)cy << 16)));
    grid.cellOf[g] = c;
}

__device__ __forceinline__ uint32_t med3u(uint32_t a, uint32_t b, uint32_t c) {
    return max(min(a, b), min(max(a, b), c));
}

// S/collision_crowd.glsl:2-15: the lower index goes first; squares summed
// without contraction. Positions are in pixels (the products x * width and
// y * height the reference forms first). Without the torus's minimum image the
// order cannot matter: fl(a - b) = -fl(b - a) exactly, so the squares are the
// same.
template <bool BOUNDED>
__device__ __forceinline__ float pair_d2(int id, float px, float py, int other, float ox,
                                         float oy, float vw, float vh) {
    if (BOUNDED) {
        float dx = px - ox;
        float dy = py - oy;
        return dx * dx + dy * dy;
    }
    bool ownFirst = id < other;
    float dx = ownFirst ? px - ox : ox - px;
    float dy = ownFirst ? py - oy : oy - py;
    dx -= floorf(dx / vw + 0.5f) * vw;
    dy -= floorf(dy / vh + 0.5f) * vh;
    return dx * dx + dy * dy;
}

__device__ __forceinline__ float pair_distance_squared(int id, float px, float py, int other,
                                                       float ox, float oy, const SimDesc& s) {
    return s.bounded ? pair_d2<true>(id, px, py, other, ox, oy, s.vw, s.vh)
                     : pair_d2<false>(id, px, py, other, ox, oy, s.vw, s.vh);
}

// The heads of the 3x3 stencil around cell (cx, cy), cell by cell; cells
// cut by a wall, or repeated on a grid narrower than 3, read as empty.
__device__ __forceinline__ void stencil_heads(const SimDesc& s, int cx, int cy,
                                              const GridBuf& grid, uint32_t (&head)[9]) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
        int y = s.cgh < 3 ? j : cy + j - 1;
        bool oky = s.cgh < 3 ? j < s.cgh : (!s.bounded || (y >= 0 && y < s.cgh));
        if (!s.bounded) y = wrap_axis(y, s.cgh);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            int x = s.cgw < 3 ? i : cx + i - 1;
            bool okx = s.cgw < 3 ? i < s.cgw : (!s.bounded || (x >= 0 && x < s.cgw));
            if (!s.bounded) x = wrap_axis(x, s.cgw);
            head[j * 3 + i] = oky && okx ? grid.heads[s.cellBase + x + y * s.cgw] : 0u;
        }
    }
}

// The heads of three neighbouring cells of a grid row, one 12-byte load.
struct HeadRow {
    uint32_t a, b, c;
};

// The heads of the 3x3 stencil around cell (cx, cy) of a grid at least 3
// cells wide and high, one 12-byte load per row, all three unconditional so
// they are in flight together. The heads array has a zero entry before its
// first cell and after its last, so a row's load stays in the allocation:
// under walls a row past the edge is loaded clamped and reads as empty, as do
// cells past a side wall. On the torus the load takes the three cells of the
// row nearest the stencil, and a stencil at the left or right edge takes its
// wrapped cell from one more load.
template <bool BOUNDED>
__device__ __forceinline__ void stencil_rows(const SimDesc& s, int cx, int cy,
                                             const uint32_t* __restrict__ heads,
                                             uint32_t (&e)[9]) {
    const int cgw = s.cgw, cgh = s.cgh;
    const bool left = cx == 0, right = cx == cgw - 1;
    HeadRow h[3];
    uint32_t wrapCell[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        int y = cy + j - 1;
        y = BOUNDED ? min(max(y, 0), cgh - 1) : wrap_axis(y, cgh);
        const uint32_t* row = heads + s.cellBase + y * cgw;
        if (BOUNDED) {
            h[j] = *reinterpret_cast<const HeadRow*>(row + cx - 1);
        } else {
            h[j] = *reinterpret_cast<const HeadRow*>(row + min(max(cx, 1), cgw - 2) - 1);
            wrapCell[j] = row[left ? cgw - 1 : 0];
        }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        uint32_t a = h[j].a, b = h[j].b, c = h[j].c;
        if (BOUNDED) {
            const int y = cy + j - 1;
            const bool oky = y >= 0 && y < cgh;
            a = oky && !left ? a : 0u;
            b = oky ? b : 0u;
            c = oky && !right ? c : 0u;
        } else {
            const uint32_t ha = h[j].a, hb = h[j].b, hc = h[j].c;
            a = left ? wrapCell[j] : (right ? hb : ha);
            b = left ? ha : (right ? hc : hb);
            c = left ? hb : (right ? wrapCell[j] : hc);
        }
        e[j * 3 + 0] = a;
        e[j * 3 + 1] = b;
        e[j * 3 + 2] = c;
    }
}

// Visits every body of the stencil around cell (cx, cy) (the callers'
// results do not depend on the order), with its position in pixels. The
// nine lists are walked together, one node of each per round: a round's node
// loads are independent, so a lane pays one memory round trip per round, and
// the rounds number the longest of its nine lists (at most four on a packed
// disc of diameter-wide cells), where walking the lists one after another
// chains a round trip per node. guardLimit bounds the nodes visited.
template <bool BOUNDED, typename F>
__device__ __forceinline__ void walk_stencil(const SimDesc& s, int cx, int cy, const GridBuf& grid,
                                             int guardLimit, F&& visit) {
    uint32_t e[9];
    if (s.cgw >= 3 && s.cgh >= 3) {
        stencil_rows<BOUNDED>(s, cx, cy, grid.heads, e);
    } else {
        stencil_heads(s, cx, cy, grid, e);
    }
    int guard = 0;
    for (;;) {
        uint32_t live = 0u;
#pragma unroll
        for (int k = 0; k < 9; k++) live |= e[k];
        if (live == 0u || guard >= guardLimit) break;
        float4 nd[9];
#pragma unroll
        for (int k = 0; k < 9; k++)
            if (e[k] != 0u) nd[k] = grid.nodes[e[k] - 1u];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            if (e[k] != 0u) {
                int other = (int)e[k] - 1;
                e[k] = __float_as_uint(nd[k].z);
                guard++;
                visit(other, nd[k].x, nd[k].y);
            }
        }
    }
}

// The same visits, one list after another: fewer registers, for the rare
// walks (a body with more contacts than slots) where the nine lists' nodes
// need not be in flight together.
template <typename F>
__device__ __forceinline__ void walk_stencil_serial(const SimDesc& s, int cx, int cy,
                                                    const GridBuf& grid, int guardLimit,
                                                    F&& visit) {
    uint32_t head[9];
    stencil_heads(s, cx, cy, grid, head);
    int guard = 0;
#pragma unroll 1
    for (int k = 0; k < 9; k++) {
        uint32_t e = head[k];
        while (e != 0u && guard < guardLimit) {
            int other = (int)e - 1;
            float4 nd = grid.nodes[other];
            e = __float_as_uint(nd.z);
            guard++;
            visit(other, nd.x, nd.y);
        }
    }
}

// ---------------------------------------------------------------------------
// Drift (S/collision_cache_drift.comp:60-91) and insertion into the grid.
// With the sleeping tiles on, the step's kick pass tells every body's state
// for the step: a body outside the awake tiles (and not a fluid body) rests,
// its velocity dropped, and no pass of the step moves it.
// With adaptive collisions a drift takes the step's substep, dt over the
// substeps chosen (S/collision_cache_drift.comp:79-80).
// `drifts` consecutive drifts of driftDt (after the kick, if any) run in one
// launch: without contact passes between them a body's drifts depend on
// nothing but itself, so the gravity-only step's substep drifts take one
// launch, each drift rounding and wrapping exactly as a launch of its own.
__global__ void __launch_bounds__(WG) k_drift_insert(const SimDesc* __restrict__ sims,
                                                     const BodyBlock* __restrict__ blocks,
                                                     float4* __restrict__ body,
                                                     const float2* __restrict__ acc,
                                                     const float4* __restrict__ mats,
                                                     float kickDt, float driftDt, int insert,
                                                     GridBuf grid, SleepBufs sleep,
                                                     const uint32_t* __restrict__ adapt,
                                                     int drifts) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    // The body and its acceleration load before the simulation's constants
    // are known (both arrays cover the padded slots; a slot past the active
    // bodies returns without using them); the fence keeps the compiler from
    // sinking the loads behind that wait.
    const int g = b.first + (int)threadIdx.x;
    float4 v = body[g];
    const float2 a = acc[g];
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const SimDesc& s = sims[b.sim];
    if (g - s.base >= s.n) return;
    if (adapt != nullptr && s.adaptive && driftDt != 0.0f)
        driftDt = s.dt / (float)max(adapt[ADAPT_WORDS * b.sim + 2], 1u);
    bool awake = true;
    if (s.sleepOn) {
        const SleepState st = sleep.state[b.sim];
        if (sleep_on(s, st)) {
            if (kickDt != 0.0f) {
                bool fluid = s.liquid && mats[g].z == 0.0f;
                awake = fluid || sleep.tileAwake[sleep_tile(s, v.x, v.y)] == st.stamp;
                sleep.bodyAwake[g] = awake ? st.stamp : 0u;
                if (!awake) {
                    v.z = 0.0f;
                    v.w = 0.0f;
                    body[g] = v;
                }
            } else {
                awake = sleep.bodyAwake[g] == st.stamp;
            }
        }
    }
    if (awake && (kickDt != 0.0f || driftDt != 0.0f)) {
        if (kickDt != 0.0f) {
            float h = 0.5f * kickDt;
            v.z += a.x * h;
            v.w += a.y * h;
        }
        for (int r = 0; r < drifts; r++) {
            v.x += v.z * driftDt;
            v.y += v.w * driftDt;
            if (s.bounded) {
                v.x = clampf(v.x, 0.0f, MAXIMUM_UNIT);
                v.y = clampf(v.y, 0.0f, MAXIMUM_UNIT);
            } else {
                v.x = fractf(v.x);
                v.y = fractf(v.y);
            }
        }
        body[g] = v;
    }
    if (insert) grid_insert(s, g, v.x, v.y, grid);
}

// ---------------------------------------------------------------------------
// Count (S/collision_cache_count.comp:61-95, 130-285 without RIGID): every
// hard pair with d^2 < D^2, strictly.
template <bool BOUNDED>
__device__ __forceinline__ uint32_t count_contacts(const SimDesc& s, int g, float4 own,
                                               const GridBuf& grid,
                                               const float4* __restrict__ mats,
                                               uint32_t* __restrict__ counts,
                                               uint32_t* __restrict__ contacts, int planeStride,
                                               int guardLimit, const ContactPool& pool) {
    const float px = own.x, py = own.y, vw = s.vw, vh = s.vh, D2 = s.diameter2;
    const uint32_t cell = __float_as_uint(own.w);
    const bool ownFluid = s.liquid && mats[g].z == 0.0f;
    uint32_t slot[CONTACT_SLOTS];
#pragma unroll
    for (int k = 0; k < CONTACT_SLOTS; k++) slot[k] = 0xffffffffu;
    uint32_t n = 0;
    const int cx = (int)(cell & 0xffffu), cy = (int)(cell >> 16);
    // On the torus, a stencil that does not wrap on a grid of at least 8
    // cells a side keeps every candidate within about a quarter of the
    // viewport, where floor(d / w + 0.5) is 0: the minimum image is the
    // difference itself, and either order gives the same square
    // (fl(a - b) = -fl(b - a)), so the divisions are not needed.
    const bool interior = BOUNDED || (cx >= 1 && cx <= s.cgw - 2 && cy >= 1 && cy <= s.cgh - 2 &&
                                      s.cgw >= 8 && s.cgh >= 8);
    walk_stencil<BOUNDED>(s, cx, cy, grid, guardLimit, [&](int other, float ox, float oy) {
        float d2;
        if (interior) {
            const float dx = px - ox, dy = py - oy;
            d2 = dx * dx + dy * dy;
        } else {
            d2 = pair_d2<BOUNDED>(g, px, py, other, ox, oy, vw, vh);
        }
        if (!(d2 < D2) || other == g) return;
        if (ownFluid && mats[other].z == 0.0f) return;
        n++;
        // Insertion keeping the CONTACT_SLOTS smallest indices, ascending:
        // with s sorted, the new slot k is the median of s[k-1], u, s[k]
        // (one v_med3_u32 each, all independent of one another). The index
        // it pushes out of the kept ones goes to the next free overflow
        // plane (planes CONTACT_SORTED..), while there is one (plane indices
        // fit 32 bits: the engine caps the slots at 2^32 / CONTACT_PLANES).
        uint32_t u = (uint32_t)other;
        const uint32_t out = max(u, slot[CONTACT_SLOTS - 1]);
#pragma unroll
        for (int k = CONTACT_SLOTS - 1; k > 0; k--) slot[k] = med3u(slot[k - 1], u, slot[k]);
        slot[0] = min(u, slot[0]);
        if (n > (uint32_t)CONTACT_SLOTS && n <= (uint32_t)CONTACT_SORTED)
            contacts[(n - 1u + (uint32_t)(CONTACT_SORTED - CONTACT_SLOTS)) * (uint32_t)planeStride +
                     (uint32_t)g] = out;
    });
    counts[g] = n;
#pragma unroll
    for (int k = 0; k < CONTACT_SLOTS; k++)
        if ((uint32_t)k < n) contacts[(size_t)k * planeStride + g] = slot[k];
    if (n > (uint32_t)CONTACT_SLOTS && n <= (uint32_t)CONTACT_SORTED) {
        // The overflow planes, ordered into the listed planes CONTACT_SLOTS
        // at a time: each round keeps the smallest above the last one listed,
        // from loads issued CONTACT_SLOTS at a time.
        const uint32_t m = n - (uint32_t)CONTACT_SLOTS;
        const uint32_t* over = contacts + (size_t)CONTACT_SORTED * planeStride + g;
        uint32_t floorId = slot[CONTACT_SLOTS - 1];
        for (uint32_t c0 = 0; c0 < m; c0 += CONTACT_SLOTS) {
#pragma unroll
            for (int k = 0; k < CONTACT_SLOTS; k++) slot[k] = 0xffffffffu;
            for (uint32_t e = 0; e < m; e += CONTACT_SLOTS) {
                uint32_t u[CONTACT_SLOTS];
#pragma unroll
                for (int q = 0; q < CONTACT_SLOTS; q++)
                    u[q] = over[(size_t)min(e + (uint32_t)q, m - 1u) * planeStride];
#pragma unroll
                for (int q = 0; q < CONTACT_SLOTS; q++) {
                    if (e + (uint32_t)q >= m || u[q] <= floorId) continue;
#pragma unroll
                    for (int k = CONTACT_SLOTS - 1; k > 0; k--) slot[k] = med3u(slot[k - 1], u[q], slot[k]);
                    slot[0] = min(u[q], slot[0]);
                }
            }
            const uint32_t take = min(m - c0, (uint32_t)CONTACT_SLOTS);
#pragma unroll
            for (int k = 0; k < CONTACT_SLOTS; k++)
                if ((uint32_t)k < take)
                    contacts[(size_t)(CONTACT_SLOTS + c0 + (uint32_t)k) * planeStride + g] = slot[k];
#pragma unroll
            for (int k = 0; k < CONTACT_SLOTS; k++)
                if ((uint32_t)k + 1u == take) floorId = slot[k];
        }
    } else if (n > (uint32_t)CONTACT_SORTED) {
        // More than the planes list: the contacts past the CONTACT_SLOTS
        // smallest, found by a second walk, go to the pool, so the solve
        // takes them without walking the grid again.
        const uint32_t need = n - (uint32_t)CONTACT_SLOTS, floorId = slot[CONTACT_SLOTS - 1];
        uint32_t b = atomicAdd(pool.top, need);
        if (b > pool.capacity || pool.capacity - b < need) {
            b = NO_POOL;
        } else {
            uint32_t e = 0u;
            walk_stencil<BOUNDED>(s, cx, cy, grid, guardLimit, [&](int other, float ox, float oy) {
                uint32_t u = (uint32_t)other;
                if (u <= floorId || other == g) return;
                float d2 = pair_d2<BOUNDED>(g, px, py, other, ox, oy, vw, vh);
                if (!(d2 < D2)) return;
                if (ownFluid && mats[other].z == 0.0f) return;
                if (e < need) pool.items[b + e] = u;
                e++;
            });
        }
        pool.base[g] = b;
    }
    return n;
}

// ADAPTIVE: some simulation of the batch has adaptive collisions; a substep
// pass at or past its chosen count does nothing (S/collision_cache_count.comp:131).
template <bool ADAPTIVE>
__global__ void __launch_bounds__(WG) k_count(const SimDesc* __restrict__ sims,
                                              const OrderedBlock* __restrict__ order,
                                              uint32_t* __restrict__ waveWork, GridBuf grid,
                                              GridBuf prev, int clearPrev,
                                              const float4* __restrict__ mats,
                                              uint32_t* __restrict__ counts,
                                              uint32_t* __restrict__ contacts, int planeStride,
                                              int guardLimit, ContactPool pool, SleepBufs sleep,
                                              int pass, const uint32_t* __restrict__ adapt) {
    const OrderedBlock ob = order[blockIdx.x];
    const int bi = ob.index;
    const BodyBlock b{ob.sim, ob.first};
    // The body's cell in the previous pass's grid and its own node here load
    // together, before the simulation's constants are known: both arrays
    // cover the padded slots of every simulation, and a slot past the active
    // bodies does not use them. The clearing store waits only for the first.
    const int g = b.first + (int)threadIdx.x;
    const int prevCell = prev.cellOf[g];
    const float4 own = grid.nodes[g];
    // A wavefront-scope fence emits no instruction; it keeps the compiler
    // from sinking the two loads into the branch below, behind the wait for
    // the simulation's constants.
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const SimDesc& s = sims[b.sim];
    uint32_t n = 0;
    bool active = true;
    if (ADAPTIVE && s.adaptive && pass >= 0)
        active = (uint32_t)pass < adapt[ADAPT_WORDS * b.sim + 2];
    if (g - s.base < s.n && active) {
        if (clearPrev) prev.heads[prevCell] = 0u;
        // A sleeping body keeps its last count (S/collision_cache_count.comp:147-160).
        bool asleep = false;
        if (s.sleepOn) {
            const SleepState st = sleep.state[b.sim];
            asleep = sleep_on(s, st) && sleep.bodyAwake[g] != st.stamp;
        }
        if (!asleep) {
            n = s.bounded ? count_contacts<true>(s, g, own, grid, mats, counts, contacts,
                                                 planeStride, guardLimit, pool)
                          : count_contacts<false>(s, g, own, grid, mats, counts, contacts,
                                                  planeStride, guardLimit, pool);
        }
    }
    // The wave's contacts, for the next step's dispatch order.
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n += (uint32_t)__shfl_xor((int)n, o);
    if ((threadIdx.x & 63u) == 0u) waveWork[bi * (WG / 64) + (int)(threadIdx.x >> 6)] = n;
}

// Once per step: per XCD, its body blocks ranked by descending contacts (the
// rank of a block is the number of blocks with a larger key; keys are
// distinct), written back to that XCD's workgroup positions. Blocks past the
// last full XCD round, and XCDs with more than ORDER_MAX blocks, keep the
// xcd_block order.
constexpr int ORDER_MAX = 512;
constexpr int CUS_PER_XCD = 32;
__device__ __forceinline__ OrderedBlock ordered(const BodyBlock* __restrict__ blocks, int i) {
    const BodyBlock b = blocks[i];
    return OrderedBlock{b.sim, b.first, i, 0};
}

__global__ void __launch_bounds__(WG) k_block_order(int nb, const BodyBlock* __restrict__ blocks,
                                                    const uint32_t* __restrict__ waveWork,
                                                    OrderedBlock* __restrict__ order) {
    __shared__ uint32_t key[ORDER_MAX];
    __shared__ int item[ORDER_MAX];
    const int x = (int)blockIdx.x;
    const int full = nb / (XCDS * XCD_CHUNK) * (XCDS * XCD_CHUNK);
    const int m = full / XCDS;
    if (x == 0)
        for (int i = full + (int)threadIdx.x; i < nb; i += WG) order[i] = ordered(blocks, i);
    if (m > ORDER_MAX) {
        for (int j = (int)threadIdx.x; j < m; j += WG)
            order[x + XCDS * j] = ordered(blocks, xcd_block(x + XCDS * j, nb));
        return;
    }
    for (int j = (int)threadIdx.x; j < m; j += WG) {
        int bb = xcd_block(x + XCDS * j, nb);
        const uint32_t* w = waveWork + (size_t)bb * (WG / 64);
        uint32_t work = w[0] + w[1] + w[2] + w[3];
        key[j] = (min(work, (1u << 20) - 1u) << 11) | (uint32_t)(2047 - j);
        item[j] = bb;
    }
    __syncthreads();
    for (int j = (int)threadIdx.x; j < m; j += WG) {
        uint32_t kj = key[j];
        int rank = 0;
        for (int k = 0; k < m; k++) rank += key[k] > kj ? 1 : 0;
        // Ranks are dealt to the XCD's workgroup positions in rounds of
        // CUS_PER_XCD, every other full round reversed, so each CU's share
        // of the heaviest blocks evens out.
        const int round = rank / CUS_PER_XCD, in = rank % CUS_PER_XCD;
        const bool reversed = (round & 1) && (round + 1) * CUS_PER_XCD <= m;
        const int pos = round * CUS_PER_XCD + (reversed ? CUS_PER_XCD - 1 - in : in);
        order[x + XCDS * pos] = ordered(blocks, item[j]);
    }
}

void launch_block_order(hipStream_t st, const LaunchShape& ls, OrderedBlock* order) {
    if (ls.bodyBlocks == 0) return;
    launch(k_block_order, dim3(XCDS), dim3(WG), 0, st, ls.bodyBlocks, ls.blocks, ls.waveWork,
                       order);
}

void identity_block_order(const BodyBlock* blocks, int count, OrderedBlock* order) {
    for (int i = 0; i < count; i++) {
        const int bi = xcd_block(i, count);
        order[i] = OrderedBlock{blocks[bi].sim, blocks[bi].first, bi, 0};
    }
}

// ---------------------------------------------------------------------------
// Solve (S/collision_cache_solve.comp:88-252, 274-478 without RIGID).

struct Accum {
    float pcx, pcy, vcx, vcy, pressure;
    float deepest;  // the deepest pair penetration of a pressure pass (adaptive state)
};

// 1/k for small contact counts: each workgroup of the solve divides once
// into LDS (the correctly rounded quotients) and reads the table per contact.
constexpr int RECIP_TABLE = 64;

// The simulation's constants a solve reads per contact, loaded once.
struct SolveK {
    float vw, vh, diameter, diameter2, eps, e1, subDt;
    bool bounded, liquid;
};
__device__ __forceinline__ SolveK solve_k(const SimDesc& s) {
    return SolveK{s.vw, s.vh, s.diameter, s.diameter2, s.coincidentEps, 1.0f + s.restitution,
                  s.subDt, s.bounded != 0, s.liquid != 0};
}

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
__device__ __forceinline__ void wall_contact(const SolveK& c, float distance, float nx, float ny,
                                             float vx, float vy, float weight, float friction,
                                             int writePressure, Accum& acc) {
    float penetration = 0.5f * c.diameter - distance;
    if (penetration <= 0.0f) return;
    acc.pcx += nx * (penetration * weight);
    acc.pcy += ny * (penetration * weight);
    float ns = vx * nx + vy * ny;
    float e1 = c.e1;
    if (ns < 0.0f) {
        float k = e1 * ns * weight;
        acc.vcx -= nx * k;
        acc.vcy -= ny * k;
    }
    if (friction > 0.0f) {
        float approach = -ns > 0.0f ? -ns : 0.0f;
        float fx, fy;
        friction_correction(vx, vy, nx, ny, ns, weight * (e1 * approach + penetration / c.subDt),
                            weight, friction, fx, fy);
        acc.vcx += fx;
        acc.vcy += fy;
    }
    if (writePressure) {
        float approach = -ns > 0.0f ? -ns : 0.0f;
        acc.pressure += weight * (penetration / c.subDt + approach);
    }
}

// One contact of the Jacobi sum, with the other body's data already loaded.
__device__ __forceinline__ void accumulate_contact(const SolveK& s, const float* recip, int id,
                                                   int other, float4 ob,
                                                   float4 om, uint32_t oc, bool otherSleeps,
                                                   const float2* __restrict__ fallback, float ppx,
                                                   float ppy, float vpx, float vpy, float4 ownMat,
                                                   uint32_t ownContacts, int writePressure,
                                                   Accum& acc) {
    if (s.liquid && !(ownMat.z != 0.0f || om.z != 0.0f)) return;
    float dx = ppx - ob.x * s.vw;
    float dy = ppy - ob.y * s.vh;
    // On the torus, a difference under a quarter of the viewport is its own
    // minimum image (floor(d / w + 0.5) is 0 there), with no division.
    if (!s.bounded && !(fabsf(dx) < 0.25f * s.vw && fabsf(dy) < 0.25f * s.vh)) {
        dx -= floorf(dx / s.vw + 0.5f) * s.vw;
        dy -= floorf(dy / s.vh + 0.5f) * s.vh;
    }
    float d2 = dx * dx + dy * dy;
    if (!(d2 < s.diameter2)) return;
    float eps = s.eps;
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
    uint32_t crowd = ownContacts > oc ? ownContacts : oc;
    uint32_t weightCount = crowd > 1u ? crowd : 1u;
    float pairWeight = weightCount < RECIP_TABLE ? recip[weightCount] : 1.0f / (float)weightCount;
    // m / (m + m) is exactly 0.5 for every normal m, and sqrt(mu * mu) is
    // exactly mu for mu = 0 and for every mu in [2^-63, 2^63) (both checked
    // over all floats of those ranges): equal materials skip the division
    // and the square root with identical results.
    // A sleeping partner stays where it is: this side takes the whole
    // correction (S/collision_cache_solve.comp:141-142, 201).
    float share;
    if (otherSleeps)
        share = 1.0f;
    else if (om.x == ownMat.x && ownMat.x >= 0x1p-126f && ownMat.x < 0x1p127f)
        share = 0.5f;
    else
        share = om.x / (ownMat.x + om.x);
    float ovx = ob.z * s.vw, ovy = ob.w * s.vh;
    float rvx = vpx - ovx, rvy = vpy - ovy;
    float rvn = rvx * nx + rvy * ny;
    float push = share * penetration * pairWeight;
    acc.pcx += nx * push;
    acc.pcy += ny * push;
    float e1 = s.e1;
    if (rvn < 0.0f) {
        float k = share * e1 * rvn * pairWeight;
        acc.vcx -= nx * k;
        acc.vcy -= ny * k;
    }
    float friction;
    if (om.y == ownMat.y && (ownMat.y == 0.0f || (ownMat.y >= 0x1p-63f && ownMat.y < 0x1p63f)))
        friction = ownMat.y;
    else
        friction = sqrtf(ownMat.y * om.y);
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
        acc.deepest = fmaxf(acc.deepest, penetration);
        float approach = -rvn > 0.0f ? -rvn : 0.0f;
        acc.pressure += 0.5f * pairWeight * (penetration / s.subDt + approach);
    }
}

// The OVERFLOW_CHUNK smallest contact indices above floorId, ascending, by
// one walk of the stencil (unused entries 0xffffffff).
constexpr int OVERFLOW_CHUNK = 6;
// One contact's loads at a time: the rare path keeps the solve at 71 VGPRs,
// seven waves per SIMD, so every wave of a 400k-body launch is resident.
constexpr int OVERFLOW_BATCH = 1;
template <bool BOUNDED>
__device__ __forceinline__ void next_contacts(const SimDesc& s, int cx, int cy, const GridBuf& grid,
                                              int guardLimit, int g, float px, float py,
                                              uint32_t floorId, bool ownFluid,
                                              const float4* __restrict__ mats,
                                              uint32_t (&next)[OVERFLOW_CHUNK]) {
#pragma unroll
    for (int q = 0; q < OVERFLOW_CHUNK; q++) next[q] = 0xffffffffu;
    const float vw = s.vw, vh = s.vh, D2 = s.diameter2;
    walk_stencil_serial(s, cx, cy, grid, guardLimit, [&](int other, float ox, float oy) {
        uint32_t u = (uint32_t)other;
        if (u <= floorId || other == g || u >= next[OVERFLOW_CHUNK - 1]) return;
        if (ownFluid && mats[other].z == 0.0f) return;
        float d2 = pair_d2<BOUNDED>(g, px, py, other, ox, oy, vw, vh);
        if (!(d2 < D2)) return;
#pragma unroll
        for (int q = OVERFLOW_CHUNK - 1; q > 0; q--) next[q] = med3u(next[q - 1], u, next[q]);
        next[0] = min(u, next[0]);
    });
}

// The same from a body's pool entries (its contacts past the CONTACT_SLOTS
// smallest, in any order): the OVERFLOW_CHUNK smallest above floorId,
// ascending. The loads go U at a time, unconditional (clamped).
__device__ __forceinline__ void pool_next(const uint32_t* __restrict__ items, uint32_t count,
                                          uint32_t floorId, uint32_t (&next)[OVERFLOW_CHUNK]) {
#pragma unroll
    for (int q = 0; q < OVERFLOW_CHUNK; q++) next[q] = 0xffffffffu;
    constexpr uint32_t U = 8;
    for (uint32_t e = 0; e < count; e += U) {
        uint32_t u[U];
#pragma unroll
        for (uint32_t q = 0; q < U; q++) u[q] = items[min(e + q, count - 1u)];
#pragma unroll
        for (uint32_t q = 0; q < U; q++)
            if (e + q >= count) u[q] = 0xffffffffu;
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            if (u[q] <= floorId || u[q] >= next[OVERFLOW_CHUNK - 1]) continue;
#pragma unroll
            for (int k = OVERFLOW_CHUNK - 1; k > 0; k--) next[k] = med3u(next[k - 1], u[q], next[k]);
            next[0] = min(u[q], next[0]);
        }
    }
}

// Contacts whose loads are issued together. The batch loop is not
// unrolled; three contacts a batch measured faster than two or four.
constexpr int SOLVE_BATCH = 3;

// ADAPTIVE: some simulation of the batch has adaptive collisions. For such a
// simulation a substep pass at or past the step's chosen count does nothing,
// the others take dt over that count as their substep, the last of them
// drifts no further and inserts for the final pass, and the final pass
// (writePressure) stores each body's deepest pair penetration in `deepest`
// for k_adapt_stats (S/collision_cache_solve.comp:239-243, 277-311).
template <bool ADAPTIVE>
__global__ void __launch_bounds__(WG) k_solve(const SimDesc* __restrict__ sims,
                                              const OrderedBlock* __restrict__ order,
                                              const float4* __restrict__ bodyIn,
                                              float4* __restrict__ bodyOut,
                                              const float4* __restrict__ mats,
                                              const uint32_t* __restrict__ counts,
                                              const uint32_t* __restrict__ contacts,
                                              int planeStride, GridBuf grid, int guardLimit,
                                              const float2* __restrict__ fallback,
                                              float* __restrict__ pressures, int writePressure,
                                              float nextDriftDt, int insertNext, GridBuf next,
                                              ContactPool pool, SleepBufs sleep, int pass,
                                              int lastInsert, const uint32_t* __restrict__ adapt,
                                              float* __restrict__ deepest) {
    // The pool's entries were all written by the count; its counter starts
    // the next count from zero (no kernel of this launch reads it).
    if (blockIdx.x == 0 && threadIdx.x == 0) *pool.top = 0u;
    // the block and its simulation's constants load while the table fills
    const OrderedBlock ob = order[blockIdx.x];
    const BodyBlock b{ob.sim, ob.first};
    const SimDesc& s = sims[b.sim];
    SolveK k = solve_k(s);
    // Each wave fills its own copy of the table (one entry a lane), so no
    // workgroup barrier holds the waves: a wave's LDS writes are complete
    // for its own later reads once the wave has waited for them.
    static_assert(RECIP_TABLE == 64, "one table entry per lane of a wave");
    __shared__ float recipWaves[WG / 64][RECIP_TABLE];
    const int lane = (int)(threadIdx.x & 63u);
    float* recip = recipWaves[threadIdx.x >> 6];
    recip[lane] = lane ? 1.0f / (float)lane : 1.0f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int g = b.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    if (ADAPTIVE && s.adaptive) {
        const uint32_t chosen = adapt[ADAPT_WORDS * b.sim + 2];
        if (pass >= 0 && (uint32_t)pass >= chosen) return;
        k.subDt = s.dt / (float)max(chosen, 1u);
        if (pass >= 0) {
            const bool lastActive = (uint32_t)pass + 1u >= chosen;
            nextDriftDt = lastActive ? 0.0f : k.subDt;
            insertNext = lastActive ? lastInsert : 1;
        }
    }
    float4 me = bodyIn[g];
    bool sleeping = false;
    uint32_t stamp = 0u;
    if (s.sleepOn) {
        const SleepState st = sleep.state[b.sim];
        sleeping = sleep_on(s, st);
        stamp = st.stamp;
    }
    if (sleeping && sleep.bodyAwake[g] != stamp) {
        // A sleeping body is carried over as it is (S/collision_cache_solve.comp:279-287).
        bodyOut[g] = me;
        if (insertNext) grid_insert(s, g, me.x, me.y, next);
        return;
    }
    float4 ownMat = mats[g];
    float ppx = me.x * k.vw, ppy = me.y * k.vh;
    float vpx = me.z * k.vw, vpy = me.w * k.vh;
    Accum acc{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t own = counts[g];
    // Up to CONTACT_SORTED contacts are listed in ascending order; a body
    // with more lists its CONTACT_SLOTS smallest and takes the rest below.
    int listed = own <= (uint32_t)CONTACT_SORTED ? (int)own : CONTACT_SLOTS;
    uint32_t last = 0u;
    const float2* __restrict__ matXY = reinterpret_cast<const float2*>(mats);
#pragma unroll 1
    for (int i0 = 0; i0 < listed; i0 += SOLVE_BATCH) {
        uint32_t id[SOLVE_BATCH];
#pragma unroll
        for (int q = 0; q < SOLVE_BATCH; q++)
            id[q] = i0 + q < listed ? contacts[(size_t)(i0 + q) * planeStride + g] : (uint32_t)g;
        float4 ob[SOLVE_BATCH];
        float2 mf[SOLVE_BATCH];
        uint32_t oc[SOLVE_BATCH];
#pragma unroll
        for (int q = 0; q < SOLVE_BATCH; q++) {
            ob[q] = bodyIn[id[q]];
            mf[q] = matXY[2 * (size_t)id[q]];  // mass, friction
            oc[q] = counts[id[q]];
        }
#pragma unroll
        for (int q = 0; q < SOLVE_BATCH; q++) {
            if (i0 + q < listed) {
                last = id[q];
                float solid = k.liquid ? mats[last].z : 1.0f;
                bool otherSleeps = sleeping && sleep.bodyAwake[last] != stamp;
                accumulate_contact(k, recip, g, (int)last, ob[q], make_float4(mf[q].x, mf[q].y, solid, 0.0f),
                                   oc[q], otherSleeps, fallback, ppx, ppy, vpx, vpy, ownMat, own,
                                   writePressure, acc);
            }
        }
    }
    if (own > (uint32_t)CONTACT_SORTED) {
        // More contacts than listed: the rest in ascending index,
        // OVERFLOW_CHUNK at a time, each chunk the smallest indices above the
        // last one taken, selected from the body's pool entries (or, when the
        // pool was full, gathered by one walk of the grid).
        const bool ownFluid = k.liquid && ownMat.z == 0.0f;
        const int cx = cell_x(s, me.x), cy = cell_y(s, me.y);
        const uint32_t pb = pool.base[g];
        uint32_t rest = own - (uint32_t)CONTACT_SLOTS;
        while (rest > 0u) {
            uint32_t next[OVERFLOW_CHUNK];
            if (pb != NO_POOL)
                pool_next(pool.items + pb, own - (uint32_t)CONTACT_SLOTS, last, next);
            else if (k.bounded)
                next_contacts<true>(s, cx, cy, grid, guardLimit, g, ppx, ppy, last, ownFluid, mats, next);
            else
                next_contacts<false>(s, cx, cy, grid, guardLimit, g, ppx, ppy, last, ownFluid, mats, next);
            uint32_t taken = 0u;
#pragma unroll
            for (int i0 = 0; i0 < OVERFLOW_CHUNK; i0 += OVERFLOW_BATCH) {
                uint32_t id[OVERFLOW_BATCH];
#pragma unroll
                for (int q = 0; q < OVERFLOW_BATCH; q++)
                    id[q] = i0 + q < OVERFLOW_CHUNK && next[i0 + q] != 0xffffffffu ? next[i0 + q]
                                                                                     : (uint32_t)g;
                float4 ob[OVERFLOW_BATCH];
                float2 mf[OVERFLOW_BATCH];
                uint32_t oc[OVERFLOW_BATCH];
#pragma unroll
                for (int q = 0; q < OVERFLOW_BATCH; q++) {
                    ob[q] = bodyIn[id[q]];
                    mf[q] = matXY[2 * (size_t)id[q]];
                    oc[q] = counts[id[q]];
                }
#pragma unroll
                for (int q = 0; q < OVERFLOW_BATCH; q++) {
                    if (i0 + q < OVERFLOW_CHUNK && next[i0 + q] != 0xffffffffu && taken < rest) {
                        last = id[q];
                        taken++;
                        float solid = k.liquid ? mats[last].z : 1.0f;
                        bool otherSleeps = sleeping && sleep.bodyAwake[last] != stamp;
                        accumulate_contact(k, recip, g, (int)last, ob[q],
                                           make_float4(mf[q].x, mf[q].y, solid, 0.0f), oc[q], otherSleeps,
                                           fallback, ppx, ppy, vpx, vpy, ownMat, own,
                                           writePressure, acc);
                    }
                }
            }
            if (taken == 0u) break;
            rest -= taken;
        }
    }
    if (k.bounded) {
        float weight = own + 1u < (uint32_t)RECIP_TABLE ? recip[own + 1u] : 1.0f / (float)(own + 1u);
        float fr = ownMat.y;
        wall_contact(k, ppx, 1.0f, 0.0f, vpx, vpy, weight, fr, writePressure, acc);
        wall_contact(k, k.vw - ppx, -1.0f, 0.0f, vpx, vpy, weight, fr, writePressure, acc);
        wall_contact(k, ppy, 0.0f, 1.0f, vpx, vpy, weight, fr, writePressure, acc);
        wall_contact(k, k.vh - ppy, 0.0f, -1.0f, vpx, vpy, weight, fr, writePressure, acc);
    }
    ppx += acc.pcx;
    ppy += acc.pcy;
    vpx += acc.vcx + acc.pcx / k.subDt;
    vpy += acc.vcy + acc.pcy / k.subDt;
    float4 o;
    if (k.bounded) {
        o.x = clampf(ppx / k.vw, 0.0f, MAXIMUM_UNIT);
        o.y = clampf(ppy / k.vh, 0.0f, MAXIMUM_UNIT);
    } else {
        ppx = glsl_mod(ppx, k.vw);
        ppy = glsl_mod(ppy, k.vh);
        o.x = ppx / k.vw;
        o.y = ppy / k.vh;
    }
    o.z = vpx / k.vw;
    o.w = vpy / k.vh;
    bool fluid = k.liquid && ownMat.z == 0.0f;
    if (writePressure && s.sleepOn) {
        // S/collision_cache_solve.comp:449-471: a body that moved this step,
        // or that nothing holds, keeps its tile awake for the next one.
        bool held = own > 0u || (s.bounded && (ppx < s.holdLo || ppy < s.holdLo ||
                                               ppx > s.holdHiX || ppy > s.holdHiY));
        if (fluid) held = true;
        float dx = vpx * s.dt, dy = vpy * s.dt;
        float d2 = dx * dx + dy * dy;
        float m2 = acc.pcx * acc.pcx + acc.pcy * acc.pcy;
        if (!held || d2 >= s.sleepDisp2 || m2 > s.settling2) {
            sleep.tileActive[sleep_tile(s, o.x, o.y)] = stamp + 1u;
            if (d2 >= s.tileMark2)
                sleep.tileActive[sleep_tile(s, o.x + dx / s.vw, o.y + dy / s.vh)] = stamp + 1u;
        }
    }
    if (writePressure && !fluid) pressures[g] = acc.pressure;
    // The next pass's drift (S/collision_cache_drift.comp:83-91; only the
    // first pass of a step kicks, so there is none here).
    if (nextDriftDt != 0.0f) {
        o.x += o.z * nextDriftDt;
        o.y += o.w * nextDriftDt;
        if (s.bounded) {
            o.x = clampf(o.x, 0.0f, MAXIMUM_UNIT);
            o.y = clampf(o.y, 0.0f, MAXIMUM_UNIT);
        } else {
            o.x = fractf(o.x);
            o.y = fractf(o.y);
        }
    }
    bodyOut[g] = o;
    if (insertNext) grid_insert(s, g, o.x, o.y, next);
    if (ADAPTIVE && writePressure && s.adaptive) deepest[g] = acc.deepest;
}

// Adaptive collisions, after the final pass (S/collision_cache_solve.comp:
// 239-243, 309-311): the deepest pair penetration, the most contacts of a
// body and the bodies at or above the dense contact count, over the bodies
// the pass solved (active, not sleeping); a wave's figures are reduced
// first, one atomic each a wave.
__global__ void __launch_bounds__(WG) k_adapt_stats(const SimDesc* __restrict__ sims,
                                                    const BodyBlock* __restrict__ blocks,
                                                    const uint32_t* __restrict__ counts,
                                                    const float* __restrict__ deepest,
                                                    SleepBufs sleep, uint32_t* __restrict__ adapt) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    if (!s.adaptive) return;
    int g = b.first + (int)threadIdx.x;
    uint32_t most = 0u, pen = 0u, dense = 0u;
    if (g - s.base < s.n) {
        bool solved = true;
        if (s.sleepOn) {
            const SleepState st = sleep.state[b.sim];
            solved = !sleep_on(s, st) || sleep.bodyAwake[g] == st.stamp;
        }
        if (solved) {
            most = counts[g];
            pen = __float_as_uint(deepest[g]);
            dense = most >= (uint32_t)s.adaptDenseContacts ? 1u : 0u;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        most = max(most, (uint32_t)__shfl_xor((int)most, o));
        pen = max(pen, (uint32_t)__shfl_xor((int)pen, o));
        dense += (uint32_t)__shfl_xor((int)dense, o);
    }
    if ((threadIdx.x & 63u) == 0u) {
        uint32_t* st = adapt + ADAPT_WORDS * b.sim;
        if (most != 0u) atomicMax(&st[3], most);
        if (pen != 0u) atomicMax(&st[1], pen);
        if (dense != 0u) atomicAdd(&st[4], dense);
    }
}

// Adaptive collisions, once per step before the first pass
// (S/collision_adaptive_reduce.comp): the largest speed in px/s of each
// adaptive simulation's active bodies, as float bits (non-negative, so their
// unsigned order is the float order), a wave's maximum first.
__global__ void __launch_bounds__(WG) k_adapt_reduce(const SimDesc* __restrict__ sims,
                                                     const BodyBlock* __restrict__ blocks,
                                                     const float4* __restrict__ body,
                                                     uint32_t* __restrict__ adapt) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    if (!s.adaptive) return;
    int g = b.first + (int)threadIdx.x;
    uint32_t bits = 0u;
    if (g - s.base < s.n) {
        const float4 v = body[g];
        const float sx = v.z * s.vw, sy = v.w * s.vh;
        bits = __float_as_uint(sqrtf(sx * sx + sy * sy));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bits = max(bits, (uint32_t)__shfl_xor((int)bits, o));
    if ((threadIdx.x & 63u) == 0u && bits != 0u) atomicMax(&adapt[ADAPT_WORDS * b.sim + 5], bits);
}

// GLSL uint(x) of a non-negative ceil: exact below 2^32; larger values and
// NaN (whose conversion GLSL leaves undefined) give the largest uint, which
// the substep clamp then takes.
__device__ __forceinline__ uint32_t ceil_to_uint(float x) {
    const float c = ceilf(x);
    if (!(c < 4294967296.0f)) return 0xffffffffu;
    return c > 0.0f ? (uint32_t)c : 0u;
}

// S/collision_adaptive_finalize.comp:16-56, one lane per simulation: the
// substeps of the step from its starting speed and the last final pass's
// penetration and dense bodies (all of them on the first step, and while the
// contact graph is dense), rounded up to the parity of the maximum so that
// the ping-pong buffers keep their roles across the skipped passes; then the
// last final pass's figures are cleared for this step's.
__global__ void __launch_bounds__(64) k_adapt_finalize(const SimDesc* __restrict__ sims,
                                                       int count, uint32_t* __restrict__ adapt) {
    const int i = (int)(blockIdx.x * 64 + threadIdx.x);
    if (i >= count) return;
    const SimDesc& s = sims[i];
    if (!s.adaptive) return;
    uint32_t* st = adapt + ADAPT_WORDS * i;
    const bool first = st[2] == 0u;
    const float speed = __uint_as_float(st[5]);
    const float deepest = __uint_as_float(st[1]);
    const uint32_t most = (uint32_t)s.adaptMax;
    const uint32_t displacementSteps = ceil_to_uint(speed * s.dt / s.adaptDisp);
    const uint32_t penetrationSteps = ceil_to_uint(deepest / s.adaptPen);
    uint32_t required = first ? most : min(max(max(displacementSteps, penetrationSteps), 1u), most);
    if (st[4] >= (uint32_t)s.adaptDenseBodies) required = most;
    if ((required & 1u) != (most & 1u)) required = min(required + 1u, most);
    st[0] = st[5];
    st[1] = 0u;
    st[2] = required;
    st[3] = 0u;
    st[4] = 0u;
    st[5] = 0u;
}

// S/stop_bodies.comp:9-15: zero the velocity of every active body in both
// ping-pong buffers.
__global__ void __launch_bounds__(WG) k_stop(const SimDesc* __restrict__ sims,
                                             const BodyBlock* __restrict__ blocks,
                                             float4* __restrict__ a, float4* __restrict__ b) {
    BodyBlock k = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[k.sim];
    int g = k.first + (int)threadIdx.x;
    if (g - s.base >= s.n) return;
    a[g].z = 0.0f; a[g].w = 0.0f;
    b[g].z = 0.0f; b[g].w = 0.0f;
}

// ---------------------------------------------------------------------------
// Host launchers

static inline GridBuf grid_of(const CollideGrid& c) { return GridBuf{c.heads, c.nodes, c.cellOf}; }

void launch_stop(hipStream_t st, const LaunchShape& ls, float4* a, float4* b) {
    if (ls.bodyBlocks == 0) return;
    launch(k_stop, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, a, b);
}

void launch_drift_insert(hipStream_t st, const LaunchShape& ls, float4* body, const float2* acc,
                         const float4* mats, float kickDt, float driftDt, int insert,
                         const CollideGrid& grid, int drifts) {
    if (ls.bodyBlocks == 0) return;
    launch(k_drift_insert, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       body, acc, mats, kickDt, driftDt, insert, grid_of(grid), ls.sleep,
                       (const uint32_t*)ls.adapt, drifts);
}

void launch_adapt(hipStream_t st, const LaunchShape& ls, int simCount, const float4* body) {
    if (ls.bodyBlocks == 0 || ls.adapt == nullptr) return;
    launch(k_adapt_reduce, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, body,
           ls.adapt);
    launch(k_adapt_finalize, dim3((simCount + 63) / 64), dim3(64), 0, st, ls.sims, simCount,
           ls.adapt);
}

void launch_count(hipStream_t st, const LaunchShape& ls, const CollideGrid& grid,
                  const CollideGrid& prev, int clearPrev, const float4* mats, uint32_t* counts,
                  uint32_t* contacts, int pass) {
    if (ls.bodyBlocks == 0) return;
    launch(ls.adapt ? k_count<true> : k_count<false>, dim3(ls.bodyBlocks), dim3(WG), 0, st,
           ls.sims, ls.order, ls.waveWork, grid_of(grid), grid_of(prev), clearPrev, mats, counts,
           contacts, ls.planeStride, ls.guardLimit, ls.pool, ls.sleep, pass,
           (const uint32_t*)ls.adapt);
}

void launch_solve(hipStream_t st, const LaunchShape& ls, const float4* bodyIn, float4* bodyOut,
                  const float4* mats, const uint32_t* counts, const uint32_t* contacts,
                  const CollideGrid& grid, const float2* fallback, float* pressures,
                  int writePressure, float nextDriftDt, int insertNext, const CollideGrid& next,
                  int pass, int lastInsert) {
    if (ls.bodyBlocks == 0) return;
    launch(ls.adapt ? k_solve<true> : k_solve<false>, dim3(ls.bodyBlocks), dim3(WG), 0, st,
           ls.sims, ls.order, bodyIn, bodyOut, mats, counts, contacts, ls.planeStride,
           grid_of(grid), ls.guardLimit, fallback, pressures, writePressure, nextDriftDt,
           insertNext, grid_of(next), ls.pool, ls.sleep, pass, lastInsert,
           (const uint32_t*)ls.adapt, ls.adaptDeepest);
}

void launch_adapt_stats(hipStream_t st, const LaunchShape& ls, const uint32_t* counts) {
    if (ls.bodyBlocks == 0 || ls.adapt == nullptr) return;
    launch(k_adapt_stats, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks, counts,
           (const float*)ls.adaptDeepest, ls.sleep, ls.adapt);
}

}  // namespace gs
