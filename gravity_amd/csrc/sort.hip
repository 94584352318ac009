// Device-wide exclusive scan and the spatial sort
// (S/spatial_sort_{key.glsl,count,scan,scatter,materials}.comp,
// K/GravityPipeline.kt:1269-1319).
//
// The sort is a counting sort on the reference's Morton tile key. The
// scatter hands out ranks by counting every tile back down to zero (so the
// counts never need clearing); a canonicalisation pass then orders each tile
// by original body index, which makes the permutation deterministic and
// stable. Positions, velocities, accelerations and materials move bit for bit.
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

constexpr int SCAN_BLOCK = 1024;  // values per workgroup (256 lanes x 4)

__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v) {
    int lane = (int)threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, (unsigned)o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive prefix of v over the workgroup; *total receives the sum.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* total) {
    __shared__ uint32_t waveSums[WG / 64];
    uint32_t inc = wave_inclusive(v);
    int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    if (lane == 63) waveSums[w] = inc;
    __syncthreads();
    uint32_t off = 0u, tot = 0u;
#pragma unroll
    for (int i = 0; i < WG / 64; i++) {
        uint32_t s = waveSums[i];
        if (i < w) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

// Local exclusive prefix of each 1024-value block; block sums to totals.
__global__ void __launch_bounds__(WG) k_scan_local(const uint32_t* __restrict__ in,
                                                   uint32_t* __restrict__ out, int len,
                                                   uint32_t* __restrict__ totals) {
    int base = (int)blockIdx.x * SCAN_BLOCK + (int)threadIdx.x * 4;
    uint32_t v[4];
    uint32_t sum = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int idx = base + i;
        v[i] = idx < len ? in[idx] : 0u;
        sum += v[i];
    }
    uint32_t total;
    uint32_t before = block_exclusive(sum, &total);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int idx = base + i;
        if (idx <= len) out[idx] = before;
        before += v[i];
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// Adds the block offsets: each workgroup sums the totals of the blocks
// before it. With rec, also writes each element's 16-byte record
// {prefix, count, 0, 0} and the first half of its 32-byte near record
// {count, 0, 0, 0} (the cell tables of the kick; the cell canon accumulates
// the mass into the third word of the one and the second of the other).
// With scanned, totals already holds each block's offset (k_scan_totals).
__global__ void __launch_bounds__(WG) k_scan_add(uint32_t* __restrict__ out, int len,
                                                 const uint32_t* __restrict__ totals, int scanned,
                                                 const uint32_t* __restrict__ in,
                                                 uint4* __restrict__ rec,
                                                 uint4* __restrict__ near) {
    // This lane's elements (SCAN_BLOCK / WG of them) and the totals of the
    // blocks before this one (four a lane per round) load unconditionally,
    // clamped, so they are in flight together.
    const int nb = (int)blockIdx.x;
    const int base = nb * SCAN_BLOCK + (int)threadIdx.x;
    constexpr int PER = SCAN_BLOCK / WG;
    uint32_t o[PER], c[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        o[q] = out[min(base + q * WG, len)];
        c[q] = in[min(base + q * WG, max(len - 1, 0))];
    }
    uint32_t add;
    if (scanned) {
        add = totals[nb];
    } else {
        uint32_t part = 0u;
        for (int i = (int)threadIdx.x; i < nb; i += 4 * WG) {
            uint32_t t[4];
#pragma unroll
            for (int q = 0; q < 4; q++) t[q] = totals[min(i + q * WG, nb - 1)];
#pragma unroll
            for (int q = 0; q < 4; q++) part += i + q * WG < nb ? t[q] : 0u;
        }
        block_exclusive(part, &add);
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int idx = base + q * WG;
        if (idx <= len) {
            uint32_t v = o[q] + add;
            out[idx] = v;
            if (rec != nullptr && idx < len) {
                rec[idx] = make_uint4(v, c[q], 0u, 0u);
                near[2 * (size_t)idx] = make_uint4(c[q], 0u, 0u, 0u);
            }
        }
    }
}

// Exclusive scan of the nb block totals in place, by one workgroup,
// TOTALS_PER consecutive totals a lane per round.
constexpr int TOTALS_PER = 16;
__global__ void __launch_bounds__(WG) k_scan_totals(uint32_t* __restrict__ totals, int nb) {
    uint32_t carry = 0u;
    for (int base = 0; base < nb; base += WG * TOTALS_PER) {
        const int first = base + (int)threadIdx.x * TOTALS_PER;
        uint32_t v[TOTALS_PER];
        uint32_t sum = 0u;
#pragma unroll
        for (int q = 0; q < TOTALS_PER; q++) {
            v[q] = first + q < nb ? totals[first + q] : 0u;
            sum += v[q];
        }
        uint32_t total;
        uint32_t before = block_exclusive(sum, &total) + carry;
#pragma unroll
        for (int q = 0; q < TOTALS_PER; q++) {
            if (first + q < nb) totals[first + q] = before;
            before += v[q];
        }
        carry += total;
    }
}

int scan_totals_size(int len) { return (len + 1 + SCAN_BLOCK - 1) / SCAN_BLOCK; }

void launch_scan(hipStream_t st, const uint32_t* in, uint32_t* out, int len, uint32_t* totals,
                 int directBlocks, uint4* rec, uint4* near) {
    int blocks = scan_totals_size(len);
    const int scanned = blocks > directBlocks ? 1 : 0;
    launch(k_scan_local, dim3(blocks), dim3(WG), 0, st, in, out, len, totals);
    if (scanned) launch(k_scan_totals, dim3(1), dim3(WG), 0, st, totals, blocks);
    launch(k_scan_add, dim3(blocks), dim3(WG), 0, st, out, len, totals, scanned, in, rec, near);
}

// ---------------------------------------------------------------------------
// Spatial sort

__device__ __forceinline__ uint32_t spread_bits(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// S/spatial_sort_key.glsl:16-25
__device__ __forceinline__ uint32_t sort_key(float x, float y, int wb, int hb) {
    int sx = 1 << wb, sy = 1 << hb;
    uint32_t tx = (uint32_t)floor_index(x * (float)sx, sx);
    uint32_t ty = (uint32_t)floor_index(y * (float)sy, sy);
    int common = wb < hb ? wb : hb;
    uint32_t low = (1u << common) - 1u;
    uint32_t morton = spread_bits(tx & low) | (spread_bits(ty & low) << 1);
    return morton | (((tx | ty) >> common) << (2 * common));
}

// S/spatial_sort_count.comp:8-12, one atomic per run of equal keys.
__global__ void __launch_bounds__(WG) k_sort_count(const SimDesc* __restrict__ sims,
                                                   const BodyBlock* __restrict__ blocks,
                                                   const float4* __restrict__ body,
                                                   uint32_t* tileCount, uint32_t* keyOf) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    const bool active = g - s.base < s.n;
    uint32_t t = 0xffffffffu;
    if (active) {
        float4 p = body[g];
        t = (uint32_t)s.tileBase + sort_key(p.x, p.y, s.sortWb, s.sortHb);
        keyOf[g] = t;
    }
    const LaneRun run = lane_run(t);
    if (active && run.rank == 0) atomicAdd(&tileCount[t], (uint32_t)run.length);
}

// S/spatial_sort_scatter.comp:14-28: counting down returns every count to 0.
// A run takes its ranks with one atomic and hands them out in lane order.
__global__ void __launch_bounds__(WG) k_sort_scatter(const SimDesc* __restrict__ sims,
                                                     const BodyBlock* __restrict__ blocks,
                                                     const uint32_t* __restrict__ keyOf,
                                                     uint32_t* tileCount,
                                                     const uint32_t* __restrict__ tilePrefix,
                                                     uint32_t* __restrict__ slotSrc,
                                                     uint32_t* __restrict__ slotKey) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int g = b.first + (int)threadIdx.x;
    const bool active = g - s.base < s.n;
    const uint32_t t = active ? keyOf[g] : 0xffffffffu;
    const LaneRun run = lane_run(t);
    uint32_t top = 0u;
    if (active && run.rank == 0) top = atomicSub(&tileCount[t], (uint32_t)run.length);
    top = (uint32_t)__shfl((int)top, run.start, 64);
    if (!active) return;
    uint32_t rank = top - 1u - (uint32_t)run.rank;
    int slot = s.base + (int)(tilePrefix[t] - (uint32_t)s.rank0 + rank);
    slotSrc[slot] = (uint32_t)g;
    slotKey[slot] = t;
}

// Orders every tile by original index and moves the bodies; a tile of more
// than big.limit bodies is listed for k_sort_canon_big, which orders it.
__global__ void __launch_bounds__(WG) k_sort_canon(
    const SimDesc* __restrict__ sims, const BodyBlock* __restrict__ blocks,
    const uint32_t* __restrict__ slotKey, const uint32_t* __restrict__ tilePrefix,
    const uint32_t* __restrict__ slotSrc, const float4* __restrict__ bodyIn,
    const float2* __restrict__ accIn, const float4* __restrict__ matIn,
    float4* __restrict__ bodyOut, float2* __restrict__ accOut, float4* __restrict__ matOut,
    uint32_t* __restrict__ dest, BigSegments big) {
    BodyBlock b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const SimDesc& s = sims[b.sim];
    int d = b.first + (int)threadIdx.x;
    if (d - s.base >= s.n) return;
    uint32_t src = slotSrc[d];
    uint32_t t = slotKey[d];
    // the moved data load with the tile's bounds, ahead of the rank
    const float4 bo = bodyIn[src];
    const float2 ac = accIn[src];
    const float4 ma = matIn[src];
    int s0 = s.base + (int)(tilePrefix[t] - (uint32_t)s.rank0);
    int s1 = s.base + (int)(tilePrefix[t + 1] - (uint32_t)s.rank0);
    if (is_big_segment(big, s1 - s0)) {
        if (d == s0) list_big_segment(big, s0, s1, s.base);
        return;
    }
    int fin = s0 + segment_rank<false>(slotSrc, s0, s1, src);
    bodyOut[fin] = bo;
    accOut[fin] = ac;
    matOut[fin] = ma;
    dest[src] = (uint32_t)(fin - s.base);
}

// The tiles k_sort_canon listed as big, each ordered by ascending original
// index by one workgroup, and their bodies moved.
__global__ void __launch_bounds__(BIG_NT) k_sort_canon_big(
    const uint32_t* __restrict__ slotSrc, const float4* __restrict__ bodyIn,
    const float2* __restrict__ accIn, const float4* __restrict__ matIn,
    float4* __restrict__ bodyOut, float2* __restrict__ accOut, float4* __restrict__ matOut,
    uint32_t* __restrict__ dest, BigSegments big) {
    order_big_segments<false>(big, slotSrc, [&](const BigSegment& seg, int rank, uint32_t src) {
        const uint32_t fin = seg.s0 + (uint32_t)rank;
        bodyOut[fin] = bodyIn[src];
        accOut[fin] = accIn[src];
        matOut[fin] = matIn[src];
        dest[src] = fin - seg.base;
    });
}

void launch_sort_count(hipStream_t st, const LaunchShape& ls, const float4* body,
                       uint32_t* tileCount, uint32_t* keyOf) {
    if (ls.bodyBlocks == 0) return;
    launch(k_sort_count, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       body, tileCount, keyOf);
}

void launch_sort_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* keyOf,
                         uint32_t* tileCount, const uint32_t* tilePrefix, uint32_t* slotSrc,
                         uint32_t* slotKey) {
    if (ls.bodyBlocks == 0) return;
    launch(k_sort_scatter, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       keyOf, tileCount, tilePrefix, slotSrc, slotKey);
}

void launch_sort_canon(hipStream_t st, const LaunchShape& ls, const uint32_t* slotKey,
                       const uint32_t* tilePrefix, const uint32_t* slotSrc,
                       const float4* bodyIn, const float2* accIn, const float4* matIn,
                       float4* bodyOut, float2* accOut, float4* matOut, uint32_t* dest,
                       const BigSegments& big) {
    if (ls.bodyBlocks == 0) return;
    launch(k_sort_canon, dim3(ls.bodyBlocks), dim3(WG), 0, st, ls.sims, ls.blocks,
                       slotKey, tilePrefix, slotSrc, bodyIn, accIn, matIn, bodyOut, accOut, matOut,
                       dest, big);
    if (big.limit != 0)
        launch(k_sort_canon_big, dim3(BIG_GRID), dim3(BIG_NT), 0, st, slotSrc, bodyIn, accIn,
               matIn, bodyOut, accOut, matOut, dest, big);
}

}  // namespace gs
