// Shared definitions of the MI355X particle step: the device-resident
// simulation descriptor, the batch block map and the exact float helpers.
//
// Every kernel serves a batch of independent simulations (a single
// simulation is a batch of one). Bodies of all simulations live in one set of
// SoA arrays; simulation s owns global indices [base, base + capacity), with
// base a multiple of 256, so a 256-wide workgroup never spans two of them.
// A launch over bodies walks a block table of int2 {sim, first body}.
//
// All HIP code is compiled with -ffp-contract=off: each float expression
// rounds exactly as written, like the shaders' `precise` contact predicate,
// and the CPU oracle reproduces it bit for bit.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace gs {

constexpr int WG = 256;               // workgroup width of every body kernel
constexpr int CONTACT_SLOTS = 16;     // contacts a count keeps in registers, ascending
constexpr int CONTACT_SORTED = 64;    // contacts a body lists in the contact array, ascending
constexpr int CONTACT_PLANES = 112;   // planes of the contact array: the listed contacts,
                                      // then a crowded body's unsorted overflow
constexpr int XCDS = 8;               // accelerator dies of an MI355X, each with its own L2
constexpr float MAXIMUM_UNIT = 0.99999994f;  // largest float below 1

// Glued blocks (S/rigid_layout.glsl): a block's partners in the contact
// memory are the other blocks, the four walls and the loose bodies as one;
// a pair belongs to its lower block.
constexpr int RIGID_MAXIMUM_GROUPS = 256;
constexpr int RIGID_CONTACT_PARTNERS = RIGID_MAXIMUM_GROUPS + 5;
constexpr int RIGID_CONTACT_KEYS = RIGID_MAXIMUM_GROUPS * RIGID_CONTACT_PARTNERS;

// RigidGroup (S/rigid_layout.glsl:4-15), 160 bytes.
struct RigidGroup {
    float4 frame;    // reference centre (px), centre (px)
    float4 motion;   // velocity (px/s), rotation (cos, sin)
    float4 scalars;  // angular velocity (rad/s), inertia sum |r0|^2, mass (members), struck this pass
    int4 tiles;      // first tile, tile count, first hull value, hull points
    float4 shape;    // distance field: rest-frame origin (px), spacing (px), smallest border value
    int4 grid;       // first shape value, grid width, grid height, partners in contact this pass
    float4 bounds;   // rest-frame extent of the member centres: min x, max x, min y, max y (px)
    int4 guard;      // shift out of other blocks asked for: +x, +y, -x, -y (1/65536 px)
    int4 support;    // shift off bodies wedged against a wall, likewise
    float4 pose;     // centre (px) and rotation after the current pass's drift
};
static_assert(sizeof(RigidGroup) == 160, "RigidGroup is the shader's 160-byte record");

// Per-simulation constants, mirrored from GravityConfig and the viewport
// (K/GravityPipeline.kt:73-241). Derived floats are computed on the host
// with the same operations the shaders apply to their uniforms.
struct SimDesc {
    // bodies
    int n;              // active bodies
    int base;           // first global body index
    int cap;            // capacity
    int bounded;        // walls instead of the torus
    int mutual;         // mutual gravity
    int liquid;
    int collide;        // contact passes (0: drift-only gravity integrator)
    int pad0;
    float vw, vh;       // viewport, px
    float aspect;       // vw / vh
    float dt;           // step
    float subDt;        // collision substep (effective substep dt)
    // collisions
    float diameter;
    float diameter2;    // diameter * diameter
    float restitution;
    float coincidentEps;
    int cgw, cgh;       // broad-phase grid (cells >= diameter + guard)
    int cellBase;       // first cell of this simulation in the broad-phase heads
    // spatial sort (S/spatial_sort_key.glsl)
    int sortWb, sortHb;
    int tileBase;       // first tile of this simulation in the tile counts
    int rank0;          // bodies of earlier simulations (scan offset)
    // P3M
    int pw, ph;         // physical grid
    int mw, mh;         // mesh (doubled under walls)
    int physBase;       // first physical cell of this simulation
    int meshBase;       // first mesh node of this simulation
    int rx, ry;         // near-field radii, cells
    int exactLimit;     // exactNeighborBodyLimit
    int homeLimit;      // maximumHomeBodyInteractions
    float G;
    float totalMass;
    float sigma;        // softeningCells / ph
    float twoSigma2;    // 2 * sigma * sigma
    float minDist2;     // (minimumDistanceCellFraction * min(cellW, cellH))^2
    float invMassScale; // 1 / massScale
    float invTotalMass; // 1 / totalMass, folded into the spectral multiply
    float fieldX, fieldY;  // uniform field, normalised units
    float massScale;    // fixed-point deposit scale, as float
    int pad1;
    // Reciprocals of the kick (FP32-tolerance path; see k_kick).
    float invAspect;    // 1 / aspect
    float invTwoSigma2; // 1 / (2 sigma^2)
    float invPw, invPh; // 1 / pw, 1 / ph
    float gOverM;       // G / totalMass
    int pad2, pad3, pad4;
    // Position-based liquid (K/LiquidKernel.kt, K/LiquidSolver.kt:64-101),
    // px units; each float is the uniform the shaders receive.
    float lh, lh2;          // kernel radius, and its square as the shaders form it
    float lpoly6, lspiky;   // 4 / (pi h^8), -30 / (pi h^5)
    float lrest, lrelax;    // rest density, relaxation
    float lrowStep;         // d sqrt(3) / 2, the wall rows' spacing
    float lwallScale;       // poly6 * 32 / 35 / d
    float lpressureScale;   // h / dt
    float lviscosity;       // viscosity / rest density
    float lradius;          // d / 2, where the walls stop a body
    int lgw, lgh;           // liquid grid, cells >= h
    int liqCellBase;        // first liquid cell of this simulation
    int lviscous;           // viscosity > 0: the XSPH pass runs
    int lpad1;
    // Sleeping tiles (S/sleep_layout.glsl, K/SleepTiles.kt:37-107)
    int sleepOn;            // CollisionConfig.sleep with the contact passes on
    int tilesX, tilesY;     // 32 px tiles, rounded up
    int sleepTileBase;      // first tile of this simulation
    float sleepDisp2;       // (sleepSpeedPixels * dt)^2
    float holdLo;           // d/2 + 0.05 d: a body this close to a wall is held
    float holdHiX, holdHiY; // viewport - d/2 - 0.05 d
    float settling2;        // (0.005 d)^2: a pushed body is not at rest
    float tileMark2;        // (16 px)^2: a faster body also wakes where it heads
    int spad0, spad1;
    // Adaptive collisions (S/collision_adaptive_finalize.comp, K/GravityPipeline.kt:1155-1211)
    int adaptive;           // the device picks each step's substeps (contact passes on)
    int adaptMax;           // maximumSubsteps
    float adaptDisp;        // diameter * allowedDisplacementFraction, px
    float adaptPen;         // diameter * allowedPenetrationFraction, px
    int adaptDenseContacts; // denseContactThreshold
    int adaptDenseBodies;   // max(minimumDenseBodies, ceil(n * denseBodyFraction))
    int apad0, apad1;
};

// Adaptive state of a simulation, ADAPT_WORDS words (the reference's five,
// then the running speed maximum of the step being measured): [0] the largest
// speed the step started with (float bits), [1] the deepest pair penetration
// of the last final pass, [2] the substeps chosen (0 before the first step),
// [3] the most contacts of a body and [4] the bodies at or above the dense
// contact count in the last final pass, [5] the running speed maximum.
constexpr int ADAPT_WORDS = 8;

// Per-simulation sleep state, advanced on the device at the end of a step
// (K/SleepTiles.kt endStep): the step stamp (never 0) and the steps left in
// which every body stays awake.
struct SleepState { uint32_t stamp; int wakeSteps; int pad0, pad1; };
constexpr int SLEEP_WAKE_STEPS = 2;  // K/SleepTiles.kt WAKE_STEPS

struct SleepBufs {
    SleepState* state;     // per simulation
    uint32_t* bodyAwake;   // per body: the stamp of the step it is awake in
    uint32_t* tileAwake;   // per tile: the stamp of the step it is awake in
    uint32_t* tileActive;  // per tile: the stamp of the step it was marked for
};

// uSleep of the current step: tiles on and the wake steps over.
__device__ __forceinline__ bool sleep_on(const SimDesc& s, const SleepState& st) {
    return s.sleepOn && st.wakeSteps == 0;
}

// S/sleep_layout.glsl sleepTile, as a global tile index.
__device__ __forceinline__ int sleep_tile(const SimDesc& s, float x, float y) {
    float fx = floorf(x * (float)s.tilesX), fy = floorf(y * (float)s.tilesY);
    int tx = fx >= 0.0f ? (fx < (float)(s.tilesX - 1) ? (int)fx : s.tilesX - 1) : 0;
    int ty = fy >= 0.0f ? (fy < (float)(s.tilesY - 1) ? (int)fy : s.tilesY - 1) : 0;
    return s.sleepTileBase + tx + ty * s.tilesX;
}

// Global work tables (built on the host whenever the batch shape changes).
struct BodyBlock { int sim; int first; };
// A contact workgroup's body block with its index, so that the dispatch
// order hands a workgroup its block in one load.
struct OrderedBlock { int sim; int first; int index; int pad; };

// Workgroups are dealt to the XCDs round robin (workgroup i runs on XCD
// i % 8). Body blocks are spatially sorted, so hand each XCD runs of
// XCD_CHUNK consecutive blocks: neighbour gathers within a run hit that
// XCD's own L2, while runs from all over the domain keep dense and sparse
// regions spread over every XCD. Blocks past the last full round keep their
// index (the map is a bijection of [0, n)).
constexpr int XCD_CHUNK = 8;
__host__ __device__ inline int xcd_block(int i, int n) {
    const int round = XCDS * XCD_CHUNK;
    if (i >= n / round * round) return i;
    int xcd = i % XCDS, local = i / XCDS;
    return ((local / XCD_CHUNK) * XCDS + xcd) * XCD_CHUNK + local % XCD_CHUNK;
}

// Every active body of a launch over nblocks body blocks, as f(sim, g).
// The grid may be smaller than nblocks (a multiple of 8, so a workgroup
// keeps its XCD): each workgroup then strides over several blocks.
template <typename F>
__device__ __forceinline__ void for_each_body(const SimDesc* __restrict__ sims,
                                              const BodyBlock* __restrict__ blocks, int nblocks,
                                              F&& f) {
    for (int lb = (int)blockIdx.x; lb < nblocks; lb += (int)gridDim.x) {
        BodyBlock b = blocks[xcd_block(lb, nblocks)];
        const SimDesc& s = sims[b.sim];
        int g = b.first + (int)threadIdx.x;
        if (g - s.base < s.n) f(s, g);
    }
}

__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    return fminf(fmaxf(v, lo), hi);
}
__device__ __forceinline__ float fractf(float v) { return v - floorf(v); }
// GLSL mod(x, y) = x - y * floor(x / y)
__device__ __forceinline__ float glsl_mod(float x, float y) { return x - y * floorf(x / y); }
__device__ __forceinline__ int wrap_axis(int c, int size) {
    if (c < 0) c += size; else if (c >= size) c -= size;
    return c;
}
// Float index of an already floored coordinate, clamped to [0, size) in float
// first so that non-finite positions land on a valid cell.
__device__ __forceinline__ int floor_index(float v, int size) {
    float f = floorf(v);
    if (!(f >= 0.0f)) return 0;
    if (f > (float)(size - 1)) return size - 1;
    return (int)f;
}

// A run of equal keys in consecutive lanes of a wave: the first lane of the
// run, its length and this lane's place in it. Every lane of the wave must
// call it; a lane without a key passes a value no key takes (0xffffffff).
// Bodies arrive spatially sorted, so the lanes of a wave share few cells:
// one atomic per run replaces one per lane on the same address.
struct LaneRun {
    int start, length, rank;
};
__device__ __forceinline__ LaneRun lane_run(uint32_t key) {
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t prev = (uint32_t)__shfl_up((int)key, 1u, 64);
    const bool first = lane == 0 || key != prev;
    const unsigned long long starts = __ballot(first);
    const unsigned long long upTo = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int start = 63 - __clzll((long long)(starts & upTo));
    const unsigned long long above = starts & ~upTo;
    const int next = above ? __ffsll((long long)above) - 1 : 64;
    return LaneRun{start, next - start, lane - start};
}

// The rank of v in the segment list[s0, s1) of distinct values that holds
// it: the number of entries above v (ABOVE) or below it. The loads go
// SEGMENT_BATCH at a time, clamped to the segment, so a long segment (a
// crowded cell) costs one memory round trip per batch, not per entry.
constexpr int SEGMENT_BATCH = 16;
template <bool ABOVE>
__device__ __forceinline__ int segment_rank(const uint32_t* __restrict__ list, int s0, int s1,
                                            uint32_t v) {
    int rank = 0;
    for (int e = s0; e < s1; e += SEGMENT_BATCH) {
        uint32_t u[SEGMENT_BATCH];
#pragma unroll
        for (int q = 0; q < SEGMENT_BATCH; q++) u[q] = list[min(e + q, s1 - 1)];
#pragma unroll
        for (int q = 0; q < SEGMENT_BATCH; q++)
            rank += e + q < s1 && (ABOVE ? u[q] > v : u[q] < v) ? 1 : 0;
    }
    return rank;
}

// Segments too long for segment_rank: a cell or tile holding more than a
// launch's bigLimit bodies (the dense core of a cluster can put thousands in
// one) is ordered by a workgroup instead, with a bitonic sort in LDS, so its
// cost grows as c log^2 c on one CU rather than c^2 loads spread over its
// lanes. The canon kernels list such segments {first slot, end slot, the
// simulation's first body} in `list`, counted by counters[0]; the segment
// kernel's last workgroup returns both counters to zero (counters[1] counts
// the workgroups that are done), so the next launch starts from an empty list.
struct BigSegment { uint32_t s0, s1, base, pad; };
struct BigSegments {
    BigSegment* list;
    uint32_t* counters;
    int limit;        // segments longer than this are big; 0: none is
    int pad;
};
constexpr int BIG_NT = 256;        // lanes of a segment workgroup
constexpr int BIG_LDS = 8192;      // longest segment sorted in LDS (32 KB)
constexpr int BIG_GRID = 1024;     // segment workgroups per launch (four per CU)
constexpr int BIG_LIMIT_MIN = 32;  // smallest limit the engine accepts

// Bitonic sort of keys[0, P) in LDS (P a power of two), ascending or, with
// DESC, descending; every lane of the workgroup calls it.
template <bool DESC>
__device__ __forceinline__ void lds_bitonic(uint32_t* keys, int P) {
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = (int)threadIdx.x; i < P / 2; i += BIG_NT) {
                const int a = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                const int b = a + j;
                const uint32_t x = keys[a], y = keys[b];
                const bool up = (a & k) == 0;  // this pair's run goes up
                const bool swap = DESC ? (up ? x < y : x > y) : (up ? x > y : x < y);
                if (swap) {
                    keys[a] = y;
                    keys[b] = x;
                }
            }
            __syncthreads();
        }
    }
}

// Every listed big segment, its distinct values list[s0, s1) in order
// (descending with DESC), as write(segment, position, value); one workgroup
// per segment, the grid striding over the list. Only segments of at most
// BIG_LDS values are listed (is_big_segment).
template <bool DESC, typename W>
__device__ __forceinline__ void order_big_segments(const BigSegments& big,
                                                   const uint32_t* __restrict__ list, W&& write) {
    __shared__ uint32_t keys[BIG_LDS];
    __shared__ uint32_t listed;
    if (threadIdx.x == 0) listed = atomicAdd(&big.counters[0], 0u);
    __syncthreads();
    for (uint32_t i = blockIdx.x; i < listed; i += gridDim.x) {
        const BigSegment seg = big.list[i];
        const int s0 = (int)seg.s0, n = (int)(seg.s1 - seg.s0);
        int P = 64;
        while (P < n) P <<= 1;
        for (int e = (int)threadIdx.x; e < P; e += BIG_NT)
            keys[e] = e < n ? list[s0 + e] : (DESC ? 0u : 0xffffffffu);
        __syncthreads();
        lds_bitonic<DESC>(keys, P);
        for (int e = (int)threadIdx.x; e < n; e += BIG_NT) write(seg, e, keys[e]);
        __syncthreads();
    }
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&big.counters[1], 1u) == gridDim.x - 1u) {
        atomicExch(&big.counters[0], 0u);
        atomicExch(&big.counters[1], 0u);
    }
}

// Whether a segment of n values goes to the workgroup sort: longer than the
// limit, and short enough for LDS (a longer one, rare, keeps the per-lane
// ranks, spread over as many lanes as it has values).
__device__ __forceinline__ bool is_big_segment(const BigSegments& big, int n) {
    return big.limit != 0 && n > big.limit && n <= BIG_LDS;
}

// Appends the segment [s0, s1) to the big list (one lane per segment).
__device__ __forceinline__ void list_big_segment(const BigSegments& big, int s0, int s1, int base) {
    const uint32_t i = atomicAdd(&big.counters[0], 1u);
    big.list[i] = BigSegment{(uint32_t)s0, (uint32_t)s1, (uint32_t)base, 0u};
}

// Every launch goes through launch(). While a profiling sink is set (the
// engine's eager profiling steps), each launch carries its own event pair,
// stamped by the kernel's dispatch itself: a kernel's time then excludes the
// queue's gaps between launches, as a kernel trace reports it.
struct LaunchEvents {
    hipEvent_t start, stop;
};
std::vector<LaunchEvents>*& launch_sink();

template <typename K, typename... A>
inline void launch(K kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t st, A... args) {
    std::vector<LaunchEvents>* sink = launch_sink();
    if (sink == nullptr) {
        hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
        return;
    }
    LaunchEvents e{};
    if (hipEventCreate(&e.start) != hipSuccess || hipEventCreate(&e.stop) != hipSuccess) {
        hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
        return;
    }
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, e.start, e.stop, 0u, args...);
    sink->push_back(e);
}

}  // namespace gs
