// Launchers of the device kernels; the engine is the only caller.
#pragma once

#include "common.hpp"

namespace gs {

// Broad-phase grid buffer of one pass parity.
struct CollideGrid {
    uint32_t* heads;   // per cell: first body + 1, or 0; heads[-1] and heads[cells] are 0
    float4* nodes;     // per body: x, y in px, next body + 1 (bits), cx | cy << 16 (bits)
    int* cellOf;       // per body: the cell of its insertion
};

// Contacts past the listed ones, for bodies with more than CONTACT_SORTED:
// the count appends each such body's contacts past its CONTACT_SLOTS
// smallest (in any order) to items[base[g], base[g] + count - CONTACT_SLOTS),
// from a counter `top` that the solve returns to zero; a body that finds the
// pool full gets NO_POOL and the solve walks the grid for its remaining
// contacts instead.
constexpr uint32_t NO_POOL = 0xffffffffu;
struct ContactPool {
    uint32_t* items;
    uint32_t* base;
    uint32_t* top;
    uint32_t capacity;
};

// Shape of the body launches of a batch (device pointers + sizes).
struct LaunchShape {
    const SimDesc* sims;
    const BodyBlock* blocks;
    int bodyBlocks;      // workgroups over all active bodies
    int planeStride;     // body slots of the batch (stride of the contact planes)
    int guardLimit;      // bound on any linked-list walk
    ContactPool pool;    // remaining contacts of bodies with more than the listed ones
    SleepBufs sleep;     // sleeping tiles of the batch
    // Dispatch order of the contact kernels: workgroup i takes body block
    // order[i] (a permutation keeping every block on the XCD xcd_block gives
    // it, with the block's data), and waveWork[4 * block + wave] receives the
    // contacts the count found in that wave, from which the order is rebuilt
    // once per step.
    const OrderedBlock* order;
    uint32_t* waveWork;
    // ADAPT_WORDS per simulation when any simulation of the batch has
    // adaptive collisions, else null (the passes then take the plain kernels).
    uint32_t* adapt;
    float* adaptDeepest;   // per body: its deepest pair penetration in the last final pass
    // Boundaries of the batch's simulations: WALLS_NONE (all on the torus),
    // WALLS_ALL, or WALLS_MIXED; kernels specialised on walls launch the
    // variants the batch needs.
    int walls;
};
enum { WALLS_NONE = 0, WALLS_ALL = 1, WALLS_MIXED = 2 };

// Mesh work item: {simulation, first line, lines}.
struct LineWork { int sim; int first; int count; int pad; };

// ---- scan: out[i] = sum of in[0..i) for i in [0, len]; totals is scratch ----
// Exclusive scan of len values (out[len] = total). rec, if given, receives
// {prefix, count, 0, 0} as each element's 16-byte record, and near (two
// uint4 per element) {count, 0, 0, 0} as the first half of its near record.
// A scan of more than directBlocks blocks of 1024 values scans the block
// totals in a launch of their own; a shorter one has each block sum the
// totals before it.
void launch_scan(hipStream_t st, const uint32_t* in, uint32_t* out, int len, uint32_t* totals,
                 int directBlocks, uint4* rec = nullptr, uint4* near = nullptr);
int scan_totals_size(int len);

// ---- spatial sort ----
void launch_sort_count(hipStream_t st, const LaunchShape& ls, const float4* body,
                       uint32_t* tileCount, uint32_t* keyOf);
// slotKey receives the tile of the body in each slot, which the canon reads
// with the slot instead of through the body.
void launch_sort_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* keyOf,
                         uint32_t* tileCount, const uint32_t* tilePrefix, uint32_t* slotSrc,
                         uint32_t* slotKey);
void launch_sort_canon(hipStream_t st, const LaunchShape& ls, const uint32_t* slotKey,
                       const uint32_t* tilePrefix, const uint32_t* slotSrc,
                       const float4* bodyIn, const float2* accIn, const float4* matIn,
                       float4* bodyOut, float2* accOut, float4* matOut, uint32_t* dest,
                       const BigSegments& big);

// ---- collisions ----
void launch_stop(hipStream_t st, const LaunchShape& ls, float4* a, float4* b);
// `drifts` drifts of driftDt after the kick, in one launch (the gravity-only
// step's substeps); a contact pass's drift is one.
void launch_drift_insert(hipStream_t st, const LaunchShape& ls, float4* body, const float2* acc,
                         const float4* mats, float kickDt, float driftDt, int insert,
                         const CollideGrid& grid, int drifts = 1);
// Rebuilds ls.order: on every XCD, its body blocks by descending contacts of
// the last count, so the heaviest blocks are dispatched first, one per CU.
void launch_block_order(hipStream_t st, const LaunchShape& ls, OrderedBlock* order);
// The order xcd_block alone gives, on the host.
void identity_block_order(const BodyBlock* blocks, int count, OrderedBlock* order);
// pass: the substep index of the pass (-1 for a pass outside the substeps,
// always active); with adaptive collisions a substep at or past the step's
// chosen count does nothing. lastInsert: whether the last active substep's
// solve inserts into the next grid (it does unless liquid passes follow).
void launch_count(hipStream_t st, const LaunchShape& ls, const CollideGrid& grid,
                  const CollideGrid& prev, int clearPrev, const float4* mats, uint32_t* counts,
                  uint32_t* contacts, int pass);
void launch_solve(hipStream_t st, const LaunchShape& ls, const float4* bodyIn, float4* bodyOut,
                  const float4* mats, const uint32_t* counts, const uint32_t* contacts,
                  const CollideGrid& grid, const float2* fallback, float* pressures,
                  int writePressure, float nextDriftDt, int insertNext, const CollideGrid& next,
                  int pass, int lastInsert);
// Adaptive collisions: the largest speed of every adaptive simulation's
// active bodies, then each one's substeps for the step
// (S/collision_adaptive_{reset,reduce,finalize}.comp).
void launch_adapt(hipStream_t st, const LaunchShape& ls, int simCount, const float4* body);
// ... and after the final pass, its penetration and contact figures.
void launch_adapt_stats(hipStream_t st, const LaunchShape& ls, const uint32_t* counts);

// ---- P3M gravity ----
void launch_bin(hipStream_t st, const LaunchShape& ls, const float4* body, uint32_t* cellCount,
                uint32_t* homeOf);
// slotCell receives the cell of the body in each slot, which the canon reads
// with the slot instead of through the body.
void launch_cell_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* homeOf,
                         uint32_t* cellCount, const uint32_t* cellPrefix, uint32_t* slotBody,
                         uint32_t* slotCell);
// cellNear: per physical cell two uint4, {count, mass in 1/256, x0, y0} and
// {x1, y1, m1, m0}: a cell of at most NEAR_INLINE bodies carries its list's
// bodies (x, y normalised, m mass) in the record itself.
constexpr int NEAR_INLINE = 2;
void launch_cell_canon(hipStream_t st, int slots, const uint32_t* slotCell,
                       const uint32_t* cellPrefix, const uint32_t* slotBody, const float4* body,
                       const float4* mats, float4* cellBody, uint4* cellRec, uint4* cellNear,
                       const BigSegments& big);
// density (nullable) receives the fixed-point deposit of every live mesh row;
// runs: sum a lane's bodies of one cell in registers before the LDS atomics.
void launch_fft_rows_forward(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                             const uint32_t* cellPrefix, const float4* cellBody, float2* spec,
                             const float2* twiddles, bool runs, uint32_t* density = nullptr);
// forwardOnly: leave the 2D density spectrum in spec (every row) instead of
// multiplying and transforming back.
void launch_fft_columns(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                        float2* spec, const float2* green, const float2* twiddles,
                        bool forwardOnly = false);
void launch_fft_rows_inverse(hipStream_t st, const SimDesc* sims, const LineWork* work, int items,
                             float2* spec, const float2* twiddles);
// The transforms' twiddles (`twiddles` above): per line length and radix, the
// stages' tables drawn from the FP64-rounded exp(-2 pi i m / 2048).
int fft_twiddle_entries();
void fft_twiddle_table(std::vector<float>& out);
// cellRec: per physical cell {first slot, body count, mass in 1/256, 0}.
void launch_kick(hipStream_t st, const LaunchShape& ls, float4* body, float2* acc,
                 const float2* force, const uint4* cellRec, const uint4* cellNear,
                 const float4* cellBody, int initial);

// ---- the trajectory preview (simulation 0 of the batch) ----
// Near-field terms one stencil chunk of 64 cells may hold (the home cell's
// homeLimit and 63 cells' exactLimit at most).
constexpr int TRAJ_TERMS = 2048;
struct TrajectoryInputs {
    const float2* force;      // the mesh force of the last force build
    const uint4* cellRec;     // its cell table
    const float4* cellBody;   // its cell lists (body ids in .w)
    const float4* body;       // the current bodies
    const float4* mats;
    int steps;                // TrajectoryConfig.steps
    float maximumDt;          // maximumTimeStep
    float segmentPx;          // targetSegmentPixels
    float maxPathPx;          // maximumPathViewportDiagonals x the viewport diagonal
    float minimumDashPx;      // minimumDashPixels
    float originX, originY;   // placed (wrapped or clamped), px
    float velocityX, velocityY;  // px/s
};
// out: steps x 6 vertices vec4(x, y, dash coordinate px, path fraction)
void launch_trajectory(hipStream_t st, const SimDesc* sims, const TrajectoryInputs& in, float4* out);

// ---- sleeping tiles: wake the next step's tiles, then advance the stamps ----
void launch_sleep_advance(hipStream_t st, const SimDesc* sims, int simCount,
                          const BodyBlock* tileBlocks, int tileBlockCount, const SleepBufs& sleep);
void launch_sleep_wake(hipStream_t st, const SleepBufs& sleep, int sim, int steps);

// ---- adjoint of the gravity step ----
void launch_fft_rows_forward_values(hipStream_t st, const SimDesc* sims, const LineWork* work,
                                    int items, const uint32_t* cellPrefix, const float4* cellBody,
                                    const float2* value, float2* spec, const float2* twiddles);
// wCell receives the adjoint weight of the body in every cell slot (scratch
// of the near-field VJP, one float2 per body slot).
void launch_adj_direct(hipStream_t st, const LaunchShape& ls, const float4* body,
                       const float4* mats, const float2* force, const uint4* cellRec,
                       const float4* cellBody, const float2* xbar, const float2* vbar, float dt,
                       float2* wCell, float2* gbar, float2* spread);
void launch_adj_finish(hipStream_t st, const LaunchShape& ls, const float4* body,
                       const float4* mats, const float2* result, const float2* gbar, float2* xbar,
                       float2* vbar, float dt, float coef);
void launch_adj_unsort(hipStream_t st, const LaunchShape& ls, const uint32_t* dest,
                       const float2* xin, const float2* vin, float2* xout, float2* vout);
void launch_adj_compose(hipStream_t st, const LaunchShape& ls, const uint32_t* dest,
                        uint32_t* where);
void launch_adj_iota(hipStream_t st, const LaunchShape& ls, uint32_t* where);

// ---- position-based liquid ----
void launch_liquid_bin(hipStream_t st, const LaunchShape& ls, const float4* body,
                       uint32_t* liqCount, uint32_t* cellOf);
void launch_liquid_scatter(hipStream_t st, const LaunchShape& ls, const uint32_t* cellOf,
                           uint32_t* liqCount, const uint32_t* prefix, uint32_t* unordered);
void launch_liquid_canon(hipStream_t st, int slots, const uint32_t* cellOf,
                         const uint32_t* prefix, const uint32_t* unordered, uint32_t* items);
void launch_liquid_lambda(hipStream_t st, const LaunchShape& ls, const float4* body,
                          const float4* mats, const uint32_t* cellOf, const uint32_t* prefix,
                          const uint32_t* items, const float2* fallback, float* lambdas,
                          float* pressures, int writePressure);
void launch_liquid_apply(hipStream_t st, const LaunchShape& ls, const float4* body, float4* out,
                         const float4* mats, const uint32_t* cellOf, const uint32_t* prefix,
                         const uint32_t* items, const float2* fallback, const float* lambdas);
void launch_liquid_viscosity(hipStream_t st, const LaunchShape& ls, const float4* body,
                             float4* out, const float4* mats, const uint32_t* cellOf,
                             const uint32_t* prefix, const uint32_t* items);

}  // namespace gs
