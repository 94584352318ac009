// The step engine: owns the device state of a batch of simulations and
// replays the reference's step schedule (K/GravityPipeline.kt:445-494) as a
// sequence of launches on one stream, captured as HIP graphs.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gravity_hip.h"
#include "common.hpp"
#include "green.hpp"
#include "kernels.hpp"

namespace gs {

// Mirrors Kotlin's IllegalArgumentException / IllegalStateException.
struct ArgumentError : std::runtime_error { using std::runtime_error::runtime_error; };
struct StateError : std::runtime_error { using std::runtime_error::runtime_error; };

// What wakes the sleeping tiles when it changes between steps.
struct SleepScene {
    int n = -1, bounded = 0;
    float fieldX = 0.0f, fieldY = 0.0f;
    int mutual = 0, liquid = 0;
    bool operator==(const SleepScene& o) const {
        return n == o.n && bounded == o.bounded && fieldX == o.fieldX && fieldY == o.fieldY &&
               mutual == o.mutual && liquid == o.liquid;
    }
};

struct HostSim {
    gs_config cfg;
    int n = 0;          // active bodies
    int cap = 0;
    int base = 0;       // first global body index
    int bounded = 0, mutual = 1, liquid = 0;
    int substeps = 16;
    float fieldX = 0.0f, fieldY = 0.0f;   // px/s^2
    long long steps = 0;                  // physicsSteps
    int cells = 0;                        // broad-phase cells
    int tiles = 0;                        // sort tiles
    int phys = 0;                         // physical P3M cells
    int meshMax = 0;                      // mesh nodes reserved (walls double)
    SimDesc desc{};
    SleepScene scene;                     // of the last step
};

// CircularProjectile (K/CircularProjectile.kt:10-60) and the walls check
// (GravityPipeline.supportsWalls, K/GravityPipeline.kt:1628), host only.
void projectile_offsets(float contactDiameter, float projectileDiameter, std::vector<float>& xy);
float projectile_diameter_for(float contactDiameter, int bodies);
// RectangularBlock.offsets (K/RectangularBlock.kt:10-77), host only
void block_offsets(float diameter, float width, float height, std::vector<float>& xy);
bool supports_walls(int gridWidth, int gridHeight);

// Launch shapes of the step that the autotuner chooses among; no choice
// changes a result (each is checked bit for bit against the others).
struct Tuning {
    int bigLimit = 0;              // cells and tiles longer than this: workgroup sort (0: never)
    int forwardItemBodies = 4096;  // bodies per forward row item, on average
    int depositRuns = 0;           // register runs in the deposit (0: an atomic per corner)
    int scanDirectBlocks = 1024;   // scans of more blocks scan their totals apart
};

class Engine {
public:
    Engine(const std::vector<gs_config>& cfgs, const std::vector<const float*>& states,
           const std::vector<int>& counts);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int sims() const { return (int)sims_.size(); }
    const HostSim& sim(int s) const { return sims_.at(s); }

    void set_active(int s, int n);
    void set_bounded(int s, bool v);
    void set_mutual(int s, bool v);
    void set_liquid(int s, bool v);
    void set_field(int s, float ax, float ay);
    void set_substeps(int substeps);
    void set_graphs(bool on) { useGraphs_ = on; }
    const Tuning& tuning() const { return tuning_; }
    void set_tuning(const Tuning& t);
    // Times `steps` graph-replayed steps of each candidate (coordinate
    // descent over the Tuning fields from the current choice, the state
    // restored before each trial and at the end) and keeps the fastest.
    // ms receives the chosen and the starting choice's ms per step.
    Tuning autotune(int steps, double* ms);

    void upload(int s, int first, int count, const float* xyvv, const float* mats);
    void upload_accel(int s, int first, int count, const float* accel);
    void download(int s, int what, int first, int count, void* dst);
    void synchronize();

    // K/GravityPipeline.kt:621-646, :180, :793-905, :509-512 (members.cpp)
    int spawn_projectile(int s, int firstBody, float centerX, float centerY, float velocityX,
                         float velocityY, float diameter);
    int projectile_body_count(int s, float diameter) const;
    gs_cluster_statistics measure_cluster(int s);
    // measureAdaptiveCollisions (K/GravityPipeline.kt:425-445)
    gs_adaptive_statistics measure_adaptive(int s);
    // updateTrajectory and its vertex buffer (K/GravityPipeline.kt:514-602),
    // simulation 0
    void update_trajectory(float originX, float originY, float velocityX, float velocityY);
    void trajectory_view(void** vertices, int* count);
    void download_trajectory(int first, int count, float* dst);
    void render_view(int s, void** bodies, void** pressures);
    void render_copy(int s, int count, void* bodies, void* pressures);

    void initialize_acceleration();
    void step();
    void stop_bodies(int s);
    void wake(int s);

    // Adjoint of the gravity step (one simulation, drift-only passes):
    // `steps` steps like step(), keeping the state every `every` steps, then
    // the vector-Jacobian product of those steps by checkpointed recomputation.
    // Buffers are allocated on the first recording of a shape (body
    // capacity, steps, interval) and kept for the next ones.
    void adjoint_record(int steps, int every);
    // count must be the recording's body count (dLdx and the outputs hold
    // 2 * count floats).
    void adjoint_vjp(int count, const float* dLdx, float* dLdx0, float* dLdv0);
    // where[i]: the index now holding the body that started the recording at i
    void adjoint_order(int count, uint32_t* where);
    // J_A(x)^T w of the acceleration field at the current positions
    void debug_accel_vjp(const float* w, float* out);
    // The fixed-point density mesh, its FP32 forward spectrum (both mesh
    // nodes, row-major) and the physical cell table (count, mass / 256 per
    // cell) of the current positions; each output nullable.
    void debug_density(uint32_t* density, float* spectrum, uint32_t* cells);
    // The same with the density staged in the library's static scratch array
    // instead of a buffer allocated on first use: no device allocation.
    void debug_density_noalloc(uint32_t* density, float* spectrum, uint32_t* cells);
    // validateFft (K/GravityPipeline.kt:378-419): the density spectrum
    // against an FP64 transform of the same deposit: out = {max abs error,
    // max abs error / max |reference|, RMS error}.
    void validate_fft(double* out);
    // Glued blocks of simulation 0 (rigid.hip): RigidBlocks.add without the
    // distance field (K/RigidBlocks.kt:280-301), project for blocks nothing
    // breaks (:309-357), the size, and access to the block buffers.
    int rigid_glue(int firstBody, int count, const float* offsets, float centerX, float centerY);
    // spawnBlock without a mask (K/GravityPipeline.kt:647-717): bodies written, 0 when they do not fit
    int spawn_block(int firstBody, float minX, float minY, float maxX, float maxY, float density,
                    float velocityX, float velocityY, float spin);
    void rigid_project(bool contacts);
    void rigid_size(int* out) const;
    void rigid_access(int what, long long first, long long count, void* host, bool write);

    void debug_collide(float kickDt, float driftDt, bool writePressure);
    void debug_sort();
    void debug_liquid();

    // Eager steps with an event pair around every launch.
    std::vector<std::tuple<std::string, int, double>> profile(int steps);
    int mesh_w(int s) const { return sims_.at(s).desc.mw; }
    int mesh_h(int s) const { return sims_.at(s).desc.mh; }

private:
    void check(hipError_t e, const char* what);
    void fill_desc(HostSim& h);
    void push_descs();
    void rebuild_tables();
    void normalize();
    void ensure_green(int s);
    void enqueue_step(bool sortDue);
    void enqueue_sort();
    void enqueue_force_and_kick(bool initial);
    void enqueue_force();
    void enqueue_kick(bool initial);
    void enqueue_pass(float kickDt, float driftDt, bool writePressure);
    void enqueue_liquid();
    void clear_grids();
    void invalidate_graphs();
    void rigid_allocate();
    void rigid_free();
    void rigid_remap();
    void refuse_blocks(const char* what) const;
    template <typename F> void timed(const char* name, F&& f);

    std::vector<HostSim> sims_;
    hipStream_t st_ = nullptr;
    int totalSlots_ = 0;      // body slots (sum of 256-rounded capacities)
    int totalCells_ = 0, totalTiles_ = 0, totalPhys_ = 0, totalMesh_ = 0, totalLiq_ = 0;
    int totalSleepTiles_ = 0;
    bool useGraphs_ = true;

    // bodies (ping-pong): float4 (x, y, vx, vy) in the reference's units
    float4* body_[2] = {nullptr, nullptr};
    float2* acc_[2] = {nullptr, nullptr};
    float4* mat_[2] = {nullptr, nullptr};
    int bodyPar_ = 0, auxPar_ = 0, gridPar_ = 0;
    float* press_ = nullptr;
    uint32_t* counts_ = nullptr;
    uint32_t* contacts_ = nullptr;
    uint32_t *poolItems_ = nullptr, *poolBase_ = nullptr, *poolTop_ = nullptr;  // contact pool
    CollideGrid grid_[2] = {};
    static constexpr size_t HEADS_PAD = 64;   // zero heads before and after the cells
    uint32_t* headsAlloc_[2] = {nullptr, nullptr};
    // sort
    uint32_t *tileCount_ = nullptr, *tilePrefix_ = nullptr, *keyOf_ = nullptr;
    uint32_t *slotSrc_ = nullptr, *dest_ = nullptr, *scanTmp_ = nullptr;
    // P3M
    uint32_t *cellCount_ = nullptr, *cellPrefix_ = nullptr, *homeOf_ = nullptr;
    uint32_t* slotBody_ = nullptr;
    uint32_t* slotCell_ = nullptr;   // per cell slot: the cell of the body in it
    uint32_t* slotKey_ = nullptr;    // per sorted slot: the tile of the body in it
    uint4* cellRec_ = nullptr;  // per physical cell: first slot, count, mass / 256, 0
    uint4* cellNear_ = nullptr;  // per physical cell: the kick's 32-byte near record
    float4* cellBody_ = nullptr;
    // liquid: bodies counting-sorted by their start cell, ascending per cell
    uint32_t *liqCount_ = nullptr, *liqPrefix_ = nullptr, *liqCellOf_ = nullptr;
    uint32_t *liqUnordered_ = nullptr, *liqItems_ = nullptr;
    float* lambdas_ = nullptr;
    // sleeping tiles
    SleepState* sleepState_ = nullptr;
    uint32_t *bodyAwake_ = nullptr, *tileAwake_ = nullptr, *tileActive_ = nullptr;
    BodyBlock* dTileBlocks_ = nullptr;
    int tileBlocks_ = 0;
    float2 *spec_ = nullptr, *green_ = nullptr, *tw_ = nullptr, *fallback_ = nullptr;
    std::vector<GreenKey> greenKeys_;
    // tables
    SimDesc* dSims_ = nullptr;
    BodyBlock* dBlocks_ = nullptr;
    OrderedBlock* order_ = nullptr;  // dispatch order of the contact kernels
    uint32_t* waveWork_ = nullptr;   // contacts per wave of the last count
    LineWork *dRowsF_ = nullptr, *dCols_ = nullptr, *dRowsI_ = nullptr;
    int bodyBlocks_ = 0, rowsF_ = 0, cols_ = 0, rowsI_ = 0, gravitySlots_ = 0, liquidSlots_ = 0;
    bool anyMutual_ = false, anyLiquid_ = false, anyCollide_ = false, anyViscous_ = false;
    bool anySleep_ = false, anyAdaptive_ = false;
    uint32_t* adapt_ = nullptr;      // ADAPT_WORDS per simulation (adaptive collisions)
    float* adaptDeepest_ = nullptr;  // per body, with adaptive collisions
    float4* trajectory_ = nullptr;   // simulation 0's preview vertices
    int trajectoryVertices_ = 0;
    LaunchShape shape_{};

    // glued blocks (K/RigidBlocks.kt:84-113), allocated by the first block
    struct RigidState {
        RigidGroup* groups = nullptr;   // RIGID_MAXIMUM_GROUPS records
        int* memory = nullptr;          // impacts, next impacts, states, idle passes, pair contacts
        float* shapes = nullptr;        // per block: its hull's rest-frame corners
        int* members = nullptr;         // per member slot: its body, -1 for padding
        float2* rest = nullptr;         // per member slot: rest offset about the centre (px)
        int* tileGroups = nullptr;      // per tile: its block
        float4* partials = nullptr;     // per tile: two sums
        int groupCount = 0, tileCount = 0, memberCap = 0, tileCap = 0;
        long long valueCap = 0, valueTop = 0;
    } rigid_;

    // graphs keyed by (bodyPar, auxPar, gridPar, sortDue), with the parities
    // the captured schedule ends in
    struct CapturedStep { hipGraphExec_t exec; int bodyPar, auxPar, gridPar; };
    std::map<int, CapturedStep> graphs_;

    // adjoint
    struct AdjCheckpoint {
        float4* body = nullptr;
        float2* acc = nullptr;
        float4* mats = nullptr;
        long long steps = 0;
    };
    struct AdjSlot {
        float4* body = nullptr;
        float4* mats = nullptr;
        float2* force = nullptr;
        uint4* rec = nullptr;
        uint32_t* prefix = nullptr;
        float4* cellBody = nullptr;
        uint32_t* dest = nullptr;
        bool sorted = false;
    };
    std::vector<AdjCheckpoint> adjCk_;
    std::vector<AdjSlot> adjSlot_;
    std::vector<char> adjSortDue_;
    int adjSteps_ = 0, adjEvery_ = 0;
    int adjN_ = -1;              // bodies of the last recording (-1: none)
    long long adjShape_[4] = {-1, -1, -1, -1};  // what the recording buffers were sized for
    float2 *adjX_ = nullptr, *adjV_ = nullptr, *adjX2_ = nullptr, *adjV2_ = nullptr;
    float2 *adjG_ = nullptr, *adjSpread_ = nullptr, *adjW_ = nullptr;
    uint32_t* adjWhere_ = nullptr;
    AdjCheckpoint adjFinal_;
    std::vector<void*> adjAllocs_;   // recording buffers (checkpoints, slots, final state)
    std::vector<void*> adjWork_;     // per-body work arrays, allocated once
    uint32_t* debugDensity_ = nullptr;   // density readback (diagnostics only)
    void density_readback(uint32_t* staging, uint32_t* density, float* spectrum, uint32_t* cells);
    void adjoint_free();
    void adjoint_work();
    void adjoint_save_slot(int j);

    // launch shapes
    Tuning tuning_;
    BigSegment* bigList_ = nullptr;
    uint32_t* bigCounters_ = nullptr;
    BigSegments big() const { return BigSegments{bigList_, bigCounters_, tuning_.bigLimit, 0}; }
    double time_steps(int steps);

    // profiling
    bool profiling_ = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> events_;
};

}  // namespace gs
