// Step engine (K/GravityPipeline.kt): device state, batch tables, the step
// schedule, graph capture and per-launch event timing.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace gs {

int fft_row_lines_per_group(int L);
int fft_columns_per_group(int H, int W);

static int round_up(int v, int m) { return (v + m - 1) / m * m; }
static bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// K/GravityPipeline.kt:1647-1648
static int sort_axis_bits(int extent) {
    int tiles = std::max(1, extent / 4);
    int bits = 31 - __builtin_clz((unsigned)tiles);
    return std::min(bits, 10);
}

// K/GravityPipeline.kt:1546-1547
static int mass_scale_for(float totalMass) {
    double v = 4000000000.0 / (double)totalMass;
    long long s = (long long)v;
    return (int)std::min<long long>(65536, std::max<long long>(1, s));
}

static void validate(const gs_config& c) {
    auto req = [](bool ok, const char* msg) { if (!ok) throw ArgumentError(msg); };
    req(c.viewport_width > 0 && c.viewport_height > 0, "viewport must be positive");
    req(c.capacity > 0, "capacity must be positive");
    req(c.spatial_sort_interval >= 0, "sort interval must be non-negative");
    req(pow2(c.grid_width) && c.grid_width >= 16 && c.grid_width <= 2048,
        "grid width must be a power of two in 16..2048");
    req(pow2(c.grid_height) && c.grid_height >= 16 && c.grid_height <= 2048,
        "grid height must be a power of two in 16..2048");
    req(c.gravitational_constant >= 0.0f && std::isfinite(c.gravitational_constant), "gravity");
    req(c.softening_cells > 0.0f && std::isfinite(c.softening_cells), "softening");
    req(c.local_search_sigma_multiplier > 0.0f, "near-radius-sigma");
    req(c.minimum_neighbor_radius_cells >= 1, "near-radius-min");
    req(c.maximum_neighbor_radius_cells >= c.minimum_neighbor_radius_cells, "near-radius-max");
    req(c.exact_neighbor_body_limit >= 1, "near-exact-bodies");
    req(c.maximum_home_body_interactions >= 1, "near-home-bodies");
    req(c.minimum_distance_cell_fraction > 0.0f, "near-min-distance-fraction");
    req(c.minimum_assignment_window_power > 0.0f && c.minimum_assignment_window_power <= 1.0f,
        "cic-window-floor");
    req(c.time_step > 0.0f && std::isfinite(c.time_step), "dt");
    req(c.maximum_substeps >= 1 && c.maximum_substeps <= 16, "collision substeps must be in 1..16");
    req(c.contact_diameter_pixels >= 0.5f && c.contact_diameter_pixels <= 16.0f,
        "contact diameter must be in 0.5..16.0 pixels");
    req(c.neighbor_cell_radius >= 1 && c.neighbor_cell_radius <= 4, "collision-neighbor-radius");
    req(c.restitution_coefficient >= 0.0f && c.restitution_coefficient <= 1.0f, "restitution");
    req(c.friction >= 0.0f && std::isfinite(c.friction), "friction");
    req(c.coincident_distance_squared_epsilon > 0.0f, "coincident-distance-epsilon");
    req(c.neighbor_skin_pixels > 0.0f, "collision-skin");
    // K/GravityConfig.kt:725-729
    req(c.allowed_displacement_fraction >= 0.0f && c.allowed_displacement_fraction <= 1.0f,
        "adaptive-displacement-fraction");
    req(c.allowed_penetration_fraction >= 0.0f && c.allowed_penetration_fraction <= 1.0f,
        "adaptive-penetration-fraction");
    req(c.dense_contact_threshold >= 1, "adaptive-dense-contact-count");
    req(c.dense_body_fraction >= 0.0f && c.dense_body_fraction <= 1.0f, "adaptive-dense-body-fraction");
    req(c.minimum_dense_bodies >= 1, "adaptive-dense-min-bodies");
    // K/GravityConfig.kt:747-756
    req(c.trajectory_steps >= 1 && c.trajectory_steps <= 65536, "trajectory-steps");
    req(c.trajectory_maximum_time_step > 0.0f && std::isfinite(c.trajectory_maximum_time_step),
        "trajectory-max-dt");
    req(c.trajectory_target_segment_pixels > 0.0f && std::isfinite(c.trajectory_target_segment_pixels),
        "trajectory-segment-px");
    req(c.trajectory_maximum_path_diagonals > 0.0f &&
            std::isfinite(c.trajectory_maximum_path_diagonals),
        "trajectory-path-diagonals");
    req(c.trajectory_minimum_dash_pixels > 0.0f && std::isfinite(c.trajectory_minimum_dash_pixels),
        "trajectory-min-dash-px");
    req(c.kernel_radius_diameters >= 1.5f && c.kernel_radius_diameters <= 4.0f,
        "liquid kernel must be 1.5..4 diameters");
    req(c.liquid_iterations >= 1 && c.liquid_iterations <= 16, "liquid iterations must be in 1..16");
    req(c.liquid_viscosity >= 0.0f && c.liquid_viscosity <= 1.0f, "liquid viscosity must be in 0..1");
    req(!c.bounded || 2 * std::max(c.grid_width, c.grid_height) <= 2048,
        "Walls need an FFT grid of at most 1024 cells per axis");
}

// Scratch of gs_debug_density_noalloc: the fixed-point density staging of the
// largest mesh it accepts (2048 x 1024 nodes, 8 MiB), one copy per GPU for
// the life of the process. Host code never uses the array's name as an
// address: scratch_density() returns the device address that
// hipGetSymbolAddress gives for the current device.
constexpr size_t DENSITY_SCRATCH_NODES = (size_t)2048 * 1024;
static __device__ __align__(256) uint32_t scratch[DENSITY_SCRATCH_NODES];

static uint32_t* scratch_density() {
    static std::mutex lock;
    static std::map<int, uint32_t*> byDevice;
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess) throw StateError("hipGetDevice failed");
    std::lock_guard<std::mutex> guard(lock);
    auto it = byDevice.find(device);
    if (it != byDevice.end()) return it->second;
    void* ptr = nullptr;
    hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || ptr == nullptr)
        throw StateError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
    byDevice[device] = (uint32_t*)ptr;
    return (uint32_t*)ptr;
}

std::vector<LaunchEvents>*& launch_sink() {
    static thread_local std::vector<LaunchEvents>* sink = nullptr;
    return sink;
}

void Engine::check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw StateError(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
static T* dalloc(size_t count) {
    T* p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess)
        throw StateError("hipMalloc failed");
    if (hipMemset(p, 0, count * sizeof(T)) != hipSuccess) throw StateError("hipMemset failed");
    return p;
}

Engine::Engine(const std::vector<gs_config>& cfgs, const std::vector<const float*>& states,
               const std::vector<int>& counts) {
    if (cfgs.empty()) throw ArgumentError("no simulation");
    int substeps = cfgs[0].maximum_substeps;
    for (const auto& c : cfgs) {
        validate(c);
        if (c.maximum_substeps != substeps || c.time_step != cfgs[0].time_step ||
            c.spatial_sort_interval != cfgs[0].spatial_sort_interval)
            throw ArgumentError("a batch shares dt, substeps and sort interval");
        // The step schedule is either contact passes or plain drifts for the
        // whole batch (no kernel reads a member's own choice).
        if ((c.collisions_enabled != 0) != (cfgs[0].collisions_enabled != 0))
            throw ArgumentError("a batch shares collisionsEnabled");
        // Adaptive members keep every substep pass in the schedule (their
        // kernels skip the ones past the count they chose); a fixed member's
        // substep count is the schedule's.
        if ((c.adaptive_collisions != 0) != (cfgs[0].adaptive_collisions != 0))
            throw ArgumentError("a batch shares adaptive collisions");
    }
    for (const auto& c : cfgs)
        for (const auto& o : cfgs)
            if (c.liquid && o.liquid && c.liquid_iterations != o.liquid_iterations)
                throw ArgumentError("liquid simulations of a batch share their iteration count");
    sims_.resize(cfgs.size());
    int slot = 0, cells = 0, tiles = 0, phys = 0, mesh = 0, liq = 0, sleepTiles = 0;
    for (size_t i = 0; i < cfgs.size(); i++) {
        HostSim& h = sims_[i];
        h.cfg = cfgs[i];
        h.cap = cfgs[i].capacity;
        h.base = slot;
        slot += round_up(h.cap, WG);
        h.bounded = cfgs[i].bounded;
        h.mutual = cfgs[i].mutual_gravity;
        h.liquid = cfgs[i].liquid;
        h.substeps = cfgs[i].maximum_substeps;
        h.fieldX = cfgs[i].field_acceleration_x;
        h.fieldY = cfgs[i].field_acceleration_y;
        int n0 = counts.size() > i ? counts[i] : 0;
        if (n0 < 0 || n0 > h.cap) throw ArgumentError("scenario exceeds capacity");
        h.n = n0;
        float vw = (float)h.cfg.viewport_width, vh = (float)h.cfg.viewport_height;
        float mx = std::max(vw, vh);
        float guard = 64.0f * (std::nextafter(mx, INFINITY) - mx);
        int cgw = std::max(1, (int)std::floor(vw / (h.cfg.contact_diameter_pixels + guard)));
        int cgh = std::max(1, (int)std::floor(vh / (h.cfg.contact_diameter_pixels + guard)));
        h.cells = cgw * cgh;
        h.desc.cellBase = cells;
        cells += h.cells;
        int wb = sort_axis_bits(h.cfg.viewport_width), hb = sort_axis_bits(h.cfg.viewport_height);
        h.tiles = 1 << (wb + hb);
        h.desc.tileBase = tiles;
        tiles += h.tiles;
        h.phys = h.cfg.grid_width * h.cfg.grid_height;
        h.desc.physBase = phys;
        phys += h.phys;
        // K/LiquidSolver.kt:45-46: liquid cells no smaller than the kernel radius
        float lh = h.cfg.contact_diameter_pixels * h.cfg.kernel_radius_diameters;
        h.desc.lgw = std::max(1, (int)(vw / lh));
        h.desc.lgh = std::max(1, (int)(vh / lh));
        h.desc.liqCellBase = liq;
        liq += h.desc.lgw * h.desc.lgh;
        // K/SleepTiles.kt:45-46: 32 px tiles, rounded up
        h.desc.tilesX = (h.cfg.viewport_width + 31) / 32;
        h.desc.tilesY = (h.cfg.viewport_height + 31) / 32;
        h.desc.sleepTileBase = sleepTiles;
        sleepTiles += h.desc.tilesX * h.desc.tilesY;
        bool wallsOk = 2 * std::max(h.cfg.grid_width, h.cfg.grid_height) <= 2048;
        h.meshMax = h.phys * (wallsOk ? 4 : 1);
        h.desc.meshBase = mesh;
        mesh += h.meshMax;
    }
    totalSlots_ = std::max(slot, WG);
    // contact planes are indexed in 32 bits
    if ((uint64_t)totalSlots_ * CONTACT_PLANES > 0xffffffffull)
        throw ArgumentError("too many bodies for the contact planes (38,347,922 at most)");
    totalCells_ = cells;
    totalTiles_ = tiles;
    totalPhys_ = phys;
    totalMesh_ = mesh;
    totalLiq_ = liq;
    totalSleepTiles_ = sleepTiles;

    check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "stream");
    size_t S = (size_t)totalSlots_;
    for (int b = 0; b < 2; b++) {
        body_[b] = dalloc<float4>(S);
        acc_[b] = dalloc<float2>(S);
        mat_[b] = dalloc<float4>(S);
        // one zero entry before the first cell and after the last, for the
        // stencil's row loads (the allocation is padded to keep alignment)
        headsAlloc_[b] = dalloc<uint32_t>((size_t)totalCells_ + 2 * HEADS_PAD);
        grid_[b].heads = headsAlloc_[b] + HEADS_PAD;
        grid_[b].nodes = dalloc<float4>(S);
        grid_[b].cellOf = dalloc<int>(S);
    }
    press_ = dalloc<float>(S);
    counts_ = dalloc<uint32_t>(S);
    contacts_ = dalloc<uint32_t>(S * CONTACT_PLANES);
    poolItems_ = dalloc<uint32_t>(S * CONTACT_SLOTS);
    poolBase_ = dalloc<uint32_t>(S);
    poolTop_ = dalloc<uint32_t>(1);
    tileCount_ = dalloc<uint32_t>((size_t)totalTiles_ + 1);
    tilePrefix_ = dalloc<uint32_t>((size_t)totalTiles_ + 1);
    keyOf_ = dalloc<uint32_t>(S);
    slotSrc_ = dalloc<uint32_t>(S);
    dest_ = dalloc<uint32_t>(S);
    int scanMax = std::max({scan_totals_size(totalTiles_), scan_totals_size(totalPhys_),
                            scan_totals_size(totalLiq_)});
    scanTmp_ = dalloc<uint32_t>((size_t)scanMax);
    bigList_ = dalloc<BigSegment>(S / (BIG_LIMIT_MIN + 1) + 1);
    bigCounters_ = dalloc<uint32_t>(2);
    cellCount_ = dalloc<uint32_t>((size_t)totalPhys_ + 1);
    cellPrefix_ = dalloc<uint32_t>((size_t)totalPhys_ + 1);
    homeOf_ = dalloc<uint32_t>(S);
    slotBody_ = dalloc<uint32_t>(S);
    slotCell_ = dalloc<uint32_t>(S);
    slotKey_ = dalloc<uint32_t>(S);
    cellRec_ = dalloc<uint4>((size_t)totalPhys_);
    cellNear_ = dalloc<uint4>(2 * (size_t)totalPhys_);
    cellBody_ = dalloc<float4>(S);
    liqCount_ = dalloc<uint32_t>((size_t)totalLiq_ + 1);
    liqPrefix_ = dalloc<uint32_t>((size_t)totalLiq_ + 1);
    liqCellOf_ = dalloc<uint32_t>(S);
    liqUnordered_ = dalloc<uint32_t>(S);
    liqItems_ = dalloc<uint32_t>(S);
    lambdas_ = dalloc<float>(S);
    sleepState_ = dalloc<SleepState>(sims_.size());
    bodyAwake_ = dalloc<uint32_t>(S);
    tileAwake_ = dalloc<uint32_t>((size_t)std::max(totalSleepTiles_, 1));
    tileActive_ = dalloc<uint32_t>((size_t)std::max(totalSleepTiles_, 1));
    {
        // K/SleepTiles.kt:56-60: stamp 1 (0 is never a stamp, so cleared
        // tiles sleep), everything awake for the first steps.
        std::vector<SleepState> st(sims_.size(), SleepState{1u, SLEEP_WAKE_STEPS, 0, 0});
        check(hipMemcpy(sleepState_, st.data(), st.size() * sizeof(SleepState),
                        hipMemcpyHostToDevice), "sleep state");
        std::vector<BodyBlock> tb;
        for (int i = 0; i < (int)sims_.size(); i++) {
            int tiles = sims_[i].desc.tilesX * sims_[i].desc.tilesY;
            for (int t = 0; t < tiles; t += WG) tb.push_back({i, t});
        }
        tileBlocks_ = (int)tb.size();
        dTileBlocks_ = dalloc<BodyBlock>(std::max<size_t>(tb.size(), 1));
        if (!tb.empty())
            check(hipMemcpy(dTileBlocks_, tb.data(), tb.size() * sizeof(BodyBlock),
                            hipMemcpyHostToDevice), "tile blocks");
    }
    spec_ = dalloc<float2>((size_t)totalMesh_);
    green_ = dalloc<float2>((size_t)totalMesh_);
    tw_ = dalloc<float2>((size_t)fft_twiddle_entries());
    fallback_ = dalloc<float2>(65536);
    dSims_ = dalloc<SimDesc>(sims_.size());
    int maxBlocks = totalSlots_ / WG;
    dBlocks_ = dalloc<BodyBlock>((size_t)maxBlocks);
    order_ = dalloc<OrderedBlock>((size_t)maxBlocks);
    waveWork_ = dalloc<uint32_t>((size_t)maxBlocks * (WG / 64));
    adapt_ = dalloc<uint32_t>(sims_.size() * ADAPT_WORDS);
    trajectoryVertices_ = sims_[0].cfg.trajectory_steps * 6;
    trajectory_ = dalloc<float4>((size_t)trajectoryVertices_);
    bool adaptiveBatch = false;
    for (const auto& h : sims_) adaptiveBatch |= h.cfg.adaptive_collisions && h.cfg.collisions_enabled;
    if (adaptiveBatch) adaptDeepest_ = dalloc<float>(S);
    int maxLines = 0;
    for (auto& h : sims_) maxLines += 2 * std::max(h.cfg.grid_width, h.cfg.grid_height) * 2 + 2;
    dRowsF_ = dalloc<LineWork>((size_t)maxLines);
    dRowsI_ = dalloc<LineWork>((size_t)maxLines);
    dCols_ = dalloc<LineWork>((size_t)maxLines);

    std::vector<float> host;
    fft_twiddle_table(host);
    check(hipMemcpy(tw_, host.data(), host.size() * 4, hipMemcpyHostToDevice), "twiddles");
    fallback_normals(host);
    check(hipMemcpy(fallback_, host.data(), host.size() * 4, hipMemcpyHostToDevice), "normals");

    // Default materials: unit mass, the configured friction, not solid
    // (K/GravityPipeline.kt:1494-1504); the start state of the scenario.
    for (size_t i = 0; i < sims_.size(); i++) {
        HostSim& h = sims_[i];
        std::vector<float> m((size_t)4 * h.cap);
        for (int k = 0; k < h.cap; k++) {
            m[4 * k] = 1.0f;
            m[4 * k + 1] = h.cfg.friction;
            m[4 * k + 2] = 0.0f;
            m[4 * k + 3] = 0.0f;
        }
        for (int b = 0; b < 2; b++)
            check(hipMemcpy(mat_[b] + h.base, m.data(), m.size() * 4, hipMemcpyHostToDevice), "mat");
        if (states.size() > i && states[i] && h.n > 0) upload((int)i, 0, h.n, states[i], nullptr);
    }
    greenKeys_.resize(sims_.size(), GreenKey{0, 0, 0, 0, 0, 0, 0, 0});
    for (auto& h : sims_) fill_desc(h);
    for (int s = 0; s < (int)sims_.size(); s++) ensure_green(s);
    push_descs();
    rebuild_tables();
    check(hipStreamSynchronize(st_), "init");
}

Engine::~Engine() {
    invalidate_graphs();
    adjoint_free();
    rigid_free();
    for (void* p : adjWork_) hipFree(p);
    if (debugDensity_) hipFree(debugDensity_);
    for (auto& e : events_) {
        hipEventDestroy(e.second.first);
        hipEventDestroy(e.second.second);
    }
    void* ptrs[] = {body_[0], body_[1], acc_[0], acc_[1], mat_[0], mat_[1],
                    headsAlloc_[0], headsAlloc_[1], grid_[0].nodes, grid_[1].nodes,
                    grid_[0].cellOf, grid_[1].cellOf, press_, counts_, contacts_, tileCount_,
                    tilePrefix_, keyOf_, slotSrc_, dest_, scanTmp_, cellCount_, cellPrefix_,
                    homeOf_, slotBody_, slotCell_, slotKey_, cellRec_, cellNear_, cellBody_, spec_, green_, tw_, fallback_,
                    liqCount_, liqPrefix_, liqCellOf_, liqUnordered_, liqItems_, lambdas_,
                    sleepState_, bodyAwake_, tileAwake_, tileActive_, dTileBlocks_,
                    dSims_, dBlocks_, dRowsF_, dRowsI_, dCols_, order_, waveWork_, poolItems_, poolBase_, poolTop_,
                    bigList_, bigCounters_, adapt_, adaptDeepest_, trajectory_};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (st_) hipStreamDestroy(st_);
}

void Engine::fill_desc(HostSim& h) {
    SimDesc& d = h.desc;
    const gs_config& c = h.cfg;
    d.n = h.n;
    d.base = h.base;
    d.cap = h.cap;
    d.bounded = h.bounded;
    d.mutual = h.mutual;
    d.liquid = h.liquid;
    d.collide = c.collisions_enabled;
    d.vw = (float)c.viewport_width;
    d.vh = (float)c.viewport_height;
    d.aspect = d.vw / d.vh;
    d.dt = c.time_step;
    d.subDt = c.time_step / (float)h.substeps;
    d.diameter = c.contact_diameter_pixels;
    d.diameter2 = d.diameter * d.diameter;
    d.restitution = c.restitution_coefficient;
    d.coincidentEps = c.coincident_distance_squared_epsilon;
    float mx = std::max(d.vw, d.vh);
    float guard = 64.0f * (std::nextafter(mx, INFINITY) - mx);
    d.cgw = std::max(1, (int)std::floor(d.vw / (d.diameter + guard)));
    d.cgh = std::max(1, (int)std::floor(d.vh / (d.diameter + guard)));
    if (d.cgw > 65535 || d.cgh > 65535)
        throw ArgumentError("viewport too large for the collision grid (65535 cells a side)");
    d.sortWb = sort_axis_bits(c.viewport_width);
    d.sortHb = sort_axis_bits(c.viewport_height);
    d.pw = c.grid_width;
    d.ph = c.grid_height;
    d.mw = h.bounded ? 2 * d.pw : d.pw;
    d.mh = h.bounded ? 2 * d.ph : d.ph;
    // K/GravityPipeline.kt:153-169
    float nx = c.local_search_sigma_multiplier * c.softening_cells * (float)d.pw /
               ((float)d.ph * d.aspect);
    float ny = c.softening_cells * c.local_search_sigma_multiplier;
    d.rx = std::min(std::max((int)std::ceil(nx), c.minimum_neighbor_radius_cells),
                    c.maximum_neighbor_radius_cells);
    d.ry = std::min(std::max((int)std::ceil(ny), c.minimum_neighbor_radius_cells),
                    c.maximum_neighbor_radius_cells);
    d.exactLimit = c.exact_neighbor_body_limit;
    d.homeLimit = c.maximum_home_body_interactions;
    d.G = c.gravitational_constant;
    d.totalMass = std::max(1.0f, (float)h.n);
    d.sigma = c.softening_cells / (float)d.ph;
    d.twoSigma2 = 2.0f * d.sigma * d.sigma;
    float cw = d.aspect / (float)d.pw, ch = 1.0f / (float)d.ph;
    float md = c.minimum_distance_cell_fraction * std::min(cw, ch);
    d.minDist2 = md * md;
    int ms = mass_scale_for(d.totalMass);
    d.massScale = (float)ms;
    d.invMassScale = 1.0f / (float)ms;
    d.invTotalMass = 1.0f / d.totalMass;
    d.fieldX = h.fieldX / d.vw;
    d.fieldY = h.fieldY / d.vh;
    d.invAspect = (float)(1.0 / (double)d.aspect);
    d.invTwoSigma2 = (float)(1.0 / (2.0 * (double)d.sigma * (double)d.sigma));
    d.invPw = 1.0f / (float)d.pw;
    d.invPh = 1.0f / (float)d.ph;
    d.gOverM = (float)((double)d.G / (double)d.totalMass);
    // K/LiquidSolver.kt:64-101: the liquid passes' uniforms
    LiquidConstants lk = liquid_constants(c.contact_diameter_pixels, c.kernel_radius_diameters);
    d.lh = lk.h;
    d.lh2 = lk.h * lk.h;
    d.lpoly6 = lk.poly6;
    d.lspiky = lk.spiky;
    d.lrest = lk.rest;
    d.lrelax = lk.relax;
    d.lrowStep = c.contact_diameter_pixels * std::sqrt(3.0f) * 0.5f;
    d.lwallScale = lk.poly6 * 32.0f / 35.0f / c.contact_diameter_pixels;
    d.lpressureScale = lk.h / c.time_step;
    d.lviscosity = c.liquid_viscosity / lk.rest;
    d.lviscous = c.liquid_viscosity > 0.0f ? 1 : 0;
    d.lradius = 0.5f * c.contact_diameter_pixels;
    // K/SleepTiles.kt:51, 71-72 and S/collision_cache_solve.comp:452-466
    d.sleepOn = c.sleep && c.collisions_enabled ? 1 : 0;
    // K/GravityPipeline.kt:1155-1211: the finalize shader's uniforms
    d.adaptive = c.adaptive_collisions && c.collisions_enabled ? 1 : 0;
    d.adaptMax = c.maximum_substeps;
    d.adaptDisp = c.contact_diameter_pixels * c.allowed_displacement_fraction;
    d.adaptPen = c.contact_diameter_pixels * c.allowed_penetration_fraction;
    d.adaptDenseContacts = c.dense_contact_threshold;
    d.adaptDenseBodies = std::max(c.minimum_dense_bodies,
                                  (int)std::ceil((float)h.n * c.dense_body_fraction));
    float disp = c.sleep_speed_pixels * c.time_step;
    d.sleepDisp2 = disp * disp;
    float radius = 0.5f * c.contact_diameter_pixels, slack = 0.05f * c.contact_diameter_pixels;
    d.holdLo = radius + slack;
    d.holdHiX = d.vw - radius - slack;
    d.holdHiY = d.vh - radius - slack;
    float settling = 0.005f * c.contact_diameter_pixels;
    d.settling2 = settling * settling;
    float mark = 0.5f * 32.0f;
    d.tileMark2 = mark * mark;
}

void Engine::push_descs() {
    int rank0 = 0;
    std::vector<SimDesc> all;
    for (auto& h : sims_) {
        h.desc.rank0 = rank0;
        rank0 += h.n;
        all.push_back(h.desc);
    }
    check(hipMemcpyAsync(dSims_, all.data(), all.size() * sizeof(SimDesc), hipMemcpyHostToDevice,
                         st_), "descriptors");
    check(hipStreamSynchronize(st_), "descriptors");
}

void Engine::ensure_green(int s) {
    HostSim& h = sims_[s];
    if (!h.mutual) return;
    const SimDesc& d = h.desc;
    GreenKey key{d.mw, d.mh, d.pw, d.ph, d.aspect, h.cfg.gravitational_constant,
                 h.cfg.softening_cells, h.cfg.minimum_assignment_window_power};
    if (greenKeys_[s] == key) return;
    std::vector<float> g;
    green_spectrum(key, g);
    check(hipMemcpy(green_ + d.meshBase, g.data(), g.size() * 4, hipMemcpyHostToDevice), "green");
    greenKeys_[s] = key;
}

void Engine::rebuild_tables() {
    std::vector<BodyBlock> blocks;
    std::vector<LineWork> rowsF, rowsI, cols;
    anyMutual_ = anyLiquid_ = anyCollide_ = anySleep_ = anyAdaptive_ = false;
    bool anyWalls = false, anyTorus = false;
    gravitySlots_ = 0;
    liquidSlots_ = 0;
    anyViscous_ = false;
    for (int s = 0; s < (int)sims_.size(); s++) {
        const HostSim& h = sims_[s];
        for (int i = 0; i < h.n; i += WG) blocks.push_back({s, h.base + i});
        anyLiquid_ |= h.liquid != 0;
        anySleep_ |= h.desc.sleepOn != 0;
        if (h.liquid) {
            liquidSlots_ += h.n;
            anyViscous_ |= h.cfg.liquid_viscosity > 0.0f;
        }
        anyCollide_ |= h.cfg.collisions_enabled != 0;
        anyAdaptive_ |= h.desc.adaptive != 0;
        (h.bounded ? anyWalls : anyTorus) = true;
        if (!h.mutual) continue;
        anyMutual_ = true;
        gravitySlots_ += h.n;
        const SimDesc& d = h.desc;
        int active = h.bounded ? d.ph + 2 : d.mh;
        int nl = fft_row_lines_per_group(d.mw);
        // A forward item deposits every body of its rows first: a crowded
        // simulation (many bodies per short row) takes fewer lines per item,
        // about tuning_.forwardItemBodies bodies each, so its deposits run side by
        // side (the grouping changes no result: the sums are integer and each
        // line is transformed alone).
        const long itemBodies = tuning_.forwardItemBodies;
        int nlF = nl;
        if (h.n > 0) nlF = (int)std::max<long>(1, std::min<long>(nl, itemBodies * active / h.n));
        for (int a = 0; a < active; a += nlF) rowsF.push_back({s, a, std::min(nlF, active - a), 0});
        for (int a = 0; a < active; a += nl) rowsI.push_back({s, a, std::min(nl, active - a), 0});
        int cb = fft_columns_per_group(d.mh, d.mw);
        for (int c = 0; c < d.mw; c += cb) cols.push_back({s, c, cb, 0});
    }
    bodyBlocks_ = (int)blocks.size();
    rowsF_ = (int)rowsF.size();
    rowsI_ = (int)rowsI.size();
    cols_ = (int)cols.size();
    if (!blocks.empty()) {
        check(hipMemcpy(dBlocks_, blocks.data(), blocks.size() * sizeof(BodyBlock),
                        hipMemcpyHostToDevice), "blocks");
        std::vector<OrderedBlock> order(blocks.size());
        identity_block_order(blocks.data(), bodyBlocks_, order.data());
        check(hipMemcpy(order_, order.data(), order.size() * sizeof(OrderedBlock),
                        hipMemcpyHostToDevice), "order");
    }
    if (!rowsF.empty()) {
        check(hipMemcpy(dRowsF_, rowsF.data(), rowsF.size() * sizeof(LineWork), hipMemcpyHostToDevice), "rows");
        check(hipMemcpy(dRowsI_, rowsI.data(), rowsI.size() * sizeof(LineWork), hipMemcpyHostToDevice), "rows");
        check(hipMemcpy(dCols_, cols.data(), cols.size() * sizeof(LineWork), hipMemcpyHostToDevice), "cols");
    }
    shape_.sims = dSims_;
    shape_.blocks = dBlocks_;
    shape_.bodyBlocks = bodyBlocks_;
    shape_.planeStride = totalSlots_;
    shape_.guardLimit = totalSlots_ + 1;
    shape_.sleep = SleepBufs{sleepState_, bodyAwake_, tileAwake_, tileActive_};
    shape_.order = order_;
    shape_.waveWork = waveWork_;
    shape_.adapt = anyAdaptive_ ? adapt_ : nullptr;
    shape_.adaptDeepest = adaptDeepest_;
    shape_.walls = anyWalls && anyTorus ? WALLS_MIXED : (anyWalls ? WALLS_ALL : WALLS_NONE);
    shape_.pool = ContactPool{poolItems_, poolBase_, poolTop_, (uint32_t)totalSlots_ * CONTACT_SLOTS};
    invalidate_graphs();
}

void Engine::invalidate_graphs() {
    for (auto& g : graphs_) hipGraphExecDestroy(g.second.exec);
    graphs_.clear();
}

void Engine::clear_grids() {
    for (int b = 0; b < 2; b++)
        check(hipMemsetAsync(grid_[b].heads, 0, sizeof(uint32_t) * (size_t)totalCells_, st_), "grid");
}

void Engine::set_active(int s, int n) {
    HostSim& h = sims_.at(s);
    if (n < 0 || n > h.cap) throw ArgumentError("active body count out of range");
    if (n == h.n) return;
    h.n = n;
    fill_desc(h);
    push_descs();
    clear_grids();
    rebuild_tables();
}

void Engine::set_bounded(int s, bool v) {
    HostSim& h = sims_.at(s);
    if ((h.bounded != 0) == v) return;
    if (v && 2 * std::max(h.cfg.grid_width, h.cfg.grid_height) > 2048)
        throw ArgumentError("Walls need an FFT grid of at most 1024 cells per axis");
    h.bounded = v;
    fill_desc(h);
    ensure_green(s);
    push_descs();
    rebuild_tables();
}

void Engine::set_mutual(int s, bool v) {
    HostSim& h = sims_.at(s);
    if ((h.mutual != 0) == v) return;
    h.mutual = v;
    fill_desc(h);
    ensure_green(s);
    push_descs();
    rebuild_tables();
}

void Engine::set_liquid(int s, bool v) {
    HostSim& h = sims_.at(s);
    if ((h.liquid != 0) == v) return;
    h.liquid = v;
    fill_desc(h);
    push_descs();
    rebuild_tables();
}

void Engine::set_field(int s, float ax, float ay) {
    HostSim& h = sims_.at(s);
    if (h.fieldX == ax && h.fieldY == ay) return;
    h.fieldX = ax;
    h.fieldY = ay;
    fill_desc(h);
    push_descs();
    invalidate_graphs();
}

void Engine::set_substeps(int substeps) {
    for (auto& h : sims_) {
        // K/GravityPipeline.kt:147-151: adaptive collisions choose their own
        int v = h.cfg.adaptive_collisions ? h.cfg.maximum_substeps
                                          : std::min(std::max(substeps, 1), h.cfg.maximum_substeps);
        if (v == h.substeps) continue;
        h.substeps = v;
        fill_desc(h);
    }
    push_descs();
    invalidate_graphs();
}

void Engine::normalize() {
    // K/GravityPipeline.kt:1326-1331: gravity follows the active mass.
    bool changed = false;
    for (auto& h : sims_) {
        float mass = std::max(1.0f, (float)h.n);
        if (mass == h.desc.totalMass) continue;
        fill_desc(h);
        changed = true;
    }
    if (changed) {
        push_descs();
        invalidate_graphs();
    }
}

void Engine::upload(int s, int first, int count, const float* xyvv, const float* mats) {
    HostSim& h = sims_.at(s);
    if (first < 0 || count < 0 || first > h.cap || count > h.cap - first)
        throw ArgumentError("upload range exceeds capacity");
    if (count == 0) return;
    // Both ping-pong buffers, as uploadBodies writes bodies and collisionBodies.
    for (int b = 0; b < 2; b++) {
        check(hipMemcpyAsync(body_[b] + h.base + first, xyvv, (size_t)count * 16, hipMemcpyHostToDevice, st_), "upload");
        if (mats)
            check(hipMemcpyAsync(mat_[b] + h.base + first, mats, (size_t)count * 16, hipMemcpyHostToDevice, st_), "upload");
    }
    check(hipStreamSynchronize(st_), "upload");
}

void Engine::upload_accel(int s, int first, int count, const float* accel) {
    HostSim& h = sims_.at(s);
    if (first < 0 || count < 0 || first > h.cap || count > h.cap - first)
        throw ArgumentError("upload range exceeds capacity");
    if (count == 0) return;
    check(hipMemcpy(acc_[auxPar_] + h.base + first, accel, (size_t)count * 8, hipMemcpyHostToDevice), "accel");
}

void Engine::download(int s, int what, int first, int count, void* dst) {
    HostSim& h = sims_.at(s);
    check(hipStreamSynchronize(st_), "download");
    if (what == GS_FORCE_MESH) {
        int nodes = h.desc.mw * h.desc.mh;
        if (first < 0 || count < 0 || first + count > nodes) throw ArgumentError("mesh range");
        check(hipMemcpy(dst, spec_ + h.desc.meshBase + first, (size_t)count * 8, hipMemcpyDeviceToHost), "mesh");
        return;
    }
    if (first < 0 || count < 0 || first > h.cap || count > h.cap - first)
        throw ArgumentError("download range exceeds capacity");
    if (count == 0) return;
    size_t o = (size_t)h.base + first;
    switch (what) {
        case GS_BODIES_XYVV:
            check(hipMemcpy(dst, body_[bodyPar_] + o, (size_t)count * 16, hipMemcpyDeviceToHost), "download");
            return;
        case GS_PRESSURES:
            check(hipMemcpy(dst, press_ + o, (size_t)count * 4, hipMemcpyDeviceToHost), "download");
            return;
        case GS_ACCEL:
            check(hipMemcpy(dst, acc_[auxPar_] + o, (size_t)count * 8, hipMemcpyDeviceToHost), "download");
            return;
        case GS_MATERIALS:
            check(hipMemcpy(dst, mat_[auxPar_] + o, (size_t)count * 16, hipMemcpyDeviceToHost), "download");
            return;
        case GS_CONTACT_COUNTS:
            check(hipMemcpy(dst, counts_ + o, (size_t)count * 4, hipMemcpyDeviceToHost), "download");
            return;
        case GS_SORT_DEST:
            check(hipMemcpy(dst, dest_ + o, (size_t)count * 4, hipMemcpyDeviceToHost), "download");
            return;
        case GS_CONTACTS: {
            std::vector<uint32_t> plane((size_t)count);
            uint32_t* out = (uint32_t*)dst;
            for (int k = 0; k < CONTACT_SLOTS; k++) {
                check(hipMemcpy(plane.data(), contacts_ + (size_t)k * totalSlots_ + o, (size_t)count * 4,
                                hipMemcpyDeviceToHost), "download");
                for (int i = 0; i < count; i++) out[(size_t)i * CONTACT_SLOTS + k] = plane[i] - (uint32_t)h.base;
            }
            return;
        }
        default:
            throw ArgumentError("unknown buffer");
    }
}

void Engine::synchronize() { check(hipStreamSynchronize(st_), "synchronize"); }

template <typename F>
void Engine::timed(const char* name, F&& f) {
    if (!profiling_) { f(); return; }
    std::vector<LaunchEvents> launched;
    launch_sink() = &launched;
    f();
    launch_sink() = nullptr;
    for (const LaunchEvents& e : launched) events_.push_back({name, {e.start, e.stop}});
}

// ---------------------------------------------------------------------------
// Schedule

void Engine::enqueue_sort() {
    // K/GravityPipeline.kt:1269-1319
    int o = bodyPar_ ^ 1, a = auxPar_ ^ 1;
    timed("sort_count", [&] { launch_sort_count(st_, shape_, body_[bodyPar_], tileCount_, keyOf_); });
    timed("sort_scan", [&] { launch_scan(st_, tileCount_, tilePrefix_, totalTiles_, scanTmp_, tuning_.scanDirectBlocks); });
    timed("sort_scatter", [&] { launch_sort_scatter(st_, shape_, keyOf_, tileCount_, tilePrefix_, slotSrc_, slotKey_); });
    timed("sort_canon", [&] {
        launch_sort_canon(st_, shape_, slotKey_, tilePrefix_, slotSrc_, body_[bodyPar_],
                          acc_[auxPar_], mat_[auxPar_], body_[o], acc_[a], mat_[a], dest_, big());
    });
    bodyPar_ = o;
    auxPar_ = a;
    // K/GravityPipeline.kt:1315: the blocks' members follow their bodies
    rigid_remap();
}

void Engine::enqueue_force_and_kick(bool initial) {
    // K/GravityPipeline.kt:907-973 then :1213-1261
    enqueue_force();
    enqueue_kick(initial);
}

void Engine::enqueue_force() {
    if (anyMutual_) {
        timed("bin", [&] { launch_bin(st_, shape_, body_[bodyPar_], cellCount_, homeOf_); });
        timed("cell_scan", [&] {
            launch_scan(st_, cellCount_, cellPrefix_, totalPhys_, scanTmp_, tuning_.scanDirectBlocks, cellRec_,
                cellNear_);
        });
        timed("cell_scatter", [&] { launch_cell_scatter(st_, shape_, homeOf_, cellCount_, cellPrefix_, slotBody_, slotCell_); });
        timed("cell_canon", [&] {
            launch_cell_canon(st_, gravitySlots_, slotCell_, cellPrefix_, slotBody_, body_[bodyPar_],
                              mat_[auxPar_], cellBody_, cellRec_, cellNear_, big());
        });
        timed("fft_rows_forward", [&] {
            launch_fft_rows_forward(st_, dSims_, dRowsF_, rowsF_, cellPrefix_, cellBody_, spec_, tw_,
                                    tuning_.depositRuns != 0);
        });
        timed("fft_columns", [&] { launch_fft_columns(st_, dSims_, dCols_, cols_, spec_, green_, tw_); });
        timed("fft_rows_inverse", [&] { launch_fft_rows_inverse(st_, dSims_, dRowsI_, rowsI_, spec_, tw_); });
    }
}

void Engine::enqueue_kick(bool initial) {
    timed("kick", [&] {
        launch_kick(st_, shape_, body_[bodyPar_], acc_[auxPar_], spec_, cellRec_, cellNear_,
                    cellBody_, initial ? 1 : 0);
    });
}

// One stand-alone pass (the test hook resolveCollisions).
void Engine::enqueue_pass(float kickDt, float driftDt, bool writePressure) {
    int g = gridPar_;
    timed("drift_insert", [&] {
        launch_drift_insert(st_, shape_, body_[bodyPar_], acc_[auxPar_], mat_[auxPar_], kickDt,
                            driftDt, 1, grid_[g]);
    });
    timed("count", [&] {
        launch_count(st_, shape_, grid_[g], grid_[g ^ 1], 1, mat_[auxPar_], counts_, contacts_, -1);
    });
    timed("solve", [&] {
        launch_solve(st_, shape_, body_[bodyPar_], body_[bodyPar_ ^ 1],
                     mat_[auxPar_], counts_, contacts_, grid_[g], fallback_, press_, writePressure ? 1 : 0,
                     0.0f, 0, grid_[g ^ 1], -1, 0);
    });
    bodyPar_ ^= 1;
    gridPar_ ^= 1;
}

void Engine::enqueue_step(bool sortDue) {
    // K/GravityPipeline.kt:445-494
    if (sortDue) enqueue_sort();
    const HostSim& h0 = sims_[0];
    const float dt = h0.cfg.time_step;
    const float sub = h0.desc.subDt;
    const int S = h0.substeps;
    // K/GravityPipeline.kt:458-461: adaptive simulations choose the step's substeps
    if (anyAdaptive_) timed("adaptive", [&] { launch_adapt(st_, shape_, (int)sims_.size(), body_[bodyPar_]); });
    if (anyCollide_) {
        timed("drift_insert", [&] {
            launch_drift_insert(st_, shape_, body_[bodyPar_], acc_[auxPar_], mat_[auxPar_], dt, sub, 1,
                                grid_[gridPar_]);
        });
        for (int k = 0; k < S; k++) {
            int g = gridPar_;
            timed("count", [&] {
                launch_count(st_, shape_, grid_[g], grid_[g ^ 1], 1, mat_[auxPar_], counts_, contacts_, k);
            });
            bool last = k == S - 1;
            float nextDrift = last ? 0.0f : sub;
            const int lastInsert = anyLiquid_ ? 0 : 1;
            int insertNext = last ? lastInsert : 1;
            timed("solve", [&] {
                launch_solve(st_, shape_, body_[bodyPar_], body_[bodyPar_ ^ 1],
                             mat_[auxPar_], counts_, contacts_, grid_[g], fallback_, press_, 0, nextDrift,
                             insertNext, grid_[g ^ 1], k, lastInsert);
            });
            bodyPar_ ^= 1;
            gridPar_ ^= 1;
        }
        if (anyLiquid_) {
            enqueue_liquid();
            timed("insert", [&] {
                launch_drift_insert(st_, shape_, body_[bodyPar_], acc_[auxPar_], mat_[auxPar_], 0.0f,
                                    0.0f, 1, grid_[gridPar_]);
            });
        }
        enqueue_force_and_kick(false);
        // The final zero-drift pass writes pressures (K/GravityPipeline.kt:480-486).
        int g = gridPar_;
        timed("count", [&] {
            launch_count(st_, shape_, grid_[g], grid_[g ^ 1], 1, mat_[auxPar_], counts_, contacts_, -1);
        });
        timed("solve", [&] {
            launch_solve(st_, shape_, body_[bodyPar_], body_[bodyPar_ ^ 1],
                         mat_[auxPar_], counts_, contacts_, grid_[g], fallback_, press_, 1, 0.0f, 0,
                         grid_[g ^ 1], -1, 0);
        });
        if (anyAdaptive_) timed("adaptive", [&] { launch_adapt_stats(st_, shape_, counts_); });
        bodyPar_ ^= 1;
        gridPar_ ^= 1;
        timed("block_order", [&] { launch_block_order(st_, shape_, order_); });
        // K/GravityPipeline.kt:487-490: the tiles awake in the next step.
        if (anySleep_) {
            timed("sleep_tiles", [&] {
                launch_sleep_advance(st_, dSims_, (int)sims_.size(), dTileBlocks_, tileBlocks_,
                                     shape_.sleep);
            });
        }
    } else {
        // Gravity-only configurations: kick-drift-kick with plain drifts, the
        // S substep drifts in one launch.
        timed("drift", [&] {
            launch_drift_insert(st_, shape_, body_[bodyPar_], acc_[auxPar_], mat_[auxPar_], dt, sub,
                                0, grid_[gridPar_], S);
        });
        enqueue_force_and_kick(false);
    }
}

// K/LiquidSolver.kt:107-173: start cells, `iterations` x (lambda, apply),
// then the viscosity pass; the bodies ping-pong at every apply.
void Engine::enqueue_liquid() {
    timed("liquid_bin", [&] {
        launch_liquid_bin(st_, shape_, body_[bodyPar_], liqCount_, liqCellOf_);
    });
    timed("liquid_scan", [&] { launch_scan(st_, liqCount_, liqPrefix_, totalLiq_, scanTmp_, tuning_.scanDirectBlocks); });
    timed("liquid_scatter", [&] {
        launch_liquid_scatter(st_, shape_, liqCellOf_, liqCount_, liqPrefix_, liqUnordered_);
    });
    timed("liquid_canon", [&] {
        launch_liquid_canon(st_, liquidSlots_, liqCellOf_, liqPrefix_, liqUnordered_, liqItems_);
    });
    int iterations = 0;
    for (const auto& h : sims_)
        if (h.liquid) iterations = std::max(iterations, h.cfg.liquid_iterations);
    for (int it = 0; it < iterations; it++) {
        timed("liquid_lambda", [&] {
            launch_liquid_lambda(st_, shape_, body_[bodyPar_], mat_[auxPar_], liqCellOf_,
                                 liqPrefix_, liqItems_, fallback_, lambdas_, press_,
                                 it == iterations - 1 ? 1 : 0);
        });
        timed("liquid_apply", [&] {
            launch_liquid_apply(st_, shape_, body_[bodyPar_], body_[bodyPar_ ^ 1], mat_[auxPar_],
                                liqCellOf_, liqPrefix_, liqItems_, fallback_, lambdas_);
        });
        bodyPar_ ^= 1;
    }
    if (anyViscous_) {
        timed("liquid_viscosity", [&] {
            launch_liquid_viscosity(st_, shape_, body_[bodyPar_], body_[bodyPar_ ^ 1],
                                    mat_[auxPar_], liqCellOf_, liqPrefix_, liqItems_);
        });
        bodyPar_ ^= 1;
    }
}

void Engine::debug_liquid() {
    refuse_blocks("a liquid pass");
    normalize();
    if (bodyBlocks_ == 0 || !anyLiquid_) return;
    enqueue_liquid();
    check(hipGetLastError(), "liquid");
}

// ---------------------------------------------------------------------------
// Adjoint of the gravity step (adjoint.hip): checkpoints every `every`
// steps while recording; the VJP recomputes one segment at a time from its
// checkpoint, keeping the force build of every step of the segment (body
// positions and masses, force mesh, cell tables, sort permutation), and
// runs the reversed leapfrog through it.

void Engine::adjoint_free() {
    for (void* p : adjAllocs_) hipFree(p);
    adjAllocs_.clear();
    adjCk_.clear();
    adjSlot_.clear();
    adjWhere_ = nullptr;
    adjFinal_ = AdjCheckpoint{};
    adjN_ = -1;
    for (long long& v : adjShape_) v = -1;
}

// The per-body work arrays of the reversed leapfrog and of the VJP test hook,
// allocated on first use and kept until the engine is destroyed.
void Engine::adjoint_work() {
    if (adjX_) return;
    const size_t cap = (size_t)sims_[0].cap;
    for (float2** p : {&adjX_, &adjV_, &adjX2_, &adjV2_, &adjG_, &adjSpread_, &adjW_}) {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(cap * sizeof(float2), 16)) != hipSuccess)
            throw StateError("hipMalloc failed");
        adjWork_.push_back(q);
        *p = (float2*)q;
    }
}

void Engine::adjoint_save_slot(int j) {
    const HostSim& h = sims_[0];
    AdjSlot& sl = adjSlot_[j];
    size_t n = (size_t)h.cap;
    auto copy = [&](void* dst, const void* src, size_t bytes) {
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st_), "adjoint slot");
    };
    copy(sl.body, body_[bodyPar_] + h.base, n * sizeof(float4));
    copy(sl.mats, mat_[auxPar_] + h.base, n * sizeof(float4));
    copy(sl.force, spec_ + h.desc.meshBase, (size_t)h.meshMax * sizeof(float2));
    copy(sl.rec, cellRec_ + h.desc.physBase, (size_t)h.phys * sizeof(uint4));
    copy(sl.prefix, cellPrefix_ + h.desc.physBase, ((size_t)h.phys + 1) * sizeof(uint32_t));
    copy(sl.cellBody, cellBody_, (size_t)gravitySlots_ * sizeof(float4));
}

// The adjoint kernels index their per-body arrays like the bodies and read
// the cell tables of the simulation's first cell slot: they serve one
// simulation whose bodies start at global index 0.
static void require_single_base0(const std::vector<HostSim>& sims, const char* what) {
    if (sims.size() != 1 || sims[0].base != 0) throw ArgumentError(what);
}

void Engine::adjoint_record(int steps, int every) {
    refuse_blocks("the adjoint");
    require_single_base0(sims_, "the adjoint runs on a single simulation");
    const HostSim& h = sims_[0];
    if (h.cfg.collisions_enabled) throw ArgumentError("the adjoint differentiates the gravity step without contact passes");
    if (!h.mutual) throw ArgumentError("the adjoint needs mutual gravity");
    if (h.liquid) throw ArgumentError("the adjoint does not differentiate the liquid");
    if (steps < 1) throw ArgumentError("steps must be positive");
    if (every < 1) throw ArgumentError("checkpoint interval must be positive");
    normalize();
    adjoint_work();
    size_t n = (size_t)h.cap;
    const long long shape[4] = {(long long)h.cap, steps, every, gravitySlots_};
    if (!std::equal(shape, shape + 4, adjShape_)) {
        // A new shape: the checkpoints, the per-step force builds of one
        // interval and the final state are (re)allocated once for it.
        adjoint_free();
        auto alloc = [&](size_t bytes) {
            void* p = nullptr;
            if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) throw StateError("hipMalloc failed");
            adjAllocs_.push_back(p);
            return p;
        };
        int segments = (steps + every - 1) / every;
        adjCk_.resize((size_t)segments);
        for (auto& c : adjCk_) {
            c.body = (float4*)alloc(n * sizeof(float4));
            c.acc = (float2*)alloc(n * sizeof(float2));
            c.mats = (float4*)alloc(n * sizeof(float4));
        }
        adjSlot_.resize((size_t)every);
        for (auto& sl : adjSlot_) {
            sl.body = (float4*)alloc(n * sizeof(float4));
            sl.mats = (float4*)alloc(n * sizeof(float4));
            sl.force = (float2*)alloc((size_t)h.meshMax * sizeof(float2));
            sl.rec = (uint4*)alloc((size_t)h.phys * sizeof(uint4));
            sl.prefix = (uint32_t*)alloc(((size_t)h.phys + 1) * sizeof(uint32_t));
            sl.cellBody = (float4*)alloc((size_t)std::max(gravitySlots_, 1) * sizeof(float4));
            sl.dest = (uint32_t*)alloc(n * sizeof(uint32_t));
        }
        adjFinal_.body = (float4*)alloc(n * sizeof(float4));
        adjFinal_.acc = (float2*)alloc(n * sizeof(float2));
        adjFinal_.mats = (float4*)alloc(n * sizeof(float4));
        adjWhere_ = (uint32_t*)alloc(n * sizeof(uint32_t));
        std::copy(shape, shape + 4, adjShape_);
    }
    launch_adj_iota(st_, shape_, adjWhere_);
    adjSortDue_.assign((size_t)steps, 0);
    adjSteps_ = steps;
    adjEvery_ = every;
    adjN_ = h.n;
    int interval = h.cfg.spatial_sort_interval;
    for (int k = 0; k < steps; k++) {
        if (k % every == 0) {
            AdjCheckpoint& c = adjCk_[(size_t)(k / every)];
            check(hipMemcpyAsync(c.body, body_[bodyPar_] + h.base, n * sizeof(float4),
                                 hipMemcpyDeviceToDevice, st_), "checkpoint");
            check(hipMemcpyAsync(c.acc, acc_[auxPar_] + h.base, n * sizeof(float2),
                                 hipMemcpyDeviceToDevice, st_), "checkpoint");
            check(hipMemcpyAsync(c.mats, mat_[auxPar_] + h.base, n * sizeof(float4),
                                 hipMemcpyDeviceToDevice, st_), "checkpoint");
            c.steps = sims_[0].steps;
        }
        bool sortDue = interval > 0 && (sims_[0].steps % interval) == 0;
        adjSortDue_[(size_t)k] = sortDue ? 1 : 0;
        step();
        if (sortDue) launch_adj_compose(st_, shape_, dest_, adjWhere_);
    }
    check(hipGetLastError(), "adjoint record");
}

void Engine::debug_accel_vjp(const float* w, float* out) {
    require_single_base0(sims_, "one simulation with mutual gravity");
    if (!sims_[0].mutual) throw ArgumentError("one simulation with mutual gravity");
    normalize();
    const HostSim& h = sims_[0];
    const size_t cap = (size_t)h.cap, n = (size_t)h.n;
    adjoint_work();
    check(hipMemsetAsync(adjX_, 0, cap * sizeof(float2), st_), "vjp");
    check(hipMemsetAsync(adjV_, 0, cap * sizeof(float2), st_), "vjp");
    check(hipMemcpyAsync(adjV_, w, n * sizeof(float2), hipMemcpyHostToDevice, st_), "vjp");
    enqueue_force();
    launch_adj_direct(st_, shape_, body_[bodyPar_], mat_[auxPar_], spec_, cellRec_, cellBody_,
                      adjX_, adjV_, 0.0f, adjW_, adjG_, adjSpread_);
    launch_fft_rows_forward_values(st_, dSims_, dRowsF_, rowsF_, cellPrefix_, cellBody_, adjSpread_,
                                   spec_, tw_);
    launch_fft_columns(st_, dSims_, dCols_, cols_, spec_, green_, tw_);
    launch_fft_rows_inverse(st_, dSims_, dRowsI_, rowsI_, spec_, tw_);
    launch_adj_finish(st_, shape_, body_[bodyPar_], mat_[auxPar_], spec_, adjG_, adjX_, adjV_, 0.0f,
                      1.0f);
    check(hipMemcpyAsync(out, adjX_, n * sizeof(float2), hipMemcpyDeviceToHost, st_), "vjp");
    check(hipStreamSynchronize(st_), "vjp");
}

void Engine::adjoint_order(int count, uint32_t* where) {
    if (!adjWhere_ || adjN_ < 0) throw StateError("no adjoint recording");
    if (count != adjN_) throw ArgumentError("count differs from the recording's body count");
    check(hipMemcpyAsync(where, adjWhere_, (size_t)adjN_ * 4, hipMemcpyDeviceToHost, st_),
          "adjoint order");
    check(hipStreamSynchronize(st_), "adjoint order");
}

void Engine::adjoint_vjp(int count, const float* dLdx, float* dLdx0, float* dLdv0) {
    if (adjCk_.empty() || adjN_ < 0) throw StateError("no adjoint recording");
    if (count != adjN_) throw ArgumentError("count differs from the recording's body count");
    HostSim& h = sims_[0];
    if (h.n != adjN_) throw StateError("the active body count changed since the recording");
    const size_t n = (size_t)h.n, cap = (size_t)h.cap;
    const float dt = h.cfg.time_step;
    const int S = h.substeps;
    const float sub = h.desc.subDt;
    // the final state, put back afterwards
    const int bp = bodyPar_, ap = auxPar_, gp = gridPar_;
    const long long finalSteps = h.steps;
    check(hipMemcpyAsync(adjFinal_.body, body_[bodyPar_] + h.base, cap * sizeof(float4),
                         hipMemcpyDeviceToDevice, st_), "adjoint");
    check(hipMemcpyAsync(adjFinal_.acc, acc_[auxPar_] + h.base, cap * sizeof(float2),
                         hipMemcpyDeviceToDevice, st_), "adjoint");
    check(hipMemcpyAsync(adjFinal_.mats, mat_[auxPar_] + h.base, cap * sizeof(float4),
                         hipMemcpyDeviceToDevice, st_), "adjoint");
    check(hipMemsetAsync(adjX_, 0, cap * sizeof(float2), st_), "adjoint");
    check(hipMemsetAsync(adjV_, 0, cap * sizeof(float2), st_), "adjoint");
    check(hipMemcpyAsync(adjX_, dLdx, n * sizeof(float2), hipMemcpyHostToDevice, st_), "adjoint");
    // The adjoint arrays are indexed like the bodies (sim 0 starts at 0).
    for (int seg = (int)adjCk_.size() - 1; seg >= 0; seg--) {
        const int c = seg * adjEvery_;
        const int k = std::min(adjEvery_, adjSteps_ - c);
        const AdjCheckpoint& ck = adjCk_[(size_t)seg];
        check(hipMemcpyAsync(body_[bodyPar_] + h.base, ck.body, cap * sizeof(float4),
                             hipMemcpyDeviceToDevice, st_), "adjoint restore");
        check(hipMemcpyAsync(acc_[auxPar_] + h.base, ck.acc, cap * sizeof(float2),
                             hipMemcpyDeviceToDevice, st_), "adjoint restore");
        check(hipMemcpyAsync(mat_[auxPar_] + h.base, ck.mats, cap * sizeof(float4),
                             hipMemcpyDeviceToDevice, st_), "adjoint restore");
        // recompute the segment, keeping each step's force build
        enqueue_force();
        adjoint_save_slot(0);
        for (int j = 0; j < k; j++) {
            AdjSlot& sl = adjSlot_[(size_t)j];
            sl.sorted = adjSortDue_[(size_t)(c + j)] != 0;
            if (sl.sorted) {
                enqueue_sort();
                check(hipMemcpyAsync(sl.dest, dest_ + h.base, cap * sizeof(uint32_t),
                                     hipMemcpyDeviceToDevice, st_), "adjoint slot");
            }
            if (j + 1 == k) break;  // the last step's build is not needed here
            launch_drift_insert(st_, shape_, body_[bodyPar_], acc_[auxPar_], mat_[auxPar_], dt, sub, 0,
                                grid_[gridPar_], S);
            enqueue_force();
            adjoint_save_slot(j + 1);
            enqueue_kick(false);
        }
        // reversed leapfrog through the segment
        for (int j = k - 1; j >= 0; j--) {
            const AdjSlot& sl = adjSlot_[(size_t)j];
            const int step = c + j;
            if (sl.sorted) {
                launch_adj_unsort(st_, shape_, sl.dest, adjX_, adjV_, adjX2_, adjV2_);
                std::swap(adjX_, adjX2_);
                std::swap(adjV_, adjV2_);
            }
            launch_adj_direct(st_, shape_, sl.body - h.base, sl.mats - h.base,
                              sl.force - h.desc.meshBase, sl.rec - h.desc.physBase,
                              sl.cellBody, adjX_, adjV_, dt, adjW_, adjG_, adjSpread_);
            launch_fft_rows_forward_values(st_, dSims_, dRowsF_, rowsF_, sl.prefix - h.desc.physBase,
                                           sl.cellBody, adjSpread_, spec_, tw_);
            launch_fft_columns(st_, dSims_, dCols_, cols_, spec_, green_, tw_);
            launch_fft_rows_inverse(st_, dSims_, dRowsI_, rowsI_, spec_, tw_);
            launch_adj_finish(st_, shape_, sl.body - h.base, sl.mats - h.base, spec_, adjG_, adjX_,
                              adjV_, dt, step == 0 ? 0.5f * dt : dt);
        }
    }
    check(hipGetLastError(), "adjoint");
    if (dLdx0)
        check(hipMemcpyAsync(dLdx0, adjX_, n * sizeof(float2), hipMemcpyDeviceToHost, st_), "adjoint");
    if (dLdv0)
        check(hipMemcpyAsync(dLdv0, adjV_, n * sizeof(float2), hipMemcpyDeviceToHost, st_), "adjoint");
    bodyPar_ = bp;
    auxPar_ = ap;
    gridPar_ = gp;
    h.steps = finalSteps;
    check(hipMemcpyAsync(body_[bodyPar_] + h.base, adjFinal_.body, cap * sizeof(float4),
                         hipMemcpyDeviceToDevice, st_), "adjoint");
    check(hipMemcpyAsync(acc_[auxPar_] + h.base, adjFinal_.acc, cap * sizeof(float2),
                         hipMemcpyDeviceToDevice, st_), "adjoint");
    check(hipMemcpyAsync(mat_[auxPar_] + h.base, adjFinal_.mats, cap * sizeof(float4),
                         hipMemcpyDeviceToDevice, st_), "adjoint");
    check(hipStreamSynchronize(st_), "adjoint");
}

void Engine::initialize_acceleration() {
    normalize();
    enqueue_force_and_kick(true);
    check(hipGetLastError(), "initialize acceleration");
}

void Engine::wake(int s) {
    launch_sleep_wake(st_, shape_.sleep, s, SLEEP_WAKE_STEPS);
    check(hipGetLastError(), "wake");
}

void Engine::refuse_blocks(const char* what) const {
    // The contact passes of a pipeline with glued blocks are the RIGID
    // variants (S/collision_rigid*.glsl), which this library does not have.
    if (rigid_.groupCount > 0)
        throw StateError(std::string(what) + " with glued blocks needs the RIGID contact passes, which are not built");
}

void Engine::step() {
    refuse_blocks("a step");
    normalize();
    // K/GravityPipeline.kt:446-452: new bodies, walls, field, gravity or
    // liquid wake everything for a while (nothing has marked the tiles).
    for (int i = 0; i < (int)sims_.size(); i++) {
        HostSim& h = sims_[i];
        SleepScene scene{h.n, h.bounded, h.fieldX, h.fieldY, h.mutual, h.liquid};
        if (!(scene == h.scene)) {
            h.scene = scene;
            if (h.desc.sleepOn) wake(i);
        }
    }
    const HostSim& h0 = sims_[0];
    int interval = h0.cfg.spatial_sort_interval;
    bool sortDue = false;
    if (interval > 0) sortDue = (sims_[0].steps % interval) == 0;
    for (auto& h : sims_) h.steps++;
    if (bodyBlocks_ == 0) return;
    if (!useGraphs_ || profiling_) {
        enqueue_step(sortDue);
        check(hipGetLastError(), "step");
        return;
    }
    int key = bodyPar_ | (auxPar_ << 1) | (gridPar_ << 2) | ((sortDue ? 1 : 0) << 3);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        int bp = bodyPar_, ap = auxPar_, gp = gridPar_;
        hipGraph_t graph;
        check(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal), "capture");
        enqueue_step(sortDue);
        check(hipStreamEndCapture(st_, &graph), "capture");
        hipGraphExec_t exec;
        check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
        hipGraphDestroy(graph);
        // Capture only recorded the launches: keep where the schedule ends
        // and start again from where it began.
        CapturedStep c{exec, bodyPar_, auxPar_, gridPar_};
        bodyPar_ = bp; auxPar_ = ap; gridPar_ = gp;
        it = graphs_.emplace(key, c).first;
    }
    check(hipGraphLaunch(it->second.exec, st_), "graph launch");
    bodyPar_ = it->second.bodyPar;
    auxPar_ = it->second.auxPar;
    gridPar_ = it->second.gridPar;
}

void Engine::stop_bodies(int s) {
    // S/stop_bodies.comp: zero velocities in both ping-pong buffers.
    HostSim& h = sims_.at(s);
    launch_stop(st_, shape_, body_[0], body_[1]);
    (void)h;
}

void Engine::debug_collide(float kickDt, float driftDt, bool writePressure) {
    refuse_blocks("a contact pass");
    normalize();
    if (bodyBlocks_ == 0) return;
    enqueue_pass(kickDt, driftDt, writePressure);
    check(hipGetLastError(), "collide");
}

void Engine::debug_density(uint32_t* density, float* spectrum, uint32_t* cells) {
    const HostSim& h = sims_[0];
    if (!h.mutual) throw ArgumentError("the density mesh needs mutual gravity");
    if (!debugDensity_) debugDensity_ = dalloc<uint32_t>((size_t)totalMesh_);
    density_readback(debugDensity_, density, spectrum, cells);
}

void Engine::debug_density_noalloc(uint32_t* density, float* spectrum, uint32_t* cells) {
    const HostSim& h = sims_[0];
    if (!h.mutual) throw ArgumentError("the density mesh needs mutual gravity");
    if ((size_t)totalMesh_ > DENSITY_SCRATCH_NODES)
        throw ArgumentError("the density scratch holds 2048 x 1024 mesh nodes at most");
    density_readback(scratch_density(), density, spectrum, cells);
}

// staging holds totalMesh_ uint32 (indexed by the simulations' mesh bases).
void Engine::density_readback(uint32_t* staging, uint32_t* density, float* spectrum, uint32_t* cells) {
    const HostSim& h = sims_[0];
    normalize();
    const size_t nodes = (size_t)h.desc.mw * h.desc.mh;
    // rows the walls leave empty are not written by the deposit, and the
    // staging keeps whatever an earlier call left in it
    check(hipMemsetAsync(staging, 0, (size_t)totalMesh_ * 4, st_), "density");
    launch_bin(st_, shape_, body_[bodyPar_], cellCount_, homeOf_);
    launch_scan(st_, cellCount_, cellPrefix_, totalPhys_, scanTmp_, tuning_.scanDirectBlocks, cellRec_,
                cellNear_);
    launch_cell_scatter(st_, shape_, homeOf_, cellCount_, cellPrefix_, slotBody_, slotCell_);
    launch_cell_canon(st_, gravitySlots_, slotCell_, cellPrefix_, slotBody_, body_[bodyPar_],
                      mat_[auxPar_], cellBody_, cellRec_, cellNear_, big());
    launch_fft_rows_forward(st_, dSims_, dRowsF_, rowsF_, cellPrefix_, cellBody_, spec_, tw_,
                            tuning_.depositRuns != 0, staging);
    launch_fft_columns(st_, dSims_, dCols_, cols_, spec_, green_, tw_, true);
    check(hipGetLastError(), "density");
    check(hipStreamSynchronize(st_), "density");
    if (density)
        check(hipMemcpy(density, staging + h.desc.meshBase, nodes * 4, hipMemcpyDeviceToHost),
              "density");
    if (spectrum)
        check(hipMemcpy(spectrum, spec_ + h.desc.meshBase, nodes * 8, hipMemcpyDeviceToHost), "spectrum");
    if (cells) {
        // the count and mass words (.y, .z) of each cell's 16-byte record
        check(hipMemcpy2D(cells, 2 * sizeof(uint32_t),
                          reinterpret_cast<const char*>(cellRec_ + h.desc.physBase) + sizeof(uint32_t),
                          sizeof(uint4), 2 * sizeof(uint32_t), (size_t)h.phys, hipMemcpyDeviceToHost),
              "cells");
    }
}

void Engine::validate_fft(double* out) {
    const HostSim& h = sims_[0];
    const int mw = h.desc.mw, mh = h.desc.mh;
    const size_t nodes = (size_t)mw * mh;
    std::vector<uint32_t> dens(nodes);
    std::vector<float> spec(2 * nodes);
    debug_density(dens.data(), spec.data(), nullptr);
    std::vector<std::complex<double>> ref(nodes);
    const double inv = 1.0 / (double)h.desc.massScale;
    for (size_t i = 0; i < nodes; i++) ref[i] = (double)dens[i] * inv;
    fft2d_f64(ref, mw, mh);
    double maxAbs = 0.0, maxRef = 0.0, sq = 0.0;
    for (size_t i = 0; i < nodes; i++) {
        for (int c = 0; c < 2; c++) {
            double e = c == 0 ? ref[i].real() : ref[i].imag();
            double d = std::fabs(e - (double)spec[2 * i + c]);
            maxAbs = std::max(maxAbs, d);
            maxRef = std::max(maxRef, std::fabs(e));
            sq += d * d;
        }
    }
    out[0] = maxAbs;
    out[1] = maxAbs / std::max(maxRef, 1e-30);
    out[2] = std::sqrt(sq / (double)(2 * nodes));
}

void Engine::debug_sort() {
    if (bodyBlocks_ == 0) return;
    enqueue_sort();
    check(hipGetLastError(), "sort");
}

std::vector<std::tuple<std::string, int, double>> Engine::profile(int steps) {
    profiling_ = true;
    try {
        for (int i = 0; i < steps; i++) step();
    } catch (...) {
        profiling_ = false;
        for (auto& e : events_) {
            hipEventDestroy(e.second.first);
            hipEventDestroy(e.second.second);
        }
        events_.clear();
        (void)hipGetLastError();  // leave no error for the next call to find
        throw;
    }
    profiling_ = false;
    check(hipStreamSynchronize(st_), "profile");
    std::map<std::string, std::pair<int, double>> acc;
    std::vector<std::string> order;
    for (auto& e : events_) {
        float ms = 0.0f;
        hipError_t err = hipEventElapsedTime(&ms, e.second.first, e.second.second);
        if (err != hipSuccess) {
            for (auto& x : events_) {
                hipEventDestroy(x.second.first);
                hipEventDestroy(x.second.second);
            }
            events_.clear();
            (void)hipGetLastError();
            throw StateError(std::string("elapsed: ") + hipGetErrorString(err));
        }
        if (!acc.count(e.first)) order.push_back(e.first);
        acc[e.first].first += 1;
        acc[e.first].second += ms;
        hipEventDestroy(e.second.first);
        hipEventDestroy(e.second.second);
    }
    events_.clear();
    std::vector<std::tuple<std::string, int, double>> out;
    for (auto& n : order) out.emplace_back(n, acc[n].first, acc[n].second);
    return out;
}

}  // namespace gs
