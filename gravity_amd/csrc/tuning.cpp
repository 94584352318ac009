// Launch-shape tuning of the step: the choices of Tuning, validated, and an
// autotuner that times each candidate on the engine's own workload.
//
// No choice changes a result: the big-segment sort orders a cell or tile as
// the per-lane ranks do, the deposit's integer sums do not depend on their
// order or grouping, forward row items only regroup whole lines, and a scan's
// totals add up the same either way. The autotuner therefore times the real
// step on the real state: it keeps a device copy of everything a step reads
// and writes across steps, and puts it back before every trial and at the end,
// so the simulation continues exactly as if it had not been tuned.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

namespace gs {

void Engine::set_tuning(const Tuning& t) {
    if (t.bigLimit != 0 && t.bigLimit < BIG_LIMIT_MIN)
        throw ArgumentError("big segment limit must be 0 or at least 32");
    if (t.forwardItemBodies < 64 || t.forwardItemBodies > (1 << 24))
        throw ArgumentError("forward item bodies must be within [64, 2^24]");
    if (t.depositRuns != 0 && t.depositRuns != 1) throw ArgumentError("deposit runs is 0 or 1");
    if (t.scanDirectBlocks < 1) throw ArgumentError("scan direct blocks must be positive");
    const bool regroup = t.forwardItemBodies != tuning_.forwardItemBodies;
    tuning_ = t;
    if (regroup) rebuild_tables();
    invalidate_graphs();
}

// Mean ms per step of `steps` graph-replayed steps, after two steps that
// capture the graphs the schedule alternates between.
double Engine::time_steps(int steps) {
    step();
    step();
    hipEvent_t a, b;
    check(hipEventCreate(&a), "event");
    check(hipEventCreate(&b), "event");
    check(hipEventRecord(a, st_), "event");
    for (int i = 0; i < steps; i++) step();
    check(hipEventRecord(b, st_), "event");
    check(hipEventSynchronize(b), "event");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, a, b), "event");
    hipEventDestroy(a);
    hipEventDestroy(b);
    return (double)ms / steps;
}

namespace {

// Device copies of the engine's state between steps.
struct Saved {
    std::vector<std::pair<void*, std::pair<const void*, size_t>>> regions;  // copy, source, bytes
    ~Saved() {
        for (auto& r : regions) hipFree(r.first);
    }
};

}  // namespace

Tuning Engine::autotune(int steps, double* ms) {
    if (steps < 1) throw ArgumentError("autotune needs at least one timed step");
    normalize();
    check(hipStreamSynchronize(st_), "autotune");
    const Tuning start = tuning_;
    if (bodyBlocks_ == 0) {
        if (ms) ms[0] = ms[1] = 0.0;
        return start;
    }
    const size_t S = (size_t)totalSlots_;
    // Everything a step carries over to the next: both parities of the
    // bodies, accelerations and materials (the step ping-pongs them), the
    // pressures, the sort destinations, and with sleeping tiles the tile and
    // body stamps, the sleep state and the contact tables a sleeping body
    // keeps. The broad-phase heads are cleared on restore; nothing else
    // outlives a step.
    Saved saved;
    auto keep = [&](const void* src, size_t bytes) {
        void* copy = nullptr;
        check(hipMalloc(&copy, bytes), "autotune copy");
        saved.regions.push_back({copy, {src, bytes}});
        check(hipMemcpyAsync(copy, src, bytes, hipMemcpyDeviceToDevice, st_), "autotune copy");
    };
    for (int b = 0; b < 2; b++) {
        keep(body_[b], S * sizeof(float4));
        keep(acc_[b], S * sizeof(float2));
        keep(mat_[b], S * sizeof(float4));
    }
    keep(press_, S * sizeof(float));
    keep(dest_, S * sizeof(uint32_t));
    if (anySleep_) {
        keep(sleepState_, sims_.size() * sizeof(SleepState));
        keep(bodyAwake_, S * sizeof(uint32_t));
        const size_t tiles = (size_t)std::max(totalSleepTiles_, 1);
        keep(tileAwake_, tiles * sizeof(uint32_t));
        keep(tileActive_, tiles * sizeof(uint32_t));
        keep(counts_, S * sizeof(uint32_t));
        keep(contacts_, S * CONTACT_PLANES * sizeof(uint32_t));
    }
    check(hipStreamSynchronize(st_), "autotune copy");
    const int parities[3] = {bodyPar_, auxPar_, gridPar_};
    std::vector<long long> hostSteps;
    std::vector<SleepScene> scenes;
    for (const auto& h : sims_) {
        hostSteps.push_back(h.steps);
        scenes.push_back(h.scene);
    }
    auto restore = [&]() {
        check(hipStreamSynchronize(st_), "autotune restore");
        for (auto& r : saved.regions)
            check(hipMemcpyAsync(const_cast<void*>(r.second.first), r.first, r.second.second,
                                 hipMemcpyDeviceToDevice, st_),
                  "autotune restore");
        clear_grids();
        check(hipStreamSynchronize(st_), "autotune restore");
        bodyPar_ = parities[0];
        auxPar_ = parities[1];
        gridPar_ = parities[2];
        for (size_t i = 0; i < sims_.size(); i++) {
            sims_[i].steps = hostSteps[i];
            sims_[i].scene = scenes[i];
        }
    };
    auto trial = [&](const Tuning& t) {
        set_tuning(t);
        restore();
        return time_steps(steps);
    };
    Tuning best = start;
    double bestMs = 0.0, startMs = 0.0;
    try {
        startMs = bestMs = trial(start);
        // Coordinate descent: each field in turn over its candidates, the
        // others held at the best so far; a candidate must win by 1 %.
        const int bigs[] = {0, 32, 64, 128, 256};
        const int items[] = {1024, 2048, 4096, 8192, 16384};
        const int runs[] = {0, 1};
        const int scans[] = {64, 256, 1024, 4096, 1 << 30};
        auto sweep = [&](const int* values, int count, int Tuning::*field) {
            for (int i = 0; i < count; i++) {
                if (values[i] == best.*field) continue;
                Tuning t = best;
                t.*field = values[i];
                const double m = trial(t);
                if (m < 0.99 * bestMs) {
                    bestMs = m;
                    best = t;
                }
            }
        };
        sweep(bigs, 5, &Tuning::bigLimit);
        sweep(items, 5, &Tuning::forwardItemBodies);
        sweep(runs, 2, &Tuning::depositRuns);
        sweep(scans, 5, &Tuning::scanDirectBlocks);
    } catch (...) {
        hipGetLastError();
        set_tuning(start);
        restore();
        throw;
    }
    set_tuning(best);
    restore();
    if (ms) {
        ms[0] = bestMs;
        ms[1] = startMs;
    }
    return best;
}

}  // namespace gs
