// Sleeping tiles, end of a step (S/sleep_tiles.comp:10-28, K/SleepTiles.kt:
// advance and endStep). The final collision pass marks the tiles of bodies
// that are not at rest with the next step's stamp; here every tile next to
// a marked one (3x3, wrapping on the torus) is woken for the next step, and
// then each simulation's stamp moves on. Stamps tag what is awake in a
// step, so nothing is ever cleared.
#include "common.hpp"
#include "kernels.hpp"

namespace gs {

__global__ void __launch_bounds__(WG) k_sleep_tiles(const SimDesc* __restrict__ sims,
                                                    const BodyBlock* __restrict__ tileBlocks,
                                                    SleepBufs sleep) {
    BodyBlock tb = tileBlocks[blockIdx.x];
    const SimDesc& s = sims[tb.sim];
    int local = tb.first + (int)threadIdx.x;
    if (!s.sleepOn || local >= s.tilesX * s.tilesY) return;
    uint32_t next = sleep.state[tb.sim].stamp + 1u;
    int x0 = local % s.tilesX, y0 = local / s.tilesX;
    bool awake = false;
    for (int dy = -1; dy <= 1; dy++) {
        for (int dx = -1; dx <= 1; dx++) {
            int x = x0 + dx, y = y0 + dy;
            if (s.bounded) {
                if (x < 0 || y < 0 || x >= s.tilesX || y >= s.tilesY) continue;
            } else {
                x = (x + s.tilesX) % s.tilesX;
                y = (y + s.tilesY) % s.tilesY;
            }
            if (sleep.tileActive[s.sleepTileBase + x + y * s.tilesX] == next) awake = true;
        }
    }
    if (awake) sleep.tileAwake[s.sleepTileBase + local] = next;
}

__global__ void k_sleep_end(const SimDesc* __restrict__ sims, int count, SleepBufs sleep) {
    int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= count || !sims[i].sleepOn) return;
    SleepState st = sleep.state[i];
    st.stamp += 1u;
    if (st.stamp == 0u) st.stamp = 1u;
    if (st.wakeSteps > 0) st.wakeSteps -= 1;
    sleep.state[i] = st;
}

// GravityPipeline.wake() (K/GravityPipeline.kt:610): everything awake for
// the next steps.
__global__ void k_sleep_wake(SleepBufs sleep, int sim, int steps) {
    if (threadIdx.x == 0) sleep.state[sim].wakeSteps = steps;
}

void launch_sleep_wake(hipStream_t st, const SleepBufs& sleep, int sim, int steps) {
    launch(k_sleep_wake, dim3(1), dim3(64), 0, st, sleep, sim, steps);
}

void launch_sleep_advance(hipStream_t st, const SimDesc* sims, int simCount,
                          const BodyBlock* tileBlocks, int tileBlockCount, const SleepBufs& sleep) {
    if (tileBlockCount > 0)
        launch(k_sleep_tiles, dim3(tileBlockCount), dim3(WG), 0, st, sims, tileBlocks,
                           sleep);
    launch(k_sleep_end, dim3((simCount + 63) / 64), dim3(64), 0, st, sims, simCount,
                       sleep);
}

}  // namespace gs
