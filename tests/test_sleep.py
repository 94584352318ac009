"""Sleeping tiles (K/SleepTiles.kt, S/sleep_layout.glsl, S/sleep_tiles.comp):
the oracle pinned by the reference's own sleep tests (T/SleepIntegrationTest.kt)
on reduced scenes, then whole steps on the MI355X bit for bit against it."""
import math

import numpy as np
import pytest

from gravity_amd import (CollisionConfig, GravityConfig, GravityFieldConfig, ParticleConfig,
                         StartConfig)
from gravity_amd import GravityPipeline

D = 1.12838      # the reference's default contact diameter, px
FIELD = 300.0    # px/s^2, downward


def pile(width, rows, spacing=1.02):
    """T/SleepIntegrationTest.kt pile(): a hexagonal bed of full rows on the
    floor between the walls, in pixels."""
    step = spacing * D
    row_step = step * math.sqrt(3.0) * 0.5
    pts = []
    for row in range(rows):
        y = (0.5 + row) * row_step
        x = 0.5 * D + (0.0 if row % 2 == 0 else 0.5 * step)
        while x <= width - 0.5 * D:
            pts.append((x, y))
            x += step
    return np.array(pts)


def bodies_of(xy, W, H, vel=None):
    b = np.zeros((len(xy), 4), np.float32)
    b[:, 0] = np.float32(xy[:, 0]) / np.float32(W)
    b[:, 1] = np.float32(xy[:, 1]) / np.float32(H)
    if vel is not None:
        b[:, 2] = np.float32(vel[:, 0]) / np.float32(W)
        b[:, 3] = np.float32(vel[:, 1]) / np.float32(H)
    return b


def mats_of(n, friction):
    m = np.zeros((n, 4), np.float32)
    m[:, 0] = 1.0
    m[:, 1] = friction
    return m


def oracle_sim(oracle, W, H, bodies, friction=0.5):
    p = oracle.params(W, H, bounded=True, mutual_gravity=False, diameter=D, grid=(32, 32),
                      field=(0.0, -FIELD), sleep=True, sleep_speed=6.0)
    sim = oracle.Sim(p, bodies, mats_of(len(bodies), friction))
    sim.initialize_acceleration()
    return sim


def test_settled_pile_sleeps_still_and_a_dropped_body_wakes_it(oracle):
    # The reference's test on a 96 px wide box, 10 rows, fewer settling steps.
    W = H = 96
    xy = pile(W, 10)
    n = len(xy)
    sim = oracle_sim(oracle, W, H, bodies_of(xy, W, H))
    for _ in range(400):
        sim.step()
    settled = sim.state()[0]
    for _ in range(120):
        sim.step()
    later = sim.state()[0]
    still = np.sum((settled[:, 0] == later[:, 0]) & (settled[:, 1] == later[:, 1]))
    moved = np.hypot((later[:, 0] - settled[:, 0]) * W, (later[:, 1] - settled[:, 1]) * H).max()
    assert still >= n * 65 // 100, f"{still} of {n} still"
    assert moved < 8.0
    sim.close()

    # a body falls onto the pile from above its middle
    top = later[:, 1].max() * H
    dropped = np.concatenate([later, [[0.5, (top + 40.0) / H, 0.0, 0.0]]]).astype(np.float32)
    sim = oracle_sim(oracle, W, H, dropped)
    for _ in range(240):
        sim.step()
    after = sim.state()[0]
    sim.close()
    # the spatial sort renames bodies: find the dropped one as the highest
    # body over the middle
    landed = after[:, 1].max() * H
    assert landed > top - 3.0 * D
    speeds = np.hypot(after[:, 2] * W, after[:, 3] * H)
    assert speeds.max() < 5.0
    # the pile moved under it: positions are no longer the settled set
    before = {(x, y) for x, y in later[:, :2]}
    woken = sum((x, y) not in before for x, y in after[:, :2])
    assert woken > 1


def test_a_slow_body_rolling_into_a_resting_one_still_pushes_it(oracle):
    W = H = 256
    xy = np.array([[150.0, 0.5 * D], [60.0, 0.5 * D]])
    vel = np.array([[0.0, 0.0], [40.0, 0.0]])
    sim = oracle_sim(oracle, W, H, bodies_of(xy, W, H, vel), friction=0.0)
    for _ in range(480):
        sim.step()
    state = sim.state()[0]
    sim.close()
    ahead, behind = max(state[0, 0], state[1, 0]) * W, min(state[0, 0], state[1, 0]) * W
    assert ahead > 160.0
    assert ahead - behind >= 0.9 * D


def test_a_resting_bed_sleeps_and_wake_moves_it_again(oracle):
    # A bed on the floor settles and sleeps: a step leaves most bodies
    # exactly as they were. wake() (K/GravityPipeline.kt:610) keeps every
    # body awake for the next steps, so the field moves all of them again.
    W = H = 64
    xy = pile(W, 4)
    sim = oracle_sim(oracle, W, H, bodies_of(xy, W, H))
    for _ in range(250):
        sim.step()
    b0 = sim.state()[0]
    sim.step()
    b1 = sim.state()[0]
    still = np.all(b0 == b1, axis=1).mean()
    assert still > 0.5
    assert (~sim.body_awake()).mean() > 0.5
    sim.wake()
    sim.step()
    b2 = sim.state()[0]
    assert np.all(b1 == b2, axis=1).mean() < 0.1
    sim.close()


# --------------------------------------------------------------------------
# whole steps on the MI355X against the oracle

def _config(n, friction=0.5, sleep=True):
    return GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                         gravity=GravityFieldConfig(gridWidth=32, gridHeight=32, minimumGrid=16),
                         collisions=CollisionConfig(friction=friction, sleep=sleep),
                         start=StartConfig(walls=True, mutualGravity=False))


@pytest.mark.gpu
def test_sleeping_steps_match_oracle_bit_for_bit(oracle):
    # A bed settles on the floor and goes to sleep, a body dropped from above
    # lands on it, then wake() moves everything again. Every body, asleep or
    # awake, matches the oracle bit for bit after every step.
    W = H = 64
    xy = np.concatenate([pile(W, 4), [[32.0, 4 * D + 20.0]]])
    n = len(xy)
    bodies = bodies_of(xy, W, H)
    mats = mats_of(n, 0.5)
    sim = oracle_sim(oracle, W, H, bodies)
    print(f"[sleep] {n} bodies, 262 steps", flush=True)
    slept = 0
    with GravityPipeline(_config(n), W, H, scenarioState=bodies, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.mutualGravity = False
        pipe.fieldAccelerationY = -FIELD
        pipe.initializeAcceleration(n)
        for step in range(262):
            if step == 260:
                pipe.wake()
                sim.wake()
            pipe.step(n)
            sim.step()
            np.testing.assert_array_equal(pipe.bodies(n), sim.state()[0], err_msg=f"step {step}")
            if 2 <= step < 260:  # the first two steps are wake steps
                slept = max(slept, int((~sim.body_awake()).sum()))
    sim.close()
    assert slept > n // 2
