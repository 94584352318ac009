"""Whole steps (K/GravityPipeline.kt:445-494) on the MI355X against the oracle,
and size-independent invariants at the benchmark's full size."""
import numpy as np
import pytest

from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
from gravity_amd import scenarios

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("substeps,bounded", [(1, False), (16, False), (16, True)])
def test_gravity_only_steps_follow_the_oracle(oracle, substeps, bounded):
    # The P3M-only workload at reduced N: positions within 1e-3 px; with 16
    # substeps the step's drifts run in one launch, each wrapping (torus) or
    # clamping (walls) as the oracle's pass does.
    W, H, n, steps = 1440, 720, 5000, 20
    print(f"[step] gravity only, {substeps} substeps, walls {bounded}", flush=True)
    bodies, mats = scenarios.uniform(n, W, H, seed=1)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=512, gridHeight=256),
                        collisions=CollisionConfig(maximumSubsteps=substeps), collisionsEnabled=False)
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=bounded) as pipe:
        pipe.initializeAcceleration(n)
        for _ in range(steps):
            pipe.step(n)
        got = pipe.bodies(n)
        dest = pipe.sortDestinations(n)
    sim = oracle.Sim(oracle.params(W, H, bounded=bounded, collide=False, substeps=substeps, grid=(512, 256)),
                     bodies, mats)
    sim.initialize_acceleration()
    for _ in range(steps):
        sim.step()
    want = sim.state()[0]
    np.testing.assert_array_equal(dest, sim.last_sort_dest())
    dx = (got[:, 0] - want[:, 0]) * W
    dy = (got[:, 1] - want[:, 1]) * H
    dx -= np.round(dx / W) * W
    dy -= np.round(dy / H) * H
    assert np.abs(np.hypot(dx, dy)).max() < 1e-3


def test_collision_gravity_step_follows_the_oracle(oracle):
    W, H, n = 640, 320, 4000
    bodies, mats = scenarios.exponential_disc(n, W, H, scale_px=25.0, gravity=0.035, friction=0.2, seed=5)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=128, gridHeight=64, minimumGrid=16),
                        collisions=CollisionConfig(friction=0.2, contactDiameterPixels=1.0))
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        pipe.step(n)
        got = pipe.bodies(n)
        press = pipe.pressures(n)
    p = oracle.params(W, H, bounded=True, grid=(128, 64), diameter=1.0)
    sim = oracle.Sim(p, bodies, mats)
    sim.initialize_acceleration()
    sim.step()
    want, _, wpress, _ = sim.state()
    err = np.abs(got[:, :2] - want[:, :2]) * np.array([W, H])
    assert err.max() < 1e-3, err.max()
    assert np.isfinite(press).all()


def test_full_size_galaxy_steps_stay_finite_and_inside_the_walls():
    W, H, n = 2880, 1440, 400_000
    bodies, mats = scenarios.exponential_disc(n, W, H, seed=7)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=1024, gridHeight=512),
                        collisions=CollisionConfig(friction=0.2, contactDiameterPixels=1.0))
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        for _ in range(4):
            pipe.step(n)
        b = pipe.bodies(n)
        dest = pipe.sortDestinations(n)
    assert np.isfinite(b).all()
    assert ((b[:, :2] >= 0) & (b[:, :2] < 1)).all()
    assert np.array_equal(np.sort(dest), np.arange(n))
