"""The trajectory preview on the MI355X (gs_update_trajectory,
K/GravityPipeline.kt:514-602, S/trajectory.comp) against the oracle's
restatement (orc_sim_trajectory, pinned by closed forms in
tests/test_trajectory.py) over the same force build and bodies, and the
reference's own case of a cleared session
(T/CollisionShaderIntegrationTest.kt:331-345).

The two force builds agree to FP32 tolerance (their FFTs and the kick's
exp differ in the last bits), so the paths are compared with a tolerance:
the same vertices used, positions within 1e-3 px over the first 256 steps
and the whole path's length within 1e-4 of each other."""
import numpy as np
import pytest

from gravity_amd import (CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline,
                         ParticleConfig, TrajectoryConfig, scenarios)

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[trajectory] {msg}", flush=True)


def scene(name):
    if name == "disc-walls":
        W, H, n = 640, 320, 4000
        b, m = scenarios.exponential_disc(n, W, H, scale_px=40.0, friction=0.2, seed=5)
        return W, H, True, b, m, (64, 32), True, (0.0, 0.0), (120.0, 150.0), (80.0, -40.0)
    if name == "clump-torus":
        # a Plummer clump; the path starts in its sparse outskirts and
        # crosses both seams
        W, H, n = 640, 320, 6000
        b, m = scenarios.plummer(n, W, H, a_px=30.0, seed=104)
        return W, H, False, b, m, (64, 32), True, (0.0, 0.0), (600.0, 20.0), (150.0, -100.0)
    if name == "field-walls":
        W, H, n = 640, 320, 100
        b, m = scenarios.uniform(n, W, H, seed=3)
        return W, H, True, b, m, (64, 32), False, (0.0, 300.0), (50.0, 50.0), (150.0, -200.0)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["disc-walls", "clump-torus", "field-walls"])
def test_preview_follows_the_oracle(oracle, name):
    W, H, bounded, b, m, grid, mutual, field, origin, velocity = scene(name)
    n = len(b)
    steps = 1024
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=grid[0], gridHeight=grid[1], minimumGrid=16),
                        collisions=CollisionConfig(restitutionCoefficient=0.5),
                        trajectory=TrajectoryConfig(steps=steps))
    say(f"{name}: {n} bodies, {steps} steps")
    with GravityPipeline(cfg, W, H, scenarioState=b, bounded=bounded) as p:
        p.mutualGravity = mutual
        p.fieldAccelerationX, p.fieldAccelerationY = field
        p.upload(b, m)
        p.initializeAcceleration(n)
        p.updateTrajectory(n, origin[0], origin[1], velocity[0], velocity[1])
        got = p.trajectory()
        ptr, count = p.trajectoryView()
        assert ptr != 0 and count == steps * 6
    sim = oracle.Sim(oracle.params(W, H, bounded=bounded, mutual_gravity=mutual, grid=grid, field=field,
                                   restitution=0.5), b, m)
    sim.initialize_acceleration()
    want = sim.trajectory(origin, velocity, steps=steps)
    used = got[:, 3] >= 0
    gap = np.hypot((got[:, 0] - want[:, 0]).astype(np.float64) * W,
                   (got[:, 1] - want[:, 1]).astype(np.float64) * H)
    say(f"{name}: {used.sum()} vertices used; max gap {gap[:256 * 6][used[:256 * 6]].max():.2e} px "
        f"(first 256 steps), {gap[used].max():.2e} px (all)")
    np.testing.assert_array_equal(used, want[:, 3] >= 0)
    assert gap[:256 * 6][used[:256 * 6]].max() <= 1e-3
    # the dash coordinate's end is the whole path's length in dash units
    last = np.nonzero(used)[0][-1]
    assert abs(float(got[last, 2]) - float(want[last, 2])) <= 1e-4 * abs(float(want[last, 2]))


def test_a_cleared_session_with_no_bodies_accepts_projectiles():
    W, H = 2000, 1000
    cfg = GravityConfig(particles=ParticleConfig(capacity=4096, initialBodies=1),
                        gravity=GravityFieldConfig(gridWidth=256, gridHeight=128))
    with GravityPipeline(cfg, W, H) as p:
        # nothing is active: gravity and the preview still run
        p.initializeAcceleration(0)
        p.updateTrajectory(0, 1000.0, 700.0, 100.0, 0.0)
        v = p.trajectory()
        assert np.isfinite(v).all()
        spawned = p.spawnProjectile(0, 1000.0, 700.0, 200.0, 0.0, 30.0)
        assert spawned > 0
        for _ in range(30):
            p.step(spawned)
        assert p.measureCluster(spawned).nonFiniteBodies == 0
