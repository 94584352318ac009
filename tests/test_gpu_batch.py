"""Batched ensembles (gs_batch_*): every member of a batch steps exactly as the
same simulation stepped alone in its own handle, bit for bit, whatever the
other members are (sizes, grids, walls or torus, liquid, sleep)."""
import numpy as np
import pytest

from gravity_amd import (CollisionConfig, GravityBatch, GravityConfig, GravityFieldConfig,
                         GravityPipeline, LiquidConfig, ParticleConfig, StartConfig, scenarios)

pytestmark = pytest.mark.gpu


def members():
    out = []
    b, m = scenarios.plummer(3000, 256, 256, a_px=30.0, seed=100)
    cfg = GravityConfig(particles=ParticleConfig(capacity=3000, initialBodies=3000),
                        gravity=GravityFieldConfig(gridWidth=64, gridHeight=64, minimumGrid=16),
                        collisions=CollisionConfig(sleep=False))
    out.append((cfg, (256, 256), False, b, m))
    b, m = scenarios.exponential_disc(5000, 640, 320, scale_px=25.0, friction=0.2, seed=5)
    cfg = GravityConfig(particles=ParticleConfig(capacity=6000, initialBodies=5000),
                        gravity=GravityFieldConfig(gridWidth=128, gridHeight=64, minimumGrid=16),
                        collisions=CollisionConfig(friction=0.2, contactDiameterPixels=1.0, sleep=True))
    out.append((cfg, (640, 320), True, b, m))
    b, m = scenarios.liquid_and_sand(2000, 120, 80, diameter=1.0746, seed=3)
    cfg = GravityConfig(particles=ParticleConfig(capacity=2000, initialBodies=2000),
                        gravity=GravityFieldConfig(gridWidth=32, gridHeight=32, minimumGrid=16),
                        collisions=CollisionConfig(contactDiameterPixels=1.0746, sleep=False),
                        liquid=LiquidConfig(), start=StartConfig(walls=True, mutualGravity=False,
                                                                 liquid=True))
    out.append((cfg, (120, 80), True, b, m))
    b, m = scenarios.uniform(700, 300, 200, seed=9)
    cfg = GravityConfig(particles=ParticleConfig(capacity=1000, initialBodies=700),
                        gravity=GravityFieldConfig(gridWidth=256, gridHeight=128, minimumGrid=16),
                        collisions=CollisionConfig(sleep=False))
    out.append((cfg, (300, 200), False, b, m))
    return out


def test_batch_members_step_like_plain_pipelines():
    ms = members()
    steps = 4
    print(f"[batch] {len(ms)} members, {steps} steps", flush=True)
    with GravityBatch([m[0] for m in ms], [m[1] for m in ms], [m[3] for m in ms],
                      bounded=[m[2] for m in ms]) as batch:
        assert batch.size == len(ms)
        for i, (_, _, _, b, mat) in enumerate(ms):
            batch.upload(i, b, mat)
            batch.setActive(i, len(b))
        batch.initializeAcceleration()
        for _ in range(steps):
            batch.step()
        got = [batch.bodies(i, len(m[3])) for i, m in enumerate(ms)]
    for i, (cfg, (W, H), bounded, b, mat) in enumerate(ms):
        n = len(b)
        with GravityPipeline(cfg, W, H, scenarioState=b, bounded=bounded) as pipe:
            pipe.upload(b, mat)
            pipe.initializeAcceleration(n)
            for _ in range(steps):
                pipe.step(n)
            want = pipe.bodies(n)
        np.testing.assert_array_equal(got[i], want, err_msg=f"member {i}")
