"""Adaptive collisions (CollisionConfig.adaptive, S/collision_adaptive_*.comp,
K/GravityPipeline.kt:1155-1211, measureAdaptiveCollisions :425-445).

The device picks each step's substep count r; the substep passes past r do
nothing and the others take dt/r. Two pins:

- the step with r chosen is bit for bit the plain step with r substeps
  (collisionSubsteps = r), whose parity with the oracle the plain tests
  show: this checks the skipped passes, the ping-pong parity kept across
  them, the substep of the drifts and the solves, and the last active pass's
  insertion for the final pass;
- r itself follows the finalize shader's formula from the figures the state
  reports: the largest speed (recomputed here from the downloaded
  velocities, bit for bit), the last final pass's penetration and dense
  bodies, all substeps on the first step and while the contact graph is
  dense, and the parity of the maximum; the contact figures are recomputed
  from the downloaded contact counts.
"""
import math

import numpy as np
import pytest

from gravity_amd import (CollisionConfig, GravityBatch, GravityConfig, GravityFieldConfig,
                         GravityPipeline, ParticleConfig, scenarios)
from gravity_amd.config import replace

pytestmark = pytest.mark.gpu
F = np.float32


def expected_substeps(cfg, first, speed, penetration, dense, n):
    """S/collision_adaptive_finalize.comp:16-40 in FP32."""
    k = cfg.collisions
    S = k.maximumSubsteps
    if first:
        required = S
    else:
        disp = F(k.contactDiameterPixels) * F(k.allowedDisplacementFraction)
        pen = F(k.contactDiameterPixels) * F(k.allowedPenetrationFraction)
        d = math.ceil(float(F(F(speed) * F(k.timeStep)) / disp))
        p = math.ceil(float(F(penetration) / pen))
        required = min(max(d, p, 1), S)
    threshold = max(k.minimumDenseBodies, int(math.ceil(float(F(n) * F(k.denseBodyFraction)))))
    if dense >= threshold:
        required = S
    if (required & 1) != (S & 1):
        required = min(required + 1, S)
    return required


def largest_speed(bodies, W, H):
    v = bodies[:, 2:4].astype(F)
    sx = v[:, 0] * F(W)
    sy = v[:, 1] * F(H)
    return F(np.max(np.sqrt(sx * sx + sy * sy)))


def scenes():
    out = []
    # the reference's scene (CollisionShaderIntegrationTest.kt:119-137): four
    # pairs closing at 40 px/s, 0.8 px apart
    n = 8
    b = np.zeros((n, 4), F)
    W, H = 1000, 200
    for i in range(n):
        b[i] = [(100 + (i // 2) * 100 + (i % 2) * 0.8) / W, 100 / H, (20 if i % 2 == 0 else -20) / W, 0]
    m = np.tile(np.array([1, 0, 1, 0], F), (n, 1))
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=64, gridHeight=16, minimumGrid=16),
                        collisions=CollisionConfig(adaptive=True, sleep=False))
    out.append(("pairs", cfg, (W, H), True, b, m))
    # a walled disc falling together under its own gravity, with sleeping
    # tiles; no dense-graph rule (every body would have to be dense) and
    # whole-diameter allowances, so the speed and the penetration choose few
    b, m = scenarios.exponential_disc(20000, 640, 320, scale_px=25.0, friction=0.2, seed=5)
    cfg = GravityConfig(particles=ParticleConfig(capacity=20000, initialBodies=20000),
                        gravity=GravityFieldConfig(gridWidth=128, gridHeight=64, minimumGrid=16),
                        collisions=CollisionConfig(friction=0.2, contactDiameterPixels=1.0,
                                                   adaptive=True, sleep=True, denseBodyFraction=1.0,
                                                   minimumDenseBodies=20000,
                                                   allowedDisplacementFraction=1.0,
                                                   allowedPenetrationFraction=1.0))
    out.append(("disc", cfg, (640, 320), True, b, m))
    # a torus Plummer sphere: its dense contact graph keeps every substep
    b, m = scenarios.plummer(6000, 512, 512, a_px=30.0, seed=104)
    cfg = GravityConfig(particles=ParticleConfig(capacity=6000, initialBodies=6000),
                        gravity=GravityFieldConfig(gridWidth=64, gridHeight=64, minimumGrid=16),
                        collisions=CollisionConfig(adaptive=True, sleep=False, minimumDenseBodies=64))
    out.append(("plummer", cfg, (512, 512), False, b, m))
    return out


@pytest.mark.parametrize("name", ["pairs", "disc", "plummer"])
def test_adaptive_step_is_the_plain_step_with_its_substeps(name):
    _, cfg, (W, H), bounded, b, m = next(s for s in scenes() if s[0] == name)
    n = len(b)
    plain_cfg = replace(cfg, collisions=replace(cfg.collisions, adaptive=False))
    steps = 12
    seen = set()
    dense_rule = False
    with GravityPipeline(cfg, W, H, scenarioState=b, bounded=bounded) as a, \
            GravityPipeline(plain_cfg, W, H, scenarioState=b, bounded=bounded) as p:
        a.upload(b, m)
        p.upload(b, m)
        a.initializeAcceleration(n)
        p.initializeAcceleration(n)
        a.collisionSubsteps = 2  # adaptive collisions choose their own
        assert a.collisionSubsteps == cfg.collisions.maximumSubsteps
        previous = None
        for step in range(steps):
            print(f"[adaptive] {name} step {step}", flush=True)
            speed = largest_speed(a.bodies(n), W, H)
            a.step(n)
            st = a.measureAdaptiveCollisions()
            assert st.maximumSpeedPixelsPerSecond == speed
            want = expected_substeps(cfg, previous is None, speed,
                                     0.0 if previous is None else previous.maximumPenetrationPixels,
                                     0 if previous is None else previous.denseContactBodies, n)
            assert st.substeps == want, (step, st, want)
            threshold = max(cfg.collisions.minimumDenseBodies,
                            int(math.ceil(float(F(n) * F(cfg.collisions.denseBodyFraction)))))
            dense_rule |= previous is not None and previous.denseContactBodies >= threshold
            counts = a.contactCounts(n).astype(np.int64)
            if not cfg.collisions.sleep:
                assert st.maximumContactCount == int(counts.max())
                assert st.denseContactBodies == int((counts >= cfg.collisions.denseContactThreshold).sum())
            assert 0.0 <= st.maximumPenetrationPixels <= cfg.collisions.contactDiameterPixels
            seen.add(st.substeps)
            p.collisionSubsteps = st.substeps
            p.step(n)
            for what, x, y in (("bodies", a.bodies(n), p.bodies(n)),
                               ("pressures", a.pressures(n), p.pressures(n))):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (name, step, what)
            previous = st
    if name == "plummer":
        assert dense_rule and seen == {cfg.collisions.maximumSubsteps}, seen
    else:
        # fewer substeps than the maximum
        assert min(seen) < cfg.collisions.maximumSubsteps, seen


def test_adaptive_batch_members_step_like_plain_pipelines():
    ms = [s for s in scenes() if s[0] != "disc"]
    with GravityBatch([s[1] for s in ms], [s[2] for s in ms], [s[4] for s in ms],
                      bounded=[s[3] for s in ms]) as batch:
        for i, s in enumerate(ms):
            batch.upload(i, s[4], s[5])
            batch.setActive(i, len(s[4]))
        batch.initializeAcceleration()
        for _ in range(6):
            batch.step()
        got = [(batch.bodies(i, len(s[4])), batch.measureAdaptiveCollisions(i)) for i, s in enumerate(ms)]
    for i, (name, cfg, (W, H), bounded, b, m) in enumerate(ms):
        with GravityPipeline(cfg, W, H, scenarioState=b, bounded=bounded) as p:
            p.upload(b, m)
            p.initializeAcceleration(len(b))
            for _ in range(6):
                p.step(len(b))
            assert np.array_equal(p.bodies(len(b)).view(np.uint32), got[i][0].view(np.uint32)), name
            assert p.measureAdaptiveCollisions() == got[i][1], name

