"""Launch-shape tuning (gs_tuning, gs_autotune): every choice steps bit for
bit like every other, on workloads whose crowded cells and tiles take the
workgroup sort (and one cell too long for it), and the autotuner leaves the
simulations exactly where they were."""
import numpy as np
import pytest

from gravity_amd import (CollisionConfig, GravityBatch, GravityConfig, GravityFieldConfig,
                         LiquidConfig, ParticleConfig, StartConfig, scenarios)
from gravity_amd.tuning import Tuning

pytestmark = pytest.mark.gpu

# the reference choice (the shapes the oracle parity tests also run), and two
# that exercise every other path
PLAIN = Tuning(bigSegmentLimit=0, forwardItemBodies=4096, depositRuns=0, scanDirectBlocks=1 << 30)
CHOICES = [Tuning(),
           Tuning(bigSegmentLimit=32, forwardItemBodies=1024, depositRuns=1, scanDirectBlocks=1),
           Tuning(bigSegmentLimit=256, forwardItemBodies=16384, depositRuns=0, scanDirectBlocks=64)]


def crowded_members():
    """Plummer spheres far denser than their grid cells: the 60,000-body one
    on a 16 x 16 grid puts more than 8192 bodies in its central cells (past
    the workgroup sort), the others put hundreds in theirs."""
    out = []
    for n, grid, a, seed in [(60000, 16, 50.0, 101), (20000, 64, 40.0, 102), (5000, 128, 20.0, 103)]:
        b, m = scenarios.plummer(n, 1024, 1024, a_px=a, seed=seed)
        cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=seed),
                            gravity=GravityFieldConfig(gridWidth=grid, gridHeight=grid, minimumGrid=16),
                            collisions=CollisionConfig(maximumSubsteps=4, sleep=False))
        out.append((cfg, (1024, 1024), False, b, m))
    return out


def sleepy_members():
    """A walled galaxy with sleeping tiles and a liquid-and-sand box with
    sleeping tiles: the state a step carries over is widest here."""
    out = []
    b, m = scenarios.exponential_disc(20000, 640, 320, scale_px=25.0, friction=0.2, seed=5)
    cfg = GravityConfig(particles=ParticleConfig(capacity=20000, initialBodies=20000),
                        gravity=GravityFieldConfig(gridWidth=128, gridHeight=64, minimumGrid=16),
                        collisions=CollisionConfig(friction=0.2, contactDiameterPixels=1.0, sleep=True))
    out.append((cfg, (640, 320), True, b, m))
    b, m = scenarios.liquid_and_sand(8000, 200, 120, diameter=1.0746, seed=3)
    cfg = GravityConfig(particles=ParticleConfig(capacity=8000, initialBodies=8000),
                        gravity=GravityFieldConfig(gridWidth=32, gridHeight=32, minimumGrid=16),
                        collisions=CollisionConfig(contactDiameterPixels=1.0746, sleep=True),
                        liquid=LiquidConfig(), start=StartConfig(walls=True, mutualGravity=False,
                                                                 liquid=True))
    out.append((cfg, (200, 120), True, b, m))
    return out


def run(ms, tuning, steps, autotune_after=None):
    with GravityBatch([m[0] for m in ms], [m[1] for m in ms], [m[3] for m in ms],
                      bounded=[m[2] for m in ms]) as batch:
        for i, (_, _, _, b, mat) in enumerate(ms):
            batch.upload(i, b, mat)
            batch.setActive(i, len(b))
        if tuning is not None:
            batch.tuning = tuning
            assert batch.tuning == tuning
        batch.initializeAcceleration()
        chosen = None
        for k in range(steps):
            if autotune_after == k:
                chosen = batch.autotune(steps=1)
            batch.step()
        out = [(batch.bodies(i, len(m[3])), batch.acceleration(i, len(m[3])),
                batch.pressures(i, len(m[3]))) for i, m in enumerate(ms)]
    return out, chosen


def same(a, b, what):
    for i, (x, y) in enumerate(zip(a, b)):
        for name, u, v in zip(("bodies", "acceleration", "pressures"), x, y):
            assert np.array_equal(u.view(np.uint32), v.view(np.uint32)), f"{what}: member {i} {name}"


def test_every_tuning_steps_alike():
    ms = crowded_members()
    print(f"[tuning] crowded batch of {len(ms)}, plain shapes", flush=True)
    ref, _ = run(ms, PLAIN, 3)
    for t in CHOICES:
        print(f"[tuning] {t}", flush=True)
        got, _ = run(ms, t, 3)
        same(ref, got, str(t))


def test_autotune_leaves_the_state_unchanged(tmp_path):
    ms = sleepy_members()
    print("[tuning] sleepy batch, 5 steps untuned", flush=True)
    ref, _ = run(ms, None, 5)
    print("[tuning] same with autotune after 2 steps", flush=True)
    got, (chosen, ms_) = run(ms, None, 5, autotune_after=2)
    same(ref, got, "autotuned")
    assert isinstance(chosen, Tuning)
    assert ms_["chosen"] > 0.0 and ms_["chosen"] <= ms_["start"]


def test_autotune_cache(tmp_path):
    ms = crowded_members()[1:]
    cache = tmp_path / "tuning.json"
    with GravityBatch([m[0] for m in ms], [m[1] for m in ms], [m[3] for m in ms],
                      bounded=[m[2] for m in ms]) as batch:
        for i, (_, _, _, b, mat) in enumerate(ms):
            batch.upload(i, b, mat)
            batch.setActive(i, len(b))
        batch.initializeAcceleration()
        first, timing = batch.autotune(steps=1, cache=cache)
        assert timing is not None and cache.exists()
        batch.tuning = PLAIN
        again, timing = batch.autotune(steps=1, cache=cache)
        assert timing is None and again == first and batch.tuning == first


def test_tuning_rejects_values_out_of_range():
    ms = crowded_members()[2:]
    with GravityBatch([m[0] for m in ms], [m[1] for m in ms], [m[3] for m in ms],
                      bounded=[m[2] for m in ms]) as batch:
        for bad in [Tuning(bigSegmentLimit=8), Tuning(forwardItemBodies=10),
                    Tuning(depositRuns=2), Tuning(scanDirectBlocks=0)]:
            with pytest.raises(ValueError):
                batch.tuning = bad
        assert batch.tuning == Tuning()
