"""P3M gravity on the MI355X against the oracle (FP32 tolerance: the FFT and
the spectral kernel are computed differently) and against closed forms
(T/CollisionShaderIntegrationTest.kt:263-375)."""
import numpy as np
import pytest

from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
from gravity_amd import scenarios

pytestmark = pytest.mark.gpu

# Maximum force error after one build, relative to the largest force.
FORCE_TOLERANCE = 1e-4


def make(count, W, H, grid, bounded=False, scenario=None):
    cfg = GravityConfig(particles=ParticleConfig(capacity=count, initialBodies=count, seed=1),
                        gravity=GravityFieldConfig(gridWidth=grid[0], gridHeight=grid[1]))
    return GravityPipeline(cfg, W, H, scenarioState=scenario, bounded=bounded)


def clustered(n, W, H, seed):
    rng = np.random.default_rng(seed)
    b, m = scenarios.uniform(n, W, H, seed=seed)
    k = n // 3
    # A dense clump: cells over the monopole threshold and home cells over 8.
    b[:k, 0] = (0.3 + rng.normal(0, 6.0, k) / W).astype(np.float32)
    b[:k, 1] = (0.6 + rng.normal(0, 6.0, k) / H).astype(np.float32)
    b[:, :2] = np.clip(b[:, :2], 0, np.float32(0.99999994))
    return b, m


@pytest.mark.parametrize("bounded", [False, True])
def test_initial_acceleration_matches_oracle(oracle, bounded):
    W, H, n = 1440, 720, 6000
    grid = (256, 128) if bounded else (512, 256)
    bodies, mats = clustered(n, W, H, 3)
    with make(n, W, H, grid, bounded, bodies) as pipe:
        pipe.initializeAcceleration(n)
        got = pipe.acceleration(n)
        mesh = pipe.forceMesh()
    sim = oracle.Sim(oracle.params(W, H, bounded=bounded, grid=grid), bodies, mats)
    sim.initialize_acceleration()
    want = sim.state()[1]
    fx, fy = sim.forces()
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    assert err < FORCE_TOLERANCE, err
    mscale = max(np.abs(fx).max(), np.abs(fy).max())
    if bounded:
        # the nodes the gather reads: rows 0..gh and the last, columns 0..gw
        # and the last (the padding is not produced)
        rows = list(range(grid[1] + 1)) + [2 * grid[1] - 1]
        cols = list(range(grid[0] + 1)) + [2 * grid[0] - 1]
        mesh, fx, fy = mesh[rows][:, cols], fx[rows][:, cols], fy[rows][:, cols]
    merr = max(np.abs(mesh[..., 0] - fx).max(), np.abs(mesh[..., 1] - fy).max()) / mscale
    assert merr < FORCE_TOLERANCE, merr


def test_wall_gravity_is_isolated_and_torus_acts_across_the_edge():
    W, H = 3440, 1440
    bodies = np.array([[20 / W, 720 / H, 0, 0], [(W - 20) / W, 720 / H, 0, 0]], np.float32)
    for bounded in (False, True):
        with make(2, W, H, (128, 128), bounded, bodies) as pipe:
            pipe.initializeAcceleration(2)
            a = pipe.acceleration(2)
        if bounded:
            assert a[0, 0] > 0
        else:
            assert a[0, 0] < 0
        assert abs(-a[0, 0] - a[1, 0]) < 5e-5


def test_switching_to_walls_and_back_restores_gravity_exactly():
    W, H, n = 3440, 1440, 256
    rng = np.random.default_rng(91)
    bodies = np.zeros((n, 4), np.float32)
    bodies[:, 0] = rng.random(n, dtype=np.float32)
    bodies[:, 1] = rng.random(n, dtype=np.float32)
    with make(n, W, H, (128, 128), False, bodies) as pipe:
        pipe.initializeAcceleration(n)
        before = pipe.acceleration(n)
        pipe.bounded = True
        pipe.initializeAcceleration(n)
        walls = pipe.acceleration(n)
        pipe.bounded = False
        pipe.initializeAcceleration(n)
        after = pipe.acceleration(n)
    np.testing.assert_array_equal(before.view(np.uint32), after.view(np.uint32))
    assert not np.array_equal(before, walls)


@pytest.mark.parametrize("bounded", [False, True])
def test_uniform_field_accelerates_a_free_body_at_its_strength(bounded):
    W, H = 3440, 1440
    body = np.array([[1720 / W, 300 / H, 0, 0]], np.float32)
    with make(1, W, H, (128, 128), bounded, body) as pipe:
        pipe.fieldAccelerationX = -120.0
        pipe.fieldAccelerationY = 300.0
        pipe.initializeAcceleration(1)
        for _ in range(60):
            pipe.step(1)
        b = pipe.bodies(1)
    elapsed = 60 / 120.0
    assert abs(b[0, 2] * W - (-120.0 * elapsed)) < 0.05
    assert abs(b[0, 3] * H - 300.0 * elapsed) < 0.05


def test_switching_off_mutual_gravity_leaves_only_the_field():
    W, H, n = 3440, 1440, 256
    rng = np.random.default_rng(17)
    bodies = np.zeros((n, 4), np.float32)
    bodies[:, 0] = rng.random(n, dtype=np.float32)
    bodies[:, 1] = rng.random(n, dtype=np.float32)
    with make(n, W, H, (128, 128), False, bodies) as pipe:
        pipe.fieldAccelerationY = -300.0
        pipe.initializeAcceleration(n)
        attracted = pipe.acceleration(n)
        pipe.mutualGravity = False
        pipe.upload(bodies)
        pipe.initializeAcceleration(n)
        free = pipe.acceleration(n)
        for _ in range(10):
            pipe.step(n)
        assert pipe.measureCluster(n).nonFiniteBodies == 0
    assert (free[:, 0] == 0).all()
    assert (free[:, 1] == np.float32(-300.0) / np.float32(H)).all()
    assert (np.abs(attracted[:, 0]) > 1e-3).any()


# Every line length of the specialised transforms (16..2048), along rows and
# along columns: (grid width, grid height) on the torus, where the mesh is the
# grid.
LINE_LENGTHS = [(16, 2048), (2048, 16), (32, 1024), (1024, 32), (64, 512), (512, 64), (128, 256),
                (256, 128)]


@pytest.mark.parametrize("grid", LINE_LENGTHS, ids=[f"{w}x{h}" for w, h in LINE_LENGTHS])
def test_every_line_length_transforms_like_fp64_and_the_oracle(oracle, grid):
    W, H, n = 1440, 720, 6000
    bodies, mats = clustered(n, W, H, 5)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=1),
                        gravity=GravityFieldConfig(gridWidth=grid[0], gridHeight=grid[1], minimumGrid=16))
    print(f"[gravity] grid {grid[0]}x{grid[1]}", flush=True)
    with GravityPipeline(cfg, W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        dens, spec, _ = pipe.densityMesh(n)
        pipe.initializeAcceleration(n)
        mesh = pipe.forceMesh()
    # the forward transforms against numpy's FP64 transform of the same deposit
    ref = np.fft.fft2(dens.astype(np.float64) / oracle.mass_scale(float(n)))
    rel = np.abs(spec - ref).max() / np.abs(ref).max()
    assert rel < 1.5e-7, rel
    # the whole force build (forward, Green multiply, inverse) against the oracle's
    sim = oracle.Sim(oracle.params(W, H, grid=grid), bodies, mats)
    sim.initialize_acceleration()
    fx, fy = sim.forces()
    scale = max(np.abs(fx).max(), np.abs(fy).max())
    merr = max(np.abs(mesh[..., 0] - fx).max(), np.abs(mesh[..., 1] - fy).max()) / scale
    assert merr < FORCE_TOLERANCE, merr
