"""Adjoint of the gravity step (gs_adjoint_*, gs_vjp_initial_velocity).

The reference has no derivatives; the gradient is checked against central
finite differences (h = 1e-3 px/s) of the oracle's FP64 gravity-only
leapfrog on a strongly self-gravitating field where gravity carries most of
the gradient, against the exact FP64 transposed Jacobian of the field
(oracle/accel64.py, itself checked against finite differences), and against
the closed form of free motion. The FP64 models are pinned to the FP32
oracle's forward steps."""
import numpy as np
import pytest

from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline
from gravity_amd import ParticleConfig, scenarios

W, H = 256, 128
GRID = (64, 32)
DT = 1.0 / 120.0


def params(oracle, G):
    return oracle.params(W, H, collide=False, substeps=1, grid=GRID, gravity=G, sort_interval=0)


def wave(G=10.0, seed=3):
    """One body per physical cell, displaced by a long density wave and
    jittered, 5 px/s velocities: a smooth, strongly self-gravitating field
    without close encounters, so the FP32 and FP64 trajectories stay within
    2e-4 px over 32 steps and the gradient is well conditioned."""
    rng = np.random.default_rng(seed)
    gx, gy = GRID
    X, Y = np.meshgrid((np.arange(gx) + 0.5) / gx, (np.arange(gy) + 0.5) / gy)
    X, Y = X.ravel(), Y.ravel()
    X = X + 4.0 * np.sin(2 * np.pi * Y) / gx + rng.uniform(-0.25, 0.25, X.size) / gx
    Y = Y + 4.0 * np.sin(2 * np.pi * X) / gy + rng.uniform(-0.25, 0.25, Y.size) / gy
    n = X.size
    bodies = np.zeros((n, 4), np.float32)
    bodies[:, 0], bodies[:, 1] = X % 1.0, Y % 1.0
    v = rng.normal(0.0, 5.0, (n, 2))
    bodies[:, 2], bodies[:, 3] = v[:, 0] / W, v[:, 1] / H
    mats = np.zeros((n, 4), np.float32)
    mats[:, 0] = 1.0
    return bodies, mats, scenarios.targets(n, 11)


def clump(n=1000, G=0.5):
    bodies, mats = scenarios.plummer(n, W, H, a_px=12.0, gravity=G, seed=3)
    return bodies, mats, scenarios.targets(n, 11)


def test_fp64_leapfrog_follows_the_fp32_oracle(oracle):
    G, steps = 0.5, 8
    bodies, mats, targets = clump(G=G)
    p = params(oracle, G)
    _, final = oracle.grav64_loss(p, bodies, mats, steps, targets, final=True)
    sim = oracle.Sim(p, bodies, mats)
    sim.initialize_acceleration()
    for _ in range(steps):
        sim.step()
    want = sim.state()[0][:, :2].astype(np.float64)
    sim.close()
    d = (final - want) * np.array([W, H])
    d -= np.round(d / np.array([W, H])) * np.array([W, H])
    assert np.abs(d).max() < 1e-3


def small_field():
    from oracle.accel64 import Accel64
    Ws, Hs, grid = 128, 64, (32, 16)
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 1.0, (120, 2))
    x[:40] = 0.5 + rng.normal(0.0, 0.03, (40, 2))  # a clump: monopoles and home caps
    x %= 1.0
    m = np.ones(len(x))
    m[::7] = 2.0
    return Accel64(Ws, Hs, grid, gravity=0.5), (Ws, Hs, grid), x, m


def test_fp64_field_matches_the_fp32_oracle_kick(oracle):
    model, (Ws, Hs, grid), x, m = small_field()
    bodies = np.zeros((len(x), 4), np.float32)
    bodies[:, :2] = x
    mats = np.zeros((len(x), 4), np.float32)
    mats[:, 0] = m
    sim = oracle.Sim(oracle.params(Ws, Hs, collide=False, substeps=1, grid=grid, gravity=0.5,
                                   sort_interval=0), bodies, mats)
    sim.initialize_acceleration()
    want = sim.state()[1].astype(np.float64)
    sim.close()
    got, _ = model.accel(bodies[:, :2].astype(np.float64), m)
    assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()


def test_fp64_transposed_jacobian_matches_finite_differences():
    model, _, x, m = small_field()
    cells = model.topology(x)
    w = np.random.default_rng(2).normal(0.0, 1.0, x.shape)
    g = model.vjp(x, m, w, cells)
    eps = 1e-6
    rng = np.random.default_rng(3)
    for i in rng.choice(len(x), 24, replace=False):
        for c in range(2):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += eps
            xm[i, c] -= eps
            fd = ((w * model.accel(xp, m, cells)[0]).sum() -
                  (w * model.accel(xm, m, cells)[0]).sum()) / (2 * eps)
            assert abs(g[i, c] - fd) <= 1e-6 * np.abs(g).max() + 1e-7 * abs(fd)


@pytest.mark.gpu
def test_acceleration_vjp_matches_the_fp64_model():
    model, (Ws, Hs, grid), x, m = small_field()
    n = len(x)
    bodies = np.zeros((n, 4), np.float32)
    bodies[:, :2] = x
    mats = np.zeros((n, 4), np.float32)
    mats[:, 0] = m
    w = np.random.default_rng(2).normal(0.0, 1.0, (n, 2)).astype(np.float32)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=grid[0], gridHeight=grid[1],
                                                   gravitationalConstant=0.5, minimumGrid=16),
                        collisions=CollisionConfig(maximumSubsteps=1), collisionsEnabled=False)
    with GravityPipeline(cfg, Ws, Hs, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        got = pipe.accelerationVjp(n, w)
    want = model.vjp(bodies[:, :2].astype(np.float64), m, w.astype(np.float64))
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[adjoint] J^T w max error {err:.2e} of the largest component", flush=True)
    assert err <= 1e-4


def _config(n, G):
    return GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                         gravity=GravityFieldConfig(gridWidth=GRID[0], gridHeight=GRID[1],
                                                    gravitationalConstant=G, minimumGrid=16),
                         collisions=CollisionConfig(maximumSubsteps=1), collisionsEnabled=False)


@pytest.mark.gpu
def test_vjp_matches_fp64_finite_differences(oracle):
    G, steps, every = 10.0, 32, 8
    bodies, mats, targets = wave(G=G)
    n = len(bodies)
    print(f"[adjoint] {n} bodies, {steps} steps, checkpoints every {every}", flush=True)
    with GravityPipeline(_config(n, G), W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        loss, grad = pipe.vjpInitialVelocity(n, steps, every, targets)
    p = params(oracle, G)
    ref, final = oracle.grav64_loss(p, bodies, mats, steps, targets, final=True)
    assert abs(loss - ref) <= 1e-4 * ref
    rng = np.random.default_rng(0)
    sample = rng.choice(n, 32, replace=False)
    h = 1e-3  # px/s
    got, fd, ballistic = [], [], []
    for j in sample:
        for c, S in ((0, W), (1, H)):
            dv = np.zeros((n, 2))
            dv[j, c] = h / S
            lp = oracle.grav64_loss(p, bodies, mats, steps, targets, dv)
            lm = oracle.grav64_loss(p, bodies, mats, steps, targets, -dv)
            fd.append((lp - lm) / (2.0 * h / S))
            got.append(grad[j, c])
            d = (final[j, c] - targets[j, c]) * S
            d -= np.floor(d / S + 0.5) * S
            ballistic.append(2.0 * d * S / n * steps * DT)
    got, fd, ballistic = map(np.array, (got, fd, ballistic))
    rel = np.linalg.norm(got - fd) / np.linalg.norm(fd)
    share = np.linalg.norm(fd - ballistic) / np.linalg.norm(fd)
    print(f"[adjoint] gradient error {rel:.2e} relative, gravity share {share:.2f}", flush=True)
    assert share > 0.5          # gravity carries most of the gradient here
    assert rel <= 1e-3


@pytest.mark.gpu
def test_vjp_of_free_motion_is_the_closed_form():
    # G = 0: the steps are pure drift, x_T = x_0 + T dt v_0 (mod 1), so the
    # gradient is (2/N) d * viewport * T dt, through every checkpoint
    # interval and every undone spatial sort.
    steps, every = 20, 6
    bodies, mats = scenarios.uniform(3000, W, H, seed=4)
    n = len(bodies)
    targets = scenarios.targets(n, 5)
    with GravityPipeline(_config(n, 0.0), W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        loss, grad = pipe.vjpInitialVelocity(n, steps, every, targets)
        final = pipe.bodies(n)
        where = pipe.adjointOrder(n)
    xT = final[where, :2].astype(np.float64)
    view = np.array([W, H], np.float64)
    d = (xT - targets) * view
    d -= np.floor(d / view + 0.5) * view
    want = 2.0 * d * view / n * steps * DT
    np.testing.assert_allclose(grad, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    assert abs(loss - (d ** 2).sum(axis=1).mean()) <= 1e-5 * loss
