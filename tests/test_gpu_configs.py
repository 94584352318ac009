"""The five benchmark workloads at their stated sizes on the MI355X, against
the oracle (and, for the adjoint, FP64 finite differences committed under
tests/golden/):

- P3M only, 50,000 bodies on a 1440x720 torus, grid 512x256, 100 steps:
  positions within 1e-3 px of the oracle's; the density mesh and cell table
  bit for bit, the density spectrum within 6.4e-8 of an FP64 transform;
- the galaxy, 400,000 bodies in walls, grid 1024x512 (2048x1024 mesh):
  the same density, cell and spectrum checks on the doubled mesh, and a
  whole step (sort, 17 contact passes, force, kick) against the oracle: the
  sort's permutation bit for bit, and all but a few bodies' positions bit
  for bit (the kick runs on the FP32-tolerance path, and where its last bit
  decides a contact the few bodies around it part by about a thousandth of
  a pixel);
- liquid and sand, 400,000 bodies: two whole steps bit for bit;
- the 64-simulation ensemble as one batch: every member finite and inside
  its torus, the four smallest members within 1e-3 px of the oracle;
- the adjoint, 100,000 bodies, 64 steps, checkpoint every 8: the VJP
  within 1e-3 of the FP64 central differences on 64 components.
"""
import os

import numpy as np
import pytest

from gravity_amd import (CollisionConfig, GravityBatch, GravityConfig, GravityFieldConfig,
                         GravityPipeline, LiquidConfig, ParticleConfig, StartConfig, scenarios)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(f"[configs] {msg}", flush=True)


def torus_gap_px(got, want, W, H):
    dx = (got[:, 0].astype(np.float64) - want[:, 0]) * W
    dy = (got[:, 1].astype(np.float64) - want[:, 1]) * H
    dx -= np.round(dx / W) * W
    dy -= np.round(dy / H) * H
    return np.hypot(dx, dy)


def p3m_only(n):
    return GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=1),
                         gravity=GravityFieldConfig(gridWidth=512, gridHeight=256),
                         collisions=CollisionConfig(maximumSubsteps=1), collisionsEnabled=False)


def galaxy(n):
    return GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=7),
                         gravity=GravityFieldConfig(gridWidth=1024, gridHeight=512),
                         collisions=CollisionConfig(contactDiameterPixels=1.0, friction=0.2, sleep=False))


# Spectrum bounds (largest error over the largest spectrum value): at
# 512x256 the reference's own measurement, 6.35e-8 (docs/performance-fft.md:49);
# at 2048x1024 the reference's FP32 radix-2 algorithm (the oracle's) reaches
# 5.5e-7 on this deposit, and the bound here is a quarter of that.
SPECTRUM_BOUND = {"p3m-50k-torus": 6.4e-8, "galaxy-400k-walls": 1.4e-7}


@pytest.mark.parametrize("case", ["p3m-50k-torus", "galaxy-400k-walls"])
def test_density_mesh_cells_and_spectrum_at_full_size(oracle, case):
    if case == "p3m-50k-torus":
        W, H, n, grid, bounded = 1440, 720, 50_000, (512, 256), False
        bodies, mats = scenarios.uniform(n, W, H, seed=1)
        cfg = p3m_only(n)
    else:
        W, H, n, grid, bounded = 2880, 1440, 400_000, (1024, 512), True
        bodies, mats = scenarios.exponential_disc(n, W, H, seed=7)
        cfg = galaxy(n)
    say(f"{case}: density readback")
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=bounded) as pipe:
        pipe.upload(bodies, mats)
        dens, spec, cells = pipe.densityMesh(n)
        check = pipe.validateFft(n)
    pw, ph = grid
    mw, mh = (2 * pw, 2 * ph) if bounded else (pw, ph)
    scale = oracle.mass_scale(float(n))
    want, count, mass = oracle.deposit(bodies, mats, mw, mh, pw, ph, scale)
    np.testing.assert_array_equal(dens, want)
    np.testing.assert_array_equal(cells[..., 0], count)
    np.testing.assert_array_equal(cells[..., 1], mass)  # in units of 1/256
    ref = np.fft.fft2(want.astype(np.float64) / scale)
    rel = np.abs(spec - ref).max() / np.abs(ref).max()
    say(f"{case}: spectrum max error {rel:.3e} of the largest value, rms "
        f"{np.sqrt((np.abs(spec - ref) ** 2).mean()):.3e}")
    assert rel < SPECTRUM_BOUND[case]
    # validateFft's measure: real and imaginary parts taken apart
    err = np.abs(np.stack([spec.real - ref.real, spec.imag - ref.imag]))
    part = err.max() / np.abs(np.stack([ref.real, ref.imag])).max()
    assert abs(check["maximumRelativeError"] - part) <= 1e-6 * part


def test_p3m_only_hundred_steps_follow_the_oracle(oracle):
    W, H, n, steps = 1440, 720, 50_000, 100
    bodies, mats = scenarios.uniform(n, W, H, seed=1)
    say(f"P3M only: {n} bodies, {steps} steps")
    with GravityPipeline(p3m_only(n), W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        for _ in range(steps):
            pipe.step(n)
        got = pipe.bodies(n)
        dest = pipe.sortDestinations(n)
    sim = oracle.Sim(oracle.params(W, H, collide=False, substeps=1, grid=(512, 256)), bodies, mats)
    sim.initialize_acceleration()
    for _ in range(steps):
        sim.step()
    want = sim.state()[0]
    np.testing.assert_array_equal(dest, sim.last_sort_dest())
    gap = torus_gap_px(got, want, W, H)
    say(f"P3M only: max drift from the oracle {gap.max():.2e} px after {steps} steps")
    assert gap.max() <= 1e-3


def test_galaxy_full_size_step_follows_the_oracle(oracle):
    W, H, n = 2880, 1440, 400_000
    bodies, mats = scenarios.exponential_disc(n, W, H, scale_px=200.0, dispersion=0.02, friction=0.2,
                                              seed=7)
    say(f"galaxy: {n} bodies, one step")
    with GravityPipeline(galaxy(n), W, H, scenarioState=bodies, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        pipe.step(n)
        got = pipe.bodies(n)
        dest = pipe.sortDestinations(n)
    sim = oracle.Sim(oracle.params(W, H, bounded=True, grid=(1024, 512), diameter=1.0), bodies, mats)
    sim.initialize_acceleration()
    sim.step()
    want = sim.state()[0]
    np.testing.assert_array_equal(dest, sim.last_sort_dest())
    gap = np.hypot((got[:, 0].astype(np.float64) - want[:, 0]) * W,
                   (got[:, 1].astype(np.float64) - want[:, 1]) * H)
    apart = np.count_nonzero(gap)
    say(f"galaxy: max drift from the oracle {gap.max():.2e} px, {apart} bodies apart")
    assert apart <= n // 4000, apart
    assert gap.max() <= 1e-2


def test_liquid_and_sand_full_size_steps_are_bit_exact(oracle):
    n, W, H, d = 400_000, 2880, 1440, 1.0746
    bodies, mats = scenarios.liquid_and_sand(n, W, H, diameter=d, seed=3)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=3),
                        gravity=GravityFieldConfig(gridWidth=1024, gridHeight=512),
                        collisions=CollisionConfig(contactDiameterPixels=d, maximumSubsteps=16),
                        liquid=LiquidConfig(iterations=12, viscosity=0.1),
                        start=StartConfig(walls=True, mutualGravity=False, liquid=True))
    say(f"liquid and sand: {n} bodies, 2 steps")
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.fieldAccelerationY = -500.0
        pipe.initializeAcceleration(n)
        pipe.step(n)
        pipe.step(n)
        got, press = pipe.bodies(n), pipe.pressures(n)
    p = oracle.params(W, H, bounded=True, mutual_gravity=False, liquid=True, diameter=d,
                      grid=(1024, 512), field=(0.0, -500.0), sleep=True)
    sim = oracle.Sim(p, bodies, mats)
    sim.initialize_acceleration()
    sim.step()
    sim.step()
    want, _, wpress, _ = sim.state()
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(press.view(np.uint32), wpress.view(np.uint32))


def test_ensemble_of_64_as_one_batch(oracle):
    import bench
    members, cfgs = bench.ensemble_configs()
    W, H = bench.ENSEMBLE_VIEWPORT
    say(f"ensemble: {len(members)} simulations, {sum(m[0] for m in members)} bodies")
    order = np.argsort([m[0] for m in members])
    checked = [int(k) for k in order[:4]]
    with GravityBatch(cfgs, [(W, H)] * len(cfgs), [m[2] for m in members],
                      bounded=[False] * len(cfgs)) as batch:
        for k, (n, _, b, mat) in enumerate(members):
            batch.upload(k, b, mat)
            batch.setActive(k, n)
        batch.initializeAcceleration()
        batch.step()
        first = {k: batch.bodies(k, members[k][0]) for k in checked}
        for _ in range(2):
            batch.step()
        final = [batch.bodies(k, m[0]) for k, m in enumerate(members)]
    for k, b in enumerate(final):
        assert np.isfinite(b).all(), k
        assert ((b[:, :2] >= 0) & (b[:, :2] < 1)).all(), k
    for k in checked:
        n, grid, b, mat = members[k]
        p = oracle.params(W, H, grid=(grid, grid), diameter=1.12838, sleep=False)
        sim = oracle.Sim(p, b, mat)
        sim.initialize_acceleration()
        sim.step()
        gap = torus_gap_px(first[k], sim.state()[0], W, H)
        say(f"ensemble member {k}: {n} bodies, grid {grid}, max gap {gap.max():.2e} px")
        assert gap.max() <= 1e-3, (k, gap.max())


def test_adjoint_100k_matches_fp64_finite_differences():
    g = np.load(os.path.join(GOLDEN, "adjoint_fd_100k.npz"))
    n, W, H, steps, every = 100_000, 1440, 720, int(g["steps"]), 8
    bodies, mats = scenarios.uniform(n, W, H, seed=1)
    targets = scenarios.targets(n, 12)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n, seed=1),
                        gravity=GravityFieldConfig(gridWidth=512, gridHeight=256),
                        collisions=CollisionConfig(maximumSubsteps=1), collisionsEnabled=False)
    say(f"adjoint: {n} bodies, {steps} steps, checkpoint every {every}")
    with GravityPipeline(cfg, W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        loss, grad = pipe.vjpInitialVelocity(n, steps, every, targets)
        final = pipe.bodies(n)
        where = pipe.adjointOrder(n)
    idx, axes = g["bodies"], g["axes"]
    got = grad[idx, axes].astype(np.float64)
    fd, ballistic = g["fd"], g["ballistic"]
    rel = np.linalg.norm(got - fd) / np.linalg.norm(fd)
    gravity = np.linalg.norm((got - ballistic) - (fd - ballistic)) / np.linalg.norm(fd - ballistic)
    xT = final[where[idx], :2].astype(np.float64)
    parted = torus_gap_px(xT, g["final"], W, H)
    say(f"adjoint: gradient error {rel:.2e} relative ({len(fd)} components), gravity part "
        f"{gravity:.2e} of its own size (share {np.linalg.norm(fd - ballistic) / np.linalg.norm(fd):.3f}); "
        f"FP32 and FP64 final positions of the sampled bodies {parted.max():.2e} px apart; "
        f"loss {loss:.4f} vs {float(g['loss']):.4f}")
    assert abs(loss - float(g["loss"])) <= 1e-4 * float(g["loss"])
    assert rel <= 1e-3
