"""Position-based liquid (K/LiquidSolver.kt, S/liquid_*.comp): the oracle
pinned by the properties the reference's own tests and kernel comments state
(rest density of the hexagonal packing, walls read as packing, momentum kept,
coincident bodies pushed apart), then the HIP passes bit for bit against it."""
import math

import numpy as np
import pytest

from gravity_amd import (CollisionConfig, GravityConfig, GravityFieldConfig, LiquidConfig,
                         ParticleConfig, StartConfig)
from gravity_amd import GravityPipeline, scenarios

D = 1.0  # contact diameter, px
DT = 1.0 / 120.0


def hex_lattice(x0, y0, cols, rows, spacing):
    """Hexagonal packing of `rows` x `cols` bodies (px), rows `spacing`*sqrt(3)/2 apart."""
    pts = []
    for r in range(rows):
        shift = 0.0 if r % 2 == 0 else 0.5 * spacing
        for c in range(cols):
            pts.append((x0 + c * spacing + shift, y0 + r * spacing * math.sqrt(3.0) * 0.5))
    return np.array(pts)


def pack(xy_px, W, H, bounded, vel_px=None):
    n = len(xy_px)
    b = np.zeros((n, 4), np.float32)
    b[:, 0] = np.float32(xy_px[:, 0]) / np.float32(W)
    b[:, 1] = np.float32(xy_px[:, 1]) / np.float32(H)
    if not bounded:
        b[:, :2] -= np.floor(b[:, :2])
    if vel_px is not None:
        b[:, 2] = np.float32(vel_px[:, 0]) / np.float32(W)
        b[:, 3] = np.float32(vel_px[:, 1]) / np.float32(H)
    return b


def fluid(n):
    m = np.zeros((n, 4), np.float32)
    m[:, 0] = 1.0
    return m


def liquid_params(oracle, W, H, bounded, **kw):
    return oracle.params(W, H, bounded=bounded, mutual_gravity=False, liquid=True, diameter=D,
                         grid=(64, 32), **kw)


def solve(oracle, p, bodies, mats):
    sim = oracle.Sim(p, bodies, mats)
    sim.resolve_liquid()
    b, _, press, _ = sim.state()
    sim.close()
    return b, press


# --------------------------------------------------------------------------
# Oracle pins (CPU)

def test_interior_of_the_hexagonal_packing_is_at_rest_density(oracle):
    # A torus tiled by the packing, its rows stretched by 0.07% to fit an
    # integer viewport (30 rows in 26 px): every body reads just below the
    # rest density, so no constraint acts and nothing moves at all.
    cols, rows, W, H = 40, 30, 40, 26
    xy = np.array([(c * D + (0.5 * D if r % 2 else 0.0), r * H / rows)
                   for r in range(rows) for c in range(cols)])
    b0 = pack(xy, W, H, bounded=False)
    b, press = solve(oracle, liquid_params(oracle, W, H, False), b0, fluid(len(xy)))
    np.testing.assert_array_equal(b, b0)
    assert np.all(press == 0.0)
    # 1% denser than the packing, the constraint acts everywhere
    xy2 = xy * 0.99
    b1 = pack(xy2, W, H, bounded=False)
    _, press2 = solve(oracle, liquid_params(oracle, W, H, False), b1, fluid(len(xy)))
    inner = (xy2[:, 0] > 4) & (xy2[:, 0] < 35) & (xy2[:, 1] > 4) & (xy2[:, 1] < 21)
    assert np.all(press2[inner] > 0.0)


def test_compressed_liquid_expands_and_keeps_momentum(oracle):
    W, H = 96, 64
    xy = hex_lattice(38.0, 22.0, 16, 16, 0.85 * D)
    b0 = pack(xy, W, H, bounded=False)
    m = fluid(len(xy))
    m[::3, 0] = 2.0  # unequal masses: momentum, not velocity, is kept
    b, press = solve(oracle, oracle.params(W, H, mutual_gravity=False, liquid=True, diameter=D,
                                           grid=(64, 32), viscosity=0.0), b0, m)
    centre = xy.mean(axis=0)
    r0 = np.hypot(xy[:, 0] - centre[0], xy[:, 1] - centre[1]).mean()
    xy1 = np.stack([b[:, 0] * W, b[:, 1] * H], axis=1)
    r1 = np.hypot(xy1[:, 0] - centre[0], xy1[:, 1] - centre[1]).mean()
    assert r1 > r0 + 0.05 * D
    p_vec = (m[:, :1] * (b[:, 2:] * np.array([W, H], np.float32))).sum(axis=0)
    p_abs = (m[:, 0] * np.hypot(b[:, 2] * W, b[:, 3] * H)).sum()
    assert np.abs(p_vec).max() < 1e-3 * p_abs
    assert press.max() > 0.0


def test_a_body_on_the_floor_reads_the_packing_through_the_wall(oracle):
    # Virtual rows continue the packing beyond a wall (S/liquid_kernel.glsl):
    # a packed layer resting on the floor feels no constraint there.
    W, H = 64, 48
    xy = hex_lattice(0.5 * D + 8.0, 0.5 * D, 40, 6, D)
    b0 = pack(xy, W, H, bounded=True)
    b, press = solve(oracle, liquid_params(oracle, W, H, True), b0, fluid(len(xy)))
    bottom = np.abs(xy[:, 1] - 0.5 * D) < 1e-6
    inner = bottom & (xy[:, 0] > 14.0) & (xy[:, 0] < 44.0)
    h = 2.0 * D
    constraint = press[inner] / (h / DT)
    assert constraint.max() < 0.02
    assert np.abs((b[inner, 1] - b0[inner, 1]) * H).max() < 0.02 * D


def test_coincident_bodies_are_pushed_apart_keeping_momentum(oracle):
    # Twelve bodies at one point read far above the rest density; each pair
    # has its fixed pseudo-random direction (S/liquid_kernel.glsl
    # fallbackDirection), so the clump spreads, and the pair terms cancel.
    W, H, k = 32, 32, 12
    xy = np.tile([[16.0, 16.0]], (k, 1))
    b0 = pack(xy, W, H, bounded=False)
    b, _ = solve(oracle, oracle.params(W, H, mutual_gravity=False, liquid=True, diameter=D,
                                       grid=(16, 16), viscosity=0.0), b0, fluid(k))
    p = np.stack([b[:, 0] * W, b[:, 1] * H], axis=1)
    spread = np.hypot(*(p - p.mean(axis=0)).T)
    assert spread.min() > 0.05 * D
    v = b[:, 2:] * np.array([W, H], np.float32)
    assert np.abs(v.sum(axis=0)).max() < 1e-3 * np.abs(v).sum()


def rest_density_fp64(diameter, radius_diameters):
    """rho_0 = sum of poly6 over the hexagonal packing, self included
    (K/LiquidKernel.kt:14-60), summed here in FP64 over a lattice large
    enough to cover the kernel."""
    h = diameter * radius_diameters
    poly6 = 4.0 / (math.pi * h ** 8)
    total = 0.0
    for r in range(-12, 13):
        shift = 0.0 if r % 2 == 0 else 0.5 * diameter
        for c in range(-12, 13):
            x, y = c * diameter + shift, r * diameter * math.sqrt(3.0) * 0.5
            q = h * h - (x * x + y * y)
            total += poly6 * q ** 3 if q > 0 else 0.0
    return total


@pytest.mark.parametrize("r_over_h", [0.3, 0.62, 0.95])
def test_xsph_viscosity_of_a_pair_is_the_closed_form(oracle, r_over_h):
    # S/liquid_viscosity.comp:27-55: two fluid bodies closer than h, too
    # sparse for any density correction (lambda = 0), exchange velocity by
    # v1' = v1 + c (v2 - v1) W(r) / rho_0 and symmetrically, so the pair's
    # momentum is kept and the relative velocity shrinks by 1 - 2 c W / rho_0.
    W = H = 200
    c = 0.1
    h = 2.0 * D
    r = r_over_h * h
    xy = np.array([[100.0, 100.0], [100.0 + r * 0.6, 100.0 + r * 0.8]])
    vel = np.array([[30.0, -10.0], [-20.0, 25.0]])
    b = pack(xy, W, H, False, vel)
    got, _ = solve(oracle, liquid_params(oracle, W, H, False, viscosity=c), b, fluid(2))
    np.testing.assert_array_equal(got[:, :2], b[:, :2])  # no density correction moved them
    # the separation the FP32 positions actually hold
    dx = (float(b[1, 0]) - float(b[0, 0])) * W
    dy = (float(b[1, 1]) - float(b[0, 1])) * H
    w = 4.0 / (math.pi * h ** 8) * (h * h - (dx * dx + dy * dy)) ** 3
    k = c * w / rest_density_fp64(D, 2.0)
    v = np.stack([got[:, 2] * W, got[:, 3] * H], axis=1).astype(np.float64)
    v0 = np.stack([b[:, 2] * np.float32(W), b[:, 3] * np.float32(H)], axis=1).astype(np.float64)
    want = v0 + k * (v0[::-1] - v0)
    change = np.abs(want - v0).max()
    assert change > 1e-3
    assert np.abs(v - want).max() < 2e-4 * change + 1e-5
    assert np.abs(v.sum(0) - v0.sum(0)).max() < 1e-4


def test_xsph_skips_grains_and_bodies_beyond_the_kernel(oracle):
    # A fluid body next to a solid grain and a fluid body just outside h keeps
    # its velocity; two fluid bodies inside h still exchange.
    W = H = 200
    h = 2.0 * D
    xy = np.array([[50.0, 50.0], [50.0 + 0.5 * h, 50.0], [120.0, 50.0], [120.0 + 1.01 * h, 50.0]])
    vel = np.array([[10.0, 0.0], [-10.0, 0.0], [5.0, 5.0], [-5.0, -5.0]])
    b = pack(xy, W, H, False, vel)
    m = fluid(4)
    m[1, 2] = 1.0   # a grain
    m[1, 3] = -1.0
    got, _ = solve(oracle, liquid_params(oracle, W, H, False, viscosity=0.1), b, m)
    np.testing.assert_array_equal(got[:, 2:], b[:, 2:])


# --------------------------------------------------------------------------
# HIP passes against the oracle (bit-exact)

def _pipe_config(n, bounded, iterations=12, viscosity=0.1, field=(0.0, 0.0), substeps=16):
    return GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                         gravity=GravityFieldConfig(gridWidth=64, gridHeight=32, minimumGrid=16),
                         collisions=CollisionConfig(contactDiameterPixels=D, maximumSubsteps=substeps),
                         liquid=LiquidConfig(iterations=iterations, viscosity=viscosity),
                         start=StartConfig(walls=bounded, mutualGravity=False, liquid=True))


def _scenes():
    rng = np.random.default_rng(11)
    out = []
    # liquid + sand in a box (the C3 scene, reduced)
    b, m = scenarios.liquid_and_sand(3000, 160, 90, diameter=1.0746, seed=3)
    out.append(("liquid_and_sand", 160, 90, True, b, m))
    # a compressed blob on the torus, mixed masses, some solids
    xy = hex_lattice(20.0, 14.0, 30, 20, 0.8 * D)
    m = fluid(len(xy))
    m[::5, 0] = 3.0
    m[::7, 2] = 1.0
    out.append(("torus_blob", 80, 48, False, pack(xy, 80, 48, False), m))
    # a clump of coincident bodies against a corner
    xy = np.tile([[0.6, 0.6]], (40, 1)) + rng.normal(0, 0.02, (40, 2))
    out.append(("corner_clump", 48, 32, True, pack(xy, 48, 32, True), fluid(40)))
    # a torus narrower than three liquid cells: wrapped cells met twice
    xy = rng.uniform(0, 1, (30, 2)) * np.array([5.0, 24.0])
    out.append(("narrow_torus", 5, 24, False, pack(xy, 5, 24, False), fluid(30)))
    # a band across the whole width of a torus: stencils on both seams, whose
    # rows wrap from the last liquid column to the first
    xy = rng.uniform(0, 1, (400, 2)) * np.array([40.0, 8.0]) + np.array([0.0, 11.0])
    out.append(("torus_seam", 40, 30, False, pack(xy, 40, 30, False), fluid(400)))
    # a box narrower than three liquid cells
    xy = rng.uniform(0, 1, (30, 2)) * np.array([5.0, 20.0]) + np.array([0.0, 2.0])
    out.append(("narrow_box", 5, 24, True, pack(xy, 5, 24, True), fluid(30)))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("scene", _scenes(), ids=lambda s: s[0])
def test_liquid_solve_matches_oracle_bit_for_bit(oracle, scene):
    name, W, H, bounded, bodies, mats = scene
    n = len(bodies)
    print(f"[liquid] scene {name}: {n} bodies", flush=True)
    with GravityPipeline(_pipe_config(n, bounded), W, H, scenarioState=bodies, bounded=bounded) as pipe:
        pipe.upload(bodies, mats)
        pipe.resolveLiquid(n)
        got = pipe.bodies(n)
        press = pipe.pressures(n)
    p = oracle.params(W, H, bounded=bounded, mutual_gravity=False, liquid=True, diameter=D,
                      grid=(64, 32))
    want, wpress = solve(oracle, p, bodies, mats)
    np.testing.assert_array_equal(got, want)
    fluid_bodies = mats[:, 2] == 0.0
    np.testing.assert_array_equal(press[fluid_bodies], wpress[fluid_bodies])


@pytest.mark.gpu
def test_liquid_and_sand_steps_match_oracle_bit_for_bit(oracle):
    # The C3 configuration, reduced: walls, a uniform downward field, no
    # mutual gravity, 16 substeps, liquid on; collisions, liquid and the kick
    # are all exact arithmetic here, so two whole steps agree bit for bit.
    W, H, n = 160, 90, 3000
    bodies, mats = scenarios.liquid_and_sand(n, W, H, diameter=1.0746, seed=3)
    cfg = _pipe_config(n, True)
    cfg = GravityConfig(particles=cfg.particles, gravity=cfg.gravity,
                        collisions=CollisionConfig(contactDiameterPixels=1.0746, maximumSubsteps=16),
                        liquid=cfg.liquid, start=cfg.start)
    with GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.fieldAccelerationY = -500.0
        pipe.initializeAcceleration(n)
        for _ in range(2):
            pipe.step(n)
        got = pipe.bodies(n)
        press = pipe.pressures(n)
    p = oracle.params(W, H, bounded=True, mutual_gravity=False, liquid=True, diameter=1.0746,
                      grid=(64, 32), field=(0.0, -500.0))
    sim = oracle.Sim(p, bodies, mats)
    sim.initialize_acceleration()
    for _ in range(2):
        sim.step()
    want, _, wpress, _ = sim.state()
    sim.close()
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(press, wpress)
