"""Pins of the oracle's rigid projection (oracle/oracle.c orc_rigid_*,
restating S/rigid_reduce.comp, S/rigid_apply.comp, K/RigidBlocks.kt:751-777
and K/GlueShape.kt:158-190) by what the mathematics fixes:

- the hull's corners are the vertices scipy's Qhull finds (collinear points
  dropped), counter-clockwise;
- the centroid and the inertia sum are the FP64 mean and sum of squares;
- a block moving rigidly is fitted to its own centre, rotation, velocity and
  spin; the spin is L / sum |r0|^2 with L taken on the previous pose's arms,
  which for a block turned by d since that pose is spin x cos d;
- the fit keeps momentum and angular momentum, and its rotation is the FP64
  least-squares rotation of the rest shape onto the members;
- a block across a torus seam is fitted as the same block away from it;
- walls hold the hull inside, and after a contact pass the block loses the
  speed it has into the wall (the push over a substep at most) and Coulomb
  friction along it, in closed form;
- guard and support requests (1/65536 px) move the centre by exactly that;
- the contact memory follows the reference's states, pair by pair.

The GPU tests (test_gpu_rigid.py) hold the HIP kernels bit for bit to this
oracle.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from blocks import KEYS, PARTNERS, Layout, hex_lattice, place, rotate
from oracle import oracle

W, H = 1920.0, 1080.0
SUB = 1.0 / 120.0 / 16
RADIUS = 0.56419


def _row_cloud(rng, n, rows):
    """Points on `rows` row heights 0.37 px apart (rows of one height, as
    member centres lie), random x."""
    ys = rng.integers(0, rows, n).astype(np.float32) * np.float32(0.37)
    xs = rng.uniform(-30, 30, n).astype(np.float32)
    return xs, ys


@pytest.mark.parametrize("seed", range(6))
def test_hull_corners_are_the_convex_hull_vertices(seed):
    rng = np.random.default_rng(seed)
    xs, ys = _row_cloud(rng, 40 + 60 * seed, 8 + 5 * seed)
    corners = oracle.rigid_hull(xs, ys)
    pts = np.stack([xs, ys], 1).astype(np.float64)
    expected = ConvexHull(pts).vertices
    got = {(float(xs[c]), float(ys[c])) for c in corners}
    want = {(float(xs[v]), float(ys[v])) for v in expected}
    assert got == want
    # counter-clockwise: every consecutive turn is to the left
    c = pts[corners]
    d1 = np.roll(c, -1, 0) - c
    d2 = np.roll(c, -2, 0) - np.roll(c, -1, 0)
    assert np.all(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0] > 0)


def test_hull_of_a_lattice_drops_collinear_points():
    xy = hex_lattice(300)
    corners = oracle.rigid_hull(xy[:, 0], xy[:, 1])
    assert set(map(tuple, xy[corners].tolist())) == set(map(tuple, xy[ConvexHull(xy.astype(np.float64)).vertices].tolist()))
    assert len(corners) <= 6 + 2   # a hexagonally packed rectangle: its corners (and row-shift notches)


def test_hull_of_few_points():
    assert list(oracle.rigid_hull(np.array([1.0], np.float32), np.array([2.0], np.float32))) == [0]
    two = oracle.rigid_hull(np.array([3.0, 1.0], np.float32), np.array([0.0, 0.0], np.float32))
    assert list(two) == [1, 0]


@pytest.mark.parametrize("n", [1, 7, 300, 5000])
def test_prepare_centroid_and_inertia_are_the_fp64_ones(n):
    rng = np.random.default_rng(n)
    offsets = (hex_lattice(n) + rng.uniform(-3, 3, 2)).astype(np.float32)
    rest, centroid, inertia, bounds = oracle.rigid_prepare(offsets)
    mean = offsets.astype(np.float64).mean(axis=0)
    assert np.all(centroid == mean.astype(np.float32))
    r64 = offsets.astype(np.float64) - centroid.astype(np.float64)
    assert inertia == pytest.approx(float((r64 ** 2).sum()), rel=1e-6, abs=1e-9)
    assert np.abs(rest.astype(np.float64).sum(axis=0)).max() <= 1e-4 * max(n, 1)
    assert bounds[0] == pytest.approx(rest[:, 0].min(), abs=1e-5)
    assert bounds[3] == pytest.approx(rest[:, 1].max(), abs=1e-5)


def _one_block(n, center, prev_theta, bounded, frame_offset=(1.5, -0.75)):
    lay = Layout(capacity=n + 10)
    offsets = hex_lattice(n)
    g = lay.glue(0, offsets, center[0] - frame_offset[0], center[1] - frame_offset[1])
    f = lay.floats(g)
    f[6:8] = (np.cos(prev_theta), np.sin(prev_theta))
    rest = lay.rest[: 2 * n].reshape(-1, 2)
    return lay, g, rest


@pytest.mark.parametrize("bounded", [False, True])
@pytest.mark.parametrize("dtheta", [0.0, 0.2])
def test_a_rigid_block_is_fitted_to_its_own_motion(bounded, dtheta):
    n, theta, velocity, spin = 700, 0.3, (30.0, -20.0), 0.7
    center = (1000.25, 500.5)
    lay, g, rest = _one_block(n, center, theta - dtheta, bounded)
    bodies = place(rest, center, theta, velocity, spin, W, H, bounded)
    lay.project(bodies, W, H, bounded, False, SUB, RADIUS)
    f = lay.floats(g)
    assert f[0] == pytest.approx(center[0], abs=2e-3) and f[1] == pytest.approx(center[1], abs=2e-3)
    assert f[6] == pytest.approx(np.cos(theta), abs=2e-6) and f[7] == pytest.approx(np.sin(theta), abs=2e-6)
    assert f[4] == pytest.approx(velocity[0], abs=2e-3) and f[5] == pytest.approx(velocity[1], abs=2e-3)
    # L on the previous pose's arms: spin x cos(dtheta) of the inertia sum
    assert f[8] == pytest.approx(spin * np.cos(dtheta), rel=2e-5)
    after = place(rest, (f[0], f[1]), np.arctan2(f[7], f[6]), (f[4], f[5]), f[8], W, H, bounded)
    assert np.abs((bodies[:, :2] - after[:, :2]) * [W, H]).max() < 3e-3
    assert np.abs(bodies[:, 2:] * [W, H] - after[:, 2:] * [W, H]).max() < 1e-3


@pytest.mark.parametrize("bounded", [False, True])
def test_the_fit_keeps_momentum_and_angular_momentum(bounded):
    rng = np.random.default_rng(11)
    n, theta_prev = 900, 0.1
    center = (700.0, 400.0)
    lay, g, rest = _one_block(n, center, theta_prev, bounded)
    bodies = place(rest, center, 0.15, (12.0, 4.0), 0.3, W, H, bounded)
    bodies[:, 0] += rng.uniform(-0.2, 0.2, n).astype(np.float32) / np.float32(W)
    bodies[:, 1] += rng.uniform(-0.2, 0.2, n).astype(np.float32) / np.float32(H)
    bodies[:, 2] += rng.uniform(-5, 5, n).astype(np.float32) / np.float32(W)
    bodies[:, 3] += rng.uniform(-5, 5, n).astype(np.float32) / np.float32(H)
    v0 = bodies[:, 2:].astype(np.float64) * [W, H]
    arm_prev = rotate(theta_prev, rest)
    L0 = float((arm_prev[:, 0] * v0[:, 1] - arm_prev[:, 1] * v0[:, 0]).sum())
    inertia = float((rest.astype(np.float64) ** 2).sum())
    lay.project(bodies, W, H, bounded, False, SUB, RADIUS)
    f = lay.floats(g)
    v1 = bodies[:, 2:].astype(np.float64) * [W, H]
    scale = np.abs(v0).sum()
    assert np.abs(v1.sum(axis=0) - v0.sum(axis=0)).max() <= 2e-6 * scale
    assert f[8] * inertia == pytest.approx(L0, rel=2e-5)
    arm_new = rotate(np.arctan2(f[7], f[6]), rest)
    L1 = float((arm_new[:, 0] * v1[:, 1] - arm_new[:, 1] * v1[:, 0]).sum())
    assert L1 == pytest.approx(L0, rel=5e-5)


def test_the_fitted_rotation_is_the_least_squares_one():
    rng = np.random.default_rng(5)
    n = 500
    center = (900.0, 300.0)
    lay, g, rest = _one_block(n, center, 0.0, True)
    bodies = place(rest, center, -0.4, (0.0, 0.0), 0.0, W, H, True)
    bodies[:, :2] += rng.normal(0, 0.3, (n, 2)).astype(np.float32) / np.float32([W, H])
    p = bodies[:, :2].astype(np.float64) * [W, H]
    q = p - p.mean(axis=0)
    r = rest.astype(np.float64)

    def cost(t):
        return float(((q - rotate(t, r)) ** 2).sum())

    grid = np.linspace(-np.pi, np.pi, 20001)
    best = grid[np.argmin([cost(t) for t in grid])]
    for step in (1e-3, 1e-5, 1e-7):
        ts = best + np.linspace(-10 * step, 10 * step, 41)
        best = ts[np.argmin([cost(t) for t in ts])]
    lay.project(bodies, W, H, True, False, SUB, RADIUS)
    f = lay.floats(g)
    assert np.arctan2(f[7], f[6]) == pytest.approx(best, abs=2e-5)


def test_a_block_across_the_torus_seam_is_fitted_as_away_from_it():
    n = 600
    results = []
    for cx in (3.0, 3.0 + W / 2):
        lay, g, rest = _one_block(n, (cx, 540.0), 0.05, False)
        bodies = place(rest, (cx, 540.0), 0.1, (25.0, 5.0), -0.4, W, H, False)
        lay.project(bodies, W, H, False, False, SUB, RADIUS)
        results.append(lay.floats(g).copy())
    a, b = results
    assert (b[0] - a[0]) % W == pytest.approx(W / 2, abs=2e-3)
    assert b[1] == pytest.approx(a[1], abs=2e-3)
    assert np.allclose(a[4:9], b[4:9], rtol=0, atol=2e-4)


def test_walls_hold_the_hull_and_stop_the_block_in_closed_form():
    n = 400
    reach = float(oracle.rigid_prepare(hex_lattice(n))[0][:, 0].max())
    cx = W - reach - RADIUS + 3.0          # three px into the right wall
    for contacts in (False, True):
        lay, g, rest = _one_block(n, (cx, 500.0), 0.0, True, frame_offset=(0.0, 0.0))
        bodies = place(rest, (cx, 500.0), 0.0, (50.0, 40.0), 0.0, W, H, True)
        lay.project(bodies, W, H, True, contacts, SUB, RADIUS, block_friction=0.5)
        f = lay.floats(g)
        hull = lay.shapes[: 2 * lay.ints(g)[15]].reshape(-1, 2)
        assert f[0] + hull[:, 0].max() + RADIUS <= W + 1e-3
        assert f[0] + hull[:, 0].max() + RADIUS == pytest.approx(W, abs=1e-3)
        if contacts:
            # stop = min(50, 3 px / substep) = 50: vx 0; friction 0.5 x 50 takes 25 of vy 40
            assert f[4] == pytest.approx(0.0, abs=1e-3) and f[5] == pytest.approx(15.0, abs=1e-3)
        else:
            assert f[4] == pytest.approx(50.0, abs=1e-3) and f[5] == pytest.approx(40.0, abs=1e-3)
    # friction stops the sliding at most: vy 10 under a 25 friction budget ends at rest
    lay, g, rest = _one_block(n, (cx, 500.0), 0.0, True, frame_offset=(0.0, 0.0))
    bodies = place(rest, (cx, 500.0), 0.0, (50.0, 10.0), 0.0, W, H, True)
    lay.project(bodies, W, H, True, True, SUB, RADIUS, block_friction=0.5)
    assert lay.floats(g)[5] == pytest.approx(0.0, abs=1e-3)
    # a block wider than the screen is not moved along that axis
    narrow = 2 * reach - 4.0
    lay, g, rest = _one_block(n, (narrow / 2 + 1.0, 500.0), 0.0, True, frame_offset=(0.0, 0.0))
    bodies = place(rest, (narrow / 2 + 1.0, 500.0), 0.0, (0.0, 0.0), 0.0, narrow, H, True)
    lay.project(bodies, narrow, H, True, True, SUB, RADIUS)
    assert lay.floats(g)[0] == pytest.approx(narrow / 2 + 1.0, abs=2e-3)
    # a push larger than the substep allows: the stop is the push over a substep
    lay, g, rest = _one_block(n, (cx, 500.0), 0.0, True, frame_offset=(0.0, 0.0))
    bodies = place(rest, (cx, 500.0), 0.0, (9000.0, 0.0), 0.0, W, H, True)
    lay.project(bodies, W, H, True, True, SUB, RADIUS, block_friction=0.5)
    assert lay.floats(g)[4] == pytest.approx(9000.0 - 3.0 / SUB, rel=2e-4)


def test_guard_and_support_requests_move_the_centre_exactly():
    n = 300
    lay, g, rest = _one_block(n, (800.0, 500.0), 0.0, True, frame_offset=(0.0, 0.0))
    bodies = place(rest, (800.0, 500.0), 0.0, (10.0, -30.0), 0.0, W, H, True)
    free = bodies.copy()
    ref = Layout(capacity=n + 10)
    ref.glue(0, hex_lattice(n), 800.0, 500.0)
    ref.project(free, W, H, True, True, SUB, RADIUS)
    i = lay.ints(g)
    i[28:32] = (2 * 65536, 0, 0, 65536)     # guard: +2 x, -1 y
    i[32:36] = (0, 3 * 65536, 0, 0)         # support: +3 y
    lay.project(bodies, W, H, True, True, SUB, RADIUS)
    f, fr = lay.floats(g), ref.floats(0)
    assert f[0] - fr[0] == pytest.approx(2.0, abs=1e-4)
    assert f[1] - fr[1] == pytest.approx(2.0, abs=1e-4)
    assert np.all(i[28:36] == 0)
    # the support stops the speed against it (vy -30 into a 3 px push), no friction
    assert f[5] == pytest.approx(0.0, abs=1e-4) and f[4] == pytest.approx(fr[4], abs=1e-5)


def test_contact_memory_settles_pair_by_pair():
    lay = Layout(capacity=2000)
    for k in range(3):
        lay.glue(400 * k, hex_lattice(300), 300.0 + 400 * k, 500.0)
    bodies = np.zeros((1200, 4), np.float32)
    imp, nxt, st, idle, pairs = (lay.memory[j * KEYS:(j + 1) * KEYS] for j in range(5))

    def key(a, b):
        return min(a, b) * PARTNERS + max(a, b)

    wall = 256 + 1
    loose = 256 + 4
    cases = {
        # (group, partner): (impact, next, state, idle) -> (impact, idle)
        (0, 1): ((100 | 3 << 24), 150, 2, 2, (150 | 4 << 24), 0),     # pushing: faster, one pass older
        (0, 2): ((100 | 32 << 24), 50, 2, 0, 0, 0),                   # pushed 32 passes: over
        (0, 0): ((100 | 31 << 24), 20, 2, 0, (100 | 32 << 24), 0),    # the 32nd pass still pushes
        (1, 2): (0, 0, 2, 0, 0, 0),                                   # pushing with no speed: over
        (0, wall): ((80 | 1 << 24), 0, 1, 3, 0, 0),                   # touching without pushing: over
        (1, wall): ((80 | 1 << 24), 0, 0, 2, (80 | 1 << 24), 3),      # apart, third pass: kept
        (2, wall): ((80 | 1 << 24), 0, 0, 3, 0, 0),                   # apart, fourth pass: over
        (2, loose): (0, 0, 0, 5, 0, 5),                               # no impact: untouched
    }
    for (a, b), (i0, n0, s0, d0, _, _) in cases.items():
        k = key(a, b)
        imp[k], nxt[k], st[k], idle[k], pairs[k] = i0, n0, s0, d0, 7
    # pairs owned by no block in use (group 3), or with one: untouched
    k3 = key(3, wall)
    imp[k3], st[k3], pairs[k3] = 9, 2, 4
    k13 = key(1, 3)
    imp[k13], st[k13], pairs[k13] = 9, 2, 4
    lay.project(bodies, W, H, True, True, SUB, RADIUS)
    for (a, b), (_, _, _, _, i1, d1) in cases.items():
        k = key(a, b)
        assert (imp[k], idle[k], nxt[k], st[k], pairs[k]) == (i1, d1, 0, 0, 0), (a, b)
    assert (imp[k3], st[k3], pairs[k3]) == (9, 2, 4)
    assert (imp[k13], st[k13], pairs[k13]) == (9, 2, 4)
    # a projection after no contact pass leaves the memory alone
    imp[key(0, 1)], st[key(0, 1)] = 5, 2
    lay.project(bodies, W, H, True, False, SUB, RADIUS)
    assert (imp[key(0, 1)], st[key(0, 1)]) == (5, 2)


# ---- RectangularBlock (the reference's T/RectangularBlockTest.kt cases) ----
def _block_offsets(width, height, d=np.float32(1.12838)):
    from gravity_amd.pipeline import block_offsets
    return block_offsets(d, width, height)


def test_block_bodies_lie_inside_the_rectangle_without_overlapping():
    d = np.float32(1.12838)
    w, h = 40.0, 25.0
    xy = _block_offsets(w, h).astype(np.float64)
    r = 0.5 * float(d)
    assert len(xy) > 0
    assert np.all(xy[:, 0] - r >= -0.5 * w - 1e-3) and np.all(xy[:, 0] + r <= 0.5 * w + 1e-3)
    assert np.all(xy[:, 1] - r >= -0.5 * h - 1e-3) and np.all(xy[:, 1] + r <= 0.5 * h + 1e-3)
    diff = xy[:, None, :] - xy[None, :, :]
    dist = np.sqrt((diff ** 2).sum(-1)) + np.eye(len(xy)) * 1e9
    assert dist.min() >= float(d) - 1e-3


def test_block_is_hexagonally_dense():
    d = 1.12838
    count = len(_block_offsets(100.0, 60.0))
    hexagonal = 100.0 * 60.0 / (d * d * 0.8660254)
    assert 0.9 * hexagonal < count <= hexagonal


def test_rectangle_narrower_than_a_body_holds_nothing():
    d = np.float32(1.12838)
    assert len(_block_offsets(0.5, 100.0)) == 0
    assert len(_block_offsets(100.0, 0.5)) == 0
    assert len(_block_offsets(d, d)) == 1


def test_block_rows_are_the_hexagonal_lattice():
    d = np.float32(1.12838)
    xy = _block_offsets(10.0, 6.0)
    ys = np.unique(xy[:, 1])
    step = np.float32(d * np.float32(np.sqrt(3.0) * 0.5))
    assert np.allclose(np.diff(ys), step, atol=1e-5)
    for k, y in enumerate(ys):
        row = np.sort(xy[xy[:, 1] == y, 0])
        assert np.allclose(np.diff(row), d, atol=1e-5)
        # odd rows start half a diameter in
        assert row[0] == pytest.approx(-0.5 * (10.0 - d) + (0.5 * d if k % 2 else 0.0), abs=1e-5)


@pytest.mark.parametrize("w,h", [(40.0, 25.0), (10.0, 6.0), (1.7, 3.1), (200.0, 55.0), (1.12838, 1.12838)])
def test_block_offsets_are_the_restated_lattice(w, h):
    """RectangularBlock.Lattice and forEachOffset (K/RectangularBlock.kt:22-77)
    restated in float32 scalars, offset for offset."""
    f = np.float32
    d = f(1.12838)
    eps = f(1e-4)
    half_x, half_y = f(f(0.5) * f(f(w) - d)), f(f(0.5) * f(f(h) - d))
    row_step = f(d * f(np.sqrt(3.0) * 0.5))
    rows = int(np.floor(f(f(f(2) * half_y) / row_step) + eps)) + 1
    even = int(np.floor(f(f(f(2) * half_x) / d) + eps)) + 1
    odd_span = f(f(f(2) * half_x) - f(f(0.5) * d))
    odd = 0 if odd_span < -eps * d else int(np.floor(f(odd_span / d) + eps)) + 1
    want = []
    for row in range(rows):
        columns = odd if row & 1 else even
        start = f(-half_x + (f(f(0.5) * d) if row & 1 else f(0)))
        y = f(-half_y + f(f(row) * row_step))
        want += [(f(start + f(f(c) * d)), y) for c in range(columns)]
    got = _block_offsets(w, h)
    assert np.array_equal(got, np.asarray(want, np.float32).reshape(-1, 2))
