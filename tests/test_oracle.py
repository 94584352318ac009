"""The oracle pinned on the CPU: against the known answers of the reference's
own tests, against closed forms and invariants the mathematics fixes, and
against independent FP64 computations (numpy), each chosen so that a wrong
index, sign or dropped term fails."""
import math

import numpy as np
import pytest

from tests.scenes import DIAMETER, HEIGHT, WIDTH, brute_force_contacts, collision_scenes, scene, sort_scene

f32 = np.float32


# ---- a: sort key -----------------------------------------------------------

def morton_by_bits(x, y, wb, hb):
    """T/SpatialSortIntegrationTest.kt:177-186: bit-by-bit interleave, a
    different formulation from the shader's magic-number spread."""
    tx = min(max(int(math.floor(f32(x) * f32(1 << wb))), 0), (1 << wb) - 1)
    ty = min(max(int(math.floor(f32(y) * f32(1 << hb))), 0), (1 << hb) - 1)
    common = min(wb, hb)
    key = 0
    for bit in range(common):
        key |= ((tx >> bit) & 1) << (2 * bit) | ((ty >> bit) & 1) << (2 * bit + 1)
    return key | (((tx | ty) >> common) << (2 * common))


def test_sort_key_matches_bitwise_interleave(oracle):
    rng = np.random.default_rng(1)
    for _ in range(5000):
        x, y = rng.random(2, dtype=np.float32)
        wb, hb = (int(v) for v in rng.integers(1, 11, 2))
        assert oracle.sort_key(x, y, wb, hb) == morton_by_bits(x, y, wb, hb)
    # edges: 0, the largest float below 1, exactly 1, out of range
    for x, y in [(0, 0), (0.99999994, 0.99999994), (1.0, 1.0), (-0.5, 2.0)]:
        assert oracle.sort_key(x, y, 9, 8) == morton_by_bits(x, y, 9, 8)


def test_sort_axis_bits_known_answers(oracle):
    # 3440 x 1440 -> 9 / 8 bits (4 px tiles), capped at 10 bits per axis
    assert oracle.sort_axis_bits(3440) == 9
    assert oracle.sort_axis_bits(1440) == 8
    assert oracle.sort_axis_bits(2880) == 9
    assert oracle.sort_axis_bits(100000) == 10
    assert oracle.sort_axis_bits(3) == 0


def test_spatial_sort_is_a_stable_key_ordered_permutation(oracle):
    bodies = sort_scene(6000)
    n = len(bodies)
    accel = np.random.default_rng(2).random((n, 2), dtype=np.float32)
    mats = oracle.default_materials(n)
    mats[:, 3] = np.arange(n)
    sb, sa, sm, dest = oracle.spatial_sort(bodies, accel, mats, WIDTH, HEIGHT)
    assert np.array_equal(np.sort(dest), np.arange(n))
    np.testing.assert_array_equal(sb[dest].view(np.uint32), bodies.view(np.uint32))
    np.testing.assert_array_equal(sa[dest].view(np.uint32), accel.view(np.uint32))
    wb, hb = oracle.sort_axis_bits(WIDTH), oracle.sort_axis_bits(HEIGHT)
    keys = np.array([morton_by_bits(x, y, wb, hb) for x, y in sb[:, :2]])
    assert (np.diff(keys) >= 0).all()
    # stable: inside a tile the original indices ascend
    orig = sm[:, 3].astype(np.int64)
    same = keys[1:] == keys[:-1]
    assert (orig[1:][same] > orig[:-1][same]).all()
    assert len(set(keys.tolist())) > 1000


# ---- c: contacts and the contact pass -------------------------------------

@pytest.mark.parametrize("name", list(collision_scenes()))
def test_contact_sets_equal_brute_force(oracle, name):
    bodies = collision_scenes()[name]
    p = oracle.params(WIDTH, HEIGHT)
    counts, lists = oracle.contacts(p, bodies, cap=160)
    expected = brute_force_contacts(bodies)
    for i, e in enumerate(expected):
        assert counts[i] == len(e)
        np.testing.assert_array_equal(lists[i, :len(e)], e)


def test_contact_known_answers(oracle):
    # T/CpuCollisionReferenceTest.kt:28-60
    p = oracle.params(800, 600)
    c, _ = oracle.contacts(p, scene([(0.25 * 800, 0.75 * 600, 0, 0)] * 12, 800, 600))
    assert c.sum() == 2 * 66
    p = oracle.params(1000, 1000)
    b = np.array([[0.0001, 0.5, 0, 0], [0.9999, 0.5, 0, 0], [0.5, 0.0001, 0, 0], [0.5, 0.9999, 0, 0],
                  [0.25, 0.25, 0, 0]], np.float32)
    c, lists = oracle.contacts(p, b)
    assert list(c) == [1, 1, 1, 1, 0]
    assert lists[0, 0] == 1 and lists[2, 0] == 3
    # a fast crossing wraps to the opposite edge without losing the contact
    crossed = (0.9998 + 2.0 / 1000) % 1.0
    c, _ = oracle.contacts(p, np.array([[crossed, 0.5, 0, 0], [0.0015, 0.5, 0, 0]], np.float32))
    assert list(c) == [1, 1]
    # 130 bodies inside one cell all touch each other
    rng = np.random.default_rng(20260730)
    crowd = np.zeros((130, 4), np.float32)
    crowd[:, :2] = 0.5 + rng.uniform(-0.00005, 0.00005, (130, 2))
    c, _ = oracle.contacts(oracle.params(3440, 1440), crowd, cap=256)
    assert (c == 129).all()


def test_minimum_image_known_answers(oracle):
    # T/ScreenBoundaryTest.kt:19-22: 0.01 and 0.99 are 0.02 apart through the edge.
    d2 = oracle.load().orc_pair_distance_squared
    a = np.array([0.01, 0.5], np.float32)
    b = np.array([0.99, 0.5], np.float32)
    got = d2(0, a.ctypes.data, 1, b.ctypes.data, 100.0, 100.0, 0)
    assert abs(got - 4.0) < 1e-4
    walls = d2(0, a.ctypes.data, 1, b.ctypes.data, 100.0, 100.0, 1)
    assert abs(walls - 98.0 ** 2) < 1e-2


def test_isolated_pair_solves_to_the_closed_form(oracle):
    # Two unit masses 0.8 px apart approaching at 10 px/s each: each takes half
    # the penetration; the approach speed is removed and the push becomes speed.
    W = H = 1000
    d = float(DIAMETER)
    b = scene([(100, 100, 10, 0), (100.8, 100, -10, 0)], W, H)
    p = oracle.params(W, H)
    out, press, counts = oracle.collide_pass(p, b, np.zeros((2, 2), np.float32), None, 0.0, 0.0, True)
    sub = (1 / 120) / 16
    pen = d - 0.8
    x0 = out[0, 0] * W
    x1 = out[1, 0] * W
    assert list(counts) == [1, 1]
    assert abs((x1 - x0) - d) < 2e-4
    assert abs(out[0, 2] * W - (0.0 - 0.5 * pen / sub)) < 1e-2
    assert abs(out[0, 2] + out[1, 2]) * W < 1e-4
    # pressure = 0.5 * (pen / dt_sub + approach)
    assert abs(press[0] - 0.5 * (pen / sub + 20.0)) < 1e-2


def test_dense_square_corrections_conserve_momentum(oracle):
    # T/CpuCollisionReferenceTest.kt:79-141, here through the real solve.
    W, H = 800, 600
    sp = float(DIAMETER) * 0.55
    rows = [(0.5 * W + (x - 5) * sp, 0.5 * H + (y - 5) * sp, 0, 0) for y in range(11) for x in range(11)]
    b = scene(rows, W, H)
    p = oracle.params(W, H)
    out, _, counts = oracle.collide_pass(p, b, np.zeros((121, 2), np.float32), None, 0.0, 0.0, False)
    assert counts.sum() % 2 == 0 and counts.sum() > 0
    assert np.isfinite(out).all()
    assert abs(out[:, 2].astype(np.float64).sum() * W) < 1e-3
    assert abs(out[:, 3].astype(np.float64).sum() * H) < 1e-3


def test_wall_contact_matches_a_mirror_body(oracle):
    # A body 0.2 px from the left wall against its mirror pair across x = 20.
    W, H = 3440, 1440
    walls = oracle.params(W, H, bounded=True)
    b = scene([(0.2, 300, -40, 0), (20.2, 900, -40, 0), (19.8, 900, 40, 0)], W, H)
    out, _, _ = oracle.collide_pass(walls, b, np.zeros((3, 2), np.float32), None, 0.0, 0.0, True)
    assert abs(out[0, 0] * W - (out[1, 0] * W - 20)) < 5e-3
    assert abs(out[0, 2] - out[1, 2]) * W < 1e-2


# ---- b: P3M ------------------------------------------------------------------

@pytest.mark.parametrize("w,h", [(16, 16), (32, 16), (16, 64), (128, 128), (512, 128), (128, 512), (2048, 128)])
def test_fft_layout_matches_fp64_dft(oracle, w, h):
    # T/Radix2FftLayoutTest.kt:18-45: < 1e-9 absolute, < 1e-10 relative.
    rng = np.random.default_rng(73 + w * 31 + h)
    a = rng.uniform(-1, 1, (h, w)) + 1j * rng.uniform(-1, 1, (h, w))
    fwd = oracle.fft2d(a, inverse=False, f64=True)
    inv = oracle.fft2d(a, inverse=True, f64=True)
    ref_f = np.fft.fft2(a)
    ref_i = np.fft.ifft2(a) * (w * h)
    for got, ref in ((fwd, ref_f), (inv, ref_i)):
        err = np.abs(got - ref).max()
        assert err < 1e-9
        assert err / np.abs(ref).max() < 1e-10


def test_fft_convolution_is_direct_circular_convolution(oracle):
    # T/Radix2FftLayoutTest.kt:47-73, 144-166
    w, h = 16, 8
    rng = np.random.default_rng(991)
    left = rng.uniform(-1, 1, (h, w))
    right = rng.uniform(-1, 1, (h, w))
    direct = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            direct[y, x] = sum(left[sy, sx] * right[(y - sy) % h, (x - sx) % w] for sy in range(h) for sx in range(w))
    prod = oracle.fft2d(left + 0j, f64=True) * oracle.fft2d(right + 0j, f64=True)
    conv = oracle.fft2d(prod, inverse=True, f64=True) / (w * h)
    assert np.abs(conv.real - direct).max() < 1e-9


def test_twiddles_are_rounded_unit_roots(oracle):
    # S/twiddle_init.comp:15-24 forms the angle -TAU * id / (2 h) in FP32, so
    # the table is cos/sin of that FP32 angle (within 1 ulp) and deviates from
    # the true roots by at most the angle's rounding (~1 ulp of pi).
    tw = oracle.twiddles(2048).reshape(-1, 4)
    tau = np.float32(6.28318530717958647692)
    h = 1
    while h < 2048:
        k = np.arange(h)
        fa = (-tau * k.astype(np.float32)) / np.float32(2 * h)
        t = tw[h - 1:h - 1 + h].astype(np.float64)
        ulp = np.spacing(np.float32(1.0)).astype(np.float64)
        assert np.abs(t[:, 0] - np.cos(fa.astype(np.float64))).max() <= ulp
        assert np.abs(t[:, 1] - np.sin(fa.astype(np.float64))).max() <= ulp
        ang = 2 * np.pi * k / (2 * h)
        assert np.abs(t[:, 0] - np.cos(ang)).max() < 4e-7
        assert np.abs(t[:, 1] + np.sin(ang)).max() < 4e-7
        assert np.abs(t[:, 3] - np.sin(ang)).max() < 4e-7
        h *= 2


def test_deposit_conserves_mass_and_follows_translation(oracle):
    rng = np.random.default_rng(3)
    n, pw, ph = 2000, 64, 32
    b = np.zeros((n, 4), np.float32)
    b[:, 0] = rng.random(n, dtype=np.float32)
    b[:, 1] = rng.random(n, dtype=np.float32)
    m = oracle.default_materials(n)
    scale = oracle.mass_scale(float(n))
    assert scale == 65536  # 4e9 / 2000 exceeds the 65536 cap
    dens, cnt, cmass = oracle.deposit(b, m, pw, ph, pw, ph, scale)
    assert abs(int(dens.sum()) - scale * n) <= 2 * n
    hist = np.zeros((ph, pw), np.int64)
    np.add.at(hist, ((b[:, 1] * ph).astype(int) % ph, (b[:, 0] * pw).astype(int) % pw), 1)
    np.testing.assert_array_equal(cnt, hist)
    np.testing.assert_array_equal(cmass, hist * 256)
    # shifting every body by exactly one cell shifts the density by one node
    shifted = b.copy()
    shifted[:, 0] = np.mod(b[:, 0] + np.float32(1.0 / pw), 1.0).astype(np.float32)
    ok = np.abs(shifted[:, 0] - (b[:, 0] + np.float32(1.0 / pw))) < 1e-9
    d1, _, _ = oracle.deposit(b[ok], m[ok], pw, ph, pw, ph, scale)
    d2, _, _ = oracle.deposit(shifted[ok], m[ok], pw, ph, pw, ph, scale)
    assert np.abs(np.roll(d1, 1, axis=1).astype(np.int64) - d2.astype(np.int64)).max() <= 4
    # a body at a node centre deposits all of its mass on that node
    one = np.array([[(10 + 0.5) / pw, (7 + 0.5) / ph, 0, 0]], np.float32)
    d, _, _ = oracle.deposit(one, oracle.default_materials(1), pw, ph, pw, ph, 1000)
    assert d[7, 10] == 1000 and d.sum() == 1000
    assert oracle.mass_scale(400000.0) == 10000 and oracle.mass_scale(1e12) == 1


def direct_forces_isolated(b, m, W, H, G, total, min_dist2):
    """FP64 direct summation with walls: the isolated 2D sum
    a_i = sum_j G m_j d / (M max(r^2, r_min^2)) in physical units (x scaled by
    the aspect ratio), as the reference's near-field term without its Gaussian."""
    aspect = W / H
    x = b[:, 0].astype(np.float64) * aspect
    y = b[:, 1].astype(np.float64)
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    r2 = np.maximum(dx * dx + dy * dy, min_dist2)
    np.fill_diagonal(r2, np.inf)
    s = G * m[None, :, 0].astype(np.float64) / (total * r2)
    return np.stack([(s * dx).sum(1) / aspect, (s * dy).sum(1)], axis=1)


def test_p3m_force_against_fp64_direct_sum_with_walls(oracle):
    W, H, n = 1440, 720, 1500
    rng = np.random.default_rng(8)
    b = np.zeros((n, 4), np.float32)
    b[:, 0] = (0.5 + rng.normal(0, 0.08, n)).clip(0.05, 0.95).astype(np.float32)
    b[:, 1] = (0.5 + rng.normal(0, 0.12, n)).clip(0.05, 0.95).astype(np.float32)
    m = oracle.default_materials(n)
    sim = oracle.Sim(oracle.params(W, H, bounded=True, grid=(256, 128)), b, m)
    sim.initialize_acceleration()
    a = sim.state()[1].astype(np.float64)
    cw, ch = (W / H) / 256, 1 / 128
    ref = direct_forces_isolated(b, m, W, H, 0.035, float(n), (0.05 * min(cw, ch)) ** 2)
    rms = np.sqrt(((a - ref) ** 2).sum(1).mean() / (ref ** 2).sum(1).mean())
    # P3M with CIC and a 0.85-cell Gaussian split: a few percent RMS.
    assert rms < 0.05, rms


def test_liquid_kernel_constants(oracle):
    # T/LiquidKernelTest.kt:10-32
    for d in (1.12838, 3.38514):
        for radius in (2.0, 3.0):
            h, poly6, spiky, rest, relax = oracle.liquid_constants(d, radius)
            packing = 1.0 / (f32(d) * f32(d) * math.sqrt(3.0) * 0.5)
            assert abs(rest - packing) < 0.05 * packing
            assert abs(poly6 * (4 / (math.pi * h ** 8)) ** -1 - 1) < 1e-6
    small = oracle.liquid_constants(1.12838, 2.0)
    large = oracle.liquid_constants(3.38514, 2.0)
    assert abs(small[4] / large[4] - 9.0) < 0.05
