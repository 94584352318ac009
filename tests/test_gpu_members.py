"""Boundary members on the MI355X: spawnProjectile (K/GravityPipeline.kt:
621-646) against a host restatement of its placement, measureCluster
(:793-905) against an FP64 numpy restatement, and the render view of
bindBodiesForRender (:509-512) against the downloaded state."""
import ctypes
import math

import numpy as np
import pytest

from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
from gravity_amd import scenarios
from gravity_amd.pipeline import projectile_offsets

pytestmark = pytest.mark.gpu
f32 = np.float32


def make(cap, n, W, H, bounded, friction=0.3):
    bodies, mats = scenarios.uniform(n, W, H, seed=5)
    cfg = GravityConfig(particles=ParticleConfig(capacity=cap, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=256, gridHeight=128),
                        collisions=CollisionConfig(friction=friction))
    pipe = GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=bounded)
    pipe.upload(bodies, mats)
    return pipe, bodies


def placed(cx, cy, vx, vy, D, W, H, bounded, d=1.12838):
    """spawnProjectile's placement restated in float32 scalars."""
    off = projectile_offsets(d, D)
    radius = f32(np.hypot(off[:, 0].astype(np.float64), off[:, 1].astype(np.float64)).max()) + f32(0.5) * f32(d)
    cx, cy, Wf, Hf = f32(cx), f32(cy), f32(W), f32(H)
    if bounded:
        cx = min(max(cx, radius), f32(Wf - radius))
        cy = min(max(cy, radius), f32(Hf - radius))
    out = np.empty((len(off), 4), np.float32)
    for k, (ox, oy) in enumerate(off):
        px, py = f32(cx + ox), f32(cy + oy)
        if bounded:
            px = min(max(px, f32(0)), f32(Wf * f32(0.99999994)))
            py = min(max(py, f32(0)), f32(Hf * f32(0.99999994)))
        else:
            px = f32(px - f32(np.floor(f32(px / Wf)) * Wf))
            py = f32(py - f32(np.floor(f32(py / Hf)) * Hf))
        out[k] = (f32(px / Wf), f32(py / Hf), f32(f32(vx) / Wf), f32(f32(vy) / Hf))
    return out


@pytest.mark.parametrize("bounded", [False, True])
def test_spawned_projectile_is_the_restated_placement(bounded):
    W, H = 1440, 720
    pipe, _ = make(5000, 1000, W, H, bounded)
    with pipe:
        k = pipe.spawnProjectile(1000, 100.3, 705.0, 2000.0, -50.0, 30.0)
        assert k == pipe.projectileBodyCount(30.0) == len(projectile_offsets(1.12838, 30.0))
        got = pipe.bodies(1000 + k)[1000:]
        mats = pipe.materials(1000 + k)[1000:]
        assert pipe.spawnProjectile(5000 - k + 1, 500.0, 300.0, 0.0, 0.0, 30.0) == 0   # does not fit
    want = placed(100.3, 705.0, 2000.0, -50.0, 30.0, W, H, bounded)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(mats, np.tile(np.array([1.0, 0.3, 0.0, 0.0], np.float32), (k, 1)))


def cluster_restated(b, W, H):
    """measureCluster (K/GravityPipeline.kt:793-905) in numpy FP64."""
    b = b.astype(np.float64)
    n = len(b)
    finite = np.isfinite(b).all(axis=1)
    x = np.where(finite, b[:, 0], 0.0)
    y = np.where(finite, b[:, 1], 0.0)
    speed = np.where(finite, np.hypot(b[:, 2] * W, b[:, 3] * H), 0.0)
    cx = math.atan2(np.sin(2 * np.pi * x).sum(), np.cos(2 * np.pi * x).sum()) / (2 * np.pi)
    cy = math.atan2(np.sin(2 * np.pi * y).sum(), np.cos(2 * np.pi * y).sum()) / (2 * np.pi)
    cx -= math.floor(cx)
    cy -= math.floor(cy)
    ux = (cx + ((x - cx) - np.floor(x - cx + 0.5))) * W
    uy = (cy + ((y - cy) - np.floor(y - cy + 0.5))) * H
    mx, my = ux.mean(), uy.mean()
    xx, yy, xy = ((ux - mx) ** 2).mean(), ((uy - my) ** 2).mean(), ((ux - mx) * (uy - my)).mean()
    area = 4 * math.pi * math.sqrt(max(xx * yy - xy * xy, 0.0))
    medx, medy = np.sort(ux)[n // 2], np.sort(uy)[n // 2]
    r2 = np.sort((ux - medx) ** 2 + (uy - medy) ** 2)

    def core(fraction):
        return math.pi * r2[min(max(int(fraction * n), 1), n) - 1] / fraction

    return [float(f32(medx) - np.floor(f32(medx) / f32(W)) * f32(W)),
            float(f32(medy) - np.floor(f32(medy) / f32(H)) * f32(H)), area, 2 * math.sqrt(area / math.pi),
            core(0.90), core(0.99), float(speed.max()), float(speed.mean()), int((~finite).sum())]


def test_measure_cluster_is_the_restated_statistic():
    W, H, n = 640, 320, 4000
    bodies, mats = scenarios.exponential_disc(n, W, H, scale_px=25.0, seed=5, bounded=False)
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=128, gridHeight=64, minimumGrid=16))
    with GravityPipeline(cfg, W, H, scenarioState=bodies) as pipe:
        pipe.upload(bodies, mats)
        pipe.initializeAcceleration(n)
        for _ in range(3):
            pipe.step(n)
        stats = pipe.measureCluster(n)
        b = pipe.bodies(n)
    want = cluster_restated(b, W, H)
    got = [stats.centerX, stats.centerY, stats.covarianceAreaPixels, stats.equivalentDiameterPixels,
           stats.core90AreaPixels, stats.core99AreaPixels, stats.maxSpeedPixelsPerSecond,
           stats.averageSpeedPixelsPerSecond, stats.nonFiniteBodies]
    np.testing.assert_allclose(got[:8], want[:8], rtol=1e-9)
    assert got[8] == want[8] == 0
    assert 0 < stats.covarianceAreaPixels < W * H


def _hip():
    """The HIP runtime the library itself links (torch may carry another)."""
    import os
    path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64.so.7" in l)
    assert os.path.exists(path)
    hip = ctypes.CDLL(path)
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_render_view_and_copy_are_the_state():
    W, H, n = 80, 40, 3000   # packed: contacts and pressures everywhere
    pipe, _ = make(n, n, W, H, True)
    with pipe:
        hip = _hip()
        pipe.initializeAcceleration(n)
        pipe.step(n)
        pipe.synchronize()
        bptr, pptr = pipe.renderView()
        view_b = np.empty((n, 4), np.float32)
        view_p = np.empty(n, np.float32)
        assert hip.hipMemcpy(view_b.ctypes.data, bptr, view_b.nbytes, 2) == 0   # device to host
        assert hip.hipMemcpy(view_p.ctypes.data, pptr, view_p.nbytes, 2) == 0
        dst_b, dst_p = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(dst_b), view_b.nbytes) == 0
        assert hip.hipMalloc(ctypes.byref(dst_p), view_p.nbytes) == 0
        try:
            pipe.renderCopy(n, dst_b.value, dst_p.value)
            pipe.synchronize()
            copy_b = np.empty((n, 4), np.float32)
            copy_p = np.empty(n, np.float32)
            assert hip.hipMemcpy(copy_b.ctypes.data, dst_b, copy_b.nbytes, 2) == 0
            assert hip.hipMemcpy(copy_p.ctypes.data, dst_p, copy_p.nbytes, 2) == 0
        finally:
            hip.hipFree(dst_b)
            hip.hipFree(dst_p)
        want_b, want_p = pipe.bodies(n), pipe.pressures(n)
    np.testing.assert_array_equal(view_b, want_b)
    np.testing.assert_array_equal(view_p, want_p)
    np.testing.assert_array_equal(copy_b, want_b)
    np.testing.assert_array_equal(copy_p, want_p)
    assert np.abs(want_p).sum() > 0
