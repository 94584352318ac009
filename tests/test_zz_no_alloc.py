"""gs_debug_density_noalloc: the density readback with its staging in the
library's static scratch array (no device allocation in the call), against
gs_debug_density, on two calls in a row, the second with a smaller input
whose walls leave rows empty that the first call's density filled."""
import numpy as np
import pytest

from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
from gravity_amd import native, scenarios

pytestmark = pytest.mark.gpu


def readback_noalloc(pipe, n):
    q = pipe.query()
    mw, mh, pw, ph = q["meshWidth"], q["meshHeight"], q["gridWidth"], q["gridHeight"]
    dens = np.full(mw * mh, 0xDEADBEEF, np.uint32)
    spec = np.full(2 * mw * mh, np.nan, np.float32)
    cells = np.full(2 * pw * ph, 0xDEADBEEF, np.uint32)
    rc = native.load().gs_debug_density_noalloc(pipe._h, n, native.ptr(dens), native.ptr(spec), native.ptr(cells))
    native.check(rc)
    spec = spec.reshape(mh, mw, 2)
    return dens.reshape(mh, mw), spec[..., 0] + 1j * spec[..., 1], cells.reshape(ph, pw, 2)


def make(n, W, H, grid, bounded, bodies):
    cfg = GravityConfig(particles=ParticleConfig(capacity=n, initialBodies=n),
                        gravity=GravityFieldConfig(gridWidth=grid[0], gridHeight=grid[1], minimumGrid=16),
                        collisions=CollisionConfig(contactDiameterPixels=1.0), collisionsEnabled=False)
    return GravityPipeline(cfg, W, H, scenarioState=bodies, bounded=bounded)


def test_noalloc_density_equals_the_allocating_readback_twice():
    first = scenarios.uniform(20_000, 1440, 720, seed=21)[0]          # torus, 512x256 mesh, every row filled
    second = scenarios.exponential_disc(6_000, 640, 320, scale_px=40.0, seed=22)[0]  # walls, 256x128 mesh
    cases = [(first, (1440, 720), (512, 256), False), (second, (640, 320), (128, 64), True)]
    for k, (bodies, (W, H), grid, bounded) in enumerate(cases):
        n = len(bodies)
        print(f"[no-alloc] call {k + 1}: {n} bodies, grid {grid}, walls {bounded}", flush=True)
        with make(n, W, H, grid, bounded, bodies) as pipe:
            got = readback_noalloc(pipe, n)
            want = pipe.densityMesh(n)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[2], want[2])
        np.testing.assert_array_equal(got[1].view(np.float32), want[1].view(np.float32))
        assert got[0].sum() > 0
        if bounded:
            mh = got[0].shape[0]
            assert not got[0][grid[1] + 2:mh - 1].any()   # rows the walls leave empty


def test_noalloc_density_rejects_what_it_cannot_hold():
    bodies = scenarios.uniform(64, 256, 256, seed=3)[0]
    with make(64, 256, 256, (64, 64), False, bodies) as pipe:
        with pytest.raises(ValueError, match="active body count"):
            readback_noalloc(pipe, 32)
    with make(64, 4096, 4096, (2048, 2048), False, bodies) as pipe:
        with pytest.raises(ValueError, match="2048 x 1024"):
            readback_noalloc(pipe, 64)
