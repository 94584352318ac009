"""Host-side members of the boundary, no GPU: CircularProjectile's packing
and diameter search (K/CircularProjectile.kt:10-60) against an independent
float32 restatement and the reference's own tests
(T/CircularProjectileTest.kt:10-75), and supportsWalls
(K/GravityPipeline.kt:1628)."""
import math

import numpy as np
import pytest

from gravity_amd import GravityFieldConfig, GravityPipeline
from gravity_amd.pipeline import projectile_diameter_for_body_count, projectile_offsets

f32 = np.float32


def offsets_restated(d, D):
    """The packing loop in numpy float32 scalars, in Kotlin's evaluation order."""
    d, D = f32(d), f32(D)
    available = f32(f32(0.5) * D) - f32(f32(0.5) * d)
    row_step = f32(d * f32(math.sqrt(3.0) * 0.5))
    out, row, y = [], 0, -available
    while y <= available:
        x = -available + (f32(0.0) if row % 2 == 0 else f32(f32(0.5) * d))
        while x <= available:
            if f32(f32(x * x) + f32(y * y)) <= f32(available * available):
                out.append((x, y))
            x = f32(x + d)
        y = f32(y + row_step)
        row += 1
    return np.array(out, np.float32).reshape(-1, 2)


@pytest.mark.parametrize("d,D", [(1.12838, 50.0), (1.0, 10.0), (1.12838, 1.12838), (3.38514, 97.3), (0.7, 23.0)])
def test_packing_is_the_restated_loop_bit_for_bit(d, D):
    got = projectile_offsets(d, D)
    want = offsets_restated(d, D)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_default_projectile_is_a_dense_fifty_pixel_disc():
    # T/CircularProjectileTest.kt:10-24
    d = 1.12838
    off = projectile_offsets(d, 50.0)
    assert 1500 <= len(off) <= 1900
    assert (np.hypot(off[:, 0].astype(np.float64), off[:, 1]) <= 0.5 * 50.0 - 0.5 * d + 1e-4).all()


def test_larger_diameters_hold_more_bodies():
    # T/CircularProjectileTest.kt:57-75
    d, previous = 1.12838, 0
    for D in range(10, 101, 10):
        off = projectile_offsets(d, float(D))
        assert len(off) > previous
        assert (np.hypot(off[:, 0].astype(np.float64), off[:, 1]) <= 0.5 * D - 0.5 * d + 1e-4).all()
        previous = len(off)


def test_plus_key_diameter_is_the_smallest_holding_the_body_count():
    # T/CircularProjectileTest.kt:34-44: ProjectileConfig().plusKeyBodyCount = 50,000
    d = 1.12838
    D = projectile_diameter_for_body_count(d, 50_000)
    assert len(projectile_offsets(d, D)) >= 50_000
    assert len(projectile_offsets(d, np.float32(D) - np.float32(0.05))) < 50_000


def test_projectile_smaller_than_a_body_is_rejected():
    with pytest.raises(ValueError):
        projectile_offsets(1.12838, 1.0)
    with pytest.raises(ValueError):
        projectile_diameter_for_body_count(1.12838, 0)


def test_supports_walls():
    assert GravityPipeline.supportsWalls(GravityFieldConfig(gridWidth=1024, gridHeight=512))
    assert not GravityPipeline.supportsWalls(GravityFieldConfig(gridWidth=2048, gridHeight=256))
