"""Spatial sort on the MI355X (T/SpatialSortIntegrationTest.kt:18-54).

Both the HIP sort and the oracle order each tile by original index, so the
permutation is pinned bit for bit, beyond the reference's multiset check."""
import numpy as np
import pytest

from gravity_amd import GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
from gravity_amd import scenarios
from tests.scenes import HEIGHT, WIDTH, sort_scene

pytestmark = pytest.mark.gpu


def config(count, interval=1):
    return GravityConfig(particles=ParticleConfig(capacity=count, initialBodies=count, seed=7235,
                                                  spatialSortInterval=interval),
                         gravity=GravityFieldConfig(gridWidth=128, gridHeight=128))


def records(bodies, extra):
    rows = np.concatenate([bodies.view(np.uint32).reshape(len(bodies), -1),
                           extra.view(np.uint32).reshape(len(bodies), -1)], axis=1)
    return rows[np.lexsort(rows.T[::-1])]


def test_sort_is_the_oracle_permutation_bit_for_bit(oracle):
    bodies = sort_scene()
    count = len(bodies)
    rng = np.random.default_rng(5)
    accel = (rng.random((count, 2), dtype=np.float32) - np.float32(0.5))
    mats = oracle.default_materials(count)
    mats[:, 3] = np.arange(count, dtype=np.float32)  # tag each body
    with GravityPipeline(config(count), WIDTH, HEIGHT) as pipe:
        pipe.upload(bodies, mats)
        pipe.uploadAcceleration(accel)
        pipe.sortBodies(count)
        sb = pipe.bodies(count)
        sa = pipe.acceleration(count)
        sm = pipe.materials(count)
        dest = pipe.sortDestinations(count)
    ob, oa, om, odest = oracle.spatial_sort(bodies, accel, mats, WIDTH, HEIGHT)
    np.testing.assert_array_equal(records(bodies, accel), records(sb, sa))
    np.testing.assert_array_equal(sb.view(np.uint32), ob.view(np.uint32))
    np.testing.assert_array_equal(sa.view(np.uint32), oa.view(np.uint32))
    np.testing.assert_array_equal(sm.view(np.uint32), om.view(np.uint32))
    np.testing.assert_array_equal(dest, odest)
    wb, hb = oracle.sort_axis_bits(WIDTH), oracle.sort_axis_bits(HEIGHT)
    keys = np.array([oracle.sort_key(x, y, wb, hb) for x, y in sb[:, :2]])
    assert (np.diff(keys.astype(np.int64)) >= 0).all()
    assert len(set(keys.tolist())) > 1000


def test_sort_twice_keeps_counts_at_zero_and_is_idempotent(oracle):
    bodies = sort_scene(3000, seed=9)
    count = len(bodies)
    with GravityPipeline(config(count), WIDTH, HEIGHT) as pipe:
        pipe.upload(bodies)
        pipe.sortBodies(count)
        once = pipe.bodies(count)
        pipe.sortBodies(count)
        twice = pipe.bodies(count)
        dest = pipe.sortDestinations(count)
    np.testing.assert_array_equal(once.view(np.uint32), twice.view(np.uint32))
    np.testing.assert_array_equal(dest, np.arange(count))


def test_full_size_sort_is_a_sorted_permutation(oracle):
    W, H = 2880, 1440
    bodies, mats = scenarios.exponential_disc(400_000, W, H, seed=7)
    count = len(bodies)
    cfg = GravityConfig(particles=ParticleConfig(capacity=count, initialBodies=count, spatialSortInterval=1),
                        gravity=GravityFieldConfig(gridWidth=1024, gridHeight=512))
    with GravityPipeline(cfg, W, H, bounded=True) as pipe:
        pipe.upload(bodies, mats)
        pipe.sortBodies(count)
        sb = pipe.bodies(count)
        dest = pipe.sortDestinations(count)
    assert np.array_equal(np.sort(dest), np.arange(count))
    np.testing.assert_array_equal(sb[dest].view(np.uint32), bodies.view(np.uint32))
    ob, _, _, odest = oracle.spatial_sort(bodies, np.zeros((count, 2), np.float32), mats, W, H)
    np.testing.assert_array_equal(dest, odest)
