"""Contact passes on the MI355X against the oracle and the reference's own
expectations (T/CollisionShaderIntegrationTest.kt, T/FrictionIntegrationTest.kt).

The HIP path and the oracle accumulate contacts in the same ascending index
order with the same float operations, so their bodies and pressures must agree
bit for bit; contact counts and IDs must equal an independent brute force."""
import numpy as np
import pytest

from gravity_amd import CollisionConfig, GravityConfig, GravityFieldConfig, GravityPipeline, ParticleConfig
from gravity_amd.native import CONTACT_SLOTS
from tests.scenes import DIAMETER, HEIGHT, WIDTH, brute_force_contacts, collision_scenes, scene

pytestmark = pytest.mark.gpu


def config(count, friction=0.0, grid=128):
    return GravityConfig(particles=ParticleConfig(capacity=count, initialBodies=count, seed=7235),
                         gravity=GravityFieldConfig(gridWidth=grid, gridHeight=grid),
                         collisions=CollisionConfig(friction=friction))


def params(oracle, bounded=False, friction=0.0, width=WIDTH, height=HEIGHT):
    return oracle.params(width, height, bounded=bounded, grid=(128, 128))


def materials(n, friction=0.0):
    m = np.zeros((n, 4), np.float32)
    m[:, 0] = 1.0
    m[:, 1] = friction
    return m


SCENES = collision_scenes()


@pytest.mark.parametrize("name", list(SCENES))
def test_replay_matches_oracle_and_brute_force(name, oracle):
    bodies = SCENES[name]
    count = len(bodies)
    passes = 4 if name == "isolated" else 12
    p = params(oracle)
    zero = np.zeros((count, 2), np.float32)
    with GravityPipeline(config(count), WIDTH, HEIGHT) as pipe:
        state = bodies
        for k in range(passes):
            pipe.upload(state)
            pipe.resolveCollisions(count, 0.0, 0.0, True, None)
            counts = pipe.contactCounts(count)
            expected = brute_force_contacts(state)
            np.testing.assert_array_equal(counts, [len(e) for e in expected], err_msg=f"{name}/{k} counts")
            slots = pipe.contacts(count)
            for i in range(count):
                want = expected[i][:CONTACT_SLOTS]
                np.testing.assert_array_equal(slots[i, :len(want)], want, err_msg=f"{name}/{k} body {i} IDs")
            got = pipe.bodies(count)
            press = pipe.pressures(count)
            ob, opr, oc = oracle.collide_pass(p, state, zero, materials(count), 0.0, 0.0, True)
            np.testing.assert_array_equal(oc, counts)
            assert np.isfinite(got).all()
            np.testing.assert_array_equal(got.view(np.uint32), ob.view(np.uint32), err_msg=f"{name}/{k} bodies")
            np.testing.assert_array_equal(press.view(np.uint32), opr.view(np.uint32), err_msg=f"{name}/{k} pressure")
            state = got


@pytest.mark.parametrize("count", [64, 65, 66, 90])
def test_crowds_at_the_contact_plane_limit_match_oracle(count, oracle):
    # Crowds whose bodies touch all others: 63 and 64 contacts fit the
    # contact planes (sixteen listed, the rest past them), 65 and 89 go to the
    # contact pool; the solve takes every contact in ascending order either way.
    rng = np.random.default_rng(count)
    rows = [(900 + 0.3 * rng.random(), 500 + 0.3 * rng.random(), rng.normal(0, 20), rng.normal(0, 20))
            for _ in range(count)]
    state = scene(rows)
    p = params(oracle)
    zero = np.zeros((count, 2), np.float32)
    with GravityPipeline(config(count), WIDTH, HEIGHT) as pipe:
        for k in range(3):
            pipe.upload(state)
            pipe.resolveCollisions(count, 0.0, 0.0, True, None)
            counts = pipe.contactCounts(count)
            if k == 0:
                np.testing.assert_array_equal(counts, np.full(count, count - 1))
            np.testing.assert_array_equal(counts, [len(e) for e in brute_force_contacts(state)])
            got = pipe.bodies(count)
            ob, opr, oc = oracle.collide_pass(p, state, zero, materials(count), 0.0, 0.0, True)
            np.testing.assert_array_equal(oc, counts)
            np.testing.assert_array_equal(got.view(np.uint32), ob.view(np.uint32), err_msg=f"{count}/{k}")
            state = got


def test_dense_lattice_counts_are_four_neighbour_counts():
    # T/CollisionShaderIntegrationTest.kt:166-171
    count = 4096
    bodies = scene([(1000 + i % 64, 600 + i // 64, 0, 0) for i in range(count)])
    expected = [(i % 64 > 0) + (i % 64 < 63) + (i // 64 > 0) + (i // 64 < 63) for i in range(count)]
    with GravityPipeline(config(count), WIDTH, HEIGHT) as pipe:
        for _ in range(3):
            pipe.upload(bodies)
            pipe.resolveCollisions(count, 0.0, 0.0, True, None)
            np.testing.assert_array_equal(pipe.contactCounts(count), expected)


def test_coincident_crowd_has_129_contacts_each():
    # T/CollisionShaderIntegrationTest.kt:142-161: every body walks past its slots.
    bodies = scene([(700 + i % 13 * 0.025, 600 + i // 13 * 0.025, 0, 0) for i in range(130)])
    with GravityPipeline(config(130), WIDTH, HEIGHT) as pipe:
        for _ in range(4):
            pipe.upload(bodies)
            pipe.resolveCollisions(130, 0.0, 0.0, True, None)
            np.testing.assert_array_equal(pipe.contactCounts(130), np.full(130, 129))


def test_drifting_walls_scene_matches_oracle_and_never_contacts_across_an_edge(oracle):
    # T/CollisionShaderIntegrationTest.kt:194-226
    count = 64
    rows = []
    for i in range(count):
        r = i % 4
        if r == 0:
            rows.append((0.2 + (i // 4) * 3, 0.3, -30, 0))
        elif r == 1:
            rows.append((WIDTH - 0.2, 0.2 + (i // 4) * 3, 25, 10))
        elif r == 2:
            rows.append((0.1 + (i // 4) * 0.8, HEIGHT - 0.1, 0, 40))
        else:
            rows.append((0.05 if i % 8 == 3 else WIDTH - 0.05, 700 + i, 0, -5))
    state = scene(rows)
    p = params(oracle, bounded=True)
    zero = np.zeros((count, 2), np.float32)
    drift = np.float32(1.0 / 120.0) / np.float32(16)
    with GravityPipeline(config(count), WIDTH, HEIGHT, bounded=True) as pipe:
        for k in range(12):
            pipe.upload(state)
            pipe.resolveCollisions(count, 0.0, float(drift), True, None)
            got = pipe.bodies(count)
            ob, opr, _ = oracle.collide_pass(p, state, zero, materials(count), 0.0, float(drift), True)
            np.testing.assert_array_equal(got.view(np.uint32), ob.view(np.uint32), err_msg=f"walls/{k}")
            assert ((got[:, :2] >= 0) & (got[:, :2] < 1)).all()
            state = got
        counts = pipe.contactCounts(count)
        assert all(counts[i] == 0 for i in range(3, count, 4))


def test_wall_acts_like_a_mirror_body():
    # T/CollisionShaderIntegrationTest.kt:228-261
    mirror = 20.0
    state = scene([(2, 300, -400, 50), (mirror + 2, 900, -400, 50), (mirror - 2, 900, 400, 50)])
    drift = float(np.float32(1.0 / 120.0) / np.float32(16))
    touched = False
    with GravityPipeline(config(3), WIDTH, HEIGHT, bounded=True) as pipe:
        for _ in range(64):
            pipe.upload(state)
            pipe.resolveCollisions(3, 0.0, drift, True, None)
            state = pipe.bodies(3)
            wall_x = state[0, 0] * WIDTH
            pair_x = state[1, 0] * WIDTH - mirror
            if state[0, 2] > 0:
                touched = True
            assert abs(wall_x - pair_x) < 5e-3
            assert abs(state[0, 2] - state[1, 2]) * WIDTH < 2.0
    assert touched
    assert 0 <= state[0, 0] < 1


def test_kick_drift_pass_matches_oracle(oracle):
    rng = np.random.default_rng(11)
    count = 3000
    rows = [(1500 + rng.normal(0, 12), 700 + rng.normal(0, 12), rng.normal(0, 30), rng.normal(0, 30))
            for _ in range(count)]
    state = scene(rows)
    accel = rng.normal(0, 1e-3, (count, 2)).astype(np.float32)
    mats = materials(count, 0.2)
    p = params(oracle)
    dt = np.float32(1.0 / 120.0)
    with GravityPipeline(config(count, friction=0.2), WIDTH, HEIGHT) as pipe:
        pipe.upload(state, mats)
        pipe.uploadAcceleration(accel)
        pipe.resolveCollisions(count, float(dt), float(dt / np.float32(16)), False, 0)
        got = pipe.bodies(count)
    ob, _, _ = oracle.collide_pass(p, state, accel, mats, float(dt), float(dt / np.float32(16)), False)
    np.testing.assert_array_equal(got.view(np.uint32), ob.view(np.uint32))


def slide(friction):
    # T/FrictionIntegrationTest.kt:36-54
    W = H = 256
    start = 50.0
    body = np.array([[start / W, 0.5 * float(DIAMETER) / H, 100.0 / W, 0.0]], np.float32)
    cfg = GravityConfig(particles=ParticleConfig(capacity=1, initialBodies=1, seed=7235),
                        gravity=GravityFieldConfig(gridWidth=128, gridHeight=128),
                        collisions=CollisionConfig(friction=friction))
    with GravityPipeline(cfg, W, H, bounded=True) as pipe:
        pipe.mutualGravity = False
        pipe.fieldAccelerationY = -300.0
        pipe.upload(body, np.array([[1.0, friction, 0, 0]], np.float32))
        pipe.initializeAcceleration(1)
        for _ in range(240):
            pipe.step(1)
        s = pipe.bodies(1)
    return s[0, 2] * W, s[0, 0] * W - start


def test_body_slides_on_smooth_floor_and_stops_on_rough_one():
    speed, _ = slide(0.0)
    assert abs(speed - 100.0) < 0.5
    speed, distance = slide(0.5)
    assert abs(speed) < 0.5
    # Coulomb: v^2 / (2 mu g) = 100^2 / (2 * 0.5 * 300) = 33.3 px
    assert abs(distance - 33.3) < 3.0


def test_friction_between_bodies_conserves_momentum():
    # T/FrictionIntegrationTest.kt:56-76
    W = H = 256

    def hit(friction):
        b = np.array([[100 / W, 100 / H, 40 / W, 30 / H],
                      [(100 + 0.9 * float(DIAMETER)) / W, 100 / H, -40 / W, -30 / H]], np.float32)
        cfg = GravityConfig(particles=ParticleConfig(capacity=2, initialBodies=2, seed=7235),
                            gravity=GravityFieldConfig(gridWidth=128, gridHeight=128),
                            collisions=CollisionConfig(friction=friction))
        with GravityPipeline(cfg, W, H) as pipe:
            pipe.upload(b, materials(2, friction))
            pipe.resolveCollisions(2, 0.0, 0.0, True, None)
            s = pipe.bodies(2)
        return (s[0, 2] + s[1, 2]) * W, (s[0, 3] + s[1, 3]) * H, abs(s[0, 3] - s[1, 3]) * H

    smooth, rough = hit(0.0), hit(1.0)
    assert abs(smooth[0] - rough[0]) < 1e-3 and abs(smooth[1] - rough[1]) < 1e-3
    assert rough[2] < smooth[2] - 1.0
