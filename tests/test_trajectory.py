"""The trajectory preview's oracle (S/trajectory.comp restated in
oracle/oracle.c: orc_sim_trajectory) pinned by closed forms, and the host
helpers around it: TrajectoryConfig and TrajectoryUpdateScheduler (the
reference's own cases, T/TrajectoryUpdateSchedulerTest.kt:1-45,
T/GravityConfigTest.kt:59-87, 148)."""
import math

import numpy as np
import pytest

from gravity_amd import GravityConfig, TrajectoryConfig, TrajectoryUpdateScheduler

F = np.float32
W, H = 640, 320


def live(v):
    return v[v[:, 3] >= 0]


def empty_sim(oracle, **kw):
    p = oracle.params(W, H, grid=(64, 32), mutual_gravity=False, **kw)
    sim = oracle.Sim(p, np.zeros((0, 4), np.float32))
    sim.initialize_acceleration()
    return sim


# ---- host helpers -------------------------------------------------------

def test_configured_update_frequency_changes_the_interval():
    s = TrajectoryUpdateScheduler(updatesPerSecond=60)
    assert s.shouldUpdate(0)
    assert not s.shouldUpdate(16_000_000)
    assert s.shouldUpdate(17_000_000)


def test_trajectory_updates_no_more_than_thirty_times_per_second():
    s = TrajectoryUpdateScheduler(TrajectoryConfig().updatesPerSecond)
    assert s.shouldUpdate(0)
    for millis in range(1, 34):
        assert not s.shouldUpdate(millis * 1_000_000)
    assert s.shouldUpdate(34_000_000)


def test_reset_allows_immediate_update_for_a_new_drag():
    s = TrajectoryUpdateScheduler(TrajectoryConfig().updatesPerSecond)
    assert s.shouldUpdate(10)
    s.reset()
    assert s.shouldUpdate(11)


def test_trajectory_options():
    cfg = GravityConfig.parse(["--trajectory-steps=2048", "--trajectory-updates-per-second=60",
                               "--trajectory-segment-px=4", "--trajectory-min-dash-px=2"], {})
    t = cfg.trajectory
    assert (t.steps, t.updatesPerSecond, t.targetSegmentPixels, t.minimumDashPixels) == (2048, 60, 4.0, 2.0)
    c = cfg.native(W, H, False)
    assert c.trajectory_steps == 2048 and c.trajectory_target_segment_pixels == 4.0
    for bad in (dict(updatesPerSecond=0), dict(steps=0), dict(steps=65537), dict(maximumTimeStep=0.0),
                dict(targetSegmentPixels=-1.0), dict(maximumPathViewportDiagonals=math.inf),
                dict(minimumDashPixels=0.0)):
        with pytest.raises(ValueError):
            TrajectoryConfig(**bad)


# ---- the oracle's preview against closed forms ----------------------------

def test_free_flight_is_a_line_of_target_segments(oracle):
    sim = empty_sim(oracle, bounded=True)
    v = sim.trajectory((100.0, 160.0), (100.0, 0.0), steps=40)
    pts = live(v)
    assert len(pts) == 80  # one piece a step, two vertices a piece
    # every segment is the target length (dt = 6 px / speed within its clamp)
    seg = (pts[1::2, 0] - pts[0::2, 0]) * W
    np.testing.assert_allclose(seg, 6.0, rtol=1e-5)
    np.testing.assert_array_equal(pts[:, 1], F(0.5))
    # consecutive pieces join; the dash coordinate advances by the step over
    # 1 + |v| (v in viewports per second); the path fraction ends at 1
    np.testing.assert_allclose(pts[2::2, 0], pts[1:-1:2, 0], rtol=0, atol=1e-7)
    dash = pts[1::2, 2] - pts[0::2, 2]
    np.testing.assert_allclose(dash, 6.0 / (1.0 + 100.0 / W), rtol=1e-5)
    assert pts[-1, 3] == 1.0 and np.all(np.diff(pts[:, 3]) >= 0)


def test_uniform_field_follows_the_kinematics(oracle):
    # constant acceleration: kick-drift-kick is exact at the step ends, so the
    # vertices lie on x0 + v0 t + a t^2 / 2 at the accumulated times
    sim = empty_sim(oracle, bounded=True, field=(0.0, 400.0))
    x0, y0, vx, vy, a = 100.0, 40.0, 120.0, -60.0, 400.0
    v = sim.trajectory((x0, y0), (vx, vy), steps=60)
    pts = live(v)
    ends = pts[1::2, :2].astype(np.float64) * [W, H]
    # recover each step's time from the x coordinate (uniform motion in x)
    t = (ends[:, 0] - x0) / vx
    free = ends[:, 1] < H - 20  # before it reaches the floor
    assert free.sum() > 30
    np.testing.assert_allclose(ends[free, 1], (y0 + vy * t + 0.5 * a * t * t)[free], rtol=0, atol=2e-3)
    # dt solves |v| dt + |a| dt^2 / 2 = 6 px, an upper bound of the chord
    # v dt + a dt^2 / 2 it draws, reached where v and a are parallel
    seg = np.hypot(*((pts[1::2, :2] - pts[0::2, :2]).astype(np.float64) * [W, H]).T)[free]
    assert np.all(seg <= 6.0 + 1e-3) and np.all(seg > 5.0)
    dt = np.diff(np.concatenate([[0.0], t]))[free]
    speed = np.hypot(vx, vy + a * (t[free] - dt))
    np.testing.assert_allclose(speed * dt + 0.5 * a * dt * dt, 6.0, rtol=2e-4)


@pytest.mark.parametrize("restitution", [0.0, 1.0])
def test_a_wall_stops_or_mirrors_the_path(oracle, restitution):
    sim = empty_sim(oracle, bounded=True, restitution=restitution)
    v = sim.trajectory((600.0, 160.0), (300.0, 0.0), steps=40)
    x = live(v)[:, 0].astype(np.float64) * W
    top = x.max()
    assert abs(top - W * 0.99999994) < 1e-3
    after = x[int(np.argmax(x)) + 1:]
    if restitution == 0.0:
        assert np.all(np.abs(after - top) < 1e-3)  # the normal motion stopped
    else:
        assert after.min() < top - 100.0  # it came back


def test_a_torus_seam_splits_the_step(oracle):
    sim = empty_sim(oracle, bounded=False)
    v = sim.trajectory((630.0, 160.0), (300.0, 0.0), steps=10)
    steps = v.reshape(10, 6, 4)
    split = [k for k in range(10) if (steps[k, :, 3] >= 0).sum() == 4]
    assert len(split) == 1
    p = steps[split[0]]
    # the first piece ends on the right edge, the second starts on the left
    assert p[1, 0] == 1.0 and p[2, 0] == 0.0
    assert p[1, 3] == p[2, 3] and p[1, 2] == p[2, 2]


def test_the_path_ends_at_its_maximum_length(oracle):
    sim = empty_sim(oracle, bounded=False)
    # 1024 steps of 6 px would be longer than the viewport diagonal
    v = sim.trajectory((100.0, 100.0), (200.0, 50.0), steps=1024)
    pts = live(v).astype(np.float64)
    seg = np.hypot(*((pts[1::2, :2] - pts[0::2, :2]) * [W, H]).T)
    # torus pieces carry their own lengths; all of them add up to the diagonal
    assert abs(seg.sum() - math.hypot(W, H)) < 0.05
    assert pts[-1, 3] == 1.0
    assert (v[:, 3] < 0).sum() > 0  # the steps after the end are empty
