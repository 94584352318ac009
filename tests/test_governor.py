"""SubstepGovernor: the reference's own cases (T/SubstepGovernorTest.kt:1-58)."""
import math

from gravity_amd import SubstepGovernor


def settle(governor, gpu_millis):
    for _ in range(SubstepGovernor.SETTLE_FRAMES):
        governor.update(gpu_millis)


def test_a_frame_over_budget_runs_proportionally_fewer_substeps():
    assert SubstepGovernor(16, 10.0).update(12.0) == 13


def test_never_fewer_than_half():
    governor = SubstepGovernor(16, 10.0)
    assert governor.update(100.0) == 8
    settle(governor, 100.0)
    assert governor.update(100.0) == 8


def test_waits_for_a_change_to_show_before_the_next():
    governor = SubstepGovernor(16, 10.0)
    governor.update(12.0)
    for _ in range(SubstepGovernor.SETTLE_FRAMES):
        assert governor.update(20.0) == 13
    assert governor.update(20.0) == 8


def test_comes_back_one_at_a_time_well_under_budget():
    governor = SubstepGovernor(16, 10.0)
    governor.update(30.0)
    settle(governor, 5.0)
    assert governor.update(5.0) == 9
    settle(governor, 9.0)
    assert governor.update(9.0) == 9


def test_within_budget_nothing_changes():
    governor = SubstepGovernor(16, 10.0)
    assert governor.update(9.9) == 16
    assert governor.update(math.nan) == 16


def test_no_budget_keeps_every_substep():
    assert SubstepGovernor(16, 0.0).update(50.0) == 16


def test_reset_restores_the_maximum():
    governor = SubstepGovernor(15, 10.0)
    assert governor.update(100.0) == 8
    governor.reset()
    assert governor.substeps == 15
    assert governor.update(9.0) == 15
