"""The benchmark's host side: its one-GPU guard, the workloads it names and
the traffic figures it reads from the committed counter summary (no GPU)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("kernel", sorted(bench.BYTES_PER_UNIT))
def test_every_roofline_kernel_has_measured_traffic(kernel):
    t = bench.measured_traffic(kernel)
    assert t is not None and t > 0.0


def test_ensemble_workload_is_the_documented_one():
    members, cfgs = bench.ensemble_configs()
    assert len(members) == len(cfgs) == 64
    assert sum(m[0] for m in members) == 1_836_607
    for n, g, b, mat in members:
        assert 1000 <= n < 100_000
        assert g in (64, 128, 256, 512)
        assert b.shape == (n, 4) and mat.shape[0] == n
