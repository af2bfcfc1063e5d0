"""The transport's launch policy (storm_amd/csrc/launch_policy.h: the on-line
tuner and the schedule watch) on made-up transport rates, without a GPU:
the tuner's candidates (schedules, then blocks, then chunks, forwards and
backwards) and its choice on a growing workload; the watch keeping the
faster schedule with probes backing off 16, 32, 64, 128 steps; switching at
the first probe after the other schedule became faster and probing again 16
steps later; no switch between equal schedules on a workload growing 3% per
step (the probe is compared with the mean of its neighbours); no probes in a
fixed launch."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_launch_policy_decisions():
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "launch_policy")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/_build/launch_policy is not built (run __graft_entry__.build())")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count(" ok") == 5, out.stdout
