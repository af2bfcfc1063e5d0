"""The C++ drop-in's plugin surface, checked without a GPU."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_plugin_surface_checks():
    """tests/cpp/plugin_checks: the reference's RadiationIMC constructor order
    (parameters, traits, position sampler, seed), EOS calls with the traits'
    tracers, rejection of opacity models whose direction overrides or
    Compton flag the GPU cannot honour, and, in multigroup, acceptance of
    sigma(T) E^-p models with p found from the model and rejection of other
    energy laws (no GPU needed: construction only)."""
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "plugin_checks")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/_build/plugin_checks is not built (run __graft_entry__.build())")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count(" ok") == 9, out.stdout
