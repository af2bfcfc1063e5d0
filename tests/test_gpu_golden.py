"""End-to-end physics against the reference's own golden data.

The reference's Marshak-wave regression examples
(examples/marshak_wave/MarshakCommon.hpp:273-418) run serial gray IMC with
the random walk and comb to t = 1 ns and compare the gas temperature with a
radiation-diffusion profile (tests/golden/marshak_wave_N_reference.txt,
copied from examples/marshak_wave_N/reference.txt) by a relative L1 that must
stay below 0.10 to print PASS (:404-413).  The engine runs the same problems,
the same time-step schedule and the same drive, and must pass the same bar.
"""
from __future__ import annotations

import os
import time

import numpy as np
import pytest

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_marshak(problem, precision, seed=1):
    spec = P.marshak(problem, seed=seed)
    t0 = time.perf_counter()
    with GpuEngine(spec, precision) as e:
        for i, (t_now, dt) in enumerate(P.marshak_schedule(problem)):
            e.set_source_temperature(P.marshak_bath_T(problem, t_now))
            e.step(dt)
            if i % 2000 == 0:
                print(f"marshak {problem}: step {i}, census {e.census_size()}, "
                      f"{time.perf_counter() - t0:.1f} s", flush=True)
        c = e.cells()
    return spec, c


@pytest.mark.parametrize("problem", [1, 2, 3, 4])
def test_marshak_wave_matches_reference_profile(problem):
    spec, c = run_marshak(problem, capi.MIXED)
    ref = np.loadtxt(os.path.join(GOLDEN, f"marshak_wave_{problem}_reference.txt"))
    x = spec.centers()[0]
    l1 = P.marshak_l1(x, c["T"], ref)
    print(f"marshak {problem}: Tgas relative L1 = {l1:.4f} (reference bar 0.10)")
    assert 0.0 <= l1 < 0.10
