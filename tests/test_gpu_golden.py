"""End-to-end physics against the reference's own golden data.

The reference's Marshak-wave regression examples
(examples/marshak_wave/MarshakCommon.hpp:273-418) run serial gray IMC with
the random walk and comb to t = 1 ns and compare the gas temperature with a
radiation-diffusion profile (tests/golden/marshak_wave_N_reference.txt,
copied from examples/marshak_wave_N/reference.txt) by a relative L1 that must
stay below 0.10 to print PASS (:404-413).  The engine runs the same problems,
the same time-step schedule and the same drive.  Problems 1, 3 and 4 must
pass the reference's bar.  Problem 2 (kappa_P = kappa_R = 100 (T/keV)^-3, the
steepest front) does not, and neither does the reference's own algorithm:
the FP64 oracle, which restates it, gives relative L1 = 0.61, 0.30 and 0.78
on seeds 1-3 (oracle/pyoracle.run_schedule over the reference's 7912 steps).
More than 95% of that L1 comes from the diffusion profile's cold foot (points
below 0.01 keV, where it divides by T_ref down to 1e-3 keV): the IMC front
is one to two 1/1280 cm cells wide at 0.05-0.1 keV where the diffusion front
has already fallen to the floor, and its position jitters by a fraction of a
cell from seed to seed.  For problem 2 the test asserts what the oracle
satisfies: the heated region within 5% of the profile and the half-maximum
front within 4 cells.
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
    """The reference prints PASS for L1 < 0.10 and WARN above it
    (examples/marshak_wave/MarshakCommon.hpp:404-413).  Problems 1, 3 and 4
    must PASS; every problem must keep the heated region within 5% and the
    half-maximum front within 4 cells (problem 2: see the module docstring)."""
    spec, c = run_marshak(problem, capi.MIXED)
    ref = np.loadtxt(os.path.join(GOLDEN, f"marshak_wave_{problem}_reference.txt"))
    x = spec.centers()[0]
    l1 = P.marshak_l1(x, c["T"], ref)
    l1_hot, shift, h = P.marshak_shape(x, c["T"], ref)
    status = "PASS" if 0.0 <= l1 < 0.10 else "WARN"
    print(f"marshak {problem}: Tgas relative L1 = {l1:.4f} ({status}, reference bar 0.10); "
          f"heated-region L1 = {l1_hot:.4f}; front shift = {shift / h:+.2f} cells", flush=True)
    assert l1 >= 0.0
    if problem != 2:
        assert l1 < 0.10, f"marshak {problem}: relative L1 {l1:.4f} above the reference's PASS bar"
    assert l1_hot < 0.05
    assert abs(shift) <= 4.0 * h


def densmore_csv():
    rows = []
    with open(os.path.join(GOLDEN, "densmore2012_fig4_mc.csv")) as fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            x, t = line.split(",")
            rows.append((float(x), float(t)))
    return np.array(rows)


def densmore_l1(x_sim, T_sim_K, ref):
    """The L1 of examples/densmore2012/main.cpp:321-354: mean |T_sim - T_ref| in
    keV over the reference points, T_sim linearly interpolated."""
    total, cnt = 0.0, 0
    for xr, Tr in ref:
        j = 0
        while j + 1 < len(x_sim) and x_sim[j + 1] < xr:
            j += 1
        if j + 1 >= len(x_sim):
            continue
        fr = (xr - x_sim[j]) / (x_sim[j + 1] - x_sim[j])
        Ts = T_sim_K[j] + fr * (T_sim_K[j + 1] - T_sim_K[j])
        total += abs(Ts / P.KEV_K - Tr)
        cnt += 1
    return total / cnt


def test_densmore_multigroup_matches_reference_profile():
    """Densmore et al. (2012) Fig. 4 (examples/densmore2012/main.cpp): 30-group
    IMC through a step in opacity, 200 steps of 5 ps; material temperature at
    1 ns against the digitised Milagro curve, absolute L1 below 0.10 keV.
    The random walk is on, partial-gray as the reference runs it
    (main.cpp:208-209; problems.densmore defaults to random_walk=True)."""
    spec = P.densmore(seed=3)
    t0 = time.perf_counter()
    with GpuEngine(spec, capi.MIXED) as e:
        for i in range(spec.steps):
            e.step()
            if i % 40 == 0:
                c = e.counters()
                print(f"densmore: step {i}, census {e.census_size()}, segments {c['segments']}, "
                      f"{time.perf_counter() - t0:.1f} s", flush=True)
        c = e.cells()
    l1 = densmore_l1(spec.centers()[0], c["T"], densmore_csv())
    print(f"densmore: Tgas L1 = {l1:.4f} keV (reference bar 0.10)")
    assert l1 < 0.10
