"""The engine's per-step limits, on the GPU:

* the two-limb tallies' bound (a cell with 2^28 or more segments in one
  step): reached on purpose by one strongly scattering cell (every kernel
  still ends normally; the check runs after the transport), the step fails
  with status 5 before postStep commits anything -- the material state is
  the one before the step -- and the engine refuses further steps;
* the full-length configurations stay below every limit: the C3 step slab
  for its 100 steps at ~2e7 packets per step and the 64^3 cube as stated
  for its 50 steps, with the exact ledger every step.
"""
from __future__ import annotations

import numpy as np
import pytest

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine, StormError
from tests.test_gpu_parity import _ledger_ok

pytestmark = pytest.mark.gpu


def test_tally_bound_stops_the_step_before_postStep():
    # one cell of 0.1 cm, sigma_s = 1e6 / cm: a boundary packet scatters
    # ~1.5e5 times in its c dt U of flight (0.15 cm on average); 5000 of them
    # make ~7.5e8 segments in the cell, above 2^28
    spec = P.ProblemSpec(
        name="bound", x_edges=np.array([0.0, 0.1]), y_edges=np.array([0.0, 0.1]), z_edges=np.array([0.0, 0.1]),
        sigma_a0=np.full(1, 1e-6), sigma_s0=np.full(1, 1e6), T0=np.full(1, 1e-3 * P.KEV_K), dt=1e-11, steps=1,
        face_kind=(capi.FACE_REFLECT,) * 6, source_face=capi.XLO, source_T=P.KEV_K, source_per_face=5000,
        new_photons_per_cell=1, comb=0, seed=3)
    with GpuEngine(spec, capi.MIXED) as e:
        before = e.cells()
        with pytest.raises(StormError, match="2\\^28"):
            e.step()
        after = e.cells()
        assert np.array_equal(before["T"], after["T"]) and np.array_equal(before["e"], after["e"])
        with pytest.raises(StormError, match="earlier step"):
            e.step()


def _run_checked(spec, steps):
    worst = 0
    with GpuEngine(spec, capi.MIXED) as e:
        for _ in range(steps):
            e.step()
            ok, detail = _ledger_ok(e.ledger())
            assert ok, detail
            worst = max(worst, int(e.segments().max()))
    return worst


def test_full_c3_stays_within_the_limits():
    spec = P.c3_step(seed=2)
    worst = _run_checked(spec, spec.steps)
    assert worst < (1 << 28), worst


def test_full_c2_stays_within_the_limits():
    spec = P.c2_cube(seed=2, literal=True, source_per_face=2400, comb_nmin=8)
    worst = _run_checked(spec, spec.steps)
    assert worst < (1 << 28), worst
