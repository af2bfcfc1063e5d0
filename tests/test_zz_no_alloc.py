"""storm_gpu_step_noalloc against storm_gpu_step, through the C ABI.

The no-allocation step carves every transient buffer out of the library's
fixed device scratch; it must give the same step.  Two engines per problem
run the same seeded problem, one with each entry point.  The mixed engine's
tallies are integers, so its per-cell results, tallies, counters, ledger and
census must be identical bit for bit over three steps.  The FP64 engine adds
its tallies with floating-point atomics, whose order (and so the last bit)
differs from run to run even between two calls of the same entry point: its
first step is compared, counts and census exactly and floating-point results
to 1e-12 relative.  The calls alternate between a larger and a smaller
problem, so each no-allocation call finds the scratch holding the other
problem's data.
"""
from __future__ import annotations

import numpy as np
import pytest

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine, StormError

pytestmark = pytest.mark.gpu


def _state(e):
    c = e.cells()
    t = e.tallies()
    k = e.counters()
    L = e.ledger()
    cen = e.export_census()
    counts = {n: k[n] for n in ("initial_particles", "prestep_particles", "start_particles", "end_particles",
                                "emitted_particles", "boundary_particles", "census_before_comb",
                                "leaving_particles", "segments", "transport_passes", "random_walk_steps",
                                "ddmc_steps", "energy_quantum_exp")}
    return c, t, counts, L, cen


def _assert_same(a, b, what, rtol=0.0):
    ca, ta, ka, La, na = a
    cb, tb, kb, Lb, nb = b
    assert ka == kb, (what, ka, kb)
    assert len(na) == len(nb) and np.array_equal(na, nb), (what, "census", len(na), len(nb))
    if rtol == 0.0:
        for name in ca:
            assert np.array_equal(ca[name], cb[name]), (what, "cells", name)
        for name in ta:
            assert np.array_equal(ta[name], tb[name]), (what, "tally", name)
        assert La == Lb, (what, La, Lb)
    else:
        for name in ca:
            assert np.allclose(ca[name], cb[name], rtol=rtol, atol=0.0), (what, "cells", name)
        assert np.array_equal(ta["segs"], tb["segs"]), (what, "segments")
        for name in La:
            x, y = float(La[name]), float(Lb[name])
            assert abs(x - y) <= rtol * max(abs(x), abs(y)) + 2.0, (what, "ledger", name, x, y)


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64])
def test_noalloc_step_matches_step_on_two_problems(precision):
    big = P.c1_slab(seed=11, packets_scale=0.3, steps=3)
    small = P.c3_step(seed=12, nx=32, nyz=2, packets_scale=0.002, steps=3)
    steps = 3 if precision == capi.MIXED else 1
    rtol = 0.0 if precision == capi.MIXED else 1e-12
    pairs = {}
    for name, spec in (("big", big), ("small", small)):
        pairs[name] = (GpuEngine(spec, precision), GpuEngine(spec, precision))
    try:
        for step in range(steps):
            for name in ("big", "small"):  # the second call of each round has the smaller input
                ref, na = pairs[name]
                ref.step()
                na.step_noalloc()
                sa, sb = _state(ref), _state(na)
                assert sa[2]["start_particles"] > 0
                _assert_same(sa, sb, f"{name} step {step + 1}", rtol)
                print(f"precision {precision} {name} step {step + 1}: {sa[2]['start_particles']} packets, "
                      f"{sa[2]['segments']} segments, identical", flush=True)
        assert pairs["big"][0].counters()["start_particles"] > pairs["small"][0].counters()["start_particles"]
    finally:
        for r, n in pairs.values():
            r.close()
            n.close()


def test_noalloc_step_rejects_a_step_beyond_the_scratch_and_keeps_the_state():
    """~5e5 boundary packets need more than the fixed scratch: the call fails
    with status 4 and the engine's next (allocating) step equals a fresh
    engine's first step."""
    spec = P.c1_slab(seed=13, packets_scale=20.0, steps=2)
    with GpuEngine(spec, capi.MIXED) as ref, GpuEngine(spec, capi.MIXED) as na:
        with pytest.raises(StormError, match="error 4"):
            na.step_noalloc()
        ref.step()
        na.step()
        _assert_same(_state(ref), _state(na), "after a rejected call")


def test_noalloc_step_rejects_the_debug_input_copy():
    spec = P.c1_slab(seed=14, packets_scale=0.05, steps=1)
    with GpuEngine(spec, capi.MIXED) as e:
        e.keep_transport_inputs(True)
        with pytest.raises(StormError, match="error 1"):
            e.step_noalloc()
