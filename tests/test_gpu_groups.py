"""GPU parity of the per-group time-averaged radiation energy (getEgTimeAvg,
radiation/RadiationIMC.hpp:433-434; storm_gpu_get_group_erad) and of the
64-group Voronoi problem in C4's form, against the FP64 oracle:

* a Voronoi slab in 8 groups whose packets take analytic, partial-gray
  random-walk and DDMC steps: per cell and group within 3 sigma of 16-seed
  ensembles, in both precisions, and summed over the groups the cell's
  Erad_time_avg;
* the C4 set-up (64 log-spaced groups over 0.01-20 keV, sigma = 10 nu^-3
  T^-1/2, rho = 1 + 0.5 sin(2 pi x)) on a 300-cell tessellation at a small
  packet budget: T, Erad, census and the per-group tally;
* the getter refuses a problem created without the tally.
"""
from __future__ import annotations

import numpy as np
import pytest

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine
from tests.stats import check_ensembles

pytestmark = pytest.mark.gpu

NSEEDS = 16
SEEDS = range(4000, 4000 + NSEEDS)


def V_MG(s):
    spec = P.voronoi_slab(seed=s, groups=8, thick=True, packets_scale=0.5)
    spec.eg_time_avg = True
    return spec


C4_MESH = None


def C4_SMALL(s):
    global C4_MESH
    if C4_MESH is None:
        C4_MESH = P.voronoi_mesh(P.voronoi_points(300, 7))
    spec = P.c4_voronoi(seed=s, mesh=C4_MESH, steps=3, packets_scale=0.005)
    spec.eg_time_avg = True
    return spec


def _oracle(oracle, make):
    res = [oracle.run(make(s)) for s in SEEDS]
    return {k: np.array([r[k].reshape(-1) for r in res]) for k in ("T", "erad", "census", "eg")}


def _gpu(make, precision, groups):
    out = {k: [] for k in ("T", "erad", "census", "eg")}
    for s in SEEDS:
        spec = make(s)
        with GpuEngine(spec, precision) as e:
            for _ in range(spec.steps):
                e.step()
            c = e.cells()
            eg = e.group_erad(groups)
        for k in ("T", "erad", "census"):
            out[k].append(c[k])
        out["eg"].append(eg.reshape(-1))
        # the per-group tally splits the cell's Erad_time_avg
        assert np.abs(eg.sum(1) - c["erad"]).max() <= 1e-9 * c["erad"].max()
    return {k: np.array(v) for k, v in out.items()}


def _check(ref, got, label, keys):
    msgs, ok_all = [], True
    for q in keys:
        summ, ok = check_ensembles(ref[q], got[q], label=f"{label}:{q}")
        msgs.append(str(summ))
        ok_all &= ok
    print("\n".join(msgs))
    assert ok_all, "\n".join(msgs)


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_group_erad_matches_the_oracle(oracle, precision):
    ref = _oracle(oracle, V_MG)
    got = _gpu(V_MG, precision, 8)
    _check(ref, got, f"voronoi-groups:{precision}", ("erad", "eg"))


def test_64_group_voronoi_matches_the_oracle(oracle):
    ref = _oracle(oracle, C4_SMALL)
    got = _gpu(C4_SMALL, capi.MIXED, 64)
    _check(ref, got, "c4-small", ("T", "erad", "census", "eg"))


def test_group_erad_needs_the_tally():
    spec = P.voronoi_slab(seed=1, groups=4)
    with GpuEngine(spec, capi.MIXED) as e:
        e.step()
        with pytest.raises(RuntimeError, match="per-group tally is off"):
            e.group_erad(4)
