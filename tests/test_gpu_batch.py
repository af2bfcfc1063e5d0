"""Batched independent instances (storm_unstructured_mesh.cell_instance) in
one engine: every instance of the batch must give, cell by cell, what the
FP64 oracle gives for that instance alone (3 sigma over 16-seed ensembles),
with its own source temperature, photon budget and comb; the mixed engine's
integer ledger stays exact for the whole batch.
"""
from __future__ import annotations

import numpy as np
import pytest

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine
from tests.test_gpu_parity import _assert_parity, _ledger_ok, oracle_ensemble

pytestmark = pytest.mark.gpu

NSEEDS = 16
SEEDS = range(3000, 3000 + NSEEDS)
STEPS = 3
INSTANCES = P.c5_instances(5, seed=2, packets_scale=0.02, cells_range=(12, 48))


def _batch_ensemble(precision):
    out = {k: [] for k in ("T", "erad", "census")}
    for s in SEEDS:
        spec = P.c5_batch(INSTANCES, seed=s, steps=STEPS)
        with GpuEngine(spec, precision) as e:
            for _ in range(STEPS):
                e.step()
            c = e.cells()
        for k in out:
            out[k].append(c[k])
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_batched_instances_match_the_oracle_instance_by_instance(oracle, precision):
    got = _batch_ensemble(precision)
    first = 0
    for k, d in enumerate(INSTANCES):
        n = d["cells"]
        ref = oracle_ensemble(oracle, lambda s: P.c5_instance_spec(d, seed=s, steps=STEPS), SEEDS)
        part = {q: v[:, first:first + n] for q, v in got.items()}
        _assert_parity(ref, part, f"instance {k} ({n} cells, sigma {d['sigma_a']:.3g}, {d['T_keV']:.2f} keV)")
        first += n


def test_batched_ledger_exact():
    spec = P.c5_batch(INSTANCES, seed=7, steps=STEPS)
    with GpuEngine(spec, capi.MIXED) as e:
        for _ in range(STEPS):
            e.step()
            assert _ledger_ok(e.ledger())
        c = e.counters()
    assert c["boundary_particles"] == sum(d["source_packets_per_face"] for d in INSTANCES)
