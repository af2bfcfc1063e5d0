"""Batched independent instances (storm_unstructured_mesh.cell_instance) in
one engine: every instance of the batch must give, cell by cell, what the
FP64 oracle gives for that instance alone (3 sigma over 16-seed ensembles),
with its own source temperature, photon budget and comb; the mixed engine's
integer ledger stays exact for the whole batch and each instance's ledger
balances in erg.  Instances whose energies differ by many orders of magnitude
keep the energy quantum they would have alone: a cold instance batched with a
hot one emits what it emits alone and still matches the oracle.
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
            ok, detail = _ledger_ok(e.ledger())
            assert ok, detail
        c = e.counters()
    assert c["boundary_particles"] == sum(d["source_packets_per_face"] for d in INSTANCES)


def _inst(cells, sigma, T_keV, packets):
    return dict(cells=cells, sigma_a=sigma, T_keV=T_keV, packets=packets,
                source_packets_per_face=max(1, int(round(0.5 * packets))),
                new_photons_per_cell=max(1, int(round(0.25 * packets / cells))),
                comb_nmin=max(1, int(round(0.25 * packets / cells))), comb_factor=1.0)


# A hot instance (5 keV source, cells of 1/20 cm) and a cold, thin one (0.3
# keV, cells of 1/100 cm, sigma_a 0.1): their boundary sources differ by 2^21.
# A cold-cell emission packet of COLD carries ~2^-24.5 erg: above COLD's own
# quantum (2^-39) and below the one quantum of the pair's total (2^-18).
HOT = _inst(20, 3.0, 5.0, 2e4)
COLD = _inst(100, 0.1, 0.3, 4e3)


def test_cold_instance_beside_a_hot_one_matches_the_oracle(oracle):
    got = {k: [] for k in ("T", "erad", "census")}
    for s in SEEDS:
        with GpuEngine(P.c5_batch([HOT, COLD], seed=s, steps=STEPS), capi.MIXED) as e:
            for _ in range(STEPS):
                e.step()
            c = e.cells()
        for k in got:
            got[k].append(c[k][HOT["cells"]:])
    got = {k: np.array(v) for k, v in got.items()}
    ref = oracle_ensemble(oracle, lambda s: P.c5_instance_spec(COLD, seed=s, steps=STEPS), SEEDS)
    _assert_parity(ref, got, "cold instance beside a hot one")


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_instances_keep_their_own_quantum(precision):
    insts = [HOT, COLD] + P.c5_instances(6, seed=5, packets_scale=0.05, cells_range=(20, 400))
    with GpuEngine(P.c5_batch(insts, seed=1, steps=1), precision) as e:
        e.step()
        batch_emitted = e.counters()["emitted_particles"]
        led = [e.ledger(k) for k in range(len(insts))]
    alone, eq_alone = 0, []
    for k, d in enumerate(insts):
        with GpuEngine(P.c5_instance_spec(d, seed=1, steps=1), precision) as e:
            e.step()
            alone += e.counters()["emitted_particles"]
            eq_alone.append(e.ledger()["eq"])
    # the photon budget of a cell is floor(share); sums taken in another order
    # may move a share across an integer
    assert abs(batch_emitted - alone) <= 1e-3 * alone, (batch_emitted, alone)
    for k in range(len(insts)):
        assert abs(led[k]["eq"] - eq_alone[k]) <= 1, (k, led[k]["eq"], eq_alone[k])


def test_instance_ledgers_balance_in_erg():
    insts = [HOT, COLD] + P.c5_instances(4, seed=6, packets_scale=0.05, cells_range=(20, 200))
    with GpuEngine(P.c5_batch(insts, seed=2, steps=STEPS), capi.MIXED) as e:
        for _ in range(STEPS):
            e.step()
            assert _ledger_ok(e.ledger())[0]
            for k in range(len(insts)):
                L = e.ledger(k)
                ok, detail = _ledger_ok(L)
                assert ok, (k, L, detail)
                Q = 2.0 ** L["eq"]
                dmat = L["material_after"] - L["material_before"]
                expect = (L["deposited"] - L["emitted"] + L["comb_rem"]) * Q + L["requant_rem"] + L["clamp"]
                scale = L["material_before"] + (L["census_in"] + L["emitted"] + L["source"]) * Q
                assert abs(dmat - expect) <= 1e-12 * scale, (k, dmat, expect, scale)
        # the exported census is in erg: per instance it sums to census_out
        rec = e.export_census()
        first = np.cumsum([0] + [d["cells"] for d in insts])
        for k in range(len(insts)):
            L = e.ledger(k)
            sel = (rec["cell"] >= first[k]) & (rec["cell"] < first[k + 1])
            want = L["census_out"] * 2.0 ** L["eq"]
            assert rec["weight"][sel].sum() == pytest.approx(want, rel=1e-12, abs=1e-300), k
