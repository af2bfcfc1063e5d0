"""GPU stratified comb (imc.comb = 2, StratifiedCombPopulationControl,
population/CombPopulationControl.hpp:256-511) against the oracle:

* the comb alone: a hand-built census imported into a problem without
  sources or opacity and stepped for 1e-20 s (no packet leaves its cell or
  changes energy), so the step's comb sees exactly the imported packets: per
  cell and photon group the GPU keeps the oracle's packet counts exactly and
  the bins' weights, and drops the same weight;
* whole problems with the stratified comb (8-group Voronoi slab, 30-group
  Densmore slab) within 3 sigma of the oracle's 16-seed ensembles;
* the exact integer ledger with the dropped bins as a term of its own.
"""
from __future__ import annotations

import numpy as np
import pytest

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine
from tests.test_gpu_parity import _assert_parity, _ledger_ok, gpu_ensemble, oracle_ensemble

pytestmark = pytest.mark.gpu

NSEEDS = 16
SEEDS = range(5000, 5000 + NSEEDS)
C_LIGHT = 2.99792458e10


def _comb_problem(G, nmin, min_per):
    b = np.geomspace(0.01, 20.0, G + 1) * P.KEV
    return P.ProblemSpec(
        name="strat_comb", x_edges=np.linspace(0.0, 1.0, 5), y_edges=np.array([0.0, 1.0]), z_edges=np.array([0.0, 1.0]),
        sigma_a0=np.zeros(4), T0=np.full(4, 1e-3 * P.KEV_K), dt=1e-20, steps=1,
        face_kind=(capi.FACE_REFLECT,) * 6, source_face=-1, source_per_face=0, new_photons_per_cell=1,
        comb=2, comb_nmin=nmin, comb_factor=2.0, comb_min_per_group=min_per, seed=9, group_bounds=b, nu_pow=0.0,
        nu_ref=P.KEV)


@pytest.mark.parametrize("G,nmin,min_per", [(8, 40, 2), (16, 3, 2)], ids=["all-fit", "dropping"])
def test_comb_alone_matches_the_oracle(oracle, G, nmin, min_per):
    spec = _comb_problem(G, nmin, min_per)
    b = spec.group_bounds
    rng = np.random.default_rng(G)
    n = 6000
    cells = rng.integers(0, 4, n)
    logf = rng.uniform(np.log(b[0]), np.log(b[-1]), n)
    freq = np.exp(logf)
    # keep photon energies away from the bounds (the GPU holds them in FP32)
    near = np.min(np.abs(np.log(freq)[:, None] - np.log(b)[None, :]), axis=1) < 1e-4
    cells, freq = cells[~near], freq[~near]
    n = len(freq)
    w = rng.exponential(1.0, n) * np.where(rng.random(n) < 0.05, 40.0, 1.0) * 1e3
    recs = np.zeros(n, dtype=np.ctypeslib.as_array((capi.PacketRecord * 1)()).dtype)
    recs["x"] = (cells + 0.5) / 4.0
    recs["y"] = recs["z"] = 0.5
    mu = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * np.pi * rng.random(n)
    st = np.sqrt(1.0 - mu * mu)
    recs["vx"], recs["vy"], recs["vz"] = C_LIGHT * mu, C_LIGHT * st * np.cos(phi), C_LIGHT * st * np.sin(phi)
    recs["weight"] = w
    recs["frequency"] = freq
    recs["cell"] = cells
    recs["id"] = np.arange(n)
    with GpuEngine(spec, capi.MIXED) as e:
        e.import_census(recs)
        e.step()
        L = e.ledger()
        ok, detail = _ledger_ok(L)
        assert ok, detail
        out = e.export_census()
    ocnt, ows, odrop = oracle.strat_comb(spec, cells, freq, w)
    g_out = np.clip(np.searchsorted(b, out["frequency"], side="right") - 1, 0, G - 1)
    cnt = np.zeros((4, G))
    ws = np.zeros((4, G))
    np.add.at(cnt, (out["cell"].astype(np.int64), g_out), 1)
    np.add.at(ws, (out["cell"].astype(np.int64), g_out), out["weight"])
    assert np.array_equal(cnt, ocnt), (cnt, ocnt)
    assert np.allclose(ws, ows, rtol=1e-9, atol=1e-9 * w.sum())
    Q = 2.0 ** L["eq"]
    assert L["comb_dropped"] * Q == pytest.approx(odrop, rel=1e-9, abs=1e-9 * w.sum())
    if G == 16:
        assert odrop > 0.0


def V_STRAT(s):
    spec = P.voronoi_slab(seed=s, groups=8)
    spec.comb, spec.comb_nmin, spec.comb_min_per_group = 2, 6, 2
    return spec


def MG_STRAT(s):
    spec = P.densmore(nx=16, seed=s, steps=3, new_photons=6, boundary_photons=300, comb_nmin=8, comb_factor=1.0,
                      random_walk=False)
    spec.T0 = np.full(16, 0.3 * P.KEV_K)
    spec.comb, spec.comb_min_per_group = 2, 2
    return spec


@pytest.mark.parametrize("make", [V_STRAT, MG_STRAT], ids=["voronoi-8g", "densmore-30g"])
@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_stratified_comb_problems_match_the_oracle(oracle, make, precision):
    ref = oracle_ensemble(oracle, make, SEEDS)
    _assert_parity(ref, gpu_ensemble(make, SEEDS, precision), f"strat:{make.__name__}:{precision}")


def test_stratified_comb_ledger_exact_with_dropped_bins():
    spec = MG_STRAT(3)
    spec.comb_nmin = 2
    dropped = 0
    with GpuEngine(spec, capi.MIXED) as e:
        for _ in range(spec.steps):
            e.step()
            L = e.ledger()
            ok, detail = _ledger_ok(L)
            assert ok, detail
            dropped += L["comb_dropped"]
            Q = 2.0 ** L["eq"]
            dmat = L["material_after"] - L["material_before"]
            expect = (L["deposited"] - L["emitted"] + L["comb_rem"]) * Q + L["requant_rem"] + L["clamp"]
            scale = L["material_before"] + (L["census_in"] + L["emitted"] + L["source"]) * Q
            assert abs(dmat - expect) <= 1e-12 * scale, (dmat, expect, scale)
    assert dropped > 0
