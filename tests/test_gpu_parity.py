"""GPU parity: the gfx950 engine (through the C ABI) against the FP64 oracle.

Gates (per the engine's contract):
  * per-cell material temperature, time-averaged radiation energy density and
    census radiation energy agree with the FP64 oracle within 3 sigma of the
    16-seed batch variance (tests/stats.py; a 1e-6 relative floor stands for
    the FP32 representation of cell opacities, see stats.zscores);
  * energy conservation per step: exact in integer quanta for the mixed
    engine, and to 1e-12 relative for the material + radiation balance;
  * the CPU replay of the packets' Philox streams reproduces the mixed
    engine's fixed-point tallies bit for bit;
  * the same seed gives bit-identical results (order-independent tallies).
"""
from __future__ import annotations

import concurrent.futures as cf
import math

import numpy as np
import pytest
from scipy import special

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine
from tests.stats import check_ensembles

pytestmark = pytest.mark.gpu

NSEEDS = 16


def oracle_ensemble(oracle, make, seeds):
    def one(s):
        return oracle.run(make(s))
    with cf.ThreadPoolExecutor(max_workers=min(16, len(seeds))) as ex:
        res = list(ex.map(one, seeds))
    return {k: np.array([r[k] for r in res]) for k in ("T", "erad", "census")}


def gpu_ensemble(make, seeds, precision):
    out = {k: [] for k in ("T", "erad", "census")}
    for s in seeds:
        spec = make(s)
        with GpuEngine(spec, precision) as e:
            for _ in range(spec.steps):
                e.step()
            c = e.cells()
        out["T"].append(c["T"])
        out["erad"].append(c["erad"])
        out["census"].append(c["census"])
    return {k: np.array(v) for k, v in out.items()}


def _assert_parity(ref, got, label):
    msgs, ok_all = [], True
    for q in ("T", "erad", "census"):
        summ, ok = check_ensembles(ref[q], got[q], label=f"{label}:{q}")
        msgs.append(str(summ))
        ok_all &= ok
    print("\n".join(msgs))
    assert ok_all, "\n".join(msgs)


C1 = lambda s: P.c1_slab(seed=s, packets_scale=0.1, steps=5)  # noqa: E731
C2 = lambda s: P.c2_cube(seed=7, n=8, packets_scale=0.2, steps=3).with_seed(s)  # noqa: E731
C3 = lambda s: P.c3_step(seed=s, nx=32, nyz=2, packets_scale=0.1, steps=3)  # noqa: E731


@pytest.fixture(scope="module")
def c1_ref(oracle):
    return oracle_ensemble(oracle, C1, range(1000, 1000 + NSEEDS))


@pytest.fixture(scope="module")
def c2_ref(oracle):
    return oracle_ensemble(oracle, C2, range(1000, 1000 + NSEEDS))


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_c1_slab_matches_oracle(c1_ref, precision):
    got = gpu_ensemble(C1, range(NSEEDS), precision)
    _assert_parity(c1_ref, got, f"c1/{precision}")


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_c2_cube_matches_oracle(c2_ref, precision):
    got = gpu_ensemble(C2, range(NSEEDS), precision)
    _assert_parity(c2_ref, got, f"c2/{precision}")


@pytest.fixture(scope="module")
def c3_ref(oracle):
    return oracle_ensemble(oracle, C3, range(1000, 1000 + NSEEDS))


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_c3_random_walk_ddmc_matches_oracle(c3_ref, precision):
    """Step-opacity slab with random walk and DDMC in the thick half."""
    got = gpu_ensemble(C3, range(NSEEDS), precision)
    _assert_parity(c3_ref, got, f"c3/{precision}")
    spec = C3(0)
    with GpuEngine(spec, precision) as e:
        e.step()
        c = e.counters()
    assert c["random_walk_steps"] > 0 and c["ddmc_steps"] > 0, c


def _ledger_ok(L):
    lhs = L["census_in"] + L["emitted"] + L["source"]
    rhs = L["deposited"] + L["escaped"] + L["census_pre"]
    return lhs == rhs and L["census_pre"] == L["census_out"] + L["comb_rem"], (lhs, rhs)


@pytest.mark.parametrize("make", [C1, C2], ids=["c1", "c2"])
def test_mixed_energy_ledger_exact(make):
    spec = make(3)
    with GpuEngine(spec, capi.MIXED) as e:
        for _ in range(spec.steps):
            e.step()
            L = e.ledger()
            ok, detail = _ledger_ok(L)
            assert ok, (L, detail)
            # material + radiation balance in erg, FP64 rounding only
            Q = 2.0 ** L["eq"]
            dmat = L["material_after"] - L["material_before"]
            expect = (L["deposited"] - L["emitted"] + L["comb_rem"]) * Q + L["requant_rem"] + L["clamp"]
            scale = L["material_before"] + (L["census_in"] + L["emitted"] + L["source"]) * Q
            assert abs(dmat - expect) <= 1e-12 * scale, (dmat, expect, scale)


def test_fp64_energy_balance():
    spec = C1(4)
    with GpuEngine(spec, capi.FP64) as e:
        for _ in range(spec.steps):
            e.step()
            L = e.ledger()
            lhs = L["census_in"] + L["emitted"] + L["source"]
            rhs = L["deposited"] + L["escaped"] + L["census_pre"]
            assert abs(lhs - rhs) <= 1e-12 * lhs
            assert L["census_pre"] == L["census_out"] + L["comb_rem"]


@pytest.mark.parametrize("make", [C1, C2, C3], ids=["c1", "c2", "c3"])
def test_replay_reproduces_tallies_bit_exactly(oracle, make):
    spec = make(5)
    with GpuEngine(spec, capi.MIXED) as e:
        for s in range(spec.steps):
            e.keep_transport_inputs(s == spec.steps - 1)
            e.step()
        pk, n, xs, consts = e.transport_inputs()
        got = e.tallies()
        accel = e.accel_data()
        cnt = e.counters()
    ref, outc = oracle.replay_transport(spec, pk, n, xs, consts, accel=accel)
    assert n > 0 and outc[:4].sum() == n
    assert outc[4] == cnt["random_walk_steps"] and outc[5] == cnt["ddmc_steps"]
    for k in ("segs", "dep", "rdep", "pint", "esc"):
        assert np.array_equal(got[k].astype(np.uint64), ref[k].astype(np.uint64)), k


def test_same_seed_is_bit_identical():
    spec = C2(9)
    outs = []
    for _ in range(2):
        with GpuEngine(spec, capi.MIXED) as e:
            for _ in range(spec.steps):
                e.step()
            outs.append((e.cells(), e.segments()))
    for k in ("T", "e", "erad", "census"):
        assert np.array_equal(outs[0][0][k], outs[1][0][k]), k
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_beer_lambert_transmission(precision):
    """Pure absorber slab (f = 1, T0 = 0: no emission) lit by the x- surface
    source: the fraction escaping through x+ is 2 E_3(tau) for the cosine-law
    source (mu = sqrt(U)), the textbook slab transmission."""
    tau = 1.0
    n = 20
    spec = P.ProblemSpec(
        name="beer", x_edges=np.linspace(0, 1.0, n + 1), y_edges=np.array([0.0, 1.0]), z_edges=np.array([0.0, 1.0]),
        sigma_a0=np.full(n, tau), T0=np.zeros(n), dt=1e-7, steps=1, eos_f=1e10, eos_beta=1.0,
        face_kind=(capi.FACE_REFLECT, capi.FACE_ESCAPE) + (capi.FACE_REFLECT,) * 4,
        source_face=capi.XLO, source_T=P.KEV_K, freeze_T4=0, source_per_face=400000,
        new_photons_per_cell=1, comb=0, seed=11)
    with GpuEngine(spec, precision) as e:
        e.step()
        L = e.ledger()
    Q = 2.0 ** L["eq"]
    frac = L["escaped"] / L["source"]
    expect = 2.0 * special.expn(3, tau)
    # census (packets still in flight at the end of the step) is below 1e-3 of the source
    assert L["census_pre"] * Q < 2e-3 * L["source"] * Q
    # binomial-like error of a weight fraction with 4e5 packets
    assert abs(frac - expect) < 4.0 * math.sqrt(expect * (1 - expect) / 4e5) + 2e-3, (frac, expect)


def test_census_export_import_roundtrip():
    spec = C1(6)
    with GpuEngine(spec, capi.MIXED) as e:
        e.step()
        recs = e.export_census()
        n = e.census_size()
        assert len(recs) == n > 0
        w_before = recs["weight"].sum()
        e.import_census(recs)
        assert e.census_size() == n
        recs2 = e.export_census()
        assert np.allclose(recs2["weight"], recs["weight"], rtol=0, atol=0)
        assert np.allclose(recs2["x"], recs["x"], rtol=1e-6, atol=1e-9)
        assert abs(recs2["weight"].sum() - w_before) == 0.0
        e.step()  # continues from the imported census
        assert e.counters()["initial_particles"] == n


def test_c2_full_size_one_step_invariants():
    spec = P.c2_cube(seed=2)
    with GpuEngine(spec, capi.MIXED) as e:
        e.step()
        c = e.counters()
        L = e.ledger()
        segs = e.segments()
    ok, detail = _ledger_ok(L)
    assert ok, detail
    assert c["segments"] == int(segs.sum()) > c["start_particles"]
    assert c["start_particles"] == c["emitted_particles"] + c["boundary_particles"] + c["initial_particles"]
    assert c["boundary_particles"] == 64 * 64 * spec.source_per_face
