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
import json
import math
import os
import subprocess

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


def MG(s):
    """16-cell, 30-group Densmore slab started at 0.3 keV, 3 steps."""
    spec = P.densmore(nx=16, seed=s, steps=3, new_photons=6, boundary_photons=300, comb_nmin=8, comb_factor=1.0,
                      random_walk=False)
    spec.T0 = np.full(16, 0.3 * P.KEV_K)
    return spec


def MG_RW(s):
    """The same with the partial-gray random walk."""
    spec = MG(s)
    spec.with_rw = 1
    return spec


def MG_DDMC(s):
    """The same with DDMC instead (Planck-mean data, photon energy kept: the
    reference without ddmcUseMultigroupPGRW)."""
    spec = MG(s)
    spec.with_ddmc = 1
    return spec


def MG_DDMC_PG(s):
    """Partial-gray DDMC (ddmcUseMultigroupPGRW): data from the diffusive
    groups, packets above the cutoff take analytic steps."""
    spec = MG_DDMC(s)
    spec.ddmc_pgrw = 1
    return spec


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


@pytest.fixture(scope="module")
def mg_ref(oracle):
    return oracle_ensemble(oracle, MG, range(1000, 1000 + NSEEDS))


@pytest.fixture(scope="module")
def mg_rw_ref(oracle):
    return oracle_ensemble(oracle, MG_RW, range(1000, 1000 + NSEEDS))


@pytest.fixture(scope="module")
def mg_ddmc_ref(oracle):
    return oracle_ensemble(oracle, MG_DDMC, range(1000, 1000 + NSEEDS))


@pytest.fixture(scope="module")
def mg_ddmc_pg_ref(oracle):
    return oracle_ensemble(oracle, MG_DDMC_PG, range(1000, 1000 + NSEEDS))


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_multigroup_matches_oracle(mg_ref, precision):
    """30-group transport with energy-dependent opacity and re-emission."""
    got = gpu_ensemble(MG, range(NSEEDS), precision)
    _assert_parity(mg_ref, got, f"mg/{precision}")


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_multigroup_random_walk_matches_oracle(mg_rw_ref, precision):
    """The same with the partial-gray random walk (RadiationIMC.hpp:1008-1060, 1065-1276)."""
    got = gpu_ensemble(MG_RW, range(NSEEDS), precision)
    _assert_parity(mg_rw_ref, got, f"mg_rw/{precision}")


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_multigroup_ddmc_matches_oracle(mg_ddmc_ref, precision):
    """The same with DDMC (RadiationIMC.hpp:1282-1612 without ddmcUseMultigroupPGRW)."""
    got = gpu_ensemble(MG_DDMC, range(NSEEDS), precision)
    _assert_parity(mg_ddmc_ref, got, f"mg_ddmc/{precision}")
    with GpuEngine(MG_DDMC(0), precision) as e:
        e.step()
        c = e.counters()
    assert c["ddmc_steps"] > 0, c


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_multigroup_partial_gray_ddmc_matches_oracle(mg_ddmc_pg_ref, precision):
    """Partial-gray DDMC (RadiationIMC.hpp:1301-1342, 1437-1450)."""
    got = gpu_ensemble(MG_DDMC_PG, range(NSEEDS), precision)
    _assert_parity(mg_ddmc_pg_ref, got, f"mg_ddmc_pg/{precision}")
    with GpuEngine(MG_DDMC_PG(0), precision) as e:
        e.step()
        c = e.counters()
    assert c["ddmc_steps"] > 0, c


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
    return lhs == rhs and L["census_pre"] == L["census_out"] + L["comb_rem"] + L["comb_dropped"], (lhs, rhs)


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


@pytest.mark.parametrize("make", [C1, C2, C3, MG, MG_RW, MG_DDMC, MG_DDMC_PG],
                         ids=["c1", "c2", "c3", "multigroup", "multigroup_rw", "multigroup_ddmc",
                              "multigroup_ddmc_pg"])
def test_replay_reproduces_tallies_bit_exactly(oracle, make):
    spec = make(5)
    with GpuEngine(spec, capi.MIXED) as e:
        for s in range(spec.steps):
            e.keep_transport_inputs(s == spec.steps - 1)
            e.step()
        pk, n, xs, consts = e.transport_inputs()
        got = e.tallies()
        accel = e.accel_data()
        spectrum = e.spectrum_data()
        cnt = e.counters()
    ref, outc = oracle.replay_transport(spec, pk, n, xs, consts, accel=accel, spectrum=spectrum)
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


def test_comb_order_in_a_crowded_cell_is_deterministic():
    """120 000 census packets in one cell of a 2^20-cell mesh.  The comb's
    32-bit sort keys keep 11 bits of a packet's rank there, so ~30-60 packets
    share a key and the tie pass orders them by their full rank (insertion
    sort for runs up to 32, heap sort above).  The transport writes the census
    in a different order from run to run; the combed census, the tallies and
    the energy ledger must come out the same every time."""
    nx, ny, nz = 128, 128, 64
    N = nx * ny * nz
    spec = P.ProblemSpec(
        name="crowd", x_edges=np.linspace(0, 1.0, nx + 1), y_edges=np.linspace(0, 1.0, ny + 1),
        z_edges=np.linspace(0, 0.5, nz + 1), sigma_a0=np.full(N, 1e-3), T0=np.zeros(N), dt=1e-13, steps=1,
        face_kind=(capi.FACE_REFLECT,) * 6, source_face=-1, source_per_face=0, new_photons_per_cell=1,
        comb=1, comb_nmin=20, comb_factor=1.0, seed=5)
    rng = np.random.default_rng(17)
    n = 120_000
    recs = np.zeros(n, dtype=np.ctypeslib.as_array((capi.PacketRecord * 1)()).dtype)
    recs["x"] = rng.random(n) / nx
    recs["y"] = rng.random(n) / ny
    recs["z"] = rng.random(n) * 0.5 / nz
    mu = 2.0 * rng.random(n) - 1.0
    phi = 2.0 * np.pi * rng.random(n)
    st = np.sqrt(1.0 - mu * mu)
    c_light = 2.99792458e10
    recs["vx"], recs["vy"], recs["vz"] = c_light * mu, c_light * st * np.cos(phi), c_light * st * np.sin(phi)
    recs["weight"] = 1e-3
    recs["cell"] = 0
    recs["id"] = np.arange(n)
    outs = []
    for _ in range(3):
        with GpuEngine(spec, capi.MIXED) as e:
            e.import_census(recs)
            e.step()
            L = e.ledger()
            ok, detail = _ledger_ok(L)
            assert ok, detail
            c = e.counters()
            assert c["census_before_comb"] > 50_000 and c["end_particles"] < c["census_before_comb"], c
            outs.append((e.export_census(), e.tallies(), L))
    for k in range(1, len(outs)):
        assert np.array_equal(outs[0][0], outs[k][0])
        for f in outs[0][1]:
            assert np.array_equal(outs[0][1][f], outs[k][1][f]), f
        assert outs[0][2] == outs[k][2]


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


def _run_tallies(spec, precision, **tuning):
    with GpuEngine(spec, precision) as e:
        if tuning:
            e.set_tuning(**tuning)
        for _ in range(spec.steps):
            e.step()
        return e.cells(), e.tallies(), e.counters()


def _voronoi_mg(s):
    return P.voronoi_slab(seed=s, groups=8)


@pytest.mark.parametrize("make", [C2, C3, P.voronoi_slab, _voronoi_mg], ids=["c2", "c3", "voronoi", "voronoi-8g"])
def test_refill_schedule_matches_pass_schedule(make):
    """The lane-refill transport and the one-packet-per-lane pass give
    bit-identical tallies and cell states in the mixed engine (Cartesian and
    polyhedral meshes, gray and multigroup): a packet's history depends only
    on its id, the step and the segment budget, never on the lane."""
    spec = make(12)
    a = _run_tallies(spec, capi.MIXED, schedule=capi.SCHEDULE_PASS)
    b = _run_tallies(spec, capi.MIXED, schedule=capi.SCHEDULE_REFILL, block=128)
    for k in ("T", "e", "erad", "census"):
        assert np.array_equal(a[0][k], b[0][k]), k
    for k in ("dep", "rdep", "pint", "esc", "segs"):
        assert np.array_equal(a[1][k], b[1][k]), k
    assert a[2]["segments"] == b[2]["segments"]
    assert a[2]["end_particles"] == b[2]["end_particles"]


def test_ddmc_default_launch_matches_fixed_launch():
    """A Cartesian problem with DDMC starts in 512-thread blocks refilling
    from 16 chunks per wavefront; its results are those of a fixed 256-thread
    pass launch, bit for bit."""
    spec = C3(12)
    assert spec.with_ddmc
    with GpuEngine(spec, capi.MIXED) as e:
        L = e.launch()
        assert (L["block"], L["chunks"], L["schedule"]) == (512, 16, capi.SCHEDULE_REFILL), L
    a = _run_tallies(spec, capi.MIXED, schedule=capi.SCHEDULE_PASS)
    b = _run_tallies(spec, capi.MIXED)
    assert b[2]["transport_block"] == 512
    for k in ("T", "e", "erad", "census"):
        assert np.array_equal(a[0][k], b[0][k]), k
    for k in ("dep", "rdep", "pint", "esc", "segs"):
        assert np.array_equal(a[1][k], b[1][k]), k
    assert a[2]["segments"] == b[2]["segments"]


@pytest.mark.parametrize("make", [C2, P.voronoi_slab], ids=["c2", "voronoi"])
def test_fp64_refill_default_matches_pass(make):
    """The FP64 engine refills lanes by default; the pass schedule gives the
    same histories, so the same segment counts and census, and tallies equal
    up to the order of their floating-point sums (one step: over several, a
    last-bit difference of the material state may send a packet another way)."""
    spec = make(12)
    spec.steps = 1
    with GpuEngine(spec, capi.FP64) as e:
        assert e.launch()["schedule"] == capi.SCHEDULE_REFILL
    a = _run_tallies(spec, capi.FP64, schedule=capi.SCHEDULE_PASS)
    b = _run_tallies(spec, capi.FP64)
    assert np.array_equal(a[1]["segs"], b[1]["segs"])
    assert a[2]["segments"] == b[2]["segments"] and a[2]["end_particles"] == b[2]["end_particles"]
    for k in ("T", "e", "erad", "census"):
        assert np.allclose(a[0][k], b[0][k], rtol=1e-9, atol=0.0), k


def _tuned_against_fixed(spec, steps):
    fixed, tuned, shapes = [], [], []
    with GpuEngine(spec, capi.MIXED) as e:
        e.set_tuning(block=256)
        for _ in range(steps):
            e.step()
            fixed.append((e.cells(), e.tallies()))
    with GpuEngine(spec, capi.MIXED) as e:
        e.set_tuning(block=0)
        for _ in range(steps):
            e.step()
            tuned.append((e.cells(), e.tallies()))
            c = e.counters()
            shapes.append((c["transport_block"], c["transport_chunks"], c["transport_schedule"], e.launch()["tuning"]))
    return fixed, tuned, shapes


@pytest.mark.parametrize("mesh", ["cartesian", "voronoi"])
def test_online_launch_tuning_keeps_results(mesh):
    """Block size 0: the transport's launch is tuned on line -- the pass and
    refill schedules, then blocks of 128, 256, 512, 512, 256, 128 threads,
    then 1, 2, 4, 8, 16, 16, 8, 4, 2, 1 chunks per wavefront -- and the best
    is kept, after which the schedule watch probes the other schedule 16
    steps later; the cell states and tallies are those of a fixed launch,
    step by step."""
    spec = C2(13) if mesh == "cartesian" else P.voronoi_slab(seed=13)
    steps = 40
    spec.steps = steps
    nphase = 20
    fixed, tuned, shapes = _tuned_against_fixed(spec, steps)
    assert [s for _, _, s, _ in shapes[:4]] == [0, 1, 1, 0], shapes
    assert [b for b, _, _, _ in shapes[4:10]] == [128, 256, 512, 512, 256, 128], shapes
    assert [k for _, k, _, _ in shapes[10:20]] == [1, 2, 4, 8, 16, 16, 8, 4, 2, 1], shapes
    assert all(t for _, _, _, t in shapes[:nphase - 1]) and not any(t for _, _, _, t in shapes[nphase - 1:]), shapes
    # after tuning: the chosen block stays; the watch runs one step in the
    # other schedule 17 steps after tuning ends
    b = shapes[nphase][0]
    assert b in (128, 256, 512) and all(x[0] == b for x in shapes[nphase:]), shapes
    sch = [s for _, _, s, _ in shapes[nphase:]]
    assert len(set(sch[:16])) == 1 and sch[16] != sch[15], shapes
    for (ca, ta), (cb, tb) in zip(fixed, tuned):
        for key in ("T", "e", "erad", "census"):
            assert np.array_equal(ca[key], cb[key]), key
        for k in ("dep", "rdep", "pint", "esc", "segs"):
            assert np.array_equal(ta[k], tb[k]), k


@pytest.mark.parametrize("schedule", [capi.SCHEDULE_PASS, capi.SCHEDULE_REFILL], ids=["pass", "refill"])
def test_small_segment_budget_replays_exactly(oracle, schedule):
    """With 3 segments per launch the packets go through many launches; the
    CPU replay with the same budget reproduces the tallies bit for bit."""
    spec = C3(8)
    with GpuEngine(spec, capi.MIXED) as e:
        e.set_tuning(block=256, segments_per_pass=3, schedule=schedule)
        for s in range(spec.steps):
            e.keep_transport_inputs(s == spec.steps - 1)
            e.step()
        pk, n, xs, consts = e.transport_inputs()
        got = e.tallies()
        accel = e.accel_data()
        cnt = e.counters()
    assert cnt["transport_launches"] > 3
    ref, outc = oracle.replay_transport(spec, pk, n, xs, consts, segments_per_pass=3, accel=accel)
    assert outc[:4].sum() == n
    for k in ("segs", "dep", "rdep", "pint", "esc"):
        assert np.array_equal(got[k].astype(np.uint64), ref[k].astype(np.uint64)), k


def test_beginning_counts_per_cell():
    """Packets per cell at transport start (one atomic per run of a cell's
    packets): they add up to the step's packets, the x- column holds the
    boundary packets, and after the first step the census of each cell."""
    spec = C2(13)
    with GpuEngine(spec, capi.MIXED) as e:
        e.step()
        b1 = e.beginning_counts()
        c1 = e.counters()
        assert int(b1.sum()) == c1["start_particles"]
        nx, ny, nz = spec.shape
        assert (b1.reshape(nx, ny, nz)[0] >= spec.source_per_face).all()
        rec = e.export_census()
        e.step()
        b2 = e.beginning_counts()
        c2 = e.counters()
    assert int(b2.sum()) == c2["start_particles"]
    cen = np.bincount(rec["cell"].astype(np.int64), minlength=spec.ncells)
    assert (b2 >= cen).all()


# ---------------------------------------------------------------- C++ drop-in manager
MANAGER_EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "_build",
                           "manager_slab")


def _run_manager(spec, precision, resident, seeds, user_boundary=False):
    """Runs tests/cpp/manager_slab.cpp (reference-style user code over
    MonteCarloManagerGPU.hpp) on the slab `spec` for each seed
    (`user_boundary`: its own boundary class that makes the x- packets on the
    host instead of SideTemperature)."""
    args = [MANAGER_EXE, str(precision), str(int(resident)), str(spec.steps), str(seeds[0]), str(len(seeds)),
            str(spec.shape[0]), str(spec.source_per_face), str(spec.new_photons_per_cell), str(spec.comb_nmin)]
    if spec.source_xhi:
        args.append(repr(spec.source_T_hi / P.KEV_K))
    elif user_boundary:
        args.append("user")
    print("manager:", " ".join(args), flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    runs = json.loads(out.stdout)["runs"]
    return {k: np.array([r[k] for r in runs]) for k in ("T", "erad", "census")}, runs


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_cpp_manager_matches_oracle(c1_ref, precision):
    """The C++ manager (host-tabulated RadiationIMC plugins, census returned
    to the caller as std::vector<Particle> and passed back every step) agrees
    with the FP64 oracle like the engine does."""
    got, _ = _run_manager(C1(0), precision, False, list(range(NSEEDS)))
    _assert_parity(c1_ref, got, f"c1-manager/{precision}")


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_cpp_manager_with_a_boundary_that_makes_its_packets(c1_ref, precision):
    """A user boundary whose generateNewBoundaryParticles makes the x- source
    packets on the host (as the reference's MarshakBoundary does): the
    manager joins them to every step's sources on the GPU, and the physics is
    the SideTemperature slab's, so it agrees with the same oracle ensemble."""
    got, _ = _run_manager(C1(0), precision, False, list(range(NSEEDS)), user_boundary=True)
    _assert_parity(c1_ref, got, f"c1-manager-user-boundary/{precision}")


GROUPS_EXE = os.path.join(os.path.dirname(MANAGER_EXE), "manager_groups")


def _run_manager_groups(precision, seeds, stratified):
    """tests/cpp/manager_groups.cpp: the MG slab (16 cells, 30 groups, T0 0.3
    keV, 3 steps, 300 boundary packets, 6 photons per cell, comb Nmin 8,
    factor 1) as multigroup user code over the C++ manager."""
    args = [GROUPS_EXE, str(precision), "3", str(seeds[0]), str(len(seeds)), "16", "300", "6", "8", "1.0",
            str(int(stratified))]
    print("manager:", " ".join(args), flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    runs = json.loads(out.stdout)["runs"]
    return {k: np.array([r[k] for r in runs]) for k in ("T", "erad", "census")}, runs


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_cpp_manager_multigroup_matches_oracle(mg_ref, precision):
    """Multigroup user code (a RadiationIMC<..., 30> with its own opacity
    model of the E^-3 family, a multigroup SideTemperature, withEgTimeAvg)
    through the C++ manager agrees with the oracle's MG slab, and its
    per-group tally sums to Erad."""
    got, runs = _run_manager_groups(precision, list(range(NSEEDS)), False)
    _assert_parity(mg_ref, got, f"mg-manager/{precision}")
    assert max(r["eg_gap"] for r in runs) < 1e-9


def test_cpp_manager_stratified_comb_matches_oracle(oracle):
    def make(s):
        spec = MG(s)
        spec.comb, spec.comb_min_per_group = 2, 2
        return spec
    ref = oracle_ensemble(oracle, make, range(1000, 1000 + NSEEDS))
    got, _ = _run_manager_groups(capi.MIXED, list(range(NSEEDS)), True)
    _assert_parity(ref, got, "mg-manager-stratified")


def test_cpp_manager_resident_is_the_engine():
    """With the census resident in HBM the manager's step is the engine's
    step on host-tabulated material data: the opacity and cv are the same
    doubles as the device material model's, and the host EOS differs from the
    device's only in the rounding of its power function, so the packet
    histories coincide and the cell results agree to rounding."""
    spec = C1(3)
    got, runs = _run_manager(spec, capi.MIXED, True, [3])
    with GpuEngine(spec, capi.MIXED) as e:
        for _ in range(spec.steps):
            e.step()
        c = e.cells()
        n = e.census_size()
    assert runs[0]["census_packets"] == n
    assert np.allclose(got["T"][0], c["T"], rtol=1e-12, atol=0)
    assert np.allclose(got["erad"][0], c["erad"], rtol=1e-12, atol=0)
    assert np.allclose(got["census"][0], c["census"], rtol=1e-12, atol=0)


def C1_TWO(s):
    """C1 slab with TwoSidesTemperature: 1 keV on x-, 0.5 keV on x+."""
    spec = C1(s)
    spec.source_xhi, spec.source_T_hi = 1, 0.5 * P.KEV_K
    spec.face_kind = (capi.FACE_ESCAPE, capi.FACE_ESCAPE) + (capi.FACE_REFLECT,) * 4
    return spec


@pytest.fixture(scope="module")
def c1_two_ref(oracle):
    return oracle_ensemble(oracle, C1_TWO, range(1000, 1000 + NSEEDS))


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_two_sided_source_matches_oracle(c1_two_ref, precision):
    """Planck sources on both x faces (TwoSidesTemperature), through the engine
    and through the C++ manager with the TwoSidesTemperature plugin."""
    got = gpu_ensemble(C1_TWO, range(NSEEDS), precision)
    _assert_parity(c1_two_ref, got, f"c1-two/{precision}")
    got_cpp, _ = _run_manager(C1_TWO(0), precision, False, list(range(NSEEDS)))
    _assert_parity(c1_two_ref, got_cpp, f"c1-two-manager/{precision}")
