"""GPU parity of the polyhedral-mesh path (Voronoi tessellations) against the
FP64 oracle on the same meshes, with the Cartesian parity tests' gates:
per-cell temperature, time-averaged radiation energy and census energy
within 3 sigma of 16-seed ensembles, the exact integer energy ledger of the
mixed engine, Beer-Lambert transmission through a Voronoi slab, and a
Cartesian mesh written as polyhedra against the Cartesian engine.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import special

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine
from tests.test_gpu_parity import _assert_parity, _ledger_ok, gpu_ensemble, oracle_ensemble

pytestmark = pytest.mark.gpu

NSEEDS = 16
SEEDS = range(2000, 2000 + NSEEDS)

V1 = lambda s: P.voronoi_slab(seed=s)  # noqa: E731
V_THICK = lambda s: P.voronoi_slab(seed=s, thick=True)  # noqa: E731
V_MG = lambda s: P.voronoi_slab(seed=s, groups=8)  # noqa: E731


@pytest.fixture(scope="module")
def v1_ref(oracle):
    return oracle_ensemble(oracle, V1, SEEDS)


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_voronoi_slab_matches_oracle(v1_ref, precision):
    _assert_parity(v1_ref, gpu_ensemble(V1, SEEDS, precision), f"voronoi:{precision}")


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_voronoi_random_walk_ddmc_matches_oracle(oracle, precision):
    ref = oracle_ensemble(oracle, V_THICK, SEEDS)
    got = gpu_ensemble(V_THICK, SEEDS, precision)
    _assert_parity(ref, got, f"voronoi-thick:{precision}")
    with GpuEngine(V_THICK(1), precision) as e:
        for _ in range(3):
            e.step()
        c = e.counters()
    assert c["random_walk_steps"] > 0 and c["ddmc_steps"] > 0, c


@pytest.mark.parametrize("precision", [capi.MIXED, capi.FP64], ids=["mixed", "fp64"])
def test_voronoi_multigroup_matches_oracle(oracle, precision):
    ref = oracle_ensemble(oracle, V_MG, SEEDS)
    _assert_parity(ref, gpu_ensemble(V_MG, SEEDS, precision), f"voronoi-mg:{precision}")


@pytest.mark.parametrize("make", [V1, V_THICK], ids=["v1", "thick"])
def test_voronoi_energy_ledger_exact(make):
    spec = make(5)
    with GpuEngine(spec, capi.MIXED) as e:
        for _ in range(spec.steps):
            e.step()
            ok, detail = _ledger_ok(e.ledger())
            assert ok, detail


@pytest.mark.parametrize("tau", [0.5, 2.0])
def test_voronoi_beer_lambert_transmission(tau):
    lo, hi = (0.0, 0.0, 0.0), (1.0, 0.4, 0.4)
    m = P.voronoi_mesh(P.voronoi_points(150, 9, lo, hi), lo, hi)
    N = m.ncells
    spec = P.ProblemSpec(
        name="beer_voronoi", x_edges=np.array([0.0, 1.0]), y_edges=np.array([0.0, 0.4]),
        z_edges=np.array([0.0, 0.4]), sigma_a0=np.full(N, tau), T0=np.zeros(N), dt=1e-7, steps=1, eos_f=1e10,
        eos_beta=1.0, face_kind=(capi.FACE_REFLECT, capi.FACE_ESCAPE) + (capi.FACE_REFLECT,) * 4,
        source_face=capi.XLO, source_T=P.KEV_K, freeze_T4=0, source_per_face=20000, new_photons_per_cell=1,
        comb=0, seed=11, umesh=m)
    with GpuEngine(spec, capi.MIXED) as e:
        e.step()
        L = e.ledger()
        n = e.counters()["boundary_particles"]
    frac = L["escaped"] / L["source"]
    expect = 2.0 * special.expn(3, tau)
    assert n > 200000
    assert abs(frac - expect) < 4.0 * math.sqrt(expect * (1 - expect) / n) + 1e-3, (frac, expect)


def test_cartesian_cells_as_polyhedra_match_the_cartesian_engine():
    def poly(s):
        spec = P.c1_slab(seed=s, packets_scale=0.1, steps=4, nx=20)
        spec.umesh = P.cartesian_umesh(spec.x_edges, spec.y_edges, spec.z_edges)
        return spec
    cart = gpu_ensemble(lambda s: P.c1_slab(seed=s, packets_scale=0.1, steps=4, nx=20), SEEDS, capi.MIXED)
    _assert_parity(cart, gpu_ensemble(poly, SEEDS, capi.MIXED), "cartesian-as-polyhedra")
