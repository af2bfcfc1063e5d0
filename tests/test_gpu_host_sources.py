"""Packets made on the host for a step (storm_gpu_queue_source_packets, what
a user boundary's generateNewBoundaryParticles returns;
radiation/RadiationIMC.hpp:1683-1692):

* a cosine-law source at x = 0 of a pure absorber made of such packets,
  each with its own time left, on a Cartesian slab and on a Voronoi slab:
  the escaped fraction at x = 1 is 2 E_3(tau);
* the exact ledger counts them as boundary source, weight for weight on the
  step's quantum, and the next step has no packets left in the queue.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import special

from storm_amd import capi
from storm_amd import problems as P
from storm_amd.engine import GpuEngine
from tests.test_gpu_parity import _ledger_ok

pytestmark = pytest.mark.gpu

C_LIGHT = 2.99792458e10


def _records(n, rng, cells_at_x0, ylo=0.0, yhi=0.4, zlo=0.0, zhi=0.4):
    recs = np.zeros(n, dtype=np.ctypeslib.as_array((capi.PacketRecord * 1)()).dtype)
    recs["x"] = 0.0
    recs["y"] = ylo + (yhi - ylo) * rng.random(n)
    recs["z"] = zlo + (zhi - zlo) * rng.random(n)
    mu = np.sqrt(rng.random(n))
    phi = 2.0 * np.pi * rng.random(n)
    st = np.sqrt(1.0 - mu * mu)
    recs["vx"], recs["vy"], recs["vz"] = C_LIGHT * mu, C_LIGHT * st * np.cos(phi), C_LIGHT * st * np.sin(phi)
    recs["weight"] = 1e3
    recs["cell"] = cells_at_x0(recs)
    recs["id"] = np.arange(n)
    return recs


def _slab(N, tau, umesh=None):
    return P.ProblemSpec(
        name="host_src", x_edges=np.linspace(0.0, 1.0, N + 1) if umesh is None else np.array([0.0, 1.0]),
        y_edges=np.array([0.0, 0.4]), z_edges=np.array([0.0, 0.4]), sigma_a0=np.full(N, tau), T0=np.zeros(N),
        dt=1e-7, steps=1, eos_f=1e10, eos_beta=1.0,
        face_kind=(capi.FACE_REFLECT, capi.FACE_ESCAPE) + (capi.FACE_REFLECT,) * 4, source_face=-1,
        source_per_face=0, new_photons_per_cell=1, comb=0, seed=11, umesh=umesh)


@pytest.mark.parametrize("mesh", ["cartesian", "voronoi"])
@pytest.mark.parametrize("tau", [0.5, 2.0])
def test_host_packets_give_beer_lambert(mesh, tau):
    rng = np.random.default_rng(int(tau * 10) + (7 if mesh == "voronoi" else 0))
    n = 200_000
    if mesh == "cartesian":
        spec = _slab(10, tau)
        recs = _records(n, rng, lambda r: np.zeros(len(r), dtype=np.uint64))
    else:
        lo, hi = (0.0, 0.0, 0.0), (1.0, 0.4, 0.4)
        m = P.voronoi_mesh(P.voronoi_points(150, 9, lo, hi), lo, hi)
        spec = _slab(m.ncells, tau, umesh=m)

        def at_x0(r):
            pts = np.stack([np.full(len(r), 1e-9), r["y"], r["z"]], axis=1)
            d = ((pts[:, None, :] - m.mesh_point[None, :, :]) ** 2).sum(-1)
            return np.argmin(d, axis=1).astype(np.uint64)
        recs = _records(n, rng, at_x0)
    # each its own time left; c dt is 3000 cm, so all but ~1e-3 of the
    # packets (those with the least time left and the most grazing angle)
    # end their flight within the step
    tl = spec.dt * rng.random(n)
    with GpuEngine(spec, capi.MIXED) as e:
        e.queue_source_packets(recs, tl)
        e.step()
        L = e.ledger()
        c = e.counters()
        ok, detail = _ledger_ok(L)
        assert ok, detail
        e.step()
        c2 = e.counters()
    assert c["boundary_particles"] == n and c2["boundary_particles"] == 0
    # every weight is a whole number of quanta: the source is their sum
    Q = 2.0 ** L["eq"]
    assert L["source"] * Q == pytest.approx(1e3 * n, rel=1e-12)
    frac = L["escaped"] / L["source"]
    expect = 2.0 * special.expn(3, tau)
    assert abs(frac - expect) < 4.0 * math.sqrt(expect * (1 - expect) / n) + 1e-3, (frac, expect)
