"""Edge cases of the engine's FP32 tracking, run through the CPU build of the
same transport code (oracle/replay.cpp compiles storm_amd/csrc/common/*.h).

A packet lying exactly on a cell's upper face (u = 1, where a reflection
leaves it) in an optically thick cell moves less than half a float unit per
collision flight.  Rounded to nearest, its coordinate would never change: it
would stay at distance 0 from the face, never become eligible for the random
walk, and run one collision flight at a time through its whole time budget
(3e8 segments at sigma = 1e13 / cm; at 1e21 / cm, the opacity of a cell that
cooled to 4e-5 keV in the reference's Marshak-wave problem 4, a flight is
below the unit of the time budget itself and the history never ends).  The
tracker moves such a packet one unit inside the cell, where the random walk
takes over: the packets must reach census in a handful of segments.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from storm_amd import capi
from storm_amd import problems as P


@pytest.fixture(scope="module")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


def thick_slab():
    # three thin cells (7e-5 cm) across x, a 2e-3 cm square in y and z, every
    # face reflecting, random walk on (the reference's default thresholds)
    spec = P.c1_slab(seed=3, nx=3, steps=1)
    spec.x_edges = np.array([0.0, 7e-5, 1.4e-4, 2.1e-4])
    spec.y_edges = np.array([0.0, 2e-3])
    spec.z_edges = np.array([0.0, 2e-3])
    spec.sigma_a0 = np.full(3, 1.0)
    spec.face_kind = (capi.FACE_REFLECT,) * 6
    spec.with_rw = 1
    return spec


# storm::CellAccel<float> (storm_amd/csrc/common/transport.h)
CELL_ACCEL_F32 = np.dtype([("sig_t", "<f4"), ("total", "<f4"), ("rate", "<f4", 6), ("flags", "<u4"), ("gcut", "<u4"),
                           ("sig_d", "<f4"), ("sig_abs", "<f4"), ("up_rate", "<f4"), ("e_cut", "<f4"),
                           ("ddmc_sig_t", "<f4"), ("ddmc_abs", "<f4"), ("ddmc_e_cut", "<f4")])


def test_cell_accel_record_layout(oracle):
    """The record above is the engine's: a stale layout would misread every
    cell after the first."""
    assert CELL_ACCEL_F32.itemsize == oracle.rlib().replay_cell_accel_bytes()


def accel_bytes(n, sig_t):
    rec = np.zeros(n, dtype=CELL_ACCEL_F32)
    rec["sig_t"] = sig_t
    rec["sig_d"] = sig_t  # gray walk: D = c / (3 sigma_t)
    rec["e_cut"] = np.inf
    rec["flags"] = 1  # random-walk cell
    rec["ddmc_sig_t"] = sig_t
    rec["ddmc_e_cut"] = np.inf
    return rec.tobytes()


@pytest.mark.parametrize("sig_ev", [1e13, 1e21])
def test_packet_on_upper_face_of_thick_cell_reaches_census(oracle, sig_ev):
    spec = thick_slab()
    n = 256
    sig_abs = 1e3  # f*sigma_a, cm^-1; sig_ev: effective scatter
    xs = np.tile(np.array([sig_abs, sig_ev], dtype=np.float32), 3)
    pk = (capi.PacketF32 * n)()
    rng = np.random.default_rng(11)
    for q in range(n):
        mu = 2.0 * rng.random() - 1.0
        ph = 2.0 * math.pi * rng.random()
        st = math.sqrt(1.0 - mu * mu)
        p = pk[q]
        p.u[0], p.u[1], p.u[2] = 0.5, 1.0, 0.5  # on the upper y face
        p.o[0], p.o[1], p.o[2] = st * math.cos(ph), mu, st * math.sin(ph)
        p.s = 3e-5  # c * 1 fs, in cm
        p.w, p.wmin = 1.0, 1e-3
        p.ijk = (1 << 20)  # cell (1, 0, 0)
        p.id, p.ctr = q, 0
    consts = [123, 1, -100, -100]  # seed, step, quantum exponents
    tal, outc = oracle.replay_transport(spec, pk, n, xs, consts, accel=accel_bytes(3, sig_ev))
    segs = int(tal["segs"].sum())
    # every packet ends in census (nothing escapes a reflecting box, and the
    # weight stays far above the roulette threshold over 1 fs)
    assert int(outc[1]) == n, outc  # kCensus
    # off the face within a few flights, then a few random-walk steps
    assert segs < 200 * n, segs


def test_path_integral_rounding_is_unbiased(oracle):
    """fx::quantize_unbiased: a multiple of 2^e next to v, rounded up with
    probability equal to the leftover fraction -- over 4096 equally spaced
    draws of the 24-bit rounding word the mean is v to within 1/4096 of a
    quantum, and exact multiples stay."""
    q = oracle.rlib().replay_quantize_unbiased
    e = -40
    Q = 2.0 ** e
    for v in (0.3 * Q, 1e-6 * Q, 17.75 * Q, 12345.0 * Q):
        fl = math.floor(v / Q) * Q
        outs = np.array([q(v, e, u) for u in range(0, 1 << 24, 1 << 12)])
        assert set(np.unique(outs)) <= {fl, fl + Q}
        assert abs(outs.mean() - v) <= Q / 4096 + 1e-300
    assert q(5 * Q, e, 0) == 5 * Q and q(5 * Q, e, (1 << 24) - 1) == 5 * Q
