"""Pins the oracle's random-walk tables (radiation/RandomWalk.hpp:20-242)
with facts the mathematics fixes, independent of the eigen series:

  * a walker starting at the centre of a unit sphere with absorbing wall
    (tau = D t / R^2) leaves at a time T whose Laplace transform is
    E[exp(-lam T)] = sqrt(lam) / sinh(sqrt(lam)) (hitting time of a 3-d Bessel
    process); so the survival S(tau) = P(T > tau) integrates to
    int_0^inf S exp(-lam tau) dtau = (1 - sqrt(lam)/sinh(sqrt(lam))) / lam,
    in particular the mean exit time 1/6;
  * S(tau) -> 2 exp(-pi^2 tau) for large tau (first eigenmode);
  * the radius CDF at small tau is the Maxwell law of a free 3-d Gaussian of
    variance 2 tau per axis, at large tau the first eigenmode's
    (sin(pi x) - pi x cos(pi x)) / pi, and the two formulas the reference
    switches between at tau = 0.01 agree there;
  * the engine's FP32 tables equal the oracle's to FP32 rounding;
  * the engine's inverse leak-time table and sampler (both precisions) against
    the dual (image) series of the same survival, obtained by Poisson
    summation and free of its cancellation at small tau:
        1 - S(tau) = 2 / sqrt(pi tau) sum_{k>=0} exp(-(2k+1)^2 / (4 tau)),
    over every float uniform the engine draws near both ends of (0, 1].
A dropped alternating sign, a wrong pi^2 factor or a transposed table fails
one of these.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle


@pytest.fixture(scope="module")
def tables():
    return pyoracle.rw_tables()


def _laplace_of_survival(tau, S, lam):
    # trapezoid on the log-spaced table; S = 1 on [0, tau_min)
    f = S * np.exp(-lam * tau)
    return tau[0] + np.trapezoid(f, tau)


@pytest.mark.parametrize("lam", [0.0, 1.0, 10.0, 100.0])
def test_survival_laplace_transform(tables, lam):
    tau, S, _ = tables
    got = _laplace_of_survival(tau, S, lam)
    if lam == 0.0:
        expect = 1.0 / 6.0
    else:
        r = np.sqrt(lam)
        expect = (1.0 - r / np.sinh(r)) / lam
    assert got == pytest.approx(expect, rel=2e-4), (lam, got, expect)


def test_survival_limits(tables):
    tau, S, _ = tables
    # non-increasing up to the rounding of the alternating series near tau = 1e-4
    assert S[0] == 1.0 and np.all(np.diff(S) <= 1e-13)
    big = tau > 0.5
    assert np.allclose(S[big], np.minimum(1.0, 2.0 * np.exp(-np.pi ** 2 * tau[big])), rtol=1e-6, atol=1e-300)


def _maxwell_cdf(x, tau):
    from scipy import special
    a = x / (2.0 * np.sqrt(tau))
    return special.erf(a) - 2.0 * a * np.exp(-a * a) / np.sqrt(np.pi)


def _cdf_from_radius_table(rad_row, x):
    # rad_row[j] = radius at CDF j/255: invert by interpolation
    xi = np.linspace(0.0, 1.0, rad_row.size)
    return np.interp(x, rad_row, xi)


def test_radius_table_small_tau_is_maxwell(tables):
    tau, _, rad = tables
    i = np.searchsorted(tau, 1e-3)
    x = np.linspace(0.02, 0.3, 15)
    got = _cdf_from_radius_table(rad[i], x)
    assert np.allclose(got, _maxwell_cdf(x, tau[i]), atol=6e-3)


def test_radius_table_large_tau_is_first_mode(tables):
    tau, _, rad = tables
    i = np.searchsorted(tau, 5.0)
    x = np.linspace(0.05, 0.95, 19)
    ground = (np.sin(np.pi * x) - np.pi * x * np.cos(np.pi * x)) / np.pi
    got = _cdf_from_radius_table(rad[i], x)
    assert np.allclose(got, ground, atol=6e-3)


def test_radius_cdf_branches_agree_at_switch(tables):
    """Eigen series and free-space Maxwell law at tau = 0.01 (where the
    reference switches, RandomWalk.hpp:152-167): sampled radii from the
    tables on both sides of the switch agree."""
    tau, _, rad = tables
    lo = np.searchsorted(tau, 0.01) - 1
    hi = lo + 1
    assert tau[lo] < 0.01 <= tau[hi]
    x = np.linspace(0.05, 0.6, 12)
    a = _cdf_from_radius_table(rad[lo], x)
    b = _cdf_from_radius_table(rad[hi], x)
    # neighbouring table rows differ only by the tau step (1.8 %)
    assert np.allclose(a, b, atol=2e-2)
    assert np.allclose(b, _maxwell_cdf(x, tau[hi]), atol=6e-3)


def test_leak_time_sampler_inverts_survival(tables):
    tau, S, _ = tables
    xi = np.linspace(0.01, 0.99, 50)
    leak, _ = pyoracle.rw_sample(xi, np.full_like(xi, 0.1))
    # P(T > leak(xi)) = 1 - xi
    Sl = np.interp(np.log(leak), np.log(tau), S)
    assert np.allclose(Sl, 1.0 - xi, atol=2e-3)


def test_engine_tables_equal_oracle_tables(tables):
    tau, S, rad = tables
    et, es, er = pyoracle.engine_rw_tables()
    assert np.allclose(et, tau, rtol=2e-7)
    assert np.allclose(es, S, rtol=2e-7, atol=1e-37)
    assert np.allclose(er, rad, atol=2e-7)


def _one_minus_survival(T):
    T = np.asarray(T, dtype=float)[..., None]
    k = np.arange(40)
    return (2.0 / np.sqrt(np.pi * T[..., 0])) * np.exp(-(2 * k + 1) ** 2 / (4.0 * T)).sum(-1)


def _exact_leak_time(xi):
    # -ln S(T) = y, solved by bisection in ln T on the image series
    y = -np.log1p(-np.asarray(xi, dtype=float))
    lo, hi = np.full_like(y, np.log(1e-3)), np.full_like(y, np.log(8.0))
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        above = -np.log1p(-_one_minus_survival(np.exp(mid))) < y
        lo, hi = np.where(above, mid, lo), np.where(above, hi, mid)
    return np.exp(0.5 * (lo + hi))


def test_engine_leak_table_is_the_inverse_of_survival():
    leak = pyoracle.engine_rw_leak_table()
    y = np.exp(-25.0 + 26.5 * np.arange(2049) / 2048)
    got = -np.log1p(-_one_minus_survival(leak))
    assert np.all(np.diff(leak) > 0)
    # below y ~ 1e-9 the eigen series' rounding (1e-16 absolute in S) is
    # the limit; above it the roots are exact
    big = y > 1e-9
    assert np.allclose(got[big], y[big], rtol=1e-8)
    assert np.allclose(got, y, rtol=1e-4)


@pytest.mark.parametrize("precision", [0, 1], ids=["float", "double"])
def test_engine_leak_sampler_inverts_survival(precision):
    # the engine's float uniforms (k + 1/2) 2^-24 rounded to float, at both
    # ends of (0, 1) and spread between; (0, 1] includes 1 by the rounding
    k = np.concatenate([np.arange(2000), np.geomspace(1, 2**24 - 1, 3000).astype(np.int64), 2**24 - 1 - np.arange(2000)])
    xi = ((k.astype(np.float32) + np.float32(0.5)) * np.float32(2.0**-24)).astype(np.float64)
    top = xi >= 1.0
    leak, _ = pyoracle.engine_rw_sample(precision, xi, np.full_like(xi, 0.1))
    assert np.all(leak[top] == 64.0)
    exact = _exact_leak_time(xi[~top])
    assert np.allclose(leak[~top], exact, rtol=3e-5)
    # past the table: the first eigenmode, T = (y + ln 2) / pi^2
    far = xi[~top] > 1.0 - np.exp(-5.0)
    assert far.sum() > 100
    assert np.allclose(leak[~top][far], (-np.log1p(-xi[~top][far]) + np.log(2.0)) / np.pi**2, rtol=3e-6)
