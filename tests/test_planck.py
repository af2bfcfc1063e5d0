"""The engine's Planck integral (storm_amd/csrc/common/planck.h, used for
group fractions of multigroup IMC) against numerical quadrature of
15/pi^4 x^3/(e^x - 1) and against the closed total 1 (the reference checks
the same sum to 1e-8, boundary/SideTemperature.hpp:75-80)."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import integrate

from oracle import pyoracle


def quad(a, b):
    f = lambda x: x ** 3 / np.expm1(x) if x > 0 else 0.0  # noqa: E731
    v, _ = integrate.quad(f, a, b, epsabs=0, epsrel=1e-13, limit=200)
    return 15.0 / np.pi ** 4 * v


@pytest.mark.parametrize("a,b", [(0.0, 0.1), (1e-4, 1e-3), (0.3, 1.2), (1.0, 1.5), (1.4, 1.6), (1.5, 3.0),
                                 (2.0, 2.5), (3.0, 10.0), (10.0, 12.0), (25.0, 30.0), (0.0, 50.0)])
def test_planck_integral_against_quadrature(a, b):
    got = pyoracle.engine_planck([a], [b])[0]
    ref = quad(a, b)
    assert got == pytest.approx(ref, rel=1e-11, abs=1e-300), (a, b, got, ref)


def test_group_fractions_sum_to_one():
    edges = np.concatenate([[0.0], np.geomspace(1e-4, 80.0, 65)])
    frac = pyoracle.engine_planck(edges[:-1], edges[1:])
    assert np.all(frac >= 0.0)
    assert abs(frac.sum() - 1.0) < 1e-12
