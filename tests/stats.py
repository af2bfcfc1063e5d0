"""Ensemble statistics shared by the parity tests."""
from __future__ import annotations

import numpy as np
from scipy import stats


def zscores(a: np.ndarray, b: np.ndarray, rel_floor: float = 0.0) -> np.ndarray:
    """Per-cell z of the difference of two seed ensembles (seeds x cells).

    Uses the pooled standard error sqrt(var_a/n_a + var_b/n_b).  `rel_floor`
    adds (rel_floor * |mean|)^2 to the variance: the representation error of
    FP32 cell opacities (~1e-7 relative) in cells whose Monte Carlo spread is
    below it (cold, optically thick cells where every packet is absorbed in
    place make the time-averaged energy density nearly deterministic).
    """
    na, nb = a.shape[0], b.shape[0]
    ma, mb = a.mean(0), b.mean(0)
    va, vb = a.var(0, ddof=1), b.var(0, ddof=1)
    scale = np.maximum(np.abs(ma), np.abs(mb))
    se2 = va / na + vb / nb + (rel_floor * scale) ** 2
    d = ma - mb
    z = np.zeros_like(d)
    nz = se2 > 0
    z[nz] = d[nz] / np.sqrt(se2[nz])
    z[~nz & (d != 0)] = np.inf
    return z


def check_ensembles(a, b, rel_floor=1e-6, label=""):
    """Per-cell agreement within 3 sigma of the seed-batch variance.

    With n seeds per side the per-cell z follows a Student t with about
    2(n-1) degrees of freedom, so a few cells beyond 3 sigma are expected.
    Pass if the count beyond 3 sigma is within 3 binomial standard deviations
    of that expectation (+1) and the mean z^2 is within 4 standard errors of
    E[t^2] = nu/(nu-2).  Returns (summary, ok)."""
    a, b = np.asarray(a), np.asarray(b)
    z = zscores(a, b, rel_floor)
    n = len(z)
    nu = max(a.shape[0] + b.shape[0] - 2, 5)
    p3 = float(2.0 * stats.t.sf(3.0, nu))
    over = int(np.sum(np.abs(z) > 3.0))
    allowed = p3 * n + 3.0 * np.sqrt(n * p3 * (1 - p3)) + 1.0
    finite = np.isfinite(z)
    z2 = float(np.mean(z[finite] ** 2)) if finite.any() else 0.0
    et2 = nu / (nu - 2.0)
    vt2 = 2.0 * nu * nu * (nu - 1.0) / ((nu - 2.0) ** 2 * (nu - 4.0))
    z2_tol = 4.0 * np.sqrt(vt2 / max(finite.sum(), 1))
    ok = (over <= allowed) and bool(finite.all()) and (n < 20 or z2 <= et2 + z2_tol)
    summary = dict(label=label, cells=n, over_3sigma=over, allowed=round(allowed, 2),
                   max_abs_z=float(np.max(np.abs(z))) if n else 0.0, mean_z2=round(z2, 3),
                   expected_z2=round(et2, 3), z2_tol=round(z2_tol, 3))
    return summary, ok
