"""The oracle's StratifiedCombPopulationControl
(population/CombPopulationControl.hpp:256-511), pinned on the CPU by hand
values and invariants:

* one cell, three occupied groups of weights 0.6 / 0.3 / 0.1, Nmin 4 (target
  4), minParticlesPerGroup 1: one packet each, the one left over to the
  largest fractional share (group 0): (2, 1, 1) packets, each group keeping
  its weight;
* more occupied groups than the target: the heaviest `target` groups get one
  packet, the rest are dropped, with their weight;
* a bin with fewer packets than its target keeps them, splitting those above
  twice the ideal weight into ceil(w / ideal) copies;
* random censuses: every kept bin conserves its weight, the dropped weight
  is what is missing, and no cell exceeds 20 Nmin;
* one group (a gray problem): the plain comb's counts, cell by cell.
"""
from __future__ import annotations

import numpy as np
import pytest

from storm_amd import problems as P


def _spec(ncells=1, groups=(1.0, 2.0, 3.0, 4.0), nmin=4, factor=1.0, min_per=1, seed=3):
    b = None if groups is None else np.asarray(groups, dtype=float) * P.KEV
    return P.ProblemSpec(
        name="strat", x_edges=np.linspace(0.0, 1.0, ncells + 1), y_edges=np.array([0.0, 1.0]),
        z_edges=np.array([0.0, 1.0]), sigma_a0=np.ones(ncells), T0=np.ones(ncells), dt=1e-11, steps=1, eos_f=1.0,
        eos_beta=1.0, comb=2, comb_nmin=nmin, comb_factor=factor, comb_min_per_group=min_per, seed=seed,
        group_bounds=b, nu_pow=0.0, nu_ref=P.KEV)


def test_targets_shared_by_largest_remainder(oracle):
    cells = np.zeros(30, dtype=np.int64)
    freq = np.repeat([1.5, 2.5, 3.5], 10) * P.KEV
    w = np.repeat([0.06, 0.03, 0.01], 10)
    cnt, ws, dropped = oracle.strat_comb(_spec(), cells, freq, w)
    assert cnt[0].tolist() == [2, 1, 1]
    assert np.allclose(ws[0], [0.6, 0.3, 0.1], rtol=1e-12)
    assert dropped == 0.0


def test_groups_beyond_the_target_are_dropped(oracle):
    # Nmin 2, one cell: target 2; four groups of 0.4 / 0.3 / 0.2 / 0.1 with a
    # minimum of 2 each do not fit, so groups 0 and 1 get one packet each
    spec = _spec(groups=(1.0, 2.0, 3.0, 4.0, 5.0), nmin=2, min_per=2)
    cells = np.zeros(40, dtype=np.int64)
    freq = np.repeat([1.5, 2.5, 3.5, 4.5], 10) * P.KEV
    w = np.repeat([0.04, 0.03, 0.02, 0.01], 10)
    cnt, ws, dropped = oracle.strat_comb(spec, cells, freq, w)
    assert cnt[0].tolist() == [1, 1, 0, 0]
    assert np.allclose(ws[0], [0.4, 0.3, 0.0, 0.0], rtol=1e-12)
    assert dropped == pytest.approx(0.3, rel=1e-12)


def test_small_bins_keep_and_split(oracle):
    # target 8 in one cell; two groups with min 2 each: 2 + 2 + 4 * share.
    # Group 0 holds one packet of weight 0.8, group 1 two of 0.1: targets
    # (2 + 3, 2 + 1) = (5, 3) (fractions 0.2 / 0.8 rounded to group 1)
    spec = _spec(groups=(1.0, 2.0, 3.0), nmin=8, min_per=2)
    cells = np.zeros(3, dtype=np.int64)
    freq = np.array([1.5, 2.5, 2.5]) * P.KEV
    w = np.array([0.8, 0.1, 0.1])
    cnt, ws, dropped = oracle.strat_comb(spec, cells, freq, w)
    # group 0: one packet above twice the ideal 0.8 / 5 -> ceil(0.8 / 0.16) = 5
    # copies; group 1: ideal 0.2 / 3, both packets below twice it -> kept
    assert cnt[0].tolist() == [5, 2]
    assert np.allclose(ws[0], [0.8, 0.2], rtol=1e-12)
    assert dropped == 0.0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_census_conserves_each_kept_bin(oracle, seed):
    rng = np.random.default_rng(seed)
    ncells, G, n = 6, 8, 4000
    spec = _spec(ncells=ncells, groups=tuple(np.geomspace(0.01, 20.0, G + 1)), nmin=3, factor=2.0, min_per=2,
                 seed=seed)
    cells = rng.integers(0, ncells, n)
    freq = np.exp(rng.uniform(np.log(0.011), np.log(19.9), n)) * P.KEV
    w = rng.exponential(1.0, n) * np.where(rng.random(n) < 0.1, 50.0, 1.0)
    cnt, ws, dropped = oracle.strat_comb(spec, cells, freq, w)
    b = spec.group_bounds
    g = np.clip(np.searchsorted(b, freq, side="right") - 1, 0, G - 1)
    win = np.zeros((ncells, G))
    np.add.at(win, (cells, g), w)
    kept = cnt > 0
    assert np.allclose(ws[kept], win[kept], rtol=1e-10)
    assert dropped == pytest.approx(win[~kept].sum(), rel=1e-10, abs=1e-12)
    assert np.all(cnt.sum(1) <= 20 * 3)
    assert cnt.sum() > 0


def test_one_group_is_the_plain_comb(oracle):
    rng = np.random.default_rng(5)
    ncells, n = 5, 3000
    cells = rng.integers(0, ncells, n)
    w = rng.exponential(1.0, n)
    strat = _spec(ncells=ncells, groups=None, nmin=7, factor=2.0, min_per=2)
    plain = _spec(ncells=ncells, groups=None, nmin=7, factor=2.0)
    plain.comb = 1
    cnt_s, ws_s, _ = oracle.strat_comb(strat, cells, np.zeros(n), w)
    cnt_p, ws_p = oracle.comb(plain, cells, w)
    assert cnt_s[:, 0].tolist() == cnt_p.tolist()
    assert np.allclose(ws_s[:, 0], ws_p, rtol=1e-12)
