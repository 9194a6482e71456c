"""Pins the oracle's PreSCRIMP restatement (ScrimpppOrig.cpp:79-283).

PreSCRIMP excludes partners with |i - j| < m/4 (ScrimpppOrig.cpp:191-200), one
diagonal fewer than SCRIMP's |i - j| <= m/4, so the brute force below runs with
exclusion m/4 - 1.  The checks are chosen so that a wrong distance formula, a
dropped or mis-signed term of the walk's recurrence, a walk that stops early or
a wrong sampled-row minimum each fail one of them.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc


def _series(seed, n):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.standard_normal(n))


CASES = [(1, 600, 32, 0.25), (2, 700, 40, 0.3), (3, 500, 16, 1.0), (4, 800, 64, 0.125)]


@pytest.mark.parametrize("seed,n,m,stride", CASES)
def test_values_are_distances_to_the_reported_partner(seed, n, m, stride):
    A = _series(seed, n)
    p, i = orc.prescrimp(A, m, stride)
    assert np.all(np.isfinite(p))
    for r in range(p.size):
        d = orc.pair_dist2(A, m, r, int(i[r]))
        assert abs(p[r] - d) <= 1e-7 * max(1.0, d), (r, p[r], d)
        assert abs(r - int(i[r])) >= m // 4


@pytest.mark.parametrize("seed,n,m,stride", CASES)
def test_bounds_and_sampled_rows(seed, n, m, stride):
    A = _series(seed, n)
    p, i = orc.prescrimp(A, m, stride)
    bf, _ = orc.brute_force(A, m, excl=m // 4 - 1)
    tol = 1e-7 * np.maximum(1.0, bf)
    assert np.all(p >= bf - tol)
    step = max(1, int(np.floor(stride * m)))
    rows = np.arange(0, p.size, step)
    assert np.all(np.abs(p[rows] - bf[rows]) <= tol[rows])


@pytest.mark.parametrize("seed,n,m,stride", CASES)
def test_walks_cover_the_diagonal_around_each_sample(seed, n, m, stride):
    A = _series(seed, n)
    p, i = orc.prescrimp(A, m, stride)
    P = p.size
    step = max(1, int(np.floor(stride * m)))
    for r in range(0, P, step):
        j = int(i[r])
        for k in range(-(step - 1), step):
            a, b = r + k, j + k
            if 0 <= a < P and 0 <= b < P:
                d = orc.pair_dist2(A, m, a, b)
                assert p[a] <= d + 1e-7 * max(1.0, d), (r, k)
                assert p[b] <= d + 1e-7 * max(1.0, d), (r, k)


def test_row_order_only_changes_ties_and_rounding():
    # every update is a minimum, so the order of the sampled rows only decides
    # ties and which walk reaches a cell first (walks from different samples
    # reach the same cell with differently rounded recurrences: ~1e-14)
    A = _series(5, 900)
    m, stride = 48, 0.25
    p, i = orc.prescrimp(A, m, stride)
    step = int(np.floor(stride * m))
    order = np.arange(0, p.size, step)
    np.random.default_rng(0).shuffle(order)
    q, j = orc.prescrimp(A, m, stride, order=order)
    assert np.all(np.abs(p - q) <= 1e-12 * np.maximum(1.0, p))
    assert np.mean(i == j) > 0.99


def test_planted_repeat_is_found_by_the_walk():
    # a pattern repeated at a distance that is not a multiple of the sampling
    # step: rows near the first copy only find the second through a walk
    rng = np.random.default_rng(9)
    A = np.cumsum(rng.standard_normal(1200))
    m = 40
    A[800:800 + m] = A[203:203 + m] + 3.0
    p, i = orc.prescrimp(A, m, 0.25)
    assert p[203] < 1e-6 and i[203] == 800
    assert p[800] < 1e-6 and i[800] == 203


@pytest.mark.parametrize("fault", [1, 2])
def test_planted_mistakes_fail_a_check(fault):
    # 1: the walk's recurrence drops its subtracted term; 2: the distance adds
    # the mean product instead of subtracting it
    A = _series(1, 600)
    m, stride = 32, 0.25
    with orc.fault(fault):
        p, i = orc.prescrimp(A, m, stride)
    ok = all(abs(p[r] - orc.pair_dist2(A, m, r, int(i[r]))) <= 1e-7 * max(1.0, p[r]) for r in range(p.size))
    assert not ok
