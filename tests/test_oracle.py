"""Pins the CPU oracle (oracle/scrimp_oracle.c) before anything is compared with it.

The reference ships no expected outputs (its only fixture is an input,
scrimppp/data_testing/test.bin) and cannot be built here (Boost missing), so
the oracle is pinned by: a long-double brute force of the z-normalised
distance (independent arithmetic), closed forms for the statistics and seeds,
invariants the reference's own validation scripts check
(scrimppp/utils/compare_profile_indices.py:101-104), special inputs whose
answer is known, the partition-coverage check of ScrimpTrivialPar.cpp:232-240,
and planted mistakes that each of these must catch.
"""
from __future__ import annotations

import numpy as np
import pytest

import thesiscode_amd as tc
from conftest import assert_indices_match, golden_cases
from oracle import oracle as orc


def _bf_close(p, bf):
    return np.all(np.abs(p - bf) <= 1e-6 * np.maximum(1.0, np.abs(bf)))


def test_window_stats_match_long_double_closed_form():
    # ScrimpSequOpt.cpp:33-79: sInv = 1/sigma (population), mus = mu*sqrt(m)/sigma
    # The naive prefix sums cancel: the error of var is bounded by a few ulps of
    # the running sum of squares divided by m, relative to var itself.
    A = tc.random_walk(5, 3000)
    eps = np.finfo(np.float64).eps
    sq_max = float(np.sum(A * A))
    for m in (2, 17, 256):
        s, mus = orc.window_stats(A, m)
        mu, sg = orc.window_moments_ld(A, m)
        bound = np.maximum(1e-12, 16 * eps * sq_max / m / sg**2)
        assert np.all(np.abs(1.0 / s - sg) <= bound * sg), m
        assert np.all(np.abs(mus - mu / sg * np.sqrt(m)) <= bound * (np.abs(mu) / sg * np.sqrt(m) + 1)), m
        # a sample (m-1) variance would be off by sqrt(m/(m-1)) - 1, far outside the bound
        assert np.any(np.abs(1.0 / s - sg * np.sqrt(m / (m - 1.0))) > bound * sg)


def test_seeds_match_direct_dot_product():
    A = tc.random_walk(6, 2000)
    m = 64
    qt = orc.seeds(A, m, 10, 900)
    Al = A.astype(np.longdouble)
    for d in (10, 11, 333, 900):
        ref = float(np.sum(Al[d:d + m] * Al[:m]))
        assert abs(qt[d] - ref) <= 1e-13 * abs(ref)


@pytest.mark.parametrize("kernel", [0, 1])
def test_triv_par_matches_long_double_brute_force(golden, kernel):
    for name, A, m in golden_cases(golden):
        bf = golden[f"{name}__bf_profile"]
        bi = golden[f"{name}__bf_index"]
        p, i, cells, _ = orc.triv_par(A, m, nranks=3, nthreads=3, blocklen=37, kernel=kernel)
        assert _bf_close(p, bf), name
        assert_indices_match(A, m, bf, bi, i, abs_tol=1e-6 * 2 * m, min_frac=0.99)
        assert cells == tc.cells(A.size, m)


def test_brute_force_fixture_is_reproduced(golden):
    # the committed long-double profiles are what the brute force computes now
    for name, A, m in golden_cases(golden)[:2]:
        p, i = orc.brute_force(A, m)
        assert np.array_equal(p, golden[f"{name}__bf_profile"])
        assert np.array_equal(i, golden[f"{name}__bf_index"])


def test_reference_fixture_profile(golden, test_series):
    # README.md:59 example: m=400 on data_testing/test.bin
    m = 400
    p, i, _, _ = orc.triv_par(test_series, m, nranks=1, nthreads=1, kernel=1)
    assert np.array_equal(p, golden["testbin_m400__oracle_profile"])
    assert np.array_equal(i, golden["testbin_m400__oracle_index"])
    rows = golden["testbin_m400__bf_rows"]
    np.testing.assert_allclose(p[rows], golden["testbin_m400__bf_row_profile"], rtol=1e-7, atol=1e-7)
    assert np.array_equal(i[rows], golden["testbin_m400__bf_row_index"])


def test_invariants_on_reference_fixture(test_series):
    m = 400
    ez = m // 4
    p, i, _, _ = orc.triv_par(test_series, m, nranks=2, nthreads=2)
    P = test_series.size - m + 1
    assert np.all(np.abs(np.arange(P) - i) > ez)
    assert np.all(p >= -1e-9) and np.all(p <= 4 * m + 1e-9)
    for r in np.linspace(0, P - 1, 25).astype(int):
        assert abs(p[r] - orc.pair_dist2(test_series, m, r, i[r])) <= 1e-7 * max(1.0, p[r])


def test_planted_repeat_is_found():
    A = tc.random_walk(21, 4000)
    m = 50
    A[3000:3000 + m] = A[500:500 + m] + 7.0  # an offset copy has the same z-normalised shape
    p, i, _, _ = orc.triv_par(A, m)
    assert i[500] == 3000 and i[3000] == 500
    assert p[500] < 1e-8 and p[3000] < 1e-8


def test_sine_matches_whole_periods():
    per = 40
    t = np.arange(2000)
    A = np.sin(2 * np.pi * t / per)
    m = 60
    p, i, _, _ = orc.triv_par(A, m)
    assert np.all(p < 1e-6)
    assert np.all((np.abs(i - np.arange(p.size)) % per) == 0)


def test_flat_windows_never_win():
    A = tc.random_walk(22, 1500)
    A[600:700] = 0.0  # windows inside the flat stretch have sigma == 0
    m = 40
    p, i, _, _ = orc.triv_par(A, m)
    flat = np.arange(600, 700 - m + 1)
    assert np.all(i[flat] == -1) and np.all(p[flat] == 2 * m + 2)
    assert not np.any(np.isin(i, flat))


def test_blocklen_and_rank_invariance(golden):
    name, A, m = golden_cases(golden)[0]
    base_p, base_i, _, _ = orc.triv_par(A, m, kernel=0)
    for bl, nr, k in ((1, 1, 1), (7, 2, 1), (500, 5, 1), (64, 8, 0)):
        p, i, _, _ = orc.triv_par(A, m, blocklen=bl, nranks=nr, nthreads=min(nr, 4), kernel=k)
        assert np.array_equal(p, base_p), (bl, nr, k)
        assert np.array_equal(i, base_i), (bl, nr, k)


def test_vert_blocked_same_values_index_zero_init():
    A = tc.random_walk(23, 900)
    m = 30
    p1, i1, _, _ = orc.triv_par(A, m)
    p2, i2 = orc.vert_blocked(A, m, blocklen=33)
    assert np.array_equal(p1, p2) and np.array_equal(i1, i2)
    # no admissible partner at all: -1 for the parallel driver, 0 for vert_blocked
    B = tc.random_walk(24, 40)
    _, ia, _, _ = orc.triv_par(B, 38)
    _, ib = orc.vert_blocked(B, 38)
    assert np.all(ia == -1) and np.all(ib == 0)


@pytest.mark.parametrize("nparts", [2, 3, 4, 8, 16, 24])
@pytest.mark.parametrize("ids", [0, 1, 5, 31, 1000, 65035])
def test_partition_coverage(nparts, ids):
    # every diagonal exactly once (ScrimpTrivialPar.cpp:232-240)
    seen = np.zeros(max(ids, 1), dtype=int)
    for r in range(nparts):
        f, l_ = orc.get_partition(r, nparts, ids)
        if l_ >= f:
            seen[f:l_ + 1] += 1
    assert np.all(seen[:ids] == 1)


@pytest.mark.parametrize("code", [1, 2, 3, 5])
def test_planted_mistakes_are_caught_by_brute_force(golden, code):
    name, A, m = golden_cases(golden)[0]
    bf = golden[f"{name}__bf_profile"]
    bi = golden[f"{name}__bf_index"]
    with orc.fault(code):
        p, i, _, _ = orc.triv_par(A, m, kernel=0)
    assert (not _bf_close(p, bf)) or np.mean(i == bi) < 0.9


def test_planted_exclusion_mistake_is_caught():
    # a smooth (integrated) walk makes the diagonal next to the exclusion zone win
    A = np.cumsum(tc.random_walk(11, 1200))
    m = 32
    bf, _ = orc.brute_force(A, m)
    with orc.fault(4):
        p, i, _, _ = orc.triv_par(A, m, kernel=0)
    assert not _bf_close(p, bf)
    assert np.min(np.abs(np.arange(p.size) - i)) <= m // 4


def test_short_series_is_rejected():
    with pytest.raises(ValueError):
        orc.triv_par(np.zeros(10), 11)
