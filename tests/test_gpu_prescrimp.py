"""PreSCRIMP (the anytime phase of scrimppp_orig, ScrimpppOrig.cpp:130-283) on
the GPU against the oracle's restatement, through the C ABI.

The GPU path forms d2 from the exact path's window statistics and direct FP64
dot products summed in matrix-core order; the oracle uses ScrimpppOrig.cpp's
means and sigmas and sequential sums (the reference: FFT).  Values therefore
agree to rounding: |d2_gpu - d2_oracle| <= 1e-9 * (d2 + 2m).  Indices agree
except at exact ties, where the reference's order is shuffled anyway.
"""
from __future__ import annotations

import numpy as np
import pytest

import thesiscode_amd as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _close(a, b, m):
    return np.abs(a - b) <= 1e-9 * (np.abs(b) + 2 * m)


@pytest.mark.parametrize("n,m,stride", [(1500, 32, 0.25), (3000, 37, 0.3), (2500, 64, 1.0),
                                        (4000, 100, 0.125), (20000, 128, 0.25)])
def test_matches_the_oracle(n, m, stride):
    A = tc.random_walk(300 + m, n)
    p, i, tm = tc.prescrimp_f64(A, m, stride)
    q, j = orc.prescrimp(A, m, stride)
    assert np.all(np.isfinite(p)) == np.all(np.isfinite(q))
    fin = np.isfinite(q)
    assert np.all(_close(p[fin], q[fin], m)), (n, m, stride)
    assert np.mean(i == j) >= 0.999, (n, m, stride, np.mean(i == j))
    step = max(1, int(np.floor(stride * m)))
    assert tm["work_items"] == (p.size + step - 1) // step


def test_values_and_sampled_rows_against_brute_force():
    A = tc.random_walk(77, 900)
    m, stride = 40, 0.25
    p, i, _ = tc.prescrimp_f64(A, m, stride)
    for r in range(p.size):
        d = orc.pair_dist2(A, m, r, int(i[r]))
        assert abs(p[r] - d) <= 1e-7 * max(1.0, d)
        assert abs(r - int(i[r])) >= m // 4
    bf, _ = orc.brute_force(A, m, excl=m // 4 - 1)
    assert np.all(p >= bf - 1e-7 * np.maximum(1.0, bf))
    rows = np.arange(0, p.size, int(stride * m))
    assert np.all(np.abs(p[rows] - bf[rows]) <= 1e-7 * np.maximum(1.0, bf[rows]))


def test_single_sampled_row_and_short_series():
    # step >= P: only row 0 is sampled; the walks stop at the series' ends
    A = tc.random_walk(5, 300)
    m = 50
    p, i, tm = tc.prescrimp_f64(A, m, 10.0)
    q, j = orc.prescrimp(A, m, 10.0)
    assert tm["work_items"] == 1
    assert np.array_equal(np.isfinite(p), np.isfinite(q))
    fin = np.isfinite(q)
    assert np.all(_close(p[fin], q[fin], m))
    assert np.array_equal(i[~fin], np.zeros((~fin).sum(), dtype=np.int64))


def test_planted_repeat_at_a_sampled_row():
    # PreSCRIMP is exact at its sampled rows (multiples of step = 16 here), and
    # the partner's entry gets the same pair through the column update
    A = tc.random_walk(9, 5000)
    m = 64
    A[3100:3100 + m] = A[1200:1200 + m] + 2.0
    p, i, _ = tc.prescrimp_f64(A, m, 0.25)
    assert p[1200] < 1e-6 and i[1200] == 3100
    assert p[3100] < 1e-6 and i[3100] == 1200
    q, _ = orc.prescrimp(A, m, 0.25)
    assert np.all(_close(p, q, m))


def test_bad_stride_is_an_error():
    A = tc.random_walk(1, 500)
    with pytest.raises(tc.ScrimpError):
        tc.prescrimp_f64(A, 32, 0.0)


def test_full_size_properties():
    # n = 2^20, m = 256, stride 0.25 (the measured workload): too large for the
    # oracle's PreSCRIMP, so size-independent properties are checked instead.
    n, m, stride = 1 << 20, 256, 0.25
    A = tc.random_walk(1, n)
    print("prescrimp n=2^20 m=256", flush=True)
    p, i, tm = tc.prescrimp_f64(A, m, stride)
    P = n - m + 1
    excl = m // 4
    step = int(np.floor(stride * m))
    assert p.size == P and tm["work_items"] == (P + step - 1) // step
    assert np.all(np.isfinite(p))
    assert np.all(np.abs(np.arange(P) - i) >= excl)
    # Tolerance: the window statistics follow the reference's prefix-sum
    # formula, whose sums of A^2 reach ~5e11 on this walk; their rounding
    # leaves ~1e-8 relative error in a window's variance, ~1e-5 in d2 (for
    # n <= 20000 the tests above hold 1e-9 * (d2 + 2m)).
    def tol(d):
        return 1e-7 * (d + 2 * m)

    # every value is the distance to its reported partner
    rng = np.random.default_rng(7)
    for r in rng.choice(P, 200, replace=False):
        d = orc.pair_dist2(A, m, int(r), int(i[r]))
        assert abs(p[r] - d) <= tol(d), (r, p[r], d)
    # sampled rows hold their full row minimum (long-double window moments,
    # direct sliding dot products)
    mu, sg = orc.window_moments_ld(A, m)
    for r in (0, step * 4093, step * ((P - 1) // step)):
        x = A[r:r + m] - mu[r]
        c = np.correlate(A, x, mode="valid") - mu * x.sum()
        d2 = 2.0 * m * (1.0 - c / (m * sg[r] * sg))
        d2[max(0, r - excl + 1):r + excl] = np.inf
        best = d2.min()
        assert abs(p[r] - best) <= tol(best), (r, p[r], best)
        assert d2[i[r]] <= best + tol(best)
