"""FP32-reseeded HIP path against the exact FP64 path (itself bit-exact with the oracle).

Tolerance: with FP64 refinement of each winner the profile must be within
1e-4 relative of FP64 (the stated target); without refinement within 1e-2.
Relative error is |d2_f32r - d2_f64| / max(d2_f64, 1e-3 * 2m) so that exact
repeats (d2 ~ 0) are judged on the distance scale of the window.
"""
from __future__ import annotations

import numpy as np
import pytest

import thesiscode_amd as tc
from conftest import assert_indices_match, assert_values_equal_reference_walk
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel_err(p, p_ref, m):
    return np.abs(p - p_ref) / np.maximum(p_ref, 1e-3 * 2 * m)


@pytest.mark.parametrize("n,m", [(4000, 32), (20000, 100), (65536, 400)])
def test_refined_profile_within_1e4_of_fp64(n, m):
    A = tc.random_walk(61 + m, n)
    p64, i64, _ = tc.selfjoin_f64(A, m)
    p, i, tm = tc.selfjoin_f32r(A, m)
    err = rel_err(p, p64, m)
    print(f"n={n} m={m} max_rel={err.max():.3e} mean_rel={err.mean():.3e} idx_agree={np.mean(i == i64):.5f}")
    assert err.max() <= 1e-4
    assert np.mean(i == i64) >= 0.98
    ez = m // 4
    assert np.all(np.abs(np.arange(p.size) - i) > ez)


# Bounds measured on the precision workload (walk + sinusoids + bumps): FP32
# accumulation drifts by up to ~1e-5 in the correlation within a tile, which
# near-periodic series turn into near-ties between partners; accumulating the
# same FP32 terms in FP64 removes the drift.
F32_ACC_MAX_REL = 2e-3
F64_ACC_MAX_REL = 1e-4


@pytest.mark.parametrize("accum64", [False, True])
@pytest.mark.parametrize("tile_len", [256, 1024, 4096])
def test_reseed_interval_variants(tile_len, accum64):
    A, _ = tc.precision_series(5, 1 << 15, 256)
    m = 256
    p64, i64, _ = tc.selfjoin_f64(A, m)
    p, i, tm = tc.selfjoin_f32r(A, m, tile_len=tile_len, accum64=accum64)
    assert tm["tile_len"] == tile_len
    err = rel_err(p, p64, m)
    pu, iu, _ = tc.selfjoin_f32r(A, m, tile_len=tile_len, refine=False, accum64=accum64)
    erru = rel_err(pu, p64, m)
    print(f"B={tile_len} accum64={accum64} refined max_rel={err.max():.3e} mean_rel={err.mean():.3e} "
          f"idx_agree={np.mean(i == i64):.5f} | unrefined max_rel={erru.max():.3e} mean_rel={erru.mean():.3e}")
    assert err.max() <= (F64_ACC_MAX_REL if accum64 else F32_ACC_MAX_REL)
    assert err.mean() <= 1e-6
    assert erru.max() <= 1e-2


@pytest.mark.parametrize("tile_len", [16384, 32768])
def test_long_reseed_intervals_with_fp64_accumulation(tile_len):
    # the automatic intervals above 8192 (FP64 accumulation only): bands of up
    # to 2^17 columns cut into several long tiles and an edge tile
    A, _ = tc.precision_series(6, 1 << 17, 256)
    m = 256
    p64, i64, _ = tc.selfjoin_f64(A, m)
    p, i, tm = tc.selfjoin_f32r(A, m, tile_len=tile_len)
    assert tm["tile_len"] == tile_len
    err = rel_err(p, p64, m)
    print(f"B={tile_len} max_rel={err.max():.3e} mean_rel={err.mean():.3e} idx_equal={np.mean(i == i64):.5f}")
    assert err.max() <= F64_ACC_MAX_REL
    assert err.mean() <= 1e-6
    assert_indices_match(A, m, p64, i64, i, abs_tol=0.0, rel_tol=1e-4, min_frac=0.999)


def test_config2_full_size_within_tolerance_of_fp64():
    # n = 2^20 random walk (seed 1), m = 256: the measured workload, default tiles
    A = tc.random_walk(1, 1 << 20)
    m = 256
    p64, i64, _ = tc.selfjoin_f64(A, m)
    for acc in (True, False):
        p, i, tm = tc.selfjoin_f32r(A, m, accum64=acc)
        # the automatic interval: the longest that leaves 16 tiles per wave
        # (FP64 accumulation) or the FP32-accumulation cap
        assert tm["tile_len"] == (16384 if acc else 1024)
        err = rel_err(p, p64, m)
        print(f"accum64={acc} T={tm['tile_len']} max_rel={err.max():.3e} mean_rel={err.mean():.3e} "
              f"idx_equal={np.mean(i == i64):.5f}")
        assert err.max() <= (F64_ACC_MAX_REL if acc else F32_ACC_MAX_REL)
        assert err.mean() <= 1e-6
        assert np.mean(i == i64) >= 0.999
        assert np.all(np.abs(np.arange(p.size) - i) > m // 4)


@pytest.fixture(scope="module")
def config3_exact():
    # n = 2^22, m = 1024, the precision workload (walk + 3 sinusoids + 20 bumps,
    # seed 2) and its exact FP64 profile
    n, m = 1 << 22, 1024
    A, _ = tc.precision_series(2, n, m)
    print("cfg3 exact", flush=True)
    p64, i64, _ = tc.selfjoin_f64(A, m)
    # the exact path's values are distances to their partners (long double);
    # the prefix-sum statistics of a 2^22-point series allow ~1e-7 of d2 + 2m
    rng = np.random.default_rng(11)
    for r in rng.choice(p64.size, 50, replace=False):
        d = orc.pair_dist2(A, m, int(r), int(i64[r]))
        assert abs(p64[r] - d) <= 1e-7 * (d + 2 * m), (r, p64[r], d)
    return A, m, p64, i64


def test_config3_exact_values_equal_reference_arithmetic(config3_exact):
    # At full size (n = 2^22, m = 1024) the exact path's value at sampled
    # entries (10 random, the smallest and the largest) equals the reference's
    # arithmetic over diagonals walked for up to 4.2M steps, bit for bit
    A, m, p64, i64 = config3_exact
    rng = np.random.default_rng(12)
    entries = list(rng.choice(p64.size, 10, replace=False)) + [int(np.argmin(p64)), int(np.argmax(p64))]
    assert_values_equal_reference_walk(A, m, p64, i64, entries)


# B = 0: the library's default re-seed interval (the longest automatic one that
# leaves 16 tiles per wave: 32768 at this size)
@pytest.mark.parametrize("tile_len", [256, 1024, 4096, 0])
def test_config3_north_star_precision(config3_exact, tile_len):
    # the FP32-reseeded path (FP64 accumulation) within 1e-4 of FP64 at the
    # configuration's re-seed intervals B = 256, 1024, 4096 and the default
    A, m, p64, i64 = config3_exact
    print(f"cfg3 f32r B={tile_len}", flush=True)
    p, i, tm = tc.selfjoin_f32r(A, m, tile_len=tile_len, accum64=True)
    assert tm["tile_len"] == (tile_len if tile_len else 32768)
    err = rel_err(p, p64, m)
    print(f"T={tm['tile_len']} max_rel={err.max():.3e} mean_rel={err.mean():.3e} idx_equal={np.mean(i == i64):.5f} "
          f"kernel_ms={tm['kernel_ms']:.1f}")
    assert err.max() <= 1e-4
    assert err.mean() <= 1e-6
    assert np.all(np.abs(np.arange(p.size) - i) > m // 4)
    # a partner that differs from the exact path's is a near tie: its true
    # distance is within the profile tolerance of the exact partner's
    assert_indices_match(A, m, p64, i64, i, abs_tol=0.0, rel_tol=1e-4, min_frac=0.999)


def test_fp64_accumulation_removes_the_drift():
    # FP64 accumulation folds each 16-step FP32 partial into an FP64 sum; without
    # the folds its partial would span the tile and drift like FP32 accumulation.
    # Measured mean unrefined errors at B = 4096: 3.0e-6 (FP64) against 2.0e-5 (FP32).
    A, _ = tc.precision_series(5, 1 << 15, 256)
    m = 256
    p64, _, _ = tc.selfjoin_f64(A, m)
    e = {}
    for acc in (False, True):
        pu, _, _ = tc.selfjoin_f32r(A, m, tile_len=4096, refine=False, accum64=acc)
        e[acc] = rel_err(pu, p64, m).mean()
    print(f"mean unrefined rel err: FP32 accumulation {e[False]:.3e}, FP64 accumulation {e[True]:.3e}")
    assert e[True] <= 0.4 * e[False]


@pytest.mark.parametrize("accum64", [False, True])
@pytest.mark.parametrize("band_k", [4, 8, 16])
def test_band_width_variants_agree(band_k, accum64):
    A = tc.random_walk(71, 30000)
    m = 64
    # main tiles start at the same columns whatever the band width (edge tiles
    # do not), so the FP32 maxima agree to the recurrence's rounding; the
    # partner is the best (in FP64) of the K diagonals of the lane holding the
    # maximum, so it may differ between band widths only among near ties
    pu_ref, _, _ = tc.selfjoin_f32r(A, m, band_k=8, tile_len=2048, accum64=accum64, refine=False)
    pu, _, _ = tc.selfjoin_f32r(A, m, band_k=band_k, tile_len=2048, accum64=accum64, refine=False)
    assert np.all(rel_err(pu, pu_ref, m) <= 1e-4)
    p_ref, i_ref, _ = tc.selfjoin_f32r(A, m, band_k=8, tile_len=2048, accum64=accum64)
    p, i, _ = tc.selfjoin_f32r(A, m, band_k=band_k, tile_len=2048, accum64=accum64)
    assert np.all(rel_err(p, p_ref, m) <= 1e-4)
    assert np.mean(i == i_ref) >= 0.99


@pytest.mark.parametrize("band_k", [4, 8, 16])
@pytest.mark.parametrize("tile_len", [16, 48, 512])
def test_tile_split_extremes(band_k, tile_len):
    # A band is cut into main tiles of whole 16-column windows that lie inside
    # the triangle and one edge tile where the triangle ends.  Cover main tiles
    # of a single window (T=16), tile lengths that are odd multiples of 16, and
    # series whose bands are shorter than 64K + 16 columns (edge tiles only) or
    # just longer.  A column or row left out of every tile would show up as a
    # worse profile value than the exact path's.
    bw = 64 * band_k
    for n, m in ((bw + 40, 16), (2 * bw + 300, 32), (7000, 64)):
        A = tc.random_walk(n + band_k + tile_len, n)
        p64, i64, _ = tc.selfjoin_f64(A, m)
        p, i, tm = tc.selfjoin_f32r(A, m, band_k=band_k, tile_len=tile_len)
        assert tm["tile_len"] == tile_len and tm["band_k"] == band_k
        assert np.all(rel_err(p, p64, m) <= 1e-4), (n, m)
        assert np.all(np.abs(np.arange(p.size) - i) > m // 4), (n, m)


def test_winner_distance_is_recomputed_exactly():
    A = tc.random_walk(72, 5000)
    m = 50
    p, i, _ = tc.selfjoin_f32r(A, m)
    for r in range(0, p.size, 97):
        assert abs(p[r] - orc.pair_dist2(A, m, r, i[r])) <= 1e-8 * max(1.0, p[r])


def test_tiny_geometries():
    for n, m in ((50, 50), (60, 50), (200, 8), (300, 100)):
        A = tc.random_walk(n + m, n)
        p64, i64, _ = tc.selfjoin_f64(A, m)
        p, i, _ = tc.selfjoin_f32r(A, m)
        assert np.all(rel_err(p, p64, m) <= 1e-4), (n, m)


def test_flat_windows_report_the_initial_score():
    # Windows of zero variance have NaN scores in FP64 (1/sigma = inf), which
    # never win: the FP64 path reports the initial score -1.0 (d^2 = 2m + 2)
    # and the initial index for them.  The FP32 path does the same.
    A = tc.random_walk(22, 3000)
    A[1200:1400] = 0.0
    m = 40
    p, i, _ = tc.selfjoin_f32r(A, m)
    flat = np.zeros(p.size, dtype=bool)
    flat[1200:1400 - m + 1] = True
    assert np.all(p[flat] == 2 * m + 2) and np.all(i[flat] == -1)
    # rows whose FP64 partner is an ordinary window agree within tolerance
    p64, i64, _ = tc.selfjoin_f64(A, m)
    ok = ~flat & np.isfinite(p64) & (p64 < 2 * m)
    assert np.all(rel_err(p[ok], p64[ok], m) <= 1e-4)
