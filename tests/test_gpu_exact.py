"""Parity of the exact FP64 HIP path with the CPU oracle (bit-exact profile values).

The HIP path restates the reference's arithmetic, so every profile value must
equal the oracle's bit for bit; indices must be identical except where the
64-bit merge key (score truncated to 64 - ceil(log2(P+1)) bits) cannot tell two
partners apart, in which case both partners are a near tie.
"""
from __future__ import annotations

import numpy as np
import pytest

import thesiscode_amd as tc
from conftest import assert_indices_match, assert_values_equal_reference_walk, cpu_threads, golden_cases
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _oracle(A, m, **kw):
    p, i, _, _ = orc.triv_par(A, m, nranks=cpu_threads(), nthreads=cpu_threads(), kernel=1, **kw)
    return p, i


def _tie_tol(m):
    return 1e-7 * 2 * m


def test_golden_small_series_bit_exact(golden):
    for name, A, m in golden_cases(golden):
        print("case", name, flush=True)
        p_ref, i_ref = _oracle(A, m)
        p, i, tm = tc.selfjoin_f64(A, m)
        assert np.array_equal(p, p_ref), name
        assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m))
        assert tm["cells"] == tc.cells(A.size, m)


def test_reference_fixture_bit_exact(golden, test_series):
    p, i, _ = tc.selfjoin_f64(test_series, 400)
    p_ref = golden["testbin_m400__oracle_profile"]
    i_ref = golden["testbin_m400__oracle_index"]
    assert np.array_equal(p, p_ref)
    assert_indices_match(test_series, 400, p_ref, i_ref, i, abs_tol=_tie_tol(400))


@pytest.mark.parametrize("band_k", [1, 2, 4, 8])
def test_band_width_does_not_change_results(band_k):
    A = tc.random_walk(31, 20000)
    m = 64
    p_ref, i_ref = _oracle(A, m)
    p, i, tm = tc.selfjoin_f64(A, m, band_k=band_k)
    assert tm["band_k"] == band_k
    assert np.array_equal(p, p_ref)
    assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m))


_FORM_CASES = {}


@pytest.mark.parametrize("walk_form", [1, 2, 3])
@pytest.mark.parametrize("n,m,seed", [(20000, 64, 31), (65536, 400, 0)])
def test_band_width_8_walker_forms_bit_exact(walk_form, n, m, seed):
    # band width 8 in each of its forms, which the automatic choice takes only
    # at n > 2^20: an LDS ring of rows at one wave per SIMD, DPP rows at two,
    # rows in place with the diagonals rotating at two
    A = tc.random_walk(seed, n)
    key = (n, m, seed)
    if key not in _FORM_CASES:
        _FORM_CASES[key] = _oracle(A, m)
    p_ref, i_ref = _FORM_CASES[key]
    print(f"form {walk_form} n={n} m={m}", flush=True)
    p, i, tm = tc.selfjoin_f64(A, m, band_k=8, walk_form=walk_form)
    assert tm["band_k"] == 8
    assert tm["walk_form"] == walk_form
    assert np.array_equal(p, p_ref)
    assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m), min_frac=0.999)


# 15 N(0,1) values (numpy default_rng(5), the 677th of a search for them): with
# m = 8, six of the eight profile entries have their best score in (-1, 0)
NEG_SCORE_SERIES = [0.4970085297607596, -0.03761213355239576, -1.4097385957119746, 0.2362400362428747,
                    -0.3585197480443938, 0.10868428288495499, 1.9872942438908143, -0.5601415338663673,
                    0.4162561732853156, -0.3859331462233166, 0.1400420355261085, -0.9092892407234737,
                    -0.13049969804198341, 0.013218314525724945, 1.8224528368681574]


def test_rotating_form_results_that_stay_negative():
    # The rotating walker tests each step's cells against the column's current
    # result and lets every cell through while that result is not positive.
    # (a) scores in (-1, 0): results that stay negative to the end;
    # (b) every correlation below -1/m (a sine, partners a quarter to three
    #     quarters of a period away): no cell beats the reference's initial
    #     score -1, and every entry keeps d^2 = 2m + 2 with index -1.
    A = np.array(NEG_SCORE_SERIES)
    m = 8
    p_ref, i_ref = _oracle(A, m)
    assert int(((p_ref > 2 * m) & (p_ref < 2 * m + 2)).sum()) == 6
    p, i, tm = tc.selfjoin_f64(A, m, band_k=8, walk_form=3)
    assert tm["walk_form"] == 3
    assert np.array_equal(p, p_ref)
    assert np.array_equal(i, i_ref)
    m, ez = 64, 16
    n = 47 + m - 1
    t = np.arange(n, dtype=np.float64)
    A = np.sin(2 * np.pi * t / (4 * ez)) + 1e-3 * tc.random_walk(77, n)
    p_ref, i_ref = _oracle(A, m)
    assert np.all(p_ref == 2 * m + 2) and np.all(i_ref == -1)
    p, i, _ = tc.selfjoin_f64(A, m, band_k=8, walk_form=3)
    assert np.array_equal(p, p_ref) and np.array_equal(i, i_ref)


def test_walk_form_out_of_range_is_refused():
    with pytest.raises(tc.ScrimpError):
        tc.selfjoin_f64(tc.random_walk(1, 2000), 32, band_k=8, walk_form=5)
    with pytest.raises(tc.ScrimpError):  # column segments: band widths 1, 2 and 4 only
        tc.selfjoin_f64(tc.random_walk(1, 2000), 32, band_k=8, walk_form=4)


def _seg_items(W, band_k, seg_cols=2048):
    bw = 64 * band_k
    r = seg_cols // bw
    return sum(-(-(W - b * bw) // (r * bw)) for b in range(-(-W // bw)))


@pytest.mark.parametrize("band_k", [1, 2, 4])
@pytest.mark.parametrize("n,m,seed", [(20000, 64, 31), (65536, 400, 0)])
def test_column_segments_bit_exact(band_k, n, m, seed):
    # walk_form 4: every band cut into 2048-column segments, walked by waves of
    # their own; a segment after the first starts from the values a pre-pass
    # advanced along its diagonals, and a row that a cut splits gets one result
    # per part, merged by the result keys
    A = tc.random_walk(seed, n)
    key = (n, m, seed)
    if key not in _FORM_CASES:
        _FORM_CASES[key] = _oracle(A, m)
    p_ref, i_ref = _FORM_CASES[key]
    print(f"segments K={band_k} n={n} m={m}", flush=True)
    p, i, tm = tc.selfjoin_f64(A, m, band_k=band_k, walk_form=4)
    W = n - m - m // 4
    assert tm["band_k"] == band_k and tm["walk_form"] == 4
    assert tm["work_items"] == _seg_items(W, band_k) > -(-W // (64 * band_k))
    assert np.array_equal(p, p_ref)
    assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m), min_frac=0.999)


@pytest.mark.parametrize("W", [2047, 2048, 2049, 4096, 4097, 6200])
def test_column_segments_at_segment_boundaries(W):
    # bands whose length is a multiple of 2048, one past it (a one-column last
    # segment) or short of it, at each band width; m = 64, exclusion 16
    m = 64
    n = W + m + m // 4
    A = tc.random_walk(W, n)
    p_ref, i_ref = _oracle(A, m)
    for band_k in (1, 2, 4):
        print(f"segments W={W} K={band_k}", flush=True)
        p, i, tm = tc.selfjoin_f64(A, m, band_k=band_k, walk_form=4)
        assert tm["work_items"] == _seg_items(W, band_k)
        assert np.array_equal(p, p_ref), band_k
        assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m))


def test_column_segments_with_ties_and_negative_results():
    # a period-64 series repeated exactly: the diagonals at multiples of the
    # period tie up to rounding; and the sine whose entries all keep the initial score, cut into segments
    m = 32
    base = tc.random_walk(9, 64)
    A = np.tile(base, 80)
    p_ref, i_ref = _oracle(A, m)
    for band_k in (1, 4):
        p, i, _ = tc.selfjoin_f64(A, m, band_k=band_k, walk_form=4)
        assert np.array_equal(p, p_ref), band_k
        # nearly every entry's partners tie beyond what the merge key resolves:
        # any partner the path reports must be such a tie
        assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m), min_frac=0.0)
    m, ez = 64, 16
    n = 4500
    t = np.arange(n, dtype=np.float64)
    A = np.sin(2 * np.pi * t / (4 * ez)) + 1e-3 * tc.random_walk(78, n)
    p_ref, i_ref = _oracle(A, m)
    p, i, _ = tc.selfjoin_f64(A, m, band_k=4, walk_form=4)
    assert np.array_equal(p, p_ref) and np.array_equal(i, i_ref)


def test_config1_random_walk_bit_exact():
    # n = 65,536 random walk (seed 0), m = 400, exclusion m/4: the CPU-runnable configuration
    A = tc.random_walk(0, 65536)
    m = 400
    print("cfg1 oracle", flush=True)
    p_ref, i_ref = _oracle(A, m)
    print("cfg1 hip", flush=True)
    p, i, tm = tc.selfjoin_f64(A, m)
    assert np.array_equal(p, p_ref)
    assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m), min_frac=0.9999)


@pytest.mark.parametrize("n,m", [(50, 50), (51, 50), (60, 50), (64, 1), (100, 2), (300, 299), (777, 13)])
def test_edge_geometries(n, m):
    A = tc.random_walk(n * 7 + m, n)
    p_ref, i_ref = _oracle(A, m)
    p, i, _ = tc.selfjoin_f64(A, m)
    assert np.array_equal(p, p_ref, equal_nan=True)
    assert np.array_equal(i, i_ref)


def test_exclusion_and_index_init_options():
    A = tc.random_walk(41, 3000)
    m = 40
    for excl in (0, 3, 25):
        p_ref, i_ref = _oracle(A, m, excl=excl)
        p, i, _ = tc.selfjoin_f64(A, m, exclusion=excl)
        assert np.array_equal(p, p_ref)
        assert_indices_match(A, m, p_ref, i_ref, i, abs_tol=_tie_tol(m))
    B = tc.random_walk(42, 40)
    p, i, _ = tc.selfjoin_f64(B, 38, index_init=0)
    assert np.all(i == 0) and np.all(p == 2 * 38 + 2)


def test_flat_windows_follow_reference_semantics():
    A = tc.random_walk(22, 1500)
    A[600:700] = 0.0
    m = 40
    p_ref, i_ref = _oracle(A, m)
    p, i, _ = tc.selfjoin_f64(A, m)
    assert np.array_equal(p, p_ref)
    assert np.array_equal(i, i_ref)


def _block_inputs(seed, n, m, base):
    # one extra element after the series: the last update of every diagonal
    # reads A_vert[trianglen-1+w], one past the series (kernels.cpp:218-222); it
    # is not 0, so the written-back tmpQ shows that the drop-in reads it too
    A = np.append(tc.random_walk(seed, n), 3.75)
    s, mu = orc.window_stats(A[:n], m)
    qt = orc.seeds(A[:n], m, base, n - m)
    return A, s, mu, qt


@pytest.mark.parametrize("blocklen", [1, 37, 500, 5000])
def test_block_drop_in_aliased_self_join_adapter(blocklen):
    # the adapter of ScrimpTrivialPar.cpp:110-140: row view = profile + base
    n, m, base = 3000, 48, 13
    A, s, mu, qt = _block_inputs(51, n, m, base)
    P = n - m + 1
    tl = P - base
    outs = []
    for which in ("oracle", "hip"):
        prof = np.full(P + 1, -1.0)
        idx = np.full(P + 1, -1, dtype=np.int64)
        prof[::7] = 0.25  # pre-existing results from earlier blocks
        idx[::7] = 123
        q = qt[base:base + min(blocklen, tl)].copy()
        args = (prof, idx, prof[base:], idx[base:], q, blocklen, tl, A, A[base:], m, s, mu, s[base:], mu[base:],
                base, 0)
        if which == "oracle":
            orc.eval_block("scalar", *args)
        else:
            tc.eval_diag_block_triangle_f64(*args)
        outs.append((prof, idx, q))
    (po, io, qo), (ph, ih, qh) = outs
    assert np.array_equal(ph, po)
    assert_indices_match(A[:n], m, po, io, ih, abs_tol=_tie_tol(m))
    assert np.array_equal(qh, qo)


def test_block_drop_in_separate_row_and_column_arrays():
    n, m, base = 2500, 32, 200
    A, s, mu, qt = _block_inputs(52, n, m, base)
    P = n - m + 1
    tl, bl = P - base, 300
    outs = []
    for which in ("oracle", "hip"):
        pc = np.full(tl, -1.0)
        ic = np.full(tl, -1, dtype=np.int64)
        pr = np.full(tl, -1.0)
        ir = np.full(tl, -1, dtype=np.int64)
        q = qt[base:base + bl].copy()
        args = (pc, ic, pr, ir, q, bl, tl, A, A[base:], m, s, mu, s[base:], mu[base:], 1000 + base, 77)
        if which == "oracle":
            orc.eval_block("scalar", *args)
        else:
            tc.eval_diag_block_triangle_f64(*args)
        outs.append((pc, ic, pr, ir))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_config2_rotating_form_equals_ring_form():
    # n = 2^20 random walk (seed 1), m = 256: the rotating form (column filter,
    # single-band items) and the LDS-ring form walk the same cells in different
    # orders and lanes, so their values must agree bit for bit at every entry;
    # indices may differ only at near ties
    A = tc.random_walk(1, 1 << 20)
    m = 256
    print("cfg2 ring", flush=True)
    p1, i1, t1 = tc.selfjoin_f64(A, m, band_k=8, walk_form=1)
    print("cfg2 rotating", flush=True)
    p3, i3, t3 = tc.selfjoin_f64(A, m, band_k=8, walk_form=3)
    assert t1["walk_form"] == 1 and t3["walk_form"] == 3
    assert np.array_equal(p1, p3)
    assert_indices_match(A, m, p1, i1, i3, abs_tol=_tie_tol(m), min_frac=0.999)


def test_config2_full_size_properties():
    # n = 2^20 random walk (seed 1), m = 256: too large for the oracle in a test;
    # checked by recomputation, exclusion, bounds, row brute force and determinism
    A = tc.random_walk(1, 1 << 20)
    m, ez = 256, 64
    print("cfg2 run 1", flush=True)
    p, i, tm = tc.selfjoin_f64(A, m)
    print("cfg2 run 2", flush=True)
    p2, i2, _ = tc.selfjoin_f64(A, m)
    assert np.array_equal(p, p2) and np.array_equal(i, i2)
    P = A.size - m + 1
    assert np.all(np.abs(np.arange(P) - i) > ez)
    assert np.all(p > -1e-6) and np.all(p < 4 * m)
    rows = np.linspace(0, P - 1, 64).astype(np.int64)
    for r in rows:
        assert abs(p[r] - orc.pair_dist2(A, m, r, i[r])) <= 1e-6 * max(1.0, p[r])
    # bit for bit at 16 rows: the reference's arithmetic along the partners' diagonals
    assert_values_equal_reference_walk(A, m, p, i, rows[::4])
    # full row brute force on a few rows: rho_j = sum_t (A[j+t]-mu_j) x_t / (m sd_j)
    mu, sd = orc.window_moments_ld(A, m)
    for r in rows[::16]:
        x = (A[r:r + m] - mu[r]) / sd[r]
        cov = np.correlate(A, x, mode="valid") - mu * x.sum()
        d2 = 2 * m * (1 - cov / (m * sd))
        d2[max(0, r - ez):r + ez + 1] = np.inf
        j = int(np.argmin(d2))
        assert abs(p[r] - d2[j]) <= 1e-6 * max(1.0, d2[j])
        if i[r] != j:
            assert abs(d2[i[r]] - d2[j]) <= 1e-6
