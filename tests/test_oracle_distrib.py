"""Pinning of the oracle's distrib_par restatement (ScrimpDistribPar.cpp:444-1116, CPU only).

distrib_par cuts the triangle into k x k tiles (k^2 ranks): per tile, slices
of the series with their own statistics and seeds, the block kernel on the
lower triangle or on the transposed upper one, then a score-max merge.  It is
pinned by:
  - the checkerboard's structure (rank -> row, column, upper) and the padding
    arithmetic, from their closed forms;
  - equality with scrimp_triv_par (itself pinned by brute force) within
    rounding when no padding is needed: the per-slice statistics and per-tile
    seeds round differently, nothing else differs; and with the long-double
    brute force;
  - with padding, the reference's own quirk: windows reaching into the zero
    padding are real windows to it, so a few entries may take such a window as
    partner; every other entry still equals scrimp_triv_par's;
  - block-length invariance (bit for bit) and planted mistakes (upper tiles
    dropped, upper seeds not transposed) that the checks catch.
"""
from __future__ import annotations

import numpy as np
import pytest

import thesiscode_amd as tc
from oracle import oracle as orc


def _tol(p, m):
    return 1e-9 * (np.abs(p) + 2 * m)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_checkerboard_structure(k):
    # row r holds ranks r^2 .. r^2 + 2r: one lower tile per column c <= r and
    # one upper tile per column c < r (checkerboard_partitioning.cpp:8-41)
    tiles = [orc.partition_coords(r) for r in range(k * k)]
    want = {(r, c, False) for r in range(k) for c in range(r + 1)} | {(r, c, True) for r in range(k) for c in range(r)}
    assert set(tiles) == want and len(tiles) == len(want)
    for rank, (r, c, up) in enumerate(tiles):
        assert r * r <= rank < (r + 1) * (r + 1)
        if c == r:
            assert not up  # diagonal tiles are lower triangles only


@pytest.mark.parametrize("n,m,k", [(3000, 48, 2), (3001, 48, 2), (3002, 48, 4), (2003, 16, 5), (100, 20, 7)])
def test_padding_geometry(n, m, k):
    pad, L = orc.distrib_geometry(n, m, k)
    ez = m // 4
    W = n + pad - m - ez  # the padded triangle's base
    assert 0 <= pad < k and W % k == 0 and L == W // k
    assert (pad == 0) == ((n - m - ez) % k == 0)


def test_too_short_series_is_refused():
    with pytest.raises(ValueError):
        orc.distrib_par(tc.random_walk(1, 30), 30, 2)


@pytest.mark.parametrize("n,m,k", [(3000, 48, 1), (3000, 48, 2), (3000, 48, 3), (2501, 40, 3), (4000, 64, 4)])
def test_unpadded_equals_triv_par_and_brute_force(n, m, k):
    A = tc.random_walk(n + k, n)
    assert orc.distrib_geometry(n, m, k)[0] == 0
    p, i = orc.distrib_par(A, m, k)
    pt, it, _, _ = orc.triv_par(A, m)
    pb, ib = orc.brute_force(A, m)
    assert np.all(np.abs(p - pt) <= _tol(pt, m))
    assert np.all(np.abs(p - pb) <= _tol(pb, m))
    # indices equal except where two partners are a near tie
    for j in np.nonzero(i != it)[0]:
        assert abs(orc.pair_dist2(A, m, j, i[j]) - orc.pair_dist2(A, m, j, it[j])) <= 1e-7 * 2 * m
    assert np.mean(i == it) >= 0.999


@pytest.mark.parametrize("n,m,k", [(3001, 48, 2), (3001, 48, 3), (3002, 48, 4), (2003, 16, 5)])
def test_padded_entries(n, m, k):
    A = tc.random_walk(n + k, n)
    pad, _ = orc.distrib_geometry(n, m, k)
    assert pad > 0
    p, i = orc.distrib_par(A, m, k)
    pt, it, _, _ = orc.triv_par(A, m)
    P = n - m + 1
    Apad = np.concatenate([A, np.zeros(pad)])
    real = i < P
    assert np.mean(real) >= 0.99
    assert np.all(np.abs(p[real] - pt[real]) <= _tol(pt[real], m))
    for j in np.nonzero(~real)[0]:
        # a window reaching into the zero padding as partner: its true distance
        # on the padded series, and never worse than the best real partner
        assert abs(p[j] - orc.pair_dist2(Apad, m, j, i[j])) <= 1e-7 * (p[j] + 2 * m)
        assert p[j] <= pt[j] + _tol(pt[j], m)


def test_block_length_does_not_change_values():
    A = tc.random_walk(7, 2600)
    m, k = 40, 3
    p500, _ = orc.distrib_par(A, m, k, blocklen=500)
    p37, _ = orc.distrib_par(A, m, k, blocklen=37)
    p1, _ = orc.distrib_par(A, m, k, blocklen=10000, kernel=0)
    assert np.array_equal(p500, p37) and np.array_equal(p500, p1)


@pytest.mark.parametrize("code", [6, 7])
def test_planted_mistakes_fail_the_brute_force_check(code):
    A = tc.random_walk(11, 2000)
    m, k = 32, 3
    pb, _ = orc.brute_force(A, m)
    with orc.fault(code):
        p, _ = orc.distrib_par(A, m, k)
    assert not np.all(np.abs(p - pb) <= _tol(pb, m))
