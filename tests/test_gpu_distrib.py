"""distrib_par's checkerboard through the kernel-level drop-in (ScrimpDistribPar.cpp:444-1116).

Every tile of the k x k checkerboard is evaluated by
mp_eval_diag_block_triangle_f64 the way eval_tile_blocked calls the reference
kernel (ScrimpDistribPar.cpp:205-247): lower tiles directly, upper tiles
transposed (:819-862), from the tile inputs distrib_par builds (slices with
their own statistics, seeds, zeroed padding; taken from the oracle's
restatement, as the reference's host code builds them).  The tiles' horizontal
and vertical results are merged with the score-max rule (:175-200) and
transformed to d2 (:1010-1015).  Profile values must equal the oracle's
distrib_par bit for bit; indices may differ only at near ties.
"""
from __future__ import annotations

import numpy as np
import pytest

import thesiscode_amd as tc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _drop_in_distrib_par(A, m, k, blocklen):
    n = A.size
    pad, L = orc.distrib_geometry(n, m, k)
    ez = m // 4
    PL = n + pad - m + 1
    score = np.full(PL, -1.0)
    sidx = np.full(PL, -1, dtype=np.int64)
    for rank in range(k * k):
        t = orc.distrib_tile_inputs(A, m, k, rank)
        ro, co = t["row_offset"], t["col_offset"]
        hor = np.full(L + ez + 1, -1.0)
        hidx = np.full(L + ez + 1, -1, dtype=np.int64)
        vert = np.full(L, -1.0)
        vidx = np.full(L, -1, dtype=np.int64)
        if t["upper"]:
            # (prof_colmin, idx_colmin, prof_rowmin, idx_rowmin, tmpQ, tilelen, A_hor, A_vert,
            #  s_hor, mu_hor, s_vert, mu_vert, baserow, basecol): the upper triangle transposed
            args = (vert, vidx, hor[1:], hidx[1:], t["dots"][1:], L - 1, t["A"], t["B"][1:],
                    t["sA"], t["muA"], t["sB"][1:], t["muB"][1:], co + 1, ro)
        else:
            args = (hor, hidx, vert, vidx, t["dots"], L, t["B"], t["A"],
                    t["sB"], t["muB"], t["sA"], t["muA"], ro, co)
        pc, ic, pr, ir, q, tl, Ah, Av, sh, mh, sv, mv, brow, bcol = args
        for bo in range(0, tl, blocklen):  # eval_tile_blocked's block loop
            nd = min(blocklen, tl - bo)
            tc.eval_diag_block_triangle_f64(pc, ic, pr[bo:], ir[bo:], q[bo:], nd, tl - bo, Ah, Av[bo:], m, sh, mh,
                                            sv[bo:], mv[bo:], brow + bo, bcol)
        for res, ridx, off in ((hor[:L], hidx[:L], co), (vert, vidx, ro)):
            take = res > score[off:off + L]
            score[off:off + L][take] = res[take]
            sidx[off:off + L][take] = ridx[take]
    P = n - m + 1
    return 2.0 * m - 2.0 * score[:P], sidx[:P]


@pytest.mark.parametrize("n,m,k,blocklen", [(3000, 48, 2, 500), (3000, 48, 3, 500), (3001, 48, 3, 333),
                                            (2003, 16, 5, 100), (4000, 64, 4, 1000)])
def test_checkerboard_through_the_drop_in(n, m, k, blocklen):
    A = tc.random_walk(n + k + 1, n)
    print(f"distrib n={n} m={m} k={k} blocklen={blocklen}", flush=True)
    p_ref, i_ref = orc.distrib_par(A, m, k, blocklen=blocklen)
    p, i = _drop_in_distrib_par(A, m, k, blocklen)
    assert np.array_equal(p, p_ref)
    pad = orc.distrib_geometry(n, m, k)[0]
    Apad = np.concatenate([A, np.zeros(pad)])
    for j in np.nonzero(i != i_ref)[0]:
        d1 = orc.pair_dist2(Apad, m, j, i[j])
        d2 = orc.pair_dist2(Apad, m, j, i_ref[j])
        assert abs(d1 - d2) <= 1e-7 * 2 * m, (j, i[j], i_ref[j])
    assert np.mean(i == i_ref) >= 0.999
