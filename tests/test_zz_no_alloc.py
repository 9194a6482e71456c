"""The allocation-free exact self-join (mp_selfjoin_f64_dev_noalloc).

It runs the exact path's kernels in the library's file-scope scratch array
instead of the context's arena; its results must equal mp_selfjoin_f64_dev's
bit for bit, also on a second call that reuses the array with a smaller
series, and series longer than the array holds are refused before any work.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import thesiscode_amd as tc
from thesiscode_amd._ffi import Timing, lib

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 19
MP_ERR_TOO_LONG = -6


def _run(dev, fn, d_ts, n, m, opts):
    P = n - m + 1
    d_p = dev.alloc(8 * P)
    d_i = dev.alloc(8 * P)
    tm = Timing()
    rc = getattr(lib(), fn)(dev._ctx, C.c_void_p(d_ts), n, m, C.c_void_p(d_p), C.c_void_p(d_i), C.byref(opts),
                            C.byref(tm))
    assert rc == 0, (fn, rc)
    p = dev.download(d_p, np.empty(P))
    i = dev.download(d_i, np.empty(P, dtype=np.int64))
    dev.free(d_p)
    dev.free(d_i)
    return p, i


def test_noalloc_equals_the_allocating_entry_point():
    # two calls in a row on the shared array, the second with a smaller series
    with tc.Device(0) as dev:
        for seed, n, m, excl in ((81, MAX_POINTS, 256, -1), (82, 30000, 100, -1), (83, 5000, 64, 3)):
            A = tc.random_walk(seed, n)
            d_ts = dev.upload(A)
            opts = tc.make_opts(exclusion=excl)
            print(f"noalloc n={n} m={m}", flush=True)
            p_ref, i_ref = _run(dev, "mp_selfjoin_f64_dev", d_ts, n, m, opts)
            p, i = _run(dev, "mp_selfjoin_f64_dev_noalloc", d_ts, n, m, opts)
            assert np.array_equal(p, p_ref, equal_nan=True), n
            assert np.array_equal(i, i_ref), n
            dev.free(d_ts)


def test_noalloc_refuses_series_beyond_its_array():
    with tc.Device(0) as dev:
        n = MAX_POINTS + 1
        d_ts = dev.alloc(8 * n)
        d_p = dev.alloc(8 * n)
        d_i = dev.alloc(8 * n)
        rc = lib().mp_selfjoin_f64_dev_noalloc(dev._ctx, C.c_void_p(d_ts), n, 64, C.c_void_p(d_p),
                                               C.c_void_p(d_i), C.byref(tc.make_opts()), None)
        assert rc == MP_ERR_TOO_LONG
