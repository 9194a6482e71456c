"""Serving: continuous batching of self-join requests on one GPU.

Each request's result is the same alone and in a mix of concurrent requests
(bit-identical: every tile is computed identically whichever launch carries
it and tiles merge by an order-independent max), requests that do not fit
the HBM budget are refused, and requests that wait for memory are admitted
when it frees.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import thesiscode_amd as tc

pytestmark = pytest.mark.gpu


def _requests(seed, count, lo=10, hi=14):
    u = tc.uniforms(seed, 2 * count, stream=11)
    ns = np.floor(2.0 ** (lo + (hi - lo) * u[:count])).astype(int)
    ms = np.array([64, 128, 256, 512])[np.minimum((u[count:] * 4).astype(int), 3)]
    ms = np.minimum(ms, ns // 4)
    return [(tc.random_walk(seed * 7919 + k, int(n), stream=5), int(m)) for k, (n, m) in enumerate(zip(ns, ms))]


def test_result_is_the_same_alone_and_in_the_mix():
    reqs = _requests(1, 60)
    probe = reqs[7]
    with tc.Device(0) as dev:
        with tc.Server(dev, hbm_budget_bytes=1 << 30, tick_cells=int(5e8)) as srv:
            rid = srv.submit(*probe)
            alone_p, alone_i, lat_alone = srv.wait(rid)
            ids = [srv.submit(ts, m) for ts, m in reqs]  # the probe again, among 59 others
            out = [srv.wait(r) for r in ids]
            st = srv.stats()
    mix_p, mix_i, _ = out[7]
    assert np.array_equal(mix_p, alone_p)
    assert np.array_equal(mix_i, alone_i)
    assert st["ticks"] >= 2 and st["completed"] == 61
    # and every request equals the FP32-reseeded self-join with the same tiling
    for (ts, m), (p, i, _) in list(zip(reqs, out))[::10]:
        p1, i1, _ = tc.selfjoin_f32r(ts, m, band_k=8, tile_len=4096)
        assert np.array_equal(p, p1) and np.array_equal(i, i1)


def test_requests_are_within_fp32_tolerance_of_exact():
    reqs = _requests(2, 8)
    with tc.Device(0) as dev:
        with tc.Server(dev, hbm_budget_bytes=1 << 29) as srv:
            ids = [srv.submit(ts, m) for ts, m in reqs]
            for (ts, m), rid in zip(reqs, ids):
                p, i, lat = srv.wait(rid)
                p64, _, _ = tc.selfjoin_f64(ts, m)
                assert np.max(np.abs(p - p64) / np.maximum(p64, 1e-3 * 2 * m)) <= 1e-4
                assert lat > 0.0


def test_admission_by_hbm_footprint():
    with tc.Device(0) as dev:
        # 64 bytes per point: 4 MB holds ~65k points
        with tc.Server(dev, hbm_budget_bytes=4 << 20) as srv:
            with pytest.raises(tc.ScrimpError) as e:
                srv.submit(tc.random_walk(3, 200000), 128)
            assert e.value.status == -9  # MP_ERR_ADMISSION
            # three 30k-point requests cannot be resident together: they queue and still finish
            reqs = [(tc.random_walk(40 + k, 30000), 128) for k in range(3)]
            ids = [srv.submit(ts, m) for ts, m in reqs]
            results = [srv.wait(r) for r in ids]
            for (ts, m), (p, i, _) in zip(reqs, results):
                p1, i1, _ = tc.selfjoin_f32r(ts, m, band_k=8, tile_len=4096)
                assert np.array_equal(p, p1)
            assert srv.stats()["hbm_used_bytes"] == 0


def test_poll_reports_progress_then_result():
    with tc.Device(0) as dev:
        with tc.Server(dev, hbm_budget_bytes=1 << 28) as srv:
            ts = tc.random_walk(77, 20000)
            rid = srv.submit(ts, 100)
            t0 = time.time()
            res = None
            while res is None and time.time() - t0 < 60:
                res = srv.poll(rid)
                time.sleep(0.001)
            assert res is not None
            p, i, lat, q = res
            assert lat >= q >= 0.0
            assert np.all(np.abs(np.arange(p.size) - i) > 25)
