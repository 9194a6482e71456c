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
from oracle import oracle as orc

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


def test_config5_poisson_trace():
    # The serving configuration, shortened to 5 s of arrivals: Poisson at 2,000
    # requests/s, lengths lognormal (median 2^14, sigma 1) clipped to
    # [2^10, 2^19], m uniform in {64, 128, 256, 512}, replayed in real time.
    # Every request completes; a sample (the longest request among them) equals
    # its own single FP32-reseeded run with the server's tiling bit for bit, and
    # is within 1e-4 of the true profile.  The reference is the long double
    # distance: the exact FP64 path's prefix-sum statistics lose up to ~2.4e-4 of
    # d2 on these requests (slices of a walk far from 0, windows of 64 points:
    # the window sums cancel), so its values are not the truth here.
    import threading

    from thesiscode_amd.workloads import serve_requests

    arrivals, lengths, windows, walk, offs = serve_requests(5, 5.0, 2000.0, walk_len=1 << 22)
    count = len(lengths)
    rng = np.random.default_rng(5)
    sample = set(rng.choice(count, 19, replace=False).tolist()) | {int(np.argmax(lengths))}
    print(f"cfg5 trace: {count} requests, longest {lengths.max()}, {lengths.sum()} points", flush=True)
    kept, lat, errors = {}, {}, []
    with tc.Device(0) as dev:
        with tc.Server(dev, hbm_budget_bytes=16 << 30) as srv:
            srv.wait(srv.submit(walk[:20000], 128))  # warm the kernels
            ids = {}
            t0 = time.perf_counter()

            def submitter():
                try:
                    for k in range(count):
                        delay = t0 + arrivals[k] - time.perf_counter()
                        if delay > 0:
                            time.sleep(delay)
                        ids[srv.submit(walk[offs[k]:offs[k] + lengths[k]], int(windows[k]))] = k
                except Exception as e:  # surfaced by the main thread
                    errors.append(e)

            th = threading.Thread(target=submitter)
            th.start()
            deadline = t0 + 120.0
            pending = {}
            while (th.is_alive() or len(lat) < count) and not errors:
                for rid, st, l_ms, _, p, i in srv.completed():
                    assert st == 0, (rid, st)
                    pending[rid] = (l_ms, p, i)
                for rid in [r for r in pending if r in ids]:
                    l_ms, p, i = pending.pop(rid)
                    k = ids[rid]
                    lat[k] = l_ms
                    if k in sample:
                        kept[k] = (p, i)
                assert time.perf_counter() < deadline, f"{len(lat)}/{count} requests done"
                time.sleep(0.002)
            th.join()
            wall = time.perf_counter() - t0
            st = srv.stats()
    assert not errors, errors
    assert len(lat) == count and st["completed"] == count + 1
    lv = np.array(list(lat.values()))
    cells = sum(tc.cells(int(n), int(m)) for n, m in zip(lengths, windows))
    print(f"served {count} requests in {wall:.2f} s ({cells / wall:.3e} cells/s), "
          f"p50 {np.percentile(lv, 50):.2f} ms, p99 {np.percentile(lv, 99):.2f} ms, ticks {st['ticks']}", flush=True)
    for k in sorted(sample):
        ts = walk[offs[k]:offs[k] + lengths[k]]
        m = int(windows[k])
        p, i = kept[k]
        p1, i1, _ = tc.selfjoin_f32r(ts, m, band_k=8, tile_len=4096)
        assert np.array_equal(p, p1) and np.array_equal(i, i1), k
        p64, i64, _ = tc.selfjoin_f64(ts, m)
        floor = 2 * m * 1e-3
        worst = 0.0
        for r in np.linspace(0, p.size - 1, 100).astype(np.int64):
            d = orc.pair_dist2(ts, m, int(r), int(i[r]))      # the served partner, long double
            d64 = orc.pair_dist2(ts, m, int(r), int(i64[r]))  # the FP64 path's partner
            # the served value is its partner's distance (FP64 refinement) ...
            assert abs(p[r] - d) <= 1e-8 * max(d, floor), (k, r, p[r], d)
            # ... and that partner is as good as the FP64 path's within 1e-4
            assert d <= d64 + 1e-4 * max(d64, floor), (k, r, d, d64)
            # and the values agree within 1e-4 up to the FP64 path's own error
            assert abs(p[r] - p64[r]) <= 1e-4 * max(d64, floor) + abs(p64[r] - d64), (k, r)
            worst = max(worst, abs(p[r] - d64) / max(d64, floor))
        print(f"request {k}: n={ts.size} m={m} max rel vs long double {worst:.2e}", flush=True)


def test_small_requests_are_not_held_behind_a_large_one():
    # Continuous batching, not a queue: a 2^19-point request (about 1.4e11 cells)
    # is submitted first and 16 small ones right after it; every tick gives each
    # active request a share of its tiles, so the small ones finish while the
    # large one still runs, and each result equals its own single run.
    big = (tc.random_walk(91, 1 << 19), 64)
    small = [(tc.random_walk(92 + k, 1 << 12), 64) for k in range(16)]
    with tc.Device(0) as dev:
        with tc.Server(dev, hbm_budget_bytes=1 << 30) as srv:
            big_id = srv.submit(*big)
            small_ids = [srv.submit(ts, m) for ts, m in small]
            order, res = [], {}
            t0 = time.time()
            while len(order) < 17 and time.time() - t0 < 60:
                for rid, status, lat, q, p, i in srv.completed():
                    assert status == 0
                    order.append(rid)
                    res[rid] = (p, i, lat)
                time.sleep(0.0005)
    assert len(order) == 17
    assert order[-1] == big_id, "a small request finished after the large one"
    assert max(res[r][2] for r in small_ids) < res[big_id][2]
    for (ts, m), rid in [(big, big_id), (small[0], small_ids[0]), (small[15], small_ids[15])]:
        p1, i1, _ = tc.selfjoin_f32r(ts, m, band_k=8, tile_len=4096)
        assert np.array_equal(res[rid][0], p1) and np.array_equal(res[rid][1], i1)


def test_wait_and_completed_on_two_threads():
    # mp_server_wait on one thread while another drains mp_server_completed:
    # every request is handed out (to the waiter, or to the drain when it took
    # it first, then the waiter gets MP_ERR_NOT_FOUND), and results are intact.
    import ctypes as C
    import threading

    from thesiscode_amd._ffi import lib

    reqs = _requests(3, 40)
    bufs = [(np.empty(ts.size - m + 1), np.empty(ts.size - m + 1, dtype=np.int64)) for ts, m in reqs]
    with tc.Device(0) as dev:
        with tc.Server(dev, hbm_budget_bytes=1 << 30) as srv:
            h = srv._h
            ids = []
            for (ts, m), (p, i) in zip(reqs, bufs):
                rid = C.c_uint64()
                assert lib().mp_server_submit(h, ts.ctypes.data_as(C.POINTER(C.c_double)), ts.size, m,
                                              p.ctypes.data_as(C.POINTER(C.c_double)),
                                              i.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(rid)) == 0
                ids.append(rid.value)
            waited, drained = {}, set()

            def waiter():
                for rid in ids:
                    waited[rid] = lib().mp_server_wait(h, rid, None)

            def drainer():
                ib = (C.c_uint64 * 64)()
                sb = (C.c_int32 * 64)()
                cnt = C.c_int32()
                while len(waited) < len(ids):
                    assert lib().mp_server_completed(h, 64, ib, sb, None, None, C.byref(cnt)) == 0
                    for k in range(cnt.value):
                        assert sb[k] == 0
                        drained.add(ib[k])
                    time.sleep(0.0005)

            ths = [threading.Thread(target=waiter), threading.Thread(target=drainer)]
            for t in ths:
                t.start()
            for t in ths:
                t.join(120)
            assert not any(t.is_alive() for t in ths)
    for rid in ids:
        assert waited[rid] == 0 or (waited[rid] == -10 and rid in drained), (rid, waited[rid])
    for (ts, m), (p, i) in list(zip(reqs, bufs))[::8]:
        p1, i1, _ = tc.selfjoin_f32r(ts, m, band_k=8, tile_len=4096)
        assert np.array_equal(p, p1) and np.array_equal(i, i1)
