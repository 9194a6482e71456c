"""Batched FP32-reseeded self-joins: every series of a batch gets exactly the
result it gets alone, replays of the captured hipGraph are identical to eager
runs (with eager work and syncs between replays), and the autotuner's choice
keeps results within the FP32 tolerance of the exact path."""
from __future__ import annotations

import numpy as np
import pytest

import thesiscode_amd as tc

pytestmark = pytest.mark.gpu


def make_batch(seed, count, lo=10, hi=13):
    u = tc.uniforms(seed, 2 * count, stream=7)
    lengths = np.floor(2.0 ** (lo + (hi - lo) * u[:count])).astype(np.int64)
    windows = np.array([32, 64, 128, 256])[np.minimum((u[count:] * 4).astype(int), 3)].astype(np.int32)
    series = [tc.random_walk(seed * 1000 + s, int(n), stream=3) for s, n in enumerate(lengths)]
    return lengths, windows, series


def run_batch(dev, lengths, windows, series, use_graph, **kw):
    with tc.Batch(dev, lengths, windows, **kw) as b:
        d_ts = dev.upload(np.concatenate(series))
        P = int(b.out_offsets[-1])
        d_p = dev.alloc(8 * P)
        d_i = dev.alloc(8 * P)
        tm = b.run(d_ts, d_p, d_i, use_graph=use_graph)
        prof = dev.download(d_p, np.empty(P))
        idx = dev.download(d_i, np.empty(P, dtype=np.int64))
        return b, prof, idx, tm


def test_batch_matches_single_series():
    lengths, windows, series = make_batch(3, 40)
    with tc.Device(0) as dev:
        b, prof, idx, tm = run_batch(dev, lengths, windows, series, False, band_k=8, tile_len=512)
        oo = b.out_offsets
        for s in range(len(series)):
            p1, i1, _ = tc.selfjoin_f32r(series[s], int(windows[s]), band_k=8, tile_len=512)
            assert np.array_equal(prof[oo[s]:oo[s + 1]], p1), s
            assert np.array_equal(idx[oo[s]:oo[s + 1]], i1), s
        assert tm["cells"] == sum(tc.cells(int(n), int(m)) for n, m in zip(lengths, windows))


def test_batch_of_short_series_matches_single_series():
    # profiles of 3 to 90 entries: one unpack chunk of 64 entries spans many series
    u = tc.uniforms(21, 120, stream=7)
    windows = np.array([16, 24, 32])[np.minimum((u[60:] * 3).astype(int), 2)].astype(np.int32)
    lengths = (windows + windows // 4 + 2 + np.floor(u[:60] * 88)).astype(np.int64)
    series = [tc.random_walk(7000 + s, int(n), stream=3) for s, n in enumerate(lengths)]
    with tc.Device(0) as dev:
        b, prof, idx, tm = run_batch(dev, lengths, windows, series, False, band_k=4, tile_len=256)
        oo = b.out_offsets
        for s in range(len(series)):
            p1, i1, _ = tc.selfjoin_f32r(series[s], int(windows[s]), band_k=4, tile_len=256)
            assert np.array_equal(prof[oo[s]:oo[s + 1]], p1), s
            assert np.array_equal(idx[oo[s]:oo[s + 1]], i1), s


def test_full_4096_series_batch_sample_matches_single_series():
    # the measured batch (4,096 walks, n log-uniform in [2^10, 2^17], m in
    # {32, 64, 128, 256}; FP32-representable values), default band and tiles,
    # replayed from its hipGraph; 24 series checked against their own runs
    count = 4096
    u = tc.uniforms(3, count, stream=0)
    v = tc.uniforms(3, count, stream=1)
    lengths = np.floor(2.0 ** (10.0 + 7.0 * u)).astype(np.int64)
    windows = np.array([32, 64, 128, 256], dtype=np.int32)[np.minimum((4 * v).astype(int), 3)]
    series = [tc.random_walk(3, int(n), stream=100 + s).astype(np.float32).astype(np.float64)
              for s, n in enumerate(lengths)]
    with tc.Device(0) as dev:
        with tc.Batch(dev, lengths, windows) as b:
            d_ts = dev.upload(np.concatenate(series))
            P = int(b.out_offsets[-1])
            d_p = dev.alloc(8 * P)
            d_i = dev.alloc(8 * P)
            b.run(d_ts, d_p, d_i, use_graph=True)
            tm = b.run(d_ts, d_p, d_i, use_graph=True)
            prof = dev.download(d_p, np.empty(P))
            idx = dev.download(d_i, np.empty(P, dtype=np.int64))
            info = b.info()
            oo = b.out_offsets
    print(f"tiles={info['tiles']} band_k={info['band_k']} T={info['tile_len']} ms={tm['total_ms']:.1f}")
    assert tm["cells"] == sum(tc.cells(int(n), int(m)) for n, m in zip(lengths, windows))
    rng = np.random.default_rng(5)
    picks = list(rng.choice(count, 22, replace=False)) + [int(np.argmax(lengths)), int(np.argmin(lengths))]
    for s in picks:
        p1, i1, _ = tc.selfjoin_f32r(series[s], int(windows[s]), band_k=info["band_k"],
                                     tile_len=info["tile_len"])
        assert np.array_equal(prof[oo[s]:oo[s + 1]], p1), s
        assert np.array_equal(idx[oo[s]:oo[s + 1]], i1), s


def test_batch_within_fp32_tolerance_of_exact():
    lengths, windows, series = make_batch(4, 12)
    with tc.Device(0) as dev:
        b, prof, idx, _ = run_batch(dev, lengths, windows, series, True)
        oo = b.out_offsets
        for s in range(len(series)):
            m = int(windows[s])
            p64, _, _ = tc.selfjoin_f64(series[s], m)
            err = np.abs(prof[oo[s]:oo[s + 1]] - p64) / np.maximum(p64, 1e-3 * 2 * m)
            assert err.max() <= 1e-4, (s, err.max())


def test_graph_replays_match_eager_runs():
    lengths, windows, series = make_batch(5, 25)
    with tc.Device(0) as dev:
        with tc.Batch(dev, lengths, windows, band_k=4, tile_len=256) as b:
            d_ts = dev.upload(np.concatenate(series))
            P = int(b.out_offsets[-1])
            d_p = dev.alloc(8 * P)
            d_i = dev.alloc(8 * P)
            b.run(d_ts, d_p, d_i, use_graph=False)
            ref_p = dev.download(d_p, np.empty(P))
            ref_i = dev.download(d_i, np.empty(P, dtype=np.int64))
            for rep in range(4):
                b.run(d_ts, d_p, d_i, use_graph=True)
                dev.sync()
                assert np.array_equal(dev.download(d_p, np.empty(P)), ref_p), rep
                assert np.array_equal(dev.download(d_i, np.empty(P, dtype=np.int64)), ref_i), rep
                # eager work between replays
                tc.selfjoin_f32r(series[rep], int(windows[rep]))
                b.run(d_ts, d_p, d_i, use_graph=False)
                assert np.array_equal(dev.download(d_p, np.empty(P)), ref_p), rep


def test_autotuner_records_and_reuses_choice(tmp_path, monkeypatch):
    monkeypatch.setenv("SCRIMP_HIP_TUNE_CACHE", str(tmp_path / "tune.txt"))
    lengths, windows, series = make_batch(6, 30, 11, 14)
    with tc.Device(0) as dev:
        with tc.Batch(dev, lengths, windows) as b:
            d_ts = dev.upload(np.concatenate(series))
            P = int(b.out_offsets[-1])
            d_p = dev.alloc(8 * P)
            d_i = dev.alloc(8 * P)
            best = b.tune(d_ts, d_p, d_i, max_tile_len=2048)
            assert best["band_k"] in (4, 8) and 256 <= best["tile_len"] <= 2048
            assert b.info()["tile_len"] == best["tile_len"]
        text = (tmp_path / "tune.txt").read_text().split()
        assert int(text[1]) == best["band_k"] and int(text[2]) == best["tile_len"]
        with tc.Batch(dev, lengths, windows) as b2:
            assert b2.info()["band_k"] == best["band_k"] and b2.info()["tile_len"] == best["tile_len"]
