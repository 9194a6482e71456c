"""MI355X-native SCRIMP matrix-profile self-join (Python view of the C ABI).

Results follow the reference's MatProfSOA (scrimppp/inc/ScrimpTrivialPar.hpp:16-20):
``profile[i]`` is the squared z-normalised distance ``2m - 2*score`` of window
i's best match and ``index[i]`` its start (-1 when nothing beat the initial
score).  The exclusion zone defaults to ``m // 4`` (ScrimpTrivialPar.cpp:150).

Every call runs on the HIP device through libscrimp_hip.so; there is no CPU
fallback, and a missing library raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import Opts, ScrimpError, Timing, check, lib

__all__ = [
    "ScrimpError", "Device", "selfjoin_f64", "selfjoin_f32r", "prescrimp_f64", "eval_diag_block_triangle_f64",
    "random_walk", "normals", "uniforms", "precision_series", "profile_length", "exclusion_zone",
    "cells", "ALGORITHMS", "make_opts", "Batch", "Server",
]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _lptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def exclusion_zone(m: int) -> int:
    return m // 4


def profile_length(n: int, m: int) -> int:
    return n - m + 1


def cells(n: int, m: int, excl: int | None = None) -> int:
    """Cells one self-join evaluates: W(W+1)/2 with W = P - excl - 1."""
    excl = m // 4 if excl is None else excl
    w = max(0, profile_length(n, m) - excl - 1)
    return w * (w + 1) // 2


def make_opts(exclusion: int = -1, band_k: int = 0, tile_len: int = 0, refine: bool = True,
              index_init: int = -1, accum64: bool = True, walk_form: int = 0) -> Opts:
    o = Opts()
    lib().mp_opts_default(C.byref(o))
    o.exclusion = exclusion
    o.band_k = band_k
    o.tile_len = tile_len
    o.refine = 1 if refine else 0
    o.index_init = index_init
    o.accum64 = 1 if accum64 else 0
    o.walk_form = walk_form
    return o


def _selfjoin(fn_name: str, ts, m: int, opts: Opts):
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    n = ts.shape[0]
    P = max(0, n - m + 1)
    prof = np.empty(P, dtype=np.float64)
    idx = np.empty(P, dtype=np.int64)
    tm = Timing()
    rc = getattr(lib(), fn_name)(_dptr(ts), n, m, _dptr(prof), _lptr(idx), C.byref(opts), C.byref(tm))
    check(rc, fn_name)
    return prof, idx, tm.as_dict()


def selfjoin_f64(ts, m: int, *, exclusion: int = -1, band_k: int = 0, index_init: int = -1, walk_form: int = 0):
    """Plain FP64 self-join, bit-identical to the reference arithmetic (walk_form: see mp_opts)."""
    return _selfjoin("mp_selfjoin_f64", ts, m,
                     make_opts(exclusion, band_k, 0, True, index_init, walk_form=walk_form))


def selfjoin_f32r(ts, m: int, *, exclusion: int = -1, band_k: int = 0, tile_len: int = 0,
                  refine: bool = True, index_init: int = -1, accum64: bool = True):
    """FP32 recurrence, FP64 statistics and per-tile FP64 seeds (tile_len = re-seed interval).

    accum64 (default) keeps each diagonal's covariance in FP64, advanced by FP32 partial
    sums over 16 steps, which keeps the profile within 1e-4 of FP64; accum64=False
    accumulates in FP32 (faster, ~1e-3 on long near-periodic series).
    """
    return _selfjoin("mp_selfjoin_f32r", ts, m, make_opts(exclusion, band_k, tile_len, refine, index_init, accum64))


def prescrimp_f64(ts, m: int, stride: float = 0.25):
    """PreSCRIMP, the anytime phase of scrimppp_orig (ScrimpppOrig.cpp:130-283): an upper
    bound of the profile, exact at the sampled rows 0, step, 2*step, ... (step = floor(stride*m)).
    Returns (d2 profile, index, timing); d2 = +inf and index 0 where no update reached."""
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    n = ts.shape[0]
    P = max(0, n - m + 1)
    prof = np.empty(P, dtype=np.float64)
    idx = np.empty(P, dtype=np.int64)
    tm = Timing()
    rc = lib().mp_prescrimp_f64(_dptr(ts), n, m, float(stride), _dptr(prof), _lptr(idx), C.byref(tm))
    check(rc, "mp_prescrimp_f64")
    return prof, idx, tm.as_dict()


def eval_diag_block_triangle_f64(prof_colmin, idx_colmin, prof_rowmin, idx_rowmin, tmpQ, blocklen,
                                 trianglen, A_hor, A_vert, window, s_hor, mu_hor, s_vert, mu_vert,
                                 baserow, basecol):
    """In-place drop-in of matrix_profile::eval_diag_block_triangle (kernels.h:7-24).

    Array arguments are float64/int64 numpy arrays or views of them; row and
    column views may alias one array exactly as the reference's adapters do.
    """
    for a, dt in ((prof_colmin, np.float64), (prof_rowmin, np.float64), (tmpQ, np.float64),
                  (idx_colmin, np.int64), (idx_rowmin, np.int64)):
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError("output arrays must be contiguous float64/int64")
    ins = [np.ascontiguousarray(x, dtype=np.float64) for x in (A_hor, A_vert, s_hor, mu_hor, s_vert, mu_vert)]
    rc = lib().mp_eval_diag_block_triangle_f64(
        _dptr(prof_colmin), _lptr(idx_colmin), _dptr(prof_rowmin), _lptr(idx_rowmin), _dptr(tmpQ),
        int(blocklen), int(trianglen), _dptr(ins[0]), _dptr(ins[1]), int(window), _dptr(ins[2]),
        _dptr(ins[3]), _dptr(ins[4]), _dptr(ins[5]), int(baserow), int(basecol))
    check(rc, "mp_eval_diag_block_triangle_f64")


ALGORITHMS = {"scrimp_hip_f64": selfjoin_f64, "scrimp_hip_f32r": selfjoin_f32r}


class Device:
    """A library context on one GPU, for inputs kept resident in HBM."""

    def __init__(self, device: int = 0):
        self._ctx = C.c_void_p()
        check(lib().mp_ctx_create(device, C.byref(self._ctx)), "mp_ctx_create")
        self._bufs: list[int] = []

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        check(lib().mp_dev_alloc(self._ctx, nbytes, C.byref(p)), "mp_dev_alloc")
        self._bufs.append(p.value)
        return p.value

    def free(self, ptr: int) -> None:
        if ptr in self._bufs:
            self._bufs.remove(ptr)
            check(lib().mp_dev_free(self._ctx, C.c_void_p(ptr)), "mp_dev_free")

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        check(lib().mp_memcpy_h2d(self._ctx, C.c_void_p(p), arr.ctypes.data_as(C.c_void_p), arr.nbytes), "h2d")
        return p

    def download(self, ptr: int, out: np.ndarray) -> np.ndarray:
        check(lib().mp_memcpy_d2h(self._ctx, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes), "d2h")
        return out

    def selfjoin(self, precision: str, d_ts: int, n: int, m: int, d_prof: int, d_idx: int,
                 opts: Opts | None = None) -> dict:
        fn = {"f64": "mp_selfjoin_f64_dev", "f32r": "mp_selfjoin_f32r_dev",
              "f64_noalloc": "mp_selfjoin_f64_dev_noalloc"}[precision]
        opts = opts or make_opts()
        tm = Timing()
        rc = getattr(lib(), fn)(self._ctx, C.c_void_p(d_ts), n, m, C.c_void_p(d_prof), C.c_void_p(d_idx),
                                C.byref(opts), C.byref(tm))
        check(rc, fn)
        return tm.as_dict()

    def sync(self) -> None:
        check(lib().mp_ctx_sync(self._ctx), "mp_ctx_sync")

    def close(self) -> None:
        if self._ctx:
            for p in self._bufs:
                lib().mp_dev_free(self._ctx, C.c_void_p(p))
            self._bufs = []
            lib().mp_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many FP32-reseeded self-joins in one persistent launch (optionally a replayed hipGraph).

    The series are concatenated back to back in device memory; outputs are
    concatenated over the profile lengths n_s - m_s + 1.
    """

    def __init__(self, dev: Device, lengths, windows, *, exclusion: int = -1, band_k: int = 0,
                 tile_len: int = 0, refine: bool = True, index_init: int = -1, accum64: bool = True):
        self.lengths = np.ascontiguousarray(lengths, dtype=np.int64)
        self.windows = np.ascontiguousarray(windows, dtype=np.int32)
        self._h = C.c_void_p()
        o = make_opts(exclusion, band_k, tile_len, refine, index_init, accum64)
        check(lib().mp_batch_create(dev._ctx, len(self.lengths), _lptr(self.lengths),
                                    self.windows.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(o),
                                    C.byref(self._h)), "mp_batch_create")
        self.profile_lengths = self.lengths - self.windows + 1
        self.offsets = np.concatenate([[0], np.cumsum(self.lengths)])
        self.out_offsets = np.concatenate([[0], np.cumsum(self.profile_lengths)])

    def info(self) -> dict:
        tn, tp, nt = C.c_int64(), C.c_int64(), C.c_int64()
        cl = C.c_double()
        bk, tl = C.c_int32(), C.c_int32()
        check(lib().mp_batch_info(self._h, C.byref(tn), C.byref(tp), C.byref(nt), C.byref(cl), C.byref(bk),
                                  C.byref(tl)), "mp_batch_info")
        return {"total_n": tn.value, "total_p": tp.value, "tiles": nt.value, "cells": cl.value,
                "band_k": bk.value, "tile_len": tl.value}

    def run(self, d_ts: int, d_prof: int, d_idx: int, use_graph: bool = True) -> dict:
        tm = Timing()
        check(lib().mp_batch_run(self._h, C.c_void_p(d_ts), C.c_void_p(d_prof), C.c_void_p(d_idx),
                                 1 if use_graph else 0, C.byref(tm)), "mp_batch_run")
        return tm.as_dict()

    def tune(self, d_ts: int, d_prof: int, d_idx: int, max_tile_len: int = 0) -> dict:
        tm = Timing()
        check(lib().mp_batch_tune(self._h, C.c_void_p(d_ts), C.c_void_p(d_prof), C.c_void_p(d_idx),
                                  max_tile_len, C.byref(tm)), "mp_batch_tune")
        return tm.as_dict()

    def close(self) -> None:
        if self._h:
            lib().mp_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching request server on one GPU (FP32-reseeded self-joins)."""

    def __init__(self, dev: Device, hbm_budget_bytes: int = 8 << 30, tick_cells: int = 0, *,
                 exclusion: int = -1, band_k: int = 0, tile_len: int = 0, refine: bool = True,
                 index_init: int = -1, accum64: bool = True):
        self._h = C.c_void_p()
        o = make_opts(exclusion, band_k, tile_len, refine, index_init, accum64)
        check(lib().mp_server_create(dev._ctx, hbm_budget_bytes, tick_cells, C.byref(o), C.byref(self._h)),
              "mp_server_create")
        self._keep: dict[int, tuple] = {}

    def submit(self, ts, m: int) -> int:
        ts = np.ascontiguousarray(ts, dtype=np.float64)
        P = ts.size - m + 1
        prof = np.empty(max(P, 0))
        idx = np.empty(max(P, 0), dtype=np.int64)
        rid = C.c_uint64()
        check(lib().mp_server_submit(self._h, _dptr(ts), ts.size, m, _dptr(prof), _lptr(idx), C.byref(rid)),
              "mp_server_submit")
        self._keep[rid.value] = (prof, idx)
        return rid.value

    def poll(self, rid: int):
        """None while running, else (profile, index, latency_ms, queue_ms)."""
        done = C.c_int32()
        lat = C.c_double()
        q = C.c_double()
        rc = lib().mp_server_poll(self._h, rid, C.byref(done), C.byref(lat), C.byref(q))
        if rc == 0 and not done.value:
            return None
        prof, idx = self._keep.pop(rid)
        check(rc, "mp_server_poll")
        return prof, idx, lat.value, q.value

    def wait(self, rid: int):
        lat = C.c_double()
        rc = lib().mp_server_wait(self._h, rid, C.byref(lat))
        prof, idx = self._keep.pop(rid)
        check(rc, "mp_server_wait")
        return prof, idx, lat.value

    def completed(self, max_items: int = 4096):
        """Drain finished requests: list of (id, status, latency_ms, queue_ms, profile, index)."""
        ids = (C.c_uint64 * max_items)()
        st = (C.c_int32 * max_items)()
        lat = (C.c_double * max_items)()
        q = (C.c_double * max_items)()
        cnt = C.c_int32()
        check(lib().mp_server_completed(self._h, max_items, ids, st, lat, q, C.byref(cnt)), "mp_server_completed")
        out = []
        for k in range(cnt.value):
            prof, idx = self._keep.pop(ids[k])
            out.append((ids[k], st[k], lat[k], q[k], prof, idx))
        return out

    def stats(self) -> dict:
        t, c, q, a, h = (C.c_int64() for _ in range(5))
        cd = C.c_double()
        check(lib().mp_server_stats(self._h, C.byref(t), C.byref(c), C.byref(cd), C.byref(q), C.byref(a),
                                    C.byref(h)), "mp_server_stats")
        return {"ticks": t.value, "completed": c.value, "cells_done": cd.value, "queued": q.value,
                "active": a.value, "hbm_used_bytes": h.value}

    def close(self) -> None:
        if self._h:
            lib().mp_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---- synthetic inputs (CPU, bit-identical on every host) --------------------
def random_walk(seed: int, n: int, stream: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().mp_gen_random_walk(seed, stream, n, _dptr(out))
    return out


def normals(seed: int, count: int, stream: int = 0, first: int = 0) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().mp_gen_normals(seed, stream, first, count, _dptr(out))
    return out


def uniforms(seed: int, count: int, stream: int = 0, first: int = 0) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().mp_gen_uniforms(seed, stream, first, count, _dptr(out))
    return out


def precision_series(seed: int, n: int, m: int):
    """Random walk + 3 sinusoids + 20 Gaussian bumps; returns (series, meta)."""
    out = np.empty(n, dtype=np.float64)
    meta = np.empty(26, dtype=np.float64)
    check(lib().mp_gen_precision_series(seed, n, m, _dptr(out), _dptr(meta)), "mp_gen_precision_series")
    return out, {"periods": meta[0:3].copy(), "phases": meta[3:6].copy(), "bump_starts": meta[6:26].astype(np.int64)}
