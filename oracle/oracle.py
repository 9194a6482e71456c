"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_D = C.POINTER(C.c_double)
_L = C.POINTER(C.c_long)
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
        h = C.CDLL(str(LIB_PATH))
        h.orc_window_stats.argtypes = [_D, C.c_long, C.c_int, _D, _D]
        h.orc_seeds.argtypes = [_D, C.c_long, C.c_int, C.c_long, C.c_long, _D]
        kern = [_D, _L, _D, _L, _D, C.c_long, C.c_long, _D, _D, C.c_long, _D, _D, _D, _D, C.c_long, C.c_long]
        h.orc_eval_block_scalar.argtypes = kern
        h.orc_eval_block_abba.argtypes = kern
        h.orc_get_partition.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        h.orc_triv_par.argtypes = [_D, C.c_long, C.c_int, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, _D, _L, _D, _D]
        h.orc_triv_par.restype = C.c_int
        h.orc_vert_blocked.argtypes = [_D, C.c_long, C.c_int, C.c_long, C.c_int, C.c_int, _D, _L]
        h.orc_vert_blocked.restype = C.c_int
        h.orc_brute_force.argtypes = [_D, C.c_long, C.c_int, C.c_long, _D, _L]
        h.orc_pair_dist2.argtypes = [_D, C.c_long, C.c_int, C.c_long, C.c_long]
        h.orc_pair_dist2.restype = C.c_double
        h.orc_window_moments_ld.argtypes = [_D, C.c_long, C.c_int, _D, _D]
        h.orc_set_fault.argtypes = [C.c_int]
        h.orc_prescrimp.argtypes = [_D, C.c_long, C.c_int, C.c_double, _L, C.c_long, _D, _L]
        h.orc_partition_coords.argtypes = [C.c_long, _L, _L, C.POINTER(C.c_int)]
        h.orc_distrib_geometry.argtypes = [C.c_long, C.c_int, C.c_int, _L, _L]
        h.orc_distrib_geometry.restype = C.c_int
        h.orc_distrib_tile_inputs.argtypes = [_D, C.c_long, C.c_int, C.c_int, C.c_long, _D, _D, _D, _D, _D, _D,
                                              _D, _L, _L, C.POINTER(C.c_int)]
        h.orc_distrib_tile_inputs.restype = C.c_int
        h.orc_distrib_par.argtypes = [_D, C.c_long, C.c_int, C.c_int, C.c_int, C.c_int, _D, _L]
        h.orc_distrib_par.restype = C.c_int
        _lib = h
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _l(a):
    return a.ctypes.data_as(_L)


def window_stats(A: np.ndarray, m: int):
    A = np.ascontiguousarray(A, dtype=np.float64)
    P = A.size - m + 1
    s = np.empty(P)
    mu = np.empty(P)
    lib().orc_window_stats(_d(A), A.size, m, _d(s), _d(mu))
    return s, mu


def seeds(A: np.ndarray, m: int, first: int, last: int) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.float64)
    qt = np.zeros(A.size + 1)
    lib().orc_seeds(_d(A), A.size, m, first, last, _d(qt))
    return qt


def triv_par(A, m, *, excl=-1, blocklen=500, nranks=1, nthreads=1, kernel=1, block_stride=1):
    """scrimp_triv_par restatement; returns (profile, index, cells, seconds)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    P = A.size - m + 1
    prof = np.empty(max(P, 0))
    idx = np.empty(max(P, 0), dtype=np.int64)
    cells = C.c_double()
    secs = C.c_double()
    rc = lib().orc_triv_par(_d(A), A.size, m, excl, blocklen, nranks, nthreads, kernel, block_stride,
                            _d(prof), _l(idx), C.byref(cells), C.byref(secs))
    if rc != 0:
        raise ValueError("profile length <= 0")
    return prof, idx, cells.value, secs.value


def vert_blocked(A, m, *, excl=-1, blocklen=500, kernel=1):
    A = np.ascontiguousarray(A, dtype=np.float64)
    P = A.size - m + 1
    prof = np.empty(P)
    idx = np.empty(P, dtype=np.int64)
    if lib().orc_vert_blocked(_d(A), A.size, m, excl, blocklen, kernel, _d(prof), _l(idx)) != 0:
        raise ValueError("profile length <= 0")
    return prof, idx


def eval_block(kernel: str, *args):
    fn = {"scalar": lib().orc_eval_block_scalar, "abba": lib().orc_eval_block_abba}[kernel]
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(_d(a) if a.dtype == np.float64 else _l(a))
        else:
            conv.append(int(a))
    fn(*conv)


def brute_force(A, m, excl=-1):
    A = np.ascontiguousarray(A, dtype=np.float64)
    P = A.size - m + 1
    prof = np.empty(P)
    idx = np.empty(P, dtype=np.int64)
    lib().orc_brute_force(_d(A), A.size, m, excl, _d(prof), _l(idx))
    return prof, idx


def prescrimp(A, m, stride=0.25, order=None):
    """PreSCRIMP phase of scrimppp_orig: (d2 profile, index); order = sampled rows in
    processing order (ascending multiples of floor(stride*m) by default)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    P = A.size - m + 1
    prof = np.empty(P)
    idx = np.empty(P, dtype=np.int64)
    if order is None:
        lib().orc_prescrimp(_d(A), A.size, m, float(stride), None, 0, _d(prof), _l(idx))
    else:
        o = np.ascontiguousarray(order, dtype=np.int64)
        lib().orc_prescrimp(_d(A), A.size, m, float(stride), _l(o), o.size, _d(prof), _l(idx))
    return prof, idx


def pair_dist2(A, m, i, j) -> float:
    A = np.ascontiguousarray(A, dtype=np.float64)
    return lib().orc_pair_dist2(_d(A), A.size, m, int(i), int(j))


def window_moments_ld(A, m):
    A = np.ascontiguousarray(A, dtype=np.float64)
    P = A.size - m + 1
    mu = np.empty(P)
    sg = np.empty(P)
    lib().orc_window_moments_ld(_d(A), A.size, m, _d(mu), _d(sg))
    return mu, sg


def get_partition(part_id, nparts, ids):
    f = C.c_int()
    l_ = C.c_int()
    lib().orc_get_partition(part_id, nparts, ids, C.byref(f), C.byref(l_))
    return f.value, l_.value


class fault:
    """Context manager planting a mistake in the oracle (for pinning tests)."""

    def __init__(self, code: int):
        self.code = code

    def __enter__(self):
        lib().orc_set_fault(self.code)

    def __exit__(self, *exc):
        lib().orc_set_fault(0)


def partition_coords(rank: int):
    """(row, col, is_upper) of a rank in the checkerboard (checkerboard_partitioning.cpp:8-41)."""
    r, c = C.c_long(), C.c_long()
    up = C.c_int()
    lib().orc_partition_coords(rank, C.byref(r), C.byref(c), C.byref(up))
    return r.value, c.value, bool(up.value)


def distrib_geometry(n: int, m: int, k: int):
    """(padding, tile length L) of distrib_par with k x k tiles; ValueError when it throws."""
    pad, L = C.c_long(), C.c_long()
    if lib().orc_distrib_geometry(n, m, k, C.byref(pad), C.byref(L)) != 0:
        raise ValueError("parameters chosen badly")
    return pad.value, L.value


def distrib_tile_inputs(A, m: int, k: int, rank: int) -> dict:
    """One tile's kernel inputs as distrib_par builds them: slices, their statistics, seeds."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    _, L = distrib_geometry(A.size, m, k)
    out = {name: np.empty(L + m) for name in ("A", "B")}
    for name in ("sA", "muA", "sB", "muB", "dots"):
        out[name] = np.empty(L)
    ro, co = C.c_long(), C.c_long()
    up = C.c_int()
    rc = lib().orc_distrib_tile_inputs(_d(A), A.size, m, k, rank, _d(out["A"]), _d(out["B"]), _d(out["sA"]),
                                       _d(out["muA"]), _d(out["sB"]), _d(out["muB"]), _d(out["dots"]),
                                       C.byref(ro), C.byref(co), C.byref(up))
    if rc != 0:
        raise ValueError("bad tile")
    out.update(L=L, row_offset=ro.value, col_offset=co.value, upper=bool(up.value))
    return out


def distrib_par(A, m: int, k: int, *, blocklen: int = 500, kernel: int = 1):
    """distrib_par with k x k tiles on one process; returns (profile, index) of the n - m + 1 entries."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    P = A.size - m + 1
    prof = np.empty(max(P, 0))
    idx = np.empty(max(P, 0), dtype=np.int64)
    if lib().orc_distrib_par(_d(A), A.size, m, k, blocklen, kernel, _d(prof), _l(idx)) != 0:
        raise ValueError("parameters chosen badly")
    return prof, idx
