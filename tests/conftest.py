"""Shared fixtures and helpers for the SCRIMP tests.

`-m "not gpu"` runs on any host: the oracle's pinning tests, the C ABI's
exports and the generators.  `-m gpu` are the parity tests proper; they call
libscrimp_hip.so through its C ABI and compare with the CPU oracle.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); the parity tests proper")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN / "golden_small.npz")


def read_ts_bin(path: Path) -> np.ndarray:
    """Reference time-series file format v2 (scrimppp/inc/biniobase.h:10-21): 100-byte header."""
    raw = path.read_bytes()
    fmt, off = np.frombuffer(raw[:8], dtype="<u4").tolist()
    n = int(np.frombuffer(raw[8:16], dtype="<u8")[0])
    assert fmt == 2 and off == 100
    return np.frombuffer(raw[off:off + 8 * n], dtype="<f8").copy()


@pytest.fixture(scope="session")
def test_series() -> np.ndarray:
    return read_ts_bin(GOLDEN / "test_series.bin")


def cpu_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def golden_cases(golden):
    names = sorted({k.split("__")[0] for k in golden.files if k.endswith("__series")})
    return [(n, golden[f"{n}__series"], int(golden[f"{n}__m"])) for n in names]


def assert_indices_match(A, m, prof_ref, idx_ref, idx_got, *, abs_tol, min_frac=0.999, rel_tol=0.0,
                         sample=200):
    """Indices must agree except where both partners are a near tie.

    A mismatch is accepted when the true (long double) d^2 of the two partners
    differ by at most abs_tol + rel_tol * max(d^2, 2m * 1e-3) (the floor judges
    exact repeats on the window's distance scale); at most 1-min_frac of the
    entries may mismatch, and up to `sample` of them (spread over the profile)
    are checked.
    """
    from oracle import oracle as orc

    bad = np.nonzero(idx_ref != idx_got)[0]
    assert bad.size <= max(1, int((1.0 - min_frac) * idx_ref.size)), f"{bad.size} index mismatches"
    if bad.size > sample:
        bad = bad[np.linspace(0, bad.size - 1, sample).astype(np.int64)]
    for i in bad:
        assert idx_got[i] >= 0, (i, idx_ref[i], idx_got[i])
        d_ref = orc.pair_dist2(A, m, i, idx_ref[i])
        d_got = orc.pair_dist2(A, m, i, idx_got[i])
        tol = abs_tol + rel_tol * max(d_ref, 2 * m * 1e-3)
        assert abs(d_ref - d_got) <= tol, (i, idx_ref[i], idx_got[i], d_ref, d_got, prof_ref[i])


def assert_values_equal_reference_walk(A, m, prof, idx, entries, *, excl=-1):
    """Exact-path values at `entries` equal the reference's arithmetic bit for bit.

    For each entry the oracle walks the 15 diagonals around the entry's partner
    from their first cells (the reference's seeds and recurrence,
    ScrimpSequOpt.cpp:81-96, kernels.cpp:310-366) through the separate-view
    adapter (row results by row, column results by column).  Those cells, on
    the partner's side of the entry, include the winning lane's candidates (at
    most 8 adjacent diagonals), so their maximum score is the profile value:
    prof[e] == 2m - 2 * max.
    """
    from oracle import oracle as orc

    n = A.size
    P = n - m + 1
    base = (m // 4 if excl < 0 else excl) + 1
    Ap = np.append(A, 0.0)  # the last update of a diagonal reads one past the series
    s, mu = orc.window_stats(A, m)
    for e in entries:
        e = int(e)
        part = int(idx[e])
        assert part >= 0, (e, part)
        d = abs(e - part)
        d0, d1 = max(base, d - 7), min(P - 1, d + 7) + 1  # diagonals [d0, d1)
        tl = P - d0
        qt = orc.seeds(A, m, d0, d1)
        pc = np.full(tl, -np.inf)
        ic = np.full(tl, -1, dtype=np.int64)
        pr = np.full(tl, -np.inf)
        ir = np.full(tl, -1, dtype=np.int64)
        q = qt[d0:d1].copy()
        orc.eval_block("scalar", pc, ic, pr, ir, q, d1 - d0, tl, Ap, Ap[d0:], m, s, mu, s[d0:], mu[d0:], d0, 0)
        v = pc[e] if part > e else pr[e - d0]  # the entry as the column, or as the row, of its cells
        assert 2.0 * m - 2.0 * v == prof[e], (e, part, prof[e], 2.0 * m - 2.0 * v)
