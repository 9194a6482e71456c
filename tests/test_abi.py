"""The C ABI library: loads, exports every declared symbol, validates arguments
before touching a device, and its generators reproduce the committed hashes
and Philox4x32-10 known answers.  No compute call reaches a GPU here."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import re
import subprocess

import numpy as np
import pytest

import thesiscode_amd as tc
from conftest import GOLDEN, ROOT
from thesiscode_amd._ffi import LIB_PATH, SIGNATURES, Opts, Timing, lib

HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(mp_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (mp_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    assert set(SIGNATURES) <= exported


def test_every_declared_symbol_is_bound():
    lib()
    assert set(declared_functions()) <= set(SIGNATURES)


def test_status_strings():
    for code in range(-10, 1):
        assert lib().mp_status_string(code)


def _call(fn, ts, n, m, opts=None):
    P = max(1, n - m + 1)
    prof = np.empty(P)
    idx = np.empty(P, dtype=np.int64)
    o = opts if opts is not None else tc.make_opts()
    return fn(ts.ctypes.data_as(C.POINTER(C.c_double)) if ts is not None else None, n, m,
              prof.ctypes.data_as(C.POINTER(C.c_double)), idx.ctypes.data_as(C.POINTER(C.c_int64)),
              C.byref(o), C.byref(Timing()))


@pytest.mark.parametrize("name", ["mp_selfjoin_f64", "mp_selfjoin_f32r"])
def test_argument_errors_are_status_codes(name):
    fn = getattr(lib(), name)
    ts = np.zeros(100)
    assert _call(fn, ts, 100, 0) == -1           # MP_ERR_BAD_WINDOW
    assert _call(fn, ts, 10, 11) == -2           # MP_ERR_SHORT_SERIES (reference throws)
    assert _call(fn, None, 100, 10) == -3        # MP_ERR_NULL
    o = Opts()
    lib().mp_opts_default(C.byref(o))
    o.exclusion = -5
    assert _call(fn, ts, 100, 10, o) == -1


def test_opts_defaults():
    o = Opts()
    lib().mp_opts_default(C.byref(o))
    assert (o.exclusion, o.refine, o.index_init, o.band_k, o.tile_len, o.accum64) == (-1, 1, -1, 0, 0, 1)


def test_philox_known_answers():
    kat = json.loads((GOLDEN / "series_sha256.json").read_text())["philox4x32_10_kat"]
    for ctr, key, want in kat:
        c = (C.c_uint32 * 4)(*ctr)
        o = (C.c_uint32 * 4)()
        lib().mp_gen_philox4x32_10(c, key[0] | (key[1] << 32), o)
        assert ["%08x" % v for v in o] == want


def test_generated_workloads_match_committed_hashes():
    ref = json.loads((GOLDEN / "series_sha256.json").read_text())

    def sha(a):
        return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()

    assert sha(tc.random_walk(0, 65536)) == ref["cfg1_walk_seed0_n65536"]
    assert sha(tc.random_walk(1, 1 << 20)) == ref["cfg2_walk_seed1_n1048576"]
    assert sha(tc.precision_series(2, 1 << 22, 1024)[0]) == ref["cfg3_precision_seed2_n4194304_m1024"]
    assert tc.normals(9, 16).tolist() == ref["normals_seed9_first16"]
    assert tc.uniforms(9, 8, stream=1).tolist() == ref["uniforms_seed9_stream1_first8"]


def test_random_walk_is_a_walk_of_standard_normals():
    x = tc.random_walk(3, 200001)
    z = np.diff(x)
    assert x[0] == 0.0
    assert np.array_equal(z[:10], np.diff(x[:11]))
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    # steps are the stream's normals in order
    np.testing.assert_allclose(z[:50], tc.normals(3, 50), rtol=0, atol=1e-12)


def test_precision_series_structure():
    n, m = 1 << 16, 256
    x, meta = tc.precision_series(7, n, m)
    assert np.all((meta["periods"] >= 200) & (meta["periods"] <= 5000))
    assert len(meta["bump_starts"]) == 20 and np.all(meta["bump_starts"] <= n - m)
    base = tc.random_walk(7, n)
    t = np.arange(n)
    sines = sum(5.0 * np.sin(2 * np.pi * t / p + ph) for p, ph in zip(meta["periods"], meta["phases"]))
    resid = x - base - sines
    s0 = meta["bump_starts"][0]
    assert resid[s0 + m // 2] > 0  # a bump sits there
    mask = np.ones(n, bool)
    for s in meta["bump_starts"]:
        mask[s:s + m] = False
    assert np.max(np.abs(resid[mask])) < 1e-8
