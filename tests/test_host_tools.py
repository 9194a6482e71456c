"""Reference file formats, converters and CLI argument handling (CPU only).

Formats (scrimppp/inc/biniobase.h:13-21, binproffile.h:11-16, ScrimpSequ.cpp:265-277):
binary = 100-byte header {u32 version, u32 offset=100, u64 count} + data;
series version 2 (doubles), profile version 1 ({double d2, uint64 index}).
"""
from __future__ import annotations

import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, read_ts_bin

TOOLS = ROOT / "thesiscode_amd"


def run(tool, *args, cwd=None):
    return subprocess.run([str(TOOLS / tool), *map(str, args)], capture_output=True, text=True, cwd=cwd)


def test_reference_fixture_header():
    raw = (GOLDEN / "test_series.bin").read_bytes()
    assert struct.unpack("<IIQ", raw[:16]) == (2, 100, 10000)
    assert len(raw) == 100 + 8 * 10000


def test_ts_ascii_to_bin_round_trip(tmp_path, test_series):
    asc = tmp_path / "s.ascii"
    asc.write_text("\n".join(repr(float(v)) for v in test_series[:500]) + "\n")
    r = run("cvt_ts_hip", "ascii-to-bin", asc, tmp_path / "s.bin")
    assert r.returncode == 0, r.stderr
    back = read_ts_bin(tmp_path / "s.bin")
    assert np.array_equal(back, test_series[:500])
    r = run("cvt_ts_hip", "bin-to-ascii", tmp_path / "s.bin", tmp_path / "t.ascii")
    assert r.returncode == 0
    assert np.array_equal(np.loadtxt(tmp_path / "t.ascii"), test_series[:500])


def test_ts_conversion_default_output_name(tmp_path, test_series):
    asc = tmp_path / "x.ascii"
    asc.write_text(" ".join(repr(float(v)) for v in test_series[:50]))
    r = run("cvt_ts_hip", "ascii-to-bin", asc, cwd=tmp_path)
    assert r.returncode == 0
    assert np.array_equal(read_ts_bin(tmp_path / "converted.bin"), test_series[:50])


def _write_profile_bin(path, prof, idx):
    with open(path, "wb") as f:
        f.write(struct.pack("<IIQ", 1, 100, prof.size) + b"\0" * 84)
        rec = np.empty(prof.size, dtype=[("d", "<f8"), ("i", "<u8")])
        rec["d"] = prof
        rec["i"] = idx.astype(np.int64).view(np.uint64)
        f.write(rec.tobytes())


def test_profile_bin_to_ascii_format(tmp_path):
    prof = np.array([4.0, 0.25, 2.0, 123.456789])
    idx = np.array([7, -1, 3, 99])
    _write_profile_bin(tmp_path / "p.bin", prof, idx)
    r = run("cvt_matprof_hip", "bin-to-ascii", tmp_path / "p.bin", tmp_path / "p.ascii")
    assert r.returncode == 0, r.stderr
    lines = (tmp_path / "p.ascii").read_text().split("\n")[:4]
    # sqrt(|d2|) with 17 significant digits, then the index as stored (unsigned on disk)
    assert lines[0] == "2 7" and lines[1] == "0.5 -1"
    assert float(lines[2].split()[0]) == np.sqrt(2.0)
    assert lines[2].split()[0] == repr(np.sqrt(2.0))[:19] or abs(float(lines[2].split()[0]) - np.sqrt(2)) == 0


def test_wrong_format_version_is_rejected(tmp_path):
    _write_profile_bin(tmp_path / "p.bin", np.ones(3), np.zeros(3))
    r = run("cvt_ts_hip", "bin-to-ascii", tmp_path / "p.bin", tmp_path / "o.ascii")
    assert r.returncode != 0 and "format" in r.stderr


def test_cli_errors(tmp_path):
    r = run("scrimppp_hip", "no_such_algo", 100, 25, GOLDEN / "test_series.bin")
    assert r.returncode == 1 and "scrimp_hip_f64" in r.stderr
    r = run("scrimppp_hip", "scrimp_hip_f64", 100, 25, tmp_path / "series.txt")
    assert r.returncode == 1 and "suffix" in r.stderr
    r = run("scrimppp_hip", "--help")
    assert r.returncode == 0 and "--window-length" in r.stdout
    r = run("scrimppp_hip", "scrimp_hip_f64", 100)
    assert r.returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", ["bin", "ascii"])
def test_cli_self_join_writes_reference_format(tmp_path, golden, test_series, fmt):
    src = tmp_path / f"series.{fmt}"
    if fmt == "bin":
        src.write_bytes((GOLDEN / "test_series.bin").read_bytes())
    else:
        src.write_text("\n".join(repr(float(v)) for v in test_series) + "\n")
    r = run("scrimppp_hip", "scrimp_hip_f64", 400, 100, src, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    out = tmp_path / f"scrimp_hip_f64_new_MatrixProfile_400series.{fmt}.matprof.{fmt}"
    assert out.exists(), sorted(p.name for p in tmp_path.iterdir())
    p_ref = golden["testbin_m400__oracle_profile"]
    if fmt == "bin":
        raw = out.read_bytes()
        assert struct.unpack("<IIQ", raw[:16]) == (1, 100, p_ref.size)
        rec = np.frombuffer(raw[100:], dtype=[("d", "<f8"), ("i", "<u8")])
        assert np.array_equal(rec["d"], p_ref)
    else:
        vals = np.loadtxt(out)
        np.testing.assert_allclose(vals[:, 0], np.sqrt(np.abs(p_ref)), rtol=1e-15)


@pytest.mark.gpu
def test_cli_orig_writes_prescrimp_and_final_profiles(tmp_path):
    # scrimppp_orig's drop-in: ASCII in, ASCII out (<out>_prescrimp, <out>);
    # step = floor(prescrimp_stride * m) with the integer stride parameter
    import thesiscode_amd as tc
    from oracle import oracle as orc

    A = tc.random_walk(17, 1500)
    m, stride = 40, 1
    src = tmp_path / "series.ascii"  # the command line resolves the type from the suffix
    src.write_text("\n".join(repr(float(v)) for v in A) + "\n")
    r = run("scrimppp_hip", "scrimp_hip_orig", m, stride, src, "--output-file", tmp_path / "mp", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    pre = np.loadtxt(tmp_path / "mp_prescrimp")
    fin = np.loadtxt(tmp_path / "mp")
    q, j = orc.prescrimp(A, m, float(stride))
    np.testing.assert_allclose(pre[:, 0] ** 2, q, rtol=1e-9, atol=1e-9 * 2 * m)
    assert np.mean(pre[:, 1] == j) >= 0.999
    # the final profile: the minimum of PreSCRIMP's and the full join's
    p64, i64, _ = tc.selfjoin_f64(A, m)
    want = np.minimum(q, p64)
    np.testing.assert_allclose(fin[:, 0] ** 2, want, rtol=1e-9, atol=1e-9 * 2 * m)
    assert np.all(fin[:, 0] ** 2 <= pre[:, 0] ** 2 * (1 + 1e-12) + 1e-12)
