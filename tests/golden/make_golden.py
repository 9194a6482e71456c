"""Writes the committed golden fixtures of the SCRIMP tests.

Run from the repository root after building (python -m thesiscode_amd.build --oracle):
    python tests/golden/make_golden.py

Fixtures (all small):
  test_series.bin        the reference's own input fixture (scrimppp/data_testing/test.bin,
                         format v2: 100-byte header {u32 2, u32 100, u64 n} + n doubles),
                         copied as data when /root/reference is present
  golden_small.npz       long-double brute-force profiles of small seeded series (pins the
                         oracle independently of its own arithmetic) and the oracle's
                         profile/index of test_series.bin at m=400 (README.md:59's example)
  series_sha256.json     SHA-256 of every generated workload series and Philox known answers
"""
from __future__ import annotations

import hashlib
import json
import shutil
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import thesiscode_amd as tc  # noqa: E402
from oracle import oracle as orc  # noqa: E402

HERE = Path(__file__).resolve().parent
REF_TEST_BIN = Path("/root/reference/scrimppp/data_testing/test.bin")


def read_ts_bin(path: Path) -> np.ndarray:
    raw = path.read_bytes()
    fmt, off, n = np.frombuffer(raw[:16], dtype="<u4", count=2).tolist() + [int(np.frombuffer(raw[8:16], "<u8")[0])]
    assert fmt == 2 and off == 100, (fmt, off)
    return np.frombuffer(raw[off:off + 8 * n], dtype="<f8").copy()


def small_cases():
    """(name, series, m) cases small enough for the O(P^2 m) long-double brute force."""
    return [
        ("walk_n1024_m32", tc.random_walk(101, 1024), 32),
        ("walk_n700_m7", tc.random_walk(102, 700), 7),
        ("iwalk_n800_m24", np.cumsum(tc.random_walk(103, 800)), 24),
        ("walk_n512_m100", tc.random_walk(104, 512), 100),
    ]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main() -> None:
    if REF_TEST_BIN.exists():
        shutil.copyfile(REF_TEST_BIN, HERE / "test_series.bin")
    ts = read_ts_bin(HERE / "test_series.bin")
    out = {}
    for name, A, m in small_cases():
        p, i = orc.brute_force(A, m)
        out[f"{name}__series"] = A
        out[f"{name}__bf_profile"] = p
        out[f"{name}__bf_index"] = i
        out[f"{name}__m"] = np.array(m)
    p, i, _, _ = orc.triv_par(ts, 400, nranks=1, nthreads=1, kernel=1)
    out["testbin_m400__oracle_profile"] = p
    out["testbin_m400__oracle_index"] = i
    rows = np.linspace(0, p.size - 1, 48).astype(np.int64)
    bfr = np.array([min((orc.pair_dist2(ts, 400, r, j), j) for j in range(p.size) if abs(j - r) > 100)
                    for r in rows])
    out["testbin_m400__bf_rows"] = rows
    out["testbin_m400__bf_row_profile"] = bfr[:, 0]
    out["testbin_m400__bf_row_index"] = bfr[:, 1].astype(np.int64)
    np.savez_compressed(HERE / "golden_small.npz", **out)

    hashes = {
        "cfg1_walk_seed0_n65536": sha(tc.random_walk(0, 65536)),
        "cfg2_walk_seed1_n1048576": sha(tc.random_walk(1, 1 << 20)),
        "cfg3_precision_seed2_n4194304_m1024": sha(tc.precision_series(2, 1 << 22, 1024)[0]),
        "normals_seed9_first16": tc.normals(9, 16).tolist(),
        "uniforms_seed9_stream1_first8": tc.uniforms(9, 8, stream=1).tolist(),
        "philox4x32_10_kat": [
            # Random123 known-answer vectors for philox4x32-10 (ctr, key lo/hi, out)
            [[0, 0, 0, 0], [0, 0], ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]],
            [[0xFFFFFFFF] * 4, [0xFFFFFFFF, 0xFFFFFFFF], ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]],
            [[0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
             ["d16cfe09", "94fdcceb", "5001e420", "24126ea1"]],
        ],
    }
    (HERE / "series_sha256.json").write_text(json.dumps(hashes, indent=1) + "\n")
    print("wrote", sorted(p.name for p in HERE.iterdir()))


if __name__ == "__main__":
    main()
