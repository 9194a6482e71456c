"""The seeded workloads the library is measured and tested on (CPU generators).

Every series comes from the library's counter-based Philox4x32-10 generator,
so it is bit-identical on every host; SHA-256 hashes of the single-series
workloads are committed with the tests.

  cfg1  random walk, seed 0, n = 65,536, m = 400 (the CPU-runnable case)
  cfg2  random walk, seed 1, n = 2^20, m = 256
  cfg3  walk + 3 sinusoids + 20 Gaussian bumps, seed 2, n = 2^22, m = 1024
  cfg4  4,096 FP32 random walks, lengths log-uniform in [2^10, 2^17] (seed 3),
        windows from {32, 64, 128, 256}
  cfg5  Poisson arrivals (2,000 requests/s), lengths lognormal with median
        2^14 and sigma 1 clipped to [2^10, 2^19], windows from {64, ..., 512}
"""
from __future__ import annotations

import numpy as np

from . import normals, precision_series, random_walk, uniforms

# name: (generator, seed, n, m)
SERIES = {
    "cfg1": ("walk", 0, 65536, 400),
    "cfg2": ("walk", 1, 1 << 20, 256),
    "cfg3": ("precision", 2, 1 << 22, 1024),
}


def make_series(kind: str, seed: int, n: int, m: int) -> np.ndarray:
    if kind == "walk":
        return random_walk(seed, n)
    return precision_series(seed, n, m)[0]


def series(name: str) -> tuple[np.ndarray, int]:
    """(series, window) of one of the single-series workloads cfg1-cfg3."""
    kind, seed, n, m = SERIES[name]
    return make_series(kind, seed, n, m), m


def batch_spec(seed: int = 3, count: int = 4096):
    """cfg4: lengths log-uniform in [2^10, 2^17], windows from {32, 64, 128, 256}."""
    u = uniforms(seed, count, stream=0)
    v = uniforms(seed, count, stream=1)
    lengths = np.floor(2.0 ** (10.0 + 7.0 * u)).astype(np.int64)
    windows = np.array([32, 64, 128, 256], dtype=np.int32)[np.minimum((4 * v).astype(int), 3)]
    return lengths, windows


def serve_spec(seed: int, seconds: float, rate: float):
    """cfg5: Poisson arrival times over `seconds` at `rate` per second; lognormal
    lengths (median 2^14, sigma 1, clipped to [2^10, 2^19]); m uniform in {64, 128, 256, 512}."""
    count = int(rate * seconds * 1.2) + 100
    u = uniforms(seed, count, stream=0)
    arrivals = np.cumsum(-np.log1p(-u) / rate)
    count = int(np.searchsorted(arrivals, seconds))
    arrivals = arrivals[:count]
    z = normals(seed, count, stream=1)
    lengths = np.clip(np.floor(np.exp(np.log(2.0 ** 14) + z)), 2 ** 10, 2 ** 19).astype(np.int64)
    windows = np.array([64, 128, 256, 512])[np.minimum((4 * uniforms(seed, count, stream=2)).astype(int), 3)]
    return arrivals, lengths, windows


def serve_requests(seed: int, seconds: float, rate: float, walk_len: int = 1 << 25):
    """cfg5 requests as slices of one long random walk at seeded offsets:
    (arrivals, lengths, windows, walk, offsets)."""
    arrivals, lengths, windows = serve_spec(seed, seconds, rate)
    walk = random_walk(seed, walk_len, stream=3)
    offs = np.floor(uniforms(seed, len(lengths), stream=4) * (walk.size - lengths)).astype(np.int64)
    return arrivals, lengths, windows, walk, offs
