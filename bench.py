"""Benchmark of the SCRIMP self-join hot path on one MI355X.

Default workload: the plain FP64 drop-in configuration -- one random walk of
n = 2^20 N(0,1) steps (seed 1), window m = 256, exclusion m/4 -- computed by the
exact FP64 path with the series already resident in HBM.  One step is one full
self-join (statistics, seeds, all diagonals, merge, unpack to profile/index).

Prints ONE JSON line.  The metric is matrix-profile cells/s with
cells = W(W+1)/2, W = n - m + 1 - m/4 - 1 (the reference logs (P-ez)^2/t, about
twice that).  `roofline` is the dominant kernel's arithmetic rate (8 flops per
cell: score 2 mul + 1 FMA, recurrence 2 FMA) over the MI355X FP64 vector peak,
from HIP events on the kernel's own stream; `cpu_baseline` times the oracle's
restatement of scrimp_triv_par (AVX2 ABBA kernel, one thread per rank) on a
bounded sample of the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_VECTOR_PEAK_TFLOPS = 78.6
FP32_VECTOR_PEAK_TFLOPS = 157.3
FP64_MATRIX_PEAK_TFLOPS = 78.6  # dense FP64 MFMA
FLOPS_PER_CELL = 8

WORKLOADS = {
    # name: (generator, n, m, precision)
    "cfg2-f64": ("walk", 1, 1 << 20, 256, "f64"),
    "cfg2-f32r": ("walk", 1, 1 << 20, 256, "f32r"),
    "cfg1-f64": ("walk", 0, 65536, 400, "f64"),
    "cfg3-f32r": ("precision", 2, 1 << 22, 1024, "f32r"),
    "cfg3-f64": ("precision", 2, 1 << 22, 1024, "f64"),
}


def make_series(kind, seed, n, m):
    import thesiscode_amd as tc

    if kind == "walk":
        return tc.random_walk(seed, n)
    return tc.precision_series(seed, n, m)[0]


def run_gpu(workload: str, steps: int, warmup: int, band_k: int, tile_len: int, accum64: bool = True):
    import thesiscode_amd as tc

    kind, seed, n, m, prec = WORKLOADS[workload]
    A = make_series(kind, seed, n, m)
    P = n - m + 1
    with tc.Device(0) as dev:
        d_ts = dev.upload(A)
        d_prof = dev.alloc(8 * P)
        d_idx = dev.alloc(8 * P)
        opts = tc.make_opts(band_k=band_k, tile_len=tile_len, accum64=accum64)
        for _ in range(warmup):
            dev.selfjoin(prec, d_ts, n, m, d_prof, d_idx, opts)
        dev.sync()
        kern_ms = []
        total_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            tm = dev.selfjoin(prec, d_ts, n, m, d_prof, d_idx, opts)
            kern_ms.append(tm["kernel_ms"])
            total_ms.append(tm["total_ms"])
        dev.sync()
        wall = time.perf_counter() - t0
        prof = dev.download(d_prof, np.empty(P))
    return {
        "n": n, "m": m, "precision": prec, "cells": tm["cells"], "wall_s": wall,
        "kernel_ms": float(np.mean(kern_ms)), "device_ms": float(np.mean(total_ms)),
        "band_k": tm["band_k"], "tile_len": tm["tile_len"], "work_items": tm["work_items"],
        "profile_checksum": float(np.sum(prof)),
    }


def batch_spec(seed: int = 3, count: int = 4096):
    """4,096 FP32 random walks, lengths log-uniform in [2^10, 2^17], windows from {32,64,128,256}."""
    import thesiscode_amd as tc

    u = tc.uniforms(seed, count, stream=0)
    v = tc.uniforms(seed, count, stream=1)
    lengths = np.floor(2.0 ** (10.0 + 7.0 * u)).astype(np.int64)
    windows = np.array([32, 64, 128, 256], dtype=np.int32)[np.minimum((4 * v).astype(int), 3)]
    return lengths, windows


def gen_batch_series(seed, lengths):
    import concurrent.futures as cf

    import thesiscode_amd as tc

    def one(s):
        return tc.random_walk(seed, int(lengths[s]), stream=100 + s).astype(np.float32).astype(np.float64)

    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, range(len(lengths))))


def run_batch_bench(steps: int, warmup: int, tune: bool):
    import thesiscode_amd as tc

    lengths, windows = batch_spec()
    series = gen_batch_series(3, lengths)
    with tc.Device(0) as dev:
        with tc.Batch(dev, lengths, windows) as b:
            d_ts = dev.upload(np.concatenate(series))
            P = int(b.out_offsets[-1])
            d_p = dev.alloc(8 * P)
            d_i = dev.alloc(8 * P)
            tuned = b.tune(d_ts, d_p, d_i, max_tile_len=8192) if tune else None
            for _ in range(warmup):
                b.run(d_ts, d_p, d_i, use_graph=True)
            dev.sync()
            t0 = time.perf_counter()
            dev_ms = []
            for _ in range(steps):
                dev_ms.append(b.run(d_ts, d_p, d_i, use_graph=True)["total_ms"])
            wall = time.perf_counter() - t0
            info = b.info()
    return {"cells": info["cells"], "wall_s": wall, "device_ms": float(np.mean(dev_ms)), "band_k": info["band_k"],
            "tile_len": info["tile_len"], "series": len(lengths), "points": int(lengths.sum()),
            "tiles": info["tiles"], "tuned": tuned}


def serve_spec(seed: int, seconds: float, rate: float):
    """Poisson arrivals; lognormal lengths (median 2^14, sigma 1, clipped [2^10, 2^19]); m in {64,...,512}."""
    import thesiscode_amd as tc

    count = int(rate * seconds * 1.2) + 100
    u = tc.uniforms(seed, count, stream=0)
    arrivals = np.cumsum(-np.log1p(-u) / rate)
    count = int(np.searchsorted(arrivals, seconds))
    arrivals = arrivals[:count]
    z = tc.normals(seed, count, stream=1)
    lengths = np.clip(np.floor(np.exp(np.log(2.0 ** 14) + z)), 2 ** 10, 2 ** 19).astype(np.int64)
    windows = np.array([64, 128, 256, 512])[np.minimum((4 * tc.uniforms(seed, count, stream=2)).astype(int), 3)]
    return arrivals, lengths, windows


def run_serve_bench(seconds: float, rate: float, budget_gb: float, tick_cells: float):
    """Requests are slices of one long random walk (seed 5) at random offsets."""
    import threading

    import thesiscode_amd as tc

    arrivals, lengths, windows = serve_spec(5, seconds, rate)
    walk = tc.random_walk(5, 1 << 25, stream=3)
    offs = np.floor(tc.uniforms(5, len(lengths), stream=4) * (walk.size - lengths)).astype(np.int64)
    cells = np.array([tc.cells(int(n), int(m)) for n, m in zip(lengths, windows)], dtype=np.float64)
    lat = {}
    queue = {}
    with tc.Device(0) as dev:
        with tc.Server(dev, hbm_budget_bytes=int(budget_gb * (1 << 30)), tick_cells=int(tick_cells)) as srv:
            # warm the kernels
            srv.wait(srv.submit(walk[:20000], 128))
            ids = {}
            t0 = time.perf_counter()

            def submitter():
                for k in range(len(lengths)):
                    delay = t0 + arrivals[k] - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                    rid = srv.submit(walk[offs[k]:offs[k] + lengths[k]], int(windows[k]))
                    ids[rid] = k

            th = threading.Thread(target=submitter)
            th.start()
            last_print = t0
            finished = {}
            deadline = t0 + seconds + 300.0
            while th.is_alive() or len(lat) < len(lengths):
                for rid, st, l_ms, q_ms, _, _ in srv.completed():
                    if st != 0:
                        raise RuntimeError(f"request failed with status {st}")
                    finished[rid] = (l_ms, q_ms)
                for rid in [r for r in finished if r in ids]:
                    lat[ids[rid]], queue[ids[rid]] = finished.pop(rid)
                time.sleep(0.002)
                now = time.perf_counter()
                if now > deadline:
                    raise RuntimeError(f"serving did not finish: {len(lat)}/{len(lengths)} done")
                if now - last_print > 10:
                    print(f"serve: {now - t0:.0f} s, {len(lat)}/{len(lengths)} done", file=sys.stderr, flush=True)
                    last_print = now
            t_end = time.perf_counter()
            th.join()
            st = srv.stats()
    lat_v = np.array([lat[k] for k in range(len(lengths))])
    q_v = np.array([queue[k] for k in range(len(lengths))])
    return {"requests": len(lengths), "offered_rate_per_s": rate, "duration_s": seconds,
            "cells": float(cells.sum()), "offered_cells_per_s": float(cells.sum() / seconds),
            "served_cells_per_s": float(cells.sum() / (t_end - t0)), "wall_s": t_end - t0,
            "p50_ms": float(np.percentile(lat_v, 50)), "p99_ms": float(np.percentile(lat_v, 99)),
            "mean_ms": float(lat_v.mean()), "max_ms": float(lat_v.max()),
            "queue_p99_ms": float(np.percentile(q_v, 99)), "ticks": st["ticks"]}


def run_precision(workload: str, tiles=(256, 1024, 4096, 8192)):
    """Profile error of the FP32-reseeded path against the exact FP64 path on one workload.

    rel = |d2 - d2_f64| / max(d2_f64, 2m * 1e-3); index "equal" is exact equality with
    the FP64 index, "near" is within the exclusion zone of it.
    """
    import thesiscode_amd as tc

    kind, seed, n, m, _ = WORKLOADS[workload]
    A = make_series(kind, seed, n, m)
    p64, i64, t64 = tc.selfjoin_f64(A, m)
    den = np.maximum(p64, 1e-3 * 2 * m)
    rows = []
    for acc in (False, True):
        for B in tiles:
            p, i, tm = tc.selfjoin_f32r(A, m, tile_len=B, accum64=acc, refine=True)
            pu, _, _ = tc.selfjoin_f32r(A, m, tile_len=B, accum64=acc, refine=False)
            e = np.abs(p - p64) / den
            eu = np.abs(pu - p64) / den
            rows.append({"B": B, "accumulation": "f64" if acc else "f32", "max_rel": float(e.max()),
                         "mean_rel": float(e.mean()), "p999_rel": float(np.quantile(e, 0.999)),
                         "max_rel_unrefined": float(eu.max()), "mean_rel_unrefined": float(eu.mean()),
                         "index_equal": float(np.mean(i == i64)),
                         "index_near": float(np.mean(np.abs(i - i64) <= m // 4)),
                         "device_ms": tm["total_ms"]})
    return {"workload": workload, "n": n, "m": m, "f64_device_ms": t64["total_ms"], "variants": rows}


def run_prescrimp(workload: str, stride: float, steps: int, warmup: int, cpu: bool):
    """PreSCRIMP (scrimppp_orig's anytime phase) on the workload's series: sampled rows'
    distance profiles on the FP64 matrix cores plus the diagonal walks.  The unit is one
    distance evaluation of a sampled row against a window (2m flops on the matrix cores).
    The CPU baseline is the oracle's sequential restatement on a bounded set of rows."""
    import thesiscode_amd as tc

    kind, seed, n, m, _ = WORKLOADS[workload]
    A = make_series(kind, seed, n, m)
    for _ in range(warmup):
        tc.prescrimp_f64(A, m, stride)
    t0 = time.perf_counter()
    tms = []
    for _ in range(steps):
        _, _, tm = tc.prescrimp_f64(A, m, stride)
        tms.append(tm)
    wall = time.perf_counter() - t0
    evals = tms[0]["cells"]
    kernel_ms = min(t["kernel_ms"] for t in tms)
    achieved = 2.0 * m * evals / (kernel_ms / 1e3) / 1e12
    out = {"metric": "PreSCRIMP distance evaluations/s", "value": evals * steps / wall, "unit": "evaluations/s",
           "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * wall / steps,
           "higher_is_better": True, "dtype": "f64",
           "data": "synthetic (project Philox4x32-10 series; host buffers, copies included)",
           "config": {"workload": workload, "n": n, "m": m, "stride": stride,
                      "step": max(1, int(np.floor(stride * m))), "sampled_rows": tms[0]["work_items"],
                      "evaluations": evals},
           "roofline": {"bound": "mfma", "kernel": "k_pre_rows + k_pre_walk", "achieved": achieved,
                        "peak": FP64_MATRIX_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_MATRIX_PEAK_TFLOPS, "kernel_ms": kernel_ms,
                        "flops_per_evaluation": 2 * m}}
    if cpu:
        from oracle import oracle as orc

        P = n - m + 1
        step = max(1, int(np.floor(stride * m)))
        rows = np.arange(0, P, step)
        k = 1
        while True:
            sample = rows[:: max(1, rows.size // k)][:k]
            t0 = time.perf_counter()
            orc.prescrimp(A, m, stride, order=sample)
            dt = time.perf_counter() - t0
            if dt >= 5.0 or k >= rows.size:
                break
            k = min(rows.size, max(k * 2, int(k * 6.0 / max(dt, 1e-3))))
        out["cpu_baseline"] = {"value": sample.size * P / dt, "unit": "evaluations/s", "cores": 1, "kind": "port",
                               "sample": f"{workload}: {sample.size} of {rows.size} sampled rows with their walks "
                                         f"({sample.size * P:.3e} evaluations, {dt:.1f} s)"}
    return out


def copy_probe(gib: int = 1, reps: int = 3):
    """A plain device-to-device copy of known size, to scale rocprof's FETCH_SIZE and
    WRITE_SIZE counters into bytes (run this under `rocprofv3 --pmc ...`)."""
    import torch

    n = gib << 30
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    return {"copy_bytes": n, "reps": reps}


def cpu_baseline(workload: str, target_s: float = 15.0):
    """Oracle restatement of scrimp_triv_par on a bounded sample (every k-th block).

    The stride is refined until the sample takes at least half of target_s: the
    first probe is dominated by the unsampled statistics and seeds.
    """
    from oracle import oracle as orc

    kind, seed, n, m, _ = WORKLOADS[workload]
    A = make_series(kind, seed, n, m)
    threads = max(1, min(16, os.cpu_count() or 1))
    stride = 256
    _, _, cells, secs = orc.triv_par(A, m, nranks=threads, nthreads=threads, kernel=1, block_stride=stride)
    total = cells * stride
    for _ in range(3):
        rate = cells / max(secs, 1e-6)
        new_stride = int(max(1, min(256, np.floor(total / (rate * target_s)))))
        if new_stride == stride and secs >= 0.5 * target_s:
            break
        stride = new_stride
        _, _, cells, secs = orc.triv_par(A, m, nranks=threads, nthreads=threads, kernel=1,
                                         block_stride=stride)
        if secs >= 0.5 * target_s or stride == 1:
            break
    return {
        "value": cells / secs, "unit": "cells/s", "cores": threads, "kind": "port",
        "sample": (f"{workload}: every {stride}th block of 500 diagonals of each of {threads} ranks "
                   f"({cells:.3e} cells, {secs:.1f} s; stats+seeds+evaluation+reduction)"),
    }


def read_traffic(workload: str):
    path = ROOT / "profiles" / f"traffic_{workload}.json"
    if path.exists():
        return json.loads(path.read_text()).get("hbm_bytes_per_launch")
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2-f64", choices=sorted(WORKLOADS))
    ap.add_argument("--band-k", type=int, default=0)
    ap.add_argument("--tile-len", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", action="store_true", help="run the 4,096-series batched workload instead")
    ap.add_argument("--tune", action="store_true", help="autotune the batch before timing it")
    ap.add_argument("--serve", action="store_true", help="run the Poisson serving workload instead")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-budget-gb", type=float, default=16.0)
    ap.add_argument("--serve-tick-cells", type=float, default=1.5e9)
    ap.add_argument("--no-variant", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 1 GiB device to device 3 times (calibration of rocprof byte counters)")
    ap.add_argument("--prescrimp", action="store_true",
                    help="run PreSCRIMP (scrimppp_orig's anytime phase) on --workload instead")
    ap.add_argument("--prescrimp-stride", type=float, default=0.25)
    ap.add_argument("--precision", action="store_true",
                    help="report the FP32-reseeded profile error against FP64 on --workload")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path yet "
              "(a self-join would shard by band pairs, one rank per GPU); run with --gpus 1", file=sys.stderr)
        return 2

    if args.copy_probe:
        print(json.dumps(copy_probe()), flush=True)
        return 0
    if args.precision:
        print(json.dumps(run_precision(args.workload)), flush=True)
        return 0
    if args.prescrimp:
        print(json.dumps(run_prescrimp(args.workload, args.prescrimp_stride, args.steps, args.warmup,
                                       not args.no_cpu_baseline)), flush=True)
        return 0
    if args.batch:
        r = run_batch_bench(args.steps, args.warmup, args.tune)
        ms = 1000.0 * r["wall_s"] / args.steps
        print(json.dumps({"metric": "matrix-profile cells/s", "value": r["cells"] / (ms / 1000.0), "unit": "cells/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
                          "higher_is_better": True, "dtype": "f32 recurrence, f64 stats/seeds",
                          "data": "synthetic FP32 random walks", "config": {"workload": "batch-4096", **{
                              k: v for k, v in r.items() if k not in ("wall_s",)}}}), flush=True)
        return 0
    if args.serve:
        r = run_serve_bench(args.serve_seconds, args.serve_rate, args.serve_budget_gb, args.serve_tick_cells)
        print(json.dumps({"metric": "matrix-profile cells/s; p50/p99 latency", "value": r["served_cells_per_s"],
                          "unit": "cells/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f32 recurrence, f64 stats/seeds", "data": "synthetic random-walk requests",
                          "config": {"workload": "serve-poisson", **r}}), flush=True)
        return 0
    r = run_gpu(args.workload, args.steps, args.warmup, args.band_k, args.tile_len)
    ms_per_step = 1000.0 * r["wall_s"] / args.steps
    value = r["cells"] / (ms_per_step / 1000.0)
    peak = FP64_VECTOR_PEAK_TFLOPS if r["precision"] == "f64" else FP32_VECTOR_PEAK_TFLOPS
    achieved = FLOPS_PER_CELL * r["cells"] / (r["kernel_ms"] / 1000.0) / 1e12
    kernel_name = "k_walk_f64" if r["precision"] == "f64" else "k_tiles_f32"
    out = {
        "metric": "matrix-profile cells/s", "value": value, "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if r["precision"] == "f64" else "f32 recurrence, f64 stats/seeds",
        "data": "synthetic (project Philox4x32-10 random walk; inputs resident in HBM)",
        "config": {"workload": args.workload, "n": r["n"], "m": r["m"], "exclusion": r["m"] // 4,
                   "cells": r["cells"], "precision": r["precision"], "band_k": r["band_k"],
                   "tile_len": r["tile_len"], "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": kernel_name, "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": read_traffic(args.workload),
                     "kernel_ms": r["kernel_ms"], "flops_per_cell": FLOPS_PER_CELL},
        "device_ms_per_step": r["device_ms"],
    }
    if not args.no_variant and args.workload == "cfg2-f64":
        # the FP32 recurrence on the same series: FP64 accumulation (the default,
        # within 1e-4 of FP64) and FP32 accumulation
        out["variants"] = {}
        vsteps = max(1, args.steps // 2)
        for name, acc in (("cfg2-f32r", True), ("cfg2-f32r-acc32", False)):
            v = run_gpu("cfg2-f32r", vsteps, 1, 0, 0, accum64=acc)
            out["variants"][name] = {"value": v["cells"] / (v["wall_s"] / vsteps), "kernel_ms": v["kernel_ms"],
                                     "device_ms": v["device_ms"], "band_k": v["band_k"],
                                     "tile_len": v["tile_len"], "accum64": acc,
                                     "frac_fp32_peak": FLOPS_PER_CELL * v["cells"] / (v["kernel_ms"] / 1e3) / 1e12
                                     / FP32_VECTOR_PEAK_TFLOPS}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
