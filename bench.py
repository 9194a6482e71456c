"""Benchmark of the SCRIMP self-join hot path on one MI355X.

Default run (no flags): the precision configuration -- n = 2^22 points of a
random walk + 3 sinusoids + 20 Gaussian bumps (seed 2), window m = 1024,
exclusion m/4 -- with the series already resident in HBM.  One step is one
full self-join (statistics, seeds, all diagonals, merge, unpack to
profile/index).

Prints ONE JSON line:
  value         cells/s of the exact FP64 path (the reference's precision, bit-identical
                to the oracle's restatement), timed over --steps steps;
  variants      the FP32-reseeded path (FP64 accumulation) at the default re-seed interval
                and at B = 256, 1024, 4096, each with its profile error against that FP64
                run (max/mean relative, index agreement);
  roofline      the walker's arithmetic rate (8 flops per cell: score 2 mul + 1 FMA,
                recurrence 2 FMA) over the MI355X FP64 vector peak, from the library's
                HIP events on the kernel's own stream; roofline_f32r: the tile kernel's
                over the FP32 vector peak;
  serving       the serving configuration for --serve-leg-seconds of Poisson arrivals
                (p50/p99 latency, served cells/s), and the same at 81% of the batched
                throughput measured in the same run (batch: the 4,096-series batch);
  cpu_baseline  the oracle's restatement of scrimp_triv_par on every visible core of the
                host (one MPI process per core when MPICH is there), on a bounded sample.
The metric counts cells = W(W+1)/2, W = n - m + 1 - m/4 - 1 (the reference logs
(P-ez)^2/t, about twice that).
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
    from thesiscode_amd import workloads as wl

    return wl.make_series(kind, seed, n, m)


def run_gpu(workload: str, steps: int, warmup: int, band_k: int = 0, tile_len: int = 0, accum64: bool = True,
            A=None, dev=None, d_ts=None, precision=None, download=False):
    """Time `steps` self-joins of one workload with the series resident in HBM.

    Wall-clock time over the steps (bracketed by device synchronisations) gives the
    rate; kernel_ms is the dominant kernel's own time from the library's HIP events
    on the stream it is launched on.  download=True also returns the last step's
    profile and index."""
    import thesiscode_amd as tc

    kind, seed, n, m, prec = WORKLOADS[workload]
    prec = precision or prec
    if A is None:
        A = make_series(kind, seed, n, m)
    P = n - m + 1
    own = dev is None
    if own:
        dev = tc.Device(0)
    try:
        if d_ts is None:
            d_ts = dev.upload(A)
        d_prof = dev.alloc(8 * P)
        d_idx = dev.alloc(8 * P)
        opts = tc.make_opts(band_k=band_k, tile_len=tile_len, accum64=accum64)
        for _ in range(warmup):
            dev.selfjoin(prec, d_ts, n, m, d_prof, d_idx, opts)
        dev.sync()
        kern_ms, total_ms = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            tm = dev.selfjoin(prec, d_ts, n, m, d_prof, d_idx, opts)
            kern_ms.append(tm["kernel_ms"])
            total_ms.append(tm["total_ms"])
        dev.sync()
        wall = time.perf_counter() - t0
        out = {
            "n": n, "m": m, "precision": prec, "cells": tm["cells"], "wall_s": wall, "steps": steps,
            "value": tm["cells"] * steps / wall,
            "kernel_ms": float(np.mean(kern_ms)), "device_ms": float(np.mean(total_ms)),
            "band_k": tm["band_k"], "tile_len": tm["tile_len"], "work_items": tm["work_items"],
            "kernel": kernel_name(prec, tm, accum64),
        }
        if download:
            out["profile"] = dev.download(d_prof, np.empty(P))
            out["index"] = dev.download(d_idx, np.empty(P, dtype=np.int64))
        dev.free(d_prof)
        dev.free(d_idx)
        return out
    finally:
        if own:
            dev.close()


def kernel_name(precision: str, tm: dict, accum64: bool = True) -> str:
    """The dominant kernel of a self-join, as rocprof names it."""
    k = tm["band_k"]
    if precision == "f32r":
        return f"k_tiles_f32<{k},{'true' if accum64 else 'false'}>" if k <= 8 else f"k_tiles_f32_wide<{k}>"
    if tm.get("walk_form", 0) == 4:  # column segments: the start-value pre-pass, then the walk
        return f"k_seg_seeds + k_walk_f64_seg<{k}>"
    form = {1: "k_walk_f64_wide", 2: "k_walk_f64_pair", 3: "k_walk_f64_rot"}.get(tm.get("walk_form", 0))
    return f"{form}<{k}>" if (k == 8 and form) else f"k_walk_f64<{k}>"


def profile_error(p, i, p64, i64, m):
    """rel = |d2 - d2_f64| / max(d2_f64, 2m * 1e-3); index equal / within the exclusion zone."""
    e = np.abs(p - p64) / np.maximum(p64, 1e-3 * 2 * m)
    return {"max_rel": float(e.max()), "mean_rel": float(e.mean()), "p999_rel": float(np.quantile(e, 0.999)),
            "index_equal": float(np.mean(i == i64)), "index_near": float(np.mean(np.abs(i - i64) <= m // 4))}


def batch_spec(seed: int = 3, count: int = 4096):
    from thesiscode_amd import workloads as wl

    return wl.batch_spec(seed, count)


def gen_batch_series(seed, lengths):
    import concurrent.futures as cf

    import thesiscode_amd as tc

    def one(s):
        return tc.random_walk(seed, int(lengths[s]), stream=100 + s).astype(np.float32).astype(np.float64)

    with cf.ThreadPoolExecutor(max_workers=min(16, visible_cores())) as ex:
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
            tuned = b.tune(d_ts, d_p, d_i) if tune else None
            for _ in range(warmup):
                b.run(d_ts, d_p, d_i, use_graph=True)
            dev.sync()
            t0 = time.perf_counter()
            dev_ms = []
            for _ in range(steps):
                dev_ms.append(b.run(d_ts, d_p, d_i, use_graph=True)["total_ms"])
            wall = time.perf_counter() - t0
            info = b.info()
    return {"cells": info["cells"], "wall_s": wall, "value": info["cells"] * steps / wall,
            "device_ms": float(np.mean(dev_ms)), "band_k": info["band_k"],
            "tile_len": info["tile_len"], "series": len(lengths), "points": int(lengths.sum()),
            "tiles": info["tiles"], "tuned": tuned}


def serve_spec(seed: int, seconds: float, rate: float):
    from thesiscode_amd import workloads as wl

    return wl.serve_spec(seed, seconds, rate)


def run_serve_bench(seconds: float, rate: float, budget_gb: float, tick_cells: float, walk=None):
    """Poisson arrivals replayed in real time; requests are slices of one long random
    walk (seed 5) at seeded offsets.  Latency is submit -> results in the caller's buffers."""
    import threading

    import thesiscode_amd as tc
    from thesiscode_amd import workloads as wl

    arrivals, lengths, windows = wl.serve_spec(5, seconds, rate)
    if walk is None:
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


def rate_for_load(cells_per_s: float, seconds: float, seed: int = 5) -> float:
    """Arrival rate at which the serving trace (seed 5, `seconds` long) offers `cells_per_s`.

    The trace's request sizes are drawn per request index, so its offered cells/s is a
    step function of the rate; a few fixed-point steps on the trace itself land on the
    target (the mean size of another sample would miss it by the heavy tail's noise).
    """
    import thesiscode_amd as tc

    rate = cells_per_s / mean_request_cells(seed)
    for _ in range(6):
        _, lengths, windows = serve_spec(seed, seconds, rate)
        offered = sum(tc.cells(int(n), int(m)) for n, m in zip(lengths, windows)) / seconds
        rate *= cells_per_s / offered
    return rate


def mean_request_cells(seed: int = 5, count: int = 20000) -> float:
    """Mean cells of a serving request (same generator as the trace)."""
    import thesiscode_amd as tc

    _, lengths, windows = serve_spec(seed, count / 2000.0, 2000.0)
    return float(np.mean([tc.cells(int(n), int(m)) for n, m in zip(lengths, windows)]))


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


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def visible_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def find_mpirun():
    import shutil

    for cand in (shutil.which("mpirun"), "/opt/conda/bin/mpirun"):
        if cand and Path(cand).exists():
            return cand
    return None


def _mpi_run(mpirun, ranks, path, m, stride):
    import subprocess

    exe = ROOT / "oracle" / "triv_par_mpi"
    cmd = [mpirun, "-n", str(ranks), str(exe), str(path), str(m), str(stride)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(workload: str, target_s: float = 20.0, A=None, launcher_choice: str = "mpi"):
    """The oracle's restatement of scrimp_triv_par on every visible core of this host.

    One MPI process per core (`mpirun -n <cores> oracle/triv_par_mpi`, the way the
    reference runs) when MPICH is present, else one POSIX thread per rank in one
    process.  Both run the AVX2 ABBA kernel with 2 partitions per rank and blocks
    of 500 diagonals, and time statistics + seeds + evaluation + the min
    reduction (the reference's IGNORE_INIT_TIME).  Only every k-th block of each
    rank is evaluated (a bounded sample; k from a first probe so that the sample
    takes about target_s) and the rate is cells evaluated / time.
    """
    import tempfile

    from oracle import oracle as orc
    from thesiscode_amd import workloads as wl

    kind, seed, n, m, _ = WORKLOADS[workload]
    if A is None:
        A = wl.make_series(kind, seed, n, m)
    cores = visible_cores()
    mpirun = find_mpirun()
    total = float(_cells(n, m))
    use_mpi = launcher_choice != "threads" and mpirun and (ROOT / "oracle" / "triv_par_mpi").exists()
    if use_mpi:
        with tempfile.TemporaryDirectory() as td:
            path = Path(td) / "series.f64"
            A.astype("<f8").tofile(path)
            probe = 256
            r = _mpi_run(mpirun, cores, path, m, probe)
            rate = r["cells"] / max(r["seconds"], 1e-6)
            stride = int(max(1, min(probe, np.floor(total / (rate * target_s)))))
            if stride != probe:
                r = _mpi_run(mpirun, cores, path, m, stride)
        launcher = f"mpirun -n {cores} oracle/triv_par_mpi (MPICH, one process per rank)"
        cells, secs = r["cells"], r["seconds"]
    else:
        stride = 256
        _, _, cells, secs = orc.triv_par(A, m, nranks=cores, nthreads=cores, kernel=1, block_stride=stride)
        rate = cells / max(secs, 1e-6)
        new_stride = int(max(1, min(256, np.floor(total / (rate * target_s)))))
        if new_stride != stride:
            stride = new_stride
            _, _, cells, secs = orc.triv_par(A, m, nranks=cores, nthreads=cores, kernel=1, block_stride=stride)
        launcher = f"{cores} POSIX threads, one rank each (no MPI on this host)"
    return {
        "value": cells / secs, "unit": "cells/s", "cores": cores, "kind": "port",
        "cpu_model": cpu_model(), "visible_cores": cores, "launcher": launcher,
        "sample": (f"{workload}: every {stride}th block of 500 diagonals of each of {cores} ranks "
                   f"({cells:.3e} of {total:.3e} cells, {secs:.1f} s; stats+seeds+evaluation+reduction)"),
    }


def _cells(n, m):
    w = max(0, n - m + 1 - m // 4 - 1)
    return w * (w + 1) // 2


def read_traffic(workload: str):
    path = ROOT / "profiles" / f"traffic_{workload}.json"
    if path.exists():
        return json.loads(path.read_text()).get("hbm_bytes_per_launch")
    return None


def default_run(args) -> dict:
    """The north-star row: cfg3 exact FP64 value, FP32-reseeded variants with their
    error against it, both kernels' rooflines, the serving legs and the CPU baseline."""
    import thesiscode_amd as tc

    workload = args.workload
    kind, seed, n, m, prec = WORKLOADS[workload]
    A = make_series(kind, seed, n, m)
    out = {}
    with tc.Device(0) as dev:
        d_ts = dev.upload(A)
        print(f"bench: {workload} exact FP64, {args.steps} steps", file=sys.stderr, flush=True)
        r = run_gpu(workload, args.steps, args.warmup, args.band_k, 0, A=A, dev=dev, d_ts=d_ts,
                    precision="f64", download=not args.no_variant)
        p64, i64 = r.pop("profile", None), r.pop("index", None)
        variants = {}
        if not args.no_variant:
            # B = 0: the library's default re-seed interval
            for B in (0, 256, 1024, 4096):
                print(f"bench: {workload} FP32-reseeded B={B or 'default'}", file=sys.stderr, flush=True)
                v = run_gpu(workload, args.variant_steps, 1, 0, B, accum64=True, A=A, dev=dev, d_ts=d_ts,
                            precision="f32r", download=True)
                err = profile_error(v.pop("profile"), v.pop("index"), p64, i64, m)
                name = f"{workload.split('-')[0]}-f32r-B{v['tile_len']}" + ("-default" if B == 0 else "")
                variants[name] = {"value": v["value"], "steps": v["steps"], "kernel": v["kernel"],
                                  "kernel_ms": v["kernel_ms"],
                                  "device_ms": v["device_ms"], "band_k": v["band_k"], "tile_len": v["tile_len"],
                                  "accum64": True, **err,
                                  "frac_fp32_peak": FLOPS_PER_CELL * v["cells"] / (v["kernel_ms"] / 1e3) / 1e12
                                  / FP32_VECTOR_PEAK_TFLOPS}
    out["f64"] = r
    out["variants"] = variants
    if not args.no_serve_leg:
        print("bench: batch (4,096 series)", file=sys.stderr, flush=True)
        b = run_batch_bench(3, 1, False)
        walk = tc.random_walk(5, 1 << 25, stream=3)
        print(f"bench: serving {args.serve_leg_seconds:.0f} s at {args.serve_rate:.0f}/s", file=sys.stderr,
              flush=True)
        sv = run_serve_bench(args.serve_leg_seconds, args.serve_rate, args.serve_budget_gb, args.serve_tick_cells,
                             walk=walk)
        rate80 = rate_for_load(0.81 * b["value"], args.serve_leg_seconds)
        print(f"bench: serving {args.serve_leg_seconds:.0f} s at {rate80:.0f}/s (81% of batch)", file=sys.stderr,
              flush=True)
        sv80 = run_serve_bench(args.serve_leg_seconds, rate80, args.serve_budget_gb, args.serve_tick_cells,
                               walk=walk)
        sv80["fraction_of_batched_throughput"] = sv80["offered_cells_per_s"] / b["value"]
        out["batch"] = {k: v for k, v in b.items() if k not in ("wall_s", "tuned")}
        out["serving"] = sv
        out["serving_80pct"] = sv80
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(workload, A=A, launcher_choice=args.cpu_launcher)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg3-f64", choices=sorted(WORKLOADS))
    ap.add_argument("--band-k", type=int, default=0)
    ap.add_argument("--tile-len", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-launcher", choices=("mpi", "threads"), default="mpi",
                    help="CPU baseline: one MPI process per core (default, when MPICH is present) or threads")
    ap.add_argument("--no-variant", action="store_true")
    ap.add_argument("--variant-steps", type=int, default=3)
    ap.add_argument("--no-serve-leg", action="store_true")
    ap.add_argument("--serve-leg-seconds", type=float, default=10.0)
    ap.add_argument("--batch", action="store_true", help="run the 4,096-series batched workload instead")
    ap.add_argument("--tune", action="store_true", help="autotune the batch before timing it")
    ap.add_argument("--serve", action="store_true", help="run the Poisson serving workload instead")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-budget-gb", type=float, default=16.0)
    ap.add_argument("--serve-tick-cells", type=float, default=1.5e9)
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy 1 GiB device to device 3 times (calibration of rocprof byte counters)")
    ap.add_argument("--prescrimp", action="store_true",
                    help="run PreSCRIMP (scrimppp_orig's anytime phase) on --workload instead")
    ap.add_argument("--prescrimp-stride", type=float, default=0.25)
    ap.add_argument("--precision", action="store_true",
                    help="report the FP32-reseeded profile error against FP64 on --workload")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU, and the library has no "
              "multi-GPU path (a self-join would shard by bands, one rank per GPU, merged by one "
              "max all-reduce of the packed keys); run with --gpus 1", file=sys.stderr)
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
        print(json.dumps({"metric": "matrix-profile cells/s", "value": r["value"], "unit": "cells/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
                          "higher_is_better": True, "dtype": "f32 recurrence, f64 stats/seeds",
                          "data": "synthetic FP32 random walks", "config": {"workload": "batch-4096", **{
                              k: v for k, v in r.items() if k not in ("wall_s", "value")}}}), flush=True)
        return 0
    if args.serve:
        r = run_serve_bench(args.serve_seconds, args.serve_rate, args.serve_budget_gb, args.serve_tick_cells)
        print(json.dumps({"metric": "matrix-profile cells/s; p50/p99 latency", "value": r["served_cells_per_s"],
                          "unit": "cells/s", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f32 recurrence, f64 stats/seeds", "data": "synthetic random-walk requests",
                          "config": {"workload": "serve-poisson", **r}}), flush=True)
        return 0

    kind, seed, n, m, prec = WORKLOADS[args.workload]
    if args.workload == "cfg3-f64":
        legs = default_run(args)
        r = legs["f64"]
    else:
        r = run_gpu(args.workload, args.steps, args.warmup, args.band_k, args.tile_len)
        legs = {"cpu_baseline": None if args.no_cpu_baseline else
                cpu_baseline(args.workload, launcher_choice=args.cpu_launcher)}
    ms_per_step = 1000.0 * r["wall_s"] / args.steps
    value = r["cells"] / (ms_per_step / 1000.0)
    f64 = r["precision"] == "f64"
    peak = FP64_VECTOR_PEAK_TFLOPS if f64 else FP32_VECTOR_PEAK_TFLOPS
    achieved = FLOPS_PER_CELL * r["cells"] / (r["kernel_ms"] / 1000.0) / 1e12
    out = {
        "metric": "matrix-profile cells/s", "value": value, "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if f64 else "f32 recurrence, f64 stats/seeds",
        "data": "synthetic (project Philox4x32-10 series; inputs resident in HBM)",
        "config": {"workload": args.workload, "n": r["n"], "m": r["m"], "exclusion": r["m"] // 4,
                   "cells": r["cells"], "precision": r["precision"], "band_k": r["band_k"],
                   "tile_len": r["tile_len"], "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": r["kernel"], "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": read_traffic(args.workload), "kernel_ms": r["kernel_ms"],
                     "flops_per_cell": FLOPS_PER_CELL},
        "device_ms_per_step": r["device_ms"],
    }
    if legs.get("variants"):
        out["variants"] = legs["variants"]
        dflt = next(v for k, v in legs["variants"].items() if k.endswith("-default"))
        a32 = FLOPS_PER_CELL * r["cells"] / (dflt["kernel_ms"] / 1000.0) / 1e12
        out["roofline_f32r"] = {"bound": "valu", "kernel": dflt["kernel"], "achieved": a32,
                                "peak": FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": a32 / FP32_VECTOR_PEAK_TFLOPS,
                                "traffic": read_traffic(args.workload.split("-")[0] + "-f32r"),
                                "kernel_ms": dflt["kernel_ms"], "flops_per_cell": FLOPS_PER_CELL}
    for k in ("batch", "serving", "serving_80pct"):
        if k in legs:
            out[k] = legs[k]
    cpu = legs.get("cpu_baseline")
    if cpu:
        out["cpu_baseline"] = cpu
        ns = {"f64_over_cpu": value / cpu["value"]}
        if legs.get("variants"):
            ns.update({"f32r_value": dflt["value"], "f32r_over_cpu": dflt["value"] / cpu["value"],
                       "f32r_max_rel_err": dflt["max_rel"], "f32r_mean_rel_err": dflt["mean_rel"],
                       "target": ">= 50x the all-core CPU with the FP32-reseeded path within 1e-4 of FP64"})
        out["north_star"] = ns
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
