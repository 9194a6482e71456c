"""Build the SCRIMP HIP library (libscrimp_hip.so) in-tree for gfx950.

Every source is compiled by hipcc into an object under thesiscode_amd/_build/,
then linked into thesiscode_amd/libscrimp_hip.so.  `python -m
thesiscode_amd.build` builds; `--oracle` also builds the CPU oracle
(test infrastructure, oracle/liboracle.so).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libscrimp_hip.so"
ARCH = os.environ.get("SCRIMP_HIP_ARCH", "gfx950")

# -ffp-contract=off: the exact path's arithmetic is written with explicit fma()
# and must not be re-fused.  Packed FP32 instructions are disabled.
# -amdgpu-mfma-vgpr-form: matrix-core accumulators live in ordinary VGPRs.
# Without it the compiler keeps them in AGPRs and copies all of them to VGPRs
# and back around every loop iteration of the MFMA loops (64 moves and a
# pipeline drain per 16 MFMAs in PreSCRIMP), and AGPRs count against
# occupancy on top of the VGPRs.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-command-line-argument",
]
SOURCES = ["exact_f64.hip", "reseeded_f32.hip", "prescrimp.hip", "api.cpp", "batch.cpp", "datagen.cpp",
           "serve.cpp"]


def _hipcc() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(cand).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libscrimp_hip.so")
    return cand


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [_hipcc(), *HIP_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST = CSRC / "host"
TOOLS = {
    "scrimppp_hip": ["scrimppp_hip.cpp", "scrimp_hip_algos.cpp", "profile_io.cpp"],
    "cvt_ts_hip": ["cvt_tools.cpp", "profile_io.cpp"],
    "cvt_matprof_hip": ["cvt_tools.cpp", "profile_io.cpp"],
}


def build_tools(verbose: bool = False) -> list[Path]:
    """Host executables (CLI + converters) on the C ABI, linked against libscrimp_hip.so."""
    gxx = shutil.which("g++") or "g++"
    outs = []
    for name, srcs in TOOLS.items():
        exe = PKG / name
        cmd = [gxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{HOST}",
               *[str(HOST / s) for s in srcs], f"-L{PKG}", "-lscrimp_hip", "-Wl,-rpath,$ORIGIN", "-o", str(exe)]
        if name.startswith("cvt_"):
            cmd = [c for c in cmd if c not in ("-lscrimp_hip", f"-L{PKG}")]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"tool build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        outs.append(exe)
        if verbose:
            print(f"built {exe}")
    return outs


def build_oracle(verbose: bool = False) -> Path:
    r = subprocess.run(["make", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    out = ROOT / "oracle" / "liboracle.so"
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build_library(verbose=True)
    build_tools(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
