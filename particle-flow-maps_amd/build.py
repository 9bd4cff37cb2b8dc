"""Build libpfm.so (HIP, gfx950) in-tree with hipcc.

Usage: python particle-flow-maps_amd/build.py [--jobs N]

Each source under csrc/ is compiled to an object with hipcc for gfx950, then
linked into particle-flow-maps_amd/pfm/libpfm.so, which the ctypes loader in
pfm/_lib.py opens.  Packed-FP32 instruction selection is disabled for all
device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "pfm")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libpfm.so")
ARCH = os.environ.get("PFM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["abi.cpp", "particles.hip", "grid.hip", "p2g.hip", "mgpcg.hip", "scene.cpp", "smoke.hip",
           "batch.cpp", "batch_mg.hip"]

CFLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# Sources whose fp32 passes must equal a sequential restatement bit for bit:
# HIP's __fmul_rn/__fadd_rn are plain operators, so only a file-wide
# -ffp-contract=off keeps the compiler from fusing them into FMAs.  The
# other sources use -ffp-contract=on: a multiply-add is fused only inside one
# source expression, the same way in every kernel that inlines it (the
# default, fast, also fuses across statements depending on the surrounding
# code, so two kernels sharing one device function could round differently).
NO_CONTRACT = {"mgpcg.hip", "smoke.hip", "batch_mg.hip"}


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith((".hip", ".cpp")) else []
    extra = ["-ffp-contract=off"] if src in NO_CONTRACT else ["-ffp-contract=on"]
    cmd = [hipcc(), *lang, *CFLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if err:
        print(err, file=sys.stderr)
    return obj


def build(jobs: int = 8, verbose: bool = True) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    build(a.jobs)
