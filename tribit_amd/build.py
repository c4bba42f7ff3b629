"""Builds libtribit_hip.so (gfx950) and the `tribit` command in-tree.

    python -m tribit_amd.build [-v]

Every HIP source is compiled by hipcc for gfx950 only; objects go to
tribit_amd/_build/, the library and command next to this file, so they travel
with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtribit_hip.so")
CLI = os.path.join(HERE, "tribit")

HIP_SOURCES = ["prep.hip", "gen.hip", "tiles.hip", "count.hip", "dense.hip", "batch.hip", "mlp.hip", "api.hip"]
HOST_SOURCES = ["mtx.cpp", "webgraph.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 ops stay off in every kernel (no float math depends on them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> str:
    """Compiles every source (always, even when outputs exist). Returns the library path."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    objs = [os.path.join(BUILD, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-pthread"], verbose)
    os.replace(tmp, LIB)
    tmpc = CLI + ".tmp"
    _run(["g++", *HOST_FLAGS, os.path.join(CSRC, "tribit_cli.cpp"), "-o", tmpc, f"-L{HERE}", "-ltribit_hip",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmpc, CLI)
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    for p in (LIB, CLI):
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
