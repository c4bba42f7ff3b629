"""Benchmark harness over a directory of Matrix Market graphs.

    python -m tribit_amd.harness --exe_path tribit_amd/tribit --data_path DIR [--out_path results.csv]
        [--denyfile_path FILE] [--n_repetitions K] [--dry_run] [--get_memory_consumption] [--max_matrices N]

Writes the CSV of the reference's single-GPU harness
(artifact/sc26/run_benchmarks/run_single_gpu.py:95-100, :170), with the same
options and columns:

    exe,graph,nrows,ncols,nnz,n_blocks,preprocessing_time,kernel_time,triangles,max_memory_consumption

one row per run of `<exe> -i <graph.mtx>`, the four middle fields parsed from the
binary's summary block (tribit.cu:807-818 prints the same labels). Graphs whose
first non-comment line is not three integers of a square matrix are skipped, as
are names matching a pattern of the deny file. max_memory_consumption is the
device memory pool's high-water mark in bytes, read from the binary's --json
record of the same run (the reference profiled a separate run under nsys for
the CUDA pool's utilised size); 0 without --get_memory_consumption. Rows are
appended to an existing file that starts with the same header.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile

HEADER = "exe,graph,nrows,ncols,nnz,n_blocks,preprocessing_time,kernel_time,triangles,max_memory_consumption"
FIELDS = (
    ("n_blocks", r"Blocks of threads\s*:\s*(\d+)"),
    ("preprocessing_time", r"Preprocessing \(ms\)\s*:\s*([\d.]+)"),
    ("kernel_time", r"Kernel \(ms\)\s*:\s*([\d.]+)"),
    ("triangles", r"Triangles\s*:\s*(\d+)"),
)


def mtx_shape(path: str):
    """(nrows, ncols, nnz) of the first non-comment line, or None if it is not three integers of a square
    matrix."""
    with open(path, "r", errors="replace") as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("%"):
                continue
            parts = line.split()
            if len(parts) != 3:
                return None
            try:
                nrows, ncols, nnz = (int(p) for p in parts)
            except ValueError:
                return None
            return (nrows, ncols, nnz) if nrows == ncols else None
    return None


def parse_summary(stdout: str) -> list[str]:
    """The four summary fields, in CSV order; ValueError if one is missing."""
    out = []
    for name, pattern in FIELDS:
        m = re.search(pattern, stdout)
        if not m:
            raise ValueError(f"no '{name}' in the output")
        out.append(m.group(1))
    return out


def graphs_under(root: str) -> list[str]:
    found = []
    for d, _, files in os.walk(root):
        for name in files:
            if name.endswith(".mtx"):
                found.append(os.path.join(d, name))
    return sorted(found)


def run_one(exe: str, path: str, want_memory: bool) -> tuple[list[str], int]:
    """Runs `exe -i path`; returns the summary fields and the pool high-water mark (bytes, 0 if not asked)."""
    cmd = [exe, "-i", path]
    record = None
    if want_memory:
        fd, record = tempfile.mkstemp(suffix=".json")
        os.close(fd)
        cmd += ["--json", record]
    try:
        res = subprocess.run(cmd, capture_output=True, text=True, check=True)
        peak = 0
        if record:
            with open(record) as f:
                peak = int(json.load(f)["peak_pool_bytes"])
        return parse_summary(res.stdout), peak
    finally:
        if record and os.path.exists(record):
            os.remove(record)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--exe_path", required=True)
    ap.add_argument("--data_path", required=True)
    ap.add_argument("--denyfile_path")
    ap.add_argument("--out_path", default="results.csv")
    ap.add_argument("--n_repetitions", type=int, default=1)
    ap.add_argument("--dry_run", action="store_true")
    ap.add_argument("--get_memory_consumption", action="store_true")
    ap.add_argument("--max_matrices", type=int)
    a = ap.parse_args(argv)

    deny = []
    if a.denyfile_path:
        with open(a.denyfile_path) as f:
            deny = [line.strip() for line in f if line.strip()]
    out_dir = os.path.dirname(a.out_path)
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
    exe_name = os.path.basename(a.exe_path)
    done = matrices = 0
    with open(a.out_path, "a+") as out:
        out.seek(0)
        if out.readline().strip() != HEADER:
            out.write(HEADER + "\n")
            out.flush()
        for path in graphs_under(a.data_path):
            if a.max_matrices is not None and matrices >= a.max_matrices:
                break
            name = os.path.basename(path)
            shape = mtx_shape(path)
            if shape is None or any(re.search(p, name) for p in deny):
                continue
            matrices += 1
            if a.dry_run:
                print(done, name, flush=True)
                done += 1
                continue
            try:
                peak = None
                for _ in range(max(1, a.n_repetitions)):
                    fields, p = run_one(a.exe_path, path, a.get_memory_consumption and peak is None)
                    if peak is None:
                        peak = p
                    out.write(",".join([exe_name, name, *map(str, shape), *fields, str(peak)]) + "\n")
                    out.flush()
                    print(done, name, flush=True)
                    done += 1
            except (subprocess.CalledProcessError, ValueError, OSError, KeyError) as e:
                print(f"Execution failed for {path}: {type(e).__name__}: {e}", flush=True)
                if isinstance(e, subprocess.CalledProcessError):
                    print("* stdout:", e.stdout, "\n* stderr:", e.stderr, "\n* returncode:", e.returncode, flush=True)
    print("Total amount of matrices:", done)
    return 0


if __name__ == "__main__":
    sys.exit(main())
