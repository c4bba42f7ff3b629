"""The CSV harness (tribit_amd/harness.py) against a stand-in executable: columns, skips and
repetitions of the reference's run_single_gpu.py (artifact/sc26/run_benchmarks/run_single_gpu.py)."""
import csv
import os
import stat
import sys

from tribit_amd import harness

FAKE = f"""#!{sys.executable}
import json, sys
a = sys.argv[1:]
path = a[a.index("-i") + 1]
tri = 1 if "square" in path else 2
print("Graph              : x")
print("Blocks of threads  : 7")
print("Preprocessing (ms) : 1.250")
print("Kernel (ms)        : 0.500")
print(f"Triangles          : {{tri}}")
if "--json" in a:
    json.dump({{"peak_pool_bytes": 12345}}, open(a[a.index("--json") + 1], "w"))
"""


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_mtx_shape(tmp_path):
    p = tmp_path / "g.mtx"
    _write(p, "%%MatrixMarket matrix coordinate pattern general\n% c\n\n4 4 2\n1 2\n2 3\n")
    assert harness.mtx_shape(str(p)) == (4, 4, 2)
    _write(p, "4 5 2\n1 2\n")
    assert harness.mtx_shape(str(p)) is None
    _write(p, "4 4\n")
    assert harness.mtx_shape(str(p)) is None


def test_parse_summary():
    out = "Blocks of threads  : 12\nPreprocessing (ms) : 3.5\nKernel (ms)        : 0.25\nTriangles          : 608389\n"
    assert harness.parse_summary(out) == ["12", "3.5", "0.25", "608389"]


def test_csv_rows_skips_and_memory(tmp_path):
    exe = tmp_path / "tribit"
    _write(exe, FAKE)
    exe.chmod(exe.stat().st_mode | stat.S_IXUSR)
    data = tmp_path / "data"
    (data / "sub").mkdir(parents=True)
    _write(data / "square.mtx", "3 3 2\n1 2\n2 3\n")
    _write(data / "rect.mtx", "3 4 1\n1 2\n")            # not square: skipped
    _write(data / "sub" / "denied.mtx", "3 3 1\n1 2\n")  # matches the deny file
    _write(data / "sub" / "other.mtx", "5 5 1\n1 2\n")
    _write(data / "notes.txt", "3 3 1\n")
    deny = tmp_path / "deny.txt"
    _write(deny, "denied\n")
    out = tmp_path / "res" / "results.csv"
    assert harness.main(["--exe_path", str(exe), "--data_path", str(data), "--denyfile_path", str(deny),
                         "--out_path", str(out), "--n_repetitions", "2", "--get_memory_consumption"]) == 0
    rows = list(csv.reader(open(out)))
    assert rows[0] == harness.HEADER.split(",")
    body = rows[1:]
    assert [r[1] for r in body] == ["square.mtx", "square.mtx", "other.mtx", "other.mtx"]
    assert body[0] == ["tribit", "square.mtx", "3", "3", "2", "7", "1.250", "0.500", "1", "12345"]
    assert body[2][2:5] == ["5", "5", "1"] and body[2][8] == "2"
    # appending keeps one header; without memory the column is 0; max_matrices bounds the graphs
    assert harness.main(["--exe_path", str(exe), "--data_path", str(data), "--out_path", str(out),
                         "--max_matrices", "1"]) == 0
    rows = list(csv.reader(open(out)))
    assert sum(r == harness.HEADER.split(",") for r in rows) == 1
    assert len(rows) == 6 and rows[-1][-1] == "0"
