"""bench.py's contract on the CPU: more than one GPU is refused with a message (exit 2) before any
device or library is touched, for every workload."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    for extra in ([], ["--workload", "mixed"], ["--workload", "batch"]):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", *extra],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 2, (extra, r.stderr)
        assert "one GPU only" in r.stderr
        assert r.stdout == ""
