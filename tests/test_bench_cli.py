"""bench.py's command line on the CPU: one GPU is measured, so a request for
more exits with status 2 and says why, before anything touches a GPU."""
import os
import subprocess
import sys

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU is measured" in r.stderr
    assert r.stdout == ""
