"""bench.py's command line on the CPU: the benchmark is one GPU, so --gpus N
with N > 1 exits (status 2) with a message before touching any device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_bench_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0
    for w in ("d2h-120", "c2v-24", "c2v-180", "c2v-batch32", "c2v-100-gradient", "d2h-120-16roots"):
        assert w in r.stdout, w
