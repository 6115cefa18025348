"""CPU: bench.py's command line. More than one GPU is refused with a message (the path is
measured on one GPU; how it would shard is a design note), and the workload list is the one the
documentation names. Nothing here touches a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in r.stderr or "--gpus 1" in r.stderr
    assert r.stdout.strip() == ""


def test_workload_choices():
    r = run_bench("--help")
    assert r.returncode == 0
    for w in ("euler_2048", "allaire_1024", "euler_200", "euler_slab", "batched", "served"):
        assert w in r.stdout
