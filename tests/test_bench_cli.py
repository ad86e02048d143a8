"""bench.py's command-line contract that needs no GPU: more than one GPU is refused with a message."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_bench_refuses_more_than_one_gpu_for_every_config():
    for cfg in ("0", "2", "3", "4"):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4", "--config", cfg],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode != 0 and "one GPU" in r.stderr, cfg
