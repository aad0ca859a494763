"""bench.py's command-line contract that needs no GPU: this build measures one
MI355X, so a multi-GPU request exits with a message instead of a number."""
import os
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2",
                        "--steps", "1", "--warmup", "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert r.stdout.strip() == ""  # no JSON line
    assert "one GPU" in r.stderr and "single MI355X" in r.stderr


def test_bench_help_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for w in ("headline", "snr-sweep", "uneven", "mixed-precision", "serving", "variants", "copy"):
        assert w in r.stdout
