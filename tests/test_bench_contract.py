"""bench.py's command-line contract that needs no GPU: the bench measures one
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
    for w in ("headline", "snr-sweep", "uneven", "mixed-precision", "serving", "variants", "copy",
              "square"):
        assert w in r.stdout


def _bench():
    sys.path.insert(0, ROOT)
    import importlib

    return importlib.import_module("bench")


def test_serving_leg_refuses_a_non_persistent_decoder():
    # the serving leg must measure the batched kernel; a decoder left on the
    # plain backend (as after the plain timing) is refused before any GPU work
    import sogrand_amd as sg

    bench = _bench()

    class FakeDecoder:
        class params:
            backend = sg.BACKEND_PLAIN

    import pytest

    with pytest.raises(ValueError):
        bench.serving_capacity(FakeDecoder(), 1)
    with pytest.raises(ValueError):
        bench.serve_load(FakeDecoder(), 1000.0, 0.5, 1.0, 1)


def test_traffic_record_matches_the_sources_or_is_refused():
    # the committed counter record carries the hash of the sources it was
    # measured on; bench.py reports its traffic only for that build
    import json

    bench = _bench()
    path = os.path.join(ROOT, "profiles", "persistent_100k_cw_2dB_traffic.json")
    rec = json.load(open(path))
    val, note = bench.measured_traffic(rec["n_cw"], rec["ebn0_db"])
    if rec.get("source_sha256") == bench.kernel_source_hash():
        assert val == round(rec["traffic_bytes_per_launch"])
    else:
        assert val is None and "stale" in note
    assert bench.measured_traffic(rec["n_cw"] + 1, rec["ebn0_db"])[0] is None


def test_cpu_baseline_uses_every_allowed_core():
    bench = _bench()
    assert bench.host_cores() == len(os.sched_getaffinity(0))


def test_source_hash_does_not_depend_on_where_the_tree_sits():
    # a copy of the tree elsewhere (as on another machine) hashes the same
    import shutil
    import tempfile

    bench = _bench()
    here = bench.kernel_source_hash()
    with tempfile.TemporaryDirectory() as d:
        dst = os.path.join(d, "copy")
        os.makedirs(dst)
        shutil.copy(os.path.join(ROOT, "bench.py"), dst)
        shutil.copytree(os.path.join(ROOT, "sogrand_amd"), os.path.join(dst, "sogrand_amd"),
                        ignore=shutil.ignore_patterns("_obj", "*.so", "cubic_decoder"))
        r = subprocess.run([sys.executable, "-c",
                            "import bench; print(bench.kernel_source_hash())"],
                           cwd=dst, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout.strip() == here
