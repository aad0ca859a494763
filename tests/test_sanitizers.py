"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the host C and C++
code, run on the CPU: the C oracle (oracle/sogrand_oracle.c, driven through its
whole API by tests/sanitize_driver.c) and the schedule builder the library's
contexts run at creation (sogrand_amd/csrc/sg_schedule.cpp, driven by
tests/pos_check.cpp). Any sanitizer report fails the run (-fno-sanitize-recover)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return r


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_driver")
    _run(["gcc", *SAN, "-std=c11", "-ffp-contract=off", "-pthread",
          "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "sanitize_driver.c"),
          os.path.join(ROOT, "oracle", "sogrand_oracle.c"), "-o", exe, "-lm"])
    r = _run([exe], timeout=600, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.stdout.strip() == "ok"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_schedule_builder_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "pos_check")
    csrc = os.path.join(ROOT, "sogrand_amd", "csrc")
    _run(["g++", *SAN, "-std=c++17", "-I", csrc, os.path.join(ROOT, "tests", "pos_check.cpp"),
          os.path.join(csrc, "sg_schedule.cpp"), "-o", exe])
    r = _run([exe], timeout=900)
    assert r.stdout.strip() == "ok"
