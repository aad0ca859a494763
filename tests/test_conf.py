"""The stopping rule's confidence test without the division
(sogrand_amd/csrc/sg_conf.h), built for the host from the same source the
kernels compile, against getLConf's own expression
(SOGRAND_C/cubic_decoder_final.c:767-774, then `conf > thres` at :634): the
same decision on inputs aimed at the rounding boundary, at exact ties, at the
epsilon cut and at non-finite and negative masses."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "sogrand_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_conf_above_is_the_division(tmp_path):
    exe = str(tmp_path / "conf_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
                    os.path.join(ROOT, "tests", "conf_check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe, "4000000"], check=True, capture_output=True, text=True).stdout
    name, n, bad, divisions = out.split()
    assert name == "conf" and int(n) == 4000000
    assert int(bad) == 0
    # the fallback is exercised (ties, NaN, thresholds and epsilons outside the
    # fast test's range) but most decisions take the fast test
    assert 0 < int(divisions) < int(n) // 2
