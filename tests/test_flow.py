"""The file-level flow tools (sogrand_amd.flow) on the CPU: formats, the encoder
step against the library encoder, the channel's sigma against its closed form,
and the comparator's counts."""
import os
import subprocess
import sys

import numpy as np

from conftest import ROOT


def _flow(*args, cwd):
    env = dict(os.environ, PYTHONPATH=ROOT)
    return subprocess.run([sys.executable, "-m", "sogrand_amd.flow", *args], cwd=cwd, env=env,
                          capture_output=True, text=True, timeout=120)


def test_flow_steps_and_formats(tmp_path):
    import sogrand_amd as sg
    from sogrand_amd.flow import compare, sigma_for, wilson

    r = _flow("data", "--n-cw", "5", "--seed", "3", "orig.bin", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    orig = np.fromfile(tmp_path / "orig.bin", np.uint8)
    assert orig.size == 5 * 64 and (orig.reshape(5, 64) == sg.gen_info_bits(3, 5)).all()
    r = _flow("encode", "orig.bin", "enc.bin", cwd=tmp_path)
    assert r.returncode == 0 and "5 block(s) encoded" in r.stdout
    enc = np.fromfile(tmp_path / "enc.bin", np.uint8)
    assert (enc.reshape(5, 512) == sg.encode_cubes(orig.reshape(5, 64))).all()
    r = _flow("channel", "enc.bin", "llr.bin", "2.0", "3", "--seed", "3", cwd=tmp_path)
    assert r.returncode == 0 and "sigma = 1.588656" in r.stdout
    llr = np.fromfile(tmp_path / "llr.bin", "<f8")
    assert llr.size == 5 * 4096
    # hard decisions of the LLRs agree with the code bits on most positions
    hd = (llr <= 0).astype(np.uint8)
    assert (hd == np.unpackbits(enc)).mean() > 0.6
    assert abs(sigma_for(2.0, 3) - sg.sigma_for_ebn0(2.0)) < 1e-15
    assert _flow("channel", "enc.bin", "x.bin", "2.0", "4", cwd=tmp_path).returncode == 1
    # comparator: identical, one flipped bit, a short file
    c = compare(orig.tobytes(), orig.tobytes())
    assert c["error_bits"] == 0 and c["block_errors"] == 0 and c["total_bits"] == 5 * 512
    bad = orig.copy()
    bad[70] ^= 0x10
    c = compare(orig.tobytes(), bad.tobytes())
    assert c["error_bits"] == 1 and c["block_errors"] == 1
    c = compare(orig.tobytes(), orig[:100].tobytes())
    assert c["short"] and c["error_bits"] == 8 * (320 - 100) and c["block_errors"] == 4
    lo, hi = wilson(0, 1000)
    assert lo == 0.0 and 0.003 < hi < 0.004
    r = _flow("compare", "orig.bin", "orig.bin", cwd=tmp_path)
    assert "SUCCESS: Files are identical." in r.stdout and "Total bit errors:    0" in r.stdout
