"""Host-side checks that need no GPU: the C ABI library, the generator, the CLI's
argument handling, and the reference's data formats."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def _sg():
    import sogrand_amd as sg

    return sg


def test_library_exports_every_declared_symbol(lib):
    hdr = open(os.path.join(ROOT, "include", "sogrand.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(sogrand_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    import sogrand_amd as sg

    assert names == set(sg.EXPORTED)


def test_sigma_closed_form():
    sg = _sg()
    for eb in (0.5, 1.0, 2.0, 3.5):
        # sigma^2 = 1 / (2 R 10^(EbN0/10)), R = 1/8 (channel_sim.c:34-39)
        assert abs(sg.sigma_for_ebn0(eb) ** 2 - 1.0 / (0.25 * 10 ** (eb / 10))) < 1e-12
    assert abs(sg.sigma_for_ebn0(2.0) - 1.588656) < 1e-6  # what channel_sim prints at 2 dB


def test_generator_chunking_and_determinism():
    sg = _sg()
    a = sg.gen_llr(123, 20, ebn0_db=2.0)
    b = np.concatenate([sg.gen_llr(123, 7, ebn0_db=2.0), sg.gen_llr(123, 13, ebn0_db=2.0,
                                                                      cw_offset=7)])
    assert np.array_equal(a, b)
    assert not np.array_equal(a, sg.gen_llr(124, 20, ebn0_db=2.0))


def test_generator_statistics():
    # LLR = 2y/sigma^2, y = (+-1) + sigma z: per code bit, (LLR * sigma^2 / 2 - x) / sigma ~ N(0, 1)
    sg = _sg()
    n = 200
    eb = 2.0
    s = sg.sigma_for_ebn0(eb)
    llr = sg.gen_llr(99, n, ebn0_db=eb)
    x = 1.0 - 2.0 * np.unpackbits(sg.encode_cubes(sg.gen_info_bits(99, n)), axis=1)
    z = ((llr * s * s / 2.0) - x) / s
    z = z.ravel()
    m = z.size
    assert abs(z.mean()) < 5 / np.sqrt(m)
    assert abs(z.var() - 1.0) < 5 * np.sqrt(2.0 / m)
    # tails: fraction beyond 3 sigma
    assert abs((np.abs(z) > 3).mean() - 0.0026998) < 5 * np.sqrt(0.0027 / m)
    # info bits are fair coins
    bits = np.unpackbits(sg.gen_info_bits(5, 500))
    assert abs(bits.mean() - 0.5) < 5 * 0.5 / np.sqrt(bits.size)


def test_encoder_matches_reference_encoder(tmp_path):
    sg = _sg()
    info = sg.gen_info_bits(2024, 30)
    code = sg.encode_cubes(info)
    ref = os.path.join(ROOT, "oracle", "_ref", "cubic_encoder")
    if os.path.exists(ref):
        fi, fo = tmp_path / "i.bin", tmp_path / "o.bin"
        info.tofile(fi)
        subprocess.run([ref, str(fi), str(fo)], check=True, capture_output=True)
        assert (np.fromfile(fo, np.uint8).reshape(30, 512) == code).all()
    # systematic: the 8x8x8 info sub-cube carries the info bits (b = a0 + 8 a1 + 64 a2)
    bits = np.unpackbits(code, axis=1).reshape(30, 16, 16, 16)  # [a2, a1, a0]
    sub = bits[:, :8, :8, :8].reshape(30, 512)
    assert (np.packbits(sub, axis=1) == info).all()


def test_encoded_cubes_are_product_codewords(oracle):
    sg = _sg()
    code = np.unpackbits(sg.encode_cubes(sg.gen_info_bits(3, 10)), axis=1)
    for c in code:
        assert oracle.is_codeword(1.0 - 2.0 * c)
        d = 1.0 - 2.0 * c
        d[1234] *= -1
        assert not oracle.is_codeword(d)


def test_cli_arguments_and_empty_input(tmp_path):
    sg = _sg()
    if not os.path.exists(sg.CLI_PATH):
        pytest.skip("CLI not built")
    r = subprocess.run([sg.CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stderr
    empty = tmp_path / "e.bin"
    empty.write_bytes(b"\0" * 100)  # less than one codeword
    r = subprocess.run([sg.CLI_PATH, str(empty), str(tmp_path / "o.bin")], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "No valid blocks" in r.stderr
    r = subprocess.run([sg.CLI_PATH, str(tmp_path / "missing.bin"), str(tmp_path / "o.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 1


def test_decoder_without_gpu_fails_loudly():
    sg = _sg()
    from sogrand_amd import hip

    if hip.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(sg.SograndError) as e:
        sg.Decoder(0)
    assert e.value.code == sg.ENODEV


def test_lanegroup_backend_parameter_checks():
    # the lane-group backend has the final decoder's FP64 arithmetic only: the
    # parameters are refused before any device is touched (so this runs without
    # a GPU), while FP64 with the final variant (any epsilon) passes the checks
    sg = _sg()
    from sogrand_amd import hip

    if hip.device_count() > 0:
        pytest.skip("a GPU is present (the GPU tests cover the accepted cases)")
    for kw in ({"precision": sg.FP32_MIXED}, {"variant": sg.VARIANT_LEGACY}):
        with pytest.raises(sg.SograndError) as e:
            sg.Decoder(0, backend=sg.BACKEND_LANEGROUP, **kw)
        assert e.value.code == sg.EINVAL, kw
    with pytest.raises(sg.SograndError) as e:
        sg.Decoder(0, backend=4)
    assert e.value.code == sg.EINVAL
    for kw in ({}, {"epsilon": 2.220446049250313e-16}, {"Tmax": 1000}, {"L": 1}):
        with pytest.raises(sg.SograndError) as e:
            sg.Decoder(0, backend=sg.BACKEND_LANEGROUP, **kw)
        assert e.value.code == sg.ENODEV, kw  # past the checks, then no device
    for kw in ({}, {"precision": sg.FP32_MIXED}, {"variant": sg.VARIANT_LEGACY}):
        with pytest.raises(sg.SograndError) as e:
            sg.Decoder(0, backend=sg.BACKEND_AUTO, **kw)
        assert e.value.code == sg.ENODEV, kw
