"""The (16,8)^2 square code: the reference ships no square decoder (its channel
simulator accepts the square code, SOGRAND_C/channel_sim.c:26-29), so the square
path is pinned by what the cubic path, pinned to the reference itself, fixes:

- a square is the plane a2 = 0 of a cube whose other planes carry a codeword at
  overwhelming reliability; until a decode stops, every phase of the cube
  decode treats plane 0 exactly as the square decode treats the square (its
  lines along axes 0 and 1 are the square's rows and columns, the other lines
  stop at their first guess), so bits, half-iterations, NG - 240 per phase, NG_p
  and plane 0's soft output must be equal wherever both stop together (the
  square decoded to the transmitted codeword within two phases);
- closed forms: encoded squares are product codewords and the plane of the
  cubic encoder; noiseless squares stop after one phase with one guess per
  line; the square's noise sigma is channel_sim.c's for dimension 2.
The GPU square decoder is compared with the oracle's square decode (the same
per-codeword loop as its cube decode, over two axes) on every square.
"""
import numpy as np
import pytest

from conftest import soft_close, soft_err


def _sg():
    import sogrand_amd as sg

    return sg


def _embed(sg, info, llr):
    """Cubes whose plane a2 = 0 is the square's LLRs and whose other planes carry
    the cubic codeword of the same info plane at magnitude 1000."""
    cube_info = np.zeros((len(info), 64), np.uint8)
    cube_info[:, :8] = info
    code = np.unpackbits(sg.encode_cubes(cube_info), axis=1).astype(np.float64)
    cubes = 1000.0 * (1.0 - 2.0 * code)
    cubes[:, :256] = llr
    return cubes


def test_square_sigma_and_encoding(oracle):
    sg = _sg()
    # channel_sim.c:32-36 with R = (8/16)^2: Es/N0 = Eb/N0 + 10 log10(1/2)
    assert sg.sigma_for_ebn0(10.0 * np.log10(2.0), dim=2) == pytest.approx(1.0, abs=1e-15)
    assert sg.sigma_for_ebn0(2.0, dim=2) == pytest.approx(10 ** (-(2.0 - 10 * np.log10(2)) / 20))
    info = np.random.default_rng(1).integers(0, 256, (64, 8), dtype=np.uint8)
    code = sg.encode_squares(info)
    for c in code:
        assert oracle.is_square_codeword(1.0 - 2.0 * c)
    # systematic: info bit b = a0 + 8 a1 (MSB first) at position a0 + 16 a1
    bits = np.unpackbits(info, axis=1)
    pos = np.array([(b & 7) + 16 * (b >> 3) for b in range(64)])
    assert (code[:, pos] == bits).all()


def test_square_oracle_noiseless(oracle):
    sg = _sg()
    info = np.random.default_rng(2).integers(0, 256, (32, 8), dtype=np.uint8)
    llr = 100.0 * (1.0 - 2.0 * sg.encode_squares(info).astype(np.float64))
    bits, st, _ = oracle.decode_squares(llr)
    assert (bits == info).all()
    assert (st["half_iters"] == 1).all() and (st["converged"] == 1).all()
    assert (st["NG"] == 16).all() and (st["NG_p"] == 1).all()


def test_square_oracle_embeds_in_cube_decode(oracle):
    sg = _sg()
    info, llr = sg.gen_llr_square(11, 240, ebn0_db=np.linspace(2.0, 4.0, 240))
    bits, st, soft = oracle.decode_squares(llr, soft=True)
    ok = (st["half_iters"] <= 2) & (st["converged"] == 1) & (bits == info).all(1)
    idx = np.flatnonzero(ok)
    assert len(idx) > 100
    cb, cst, csoft = oracle.decode(_embed(sg, info[idx], llr[idx]), soft=True)
    assert (cb[:, :8] == bits[idx]).all() and (cb[:, 8:] == 0).all()
    assert (cst["half_iters"] == st["half_iters"][idx]).all()
    assert (cst["NG"] == st["NG"][idx] + 240 * st["half_iters"][idx]).all()
    assert (cst["NG_p"] == st["NG_p"][idx]).all()
    assert (csoft[:, :256] == soft[idx]).all()


@pytest.mark.gpu
def test_square_gpu_matches_oracle(decoder, oracle):
    sg = _sg()
    ebn0 = np.random.default_rng(5).uniform(0.5, 4.0, 800)
    info, llr = sg.gen_llr_square(21, 800, ebn0_db=ebn0)
    bits, st, soft = oracle.decode_squares(llr, soft=True)
    r = decoder.decode_square(llr, soft=True)
    assert (r.bits == bits).all()
    assert (r.stats == st).all()
    assert soft_close(r.soft, soft), soft_err(r.soft, soft)
    print("squares: BER", np.unpackbits(r.bits ^ info).mean(), "mean half-iterations",
          st["half_iters"].mean(), flush=True)


@pytest.mark.gpu
def test_square_gpu_embeds_in_gpu_cube_decode(decoder):
    sg = _sg()
    info, llr = sg.gen_llr_square(12, 320, ebn0_db=np.linspace(2.0, 4.5, 320))
    r = decoder.decode_square(llr, soft=True)
    ok = (r.stats["half_iters"] <= 2) & (r.stats["converged"] == 1) & (r.bits == info).all(1)
    idx = np.flatnonzero(ok)
    assert len(idx) > 150
    for backend in (0, 1):
        decoder.set_params(backend=backend)
        c = decoder.decode(_embed(sg, info[idx], llr[idx]), soft=True)
        assert (c.bits[:, :8] == r.bits[idx]).all() and (c.bits[:, 8:] == 0).all()
        assert (c.stats["half_iters"] == r.stats["half_iters"][idx]).all()
        assert (c.stats["NG"] == r.stats["NG"][idx] + 240 * r.stats["half_iters"][idx]).all()
        assert (c.stats["NG_p"] == r.stats["NG_p"][idx]).all()
        assert (c.soft[:, :256] == r.soft[idx]).all()


@pytest.mark.gpu
def test_square_gpu_budget_lists_and_device_buffers(decoder, oracle):
    sg = _sg()
    from sogrand_amd import hip

    ebn0 = np.random.default_rng(6).uniform(0.5, 3.0, 300)
    info, llr = sg.gen_llr_square(22, 300, ebn0_db=ebn0)
    try:
        for kw in ({"Tmax": 1000}, {"L": 1}, {"L": 2, "Tmax": 20}):
            bits, st, soft = oracle.decode_squares(llr, soft=True, **kw)
            decoder.set_params(**kw)
            r = decoder.decode_square(llr, soft=True)
            assert (r.bits == bits).all(), kw
            assert (r.stats == st).all(), kw
            assert soft_close(r.soft, soft), (kw, soft_err(r.soft, soft))
    finally:
        decoder.set_params(Tmax=sg.UINT64_MAX, L=3)
    h = decoder.decode_square(llr, soft=True)
    d_llr = hip.DeviceArray.from_host(llr)
    d_bits = hip.DeviceArray((300, 8), np.uint8)
    d_st = hip.DeviceArray((300,), sg.CW_STATS_DTYPE)
    d_soft = hip.DeviceArray((300, 256))
    try:
        decoder.decode_square(d_llr, out_bits=d_bits, out_stats=d_st, out_soft=d_soft)
        assert (d_bits.to_host() == h.bits).all()
        assert (d_st.to_host() == h.stats).all()
        assert (d_soft.to_host() == h.soft).all()
    finally:
        for a in (d_llr, d_bits, d_st, d_soft):
            a.free()
    # an empty batch is a no-op
    assert decoder.decode_square(np.zeros((0, 256))).bits.shape == (0, 8)


@pytest.mark.gpu
def test_square_cli(tmp_path, oracle):
    # the drop-in CLI with --dim 2: 256 LLRs in and 8 info bytes out per square
    import subprocess

    sg = _sg()
    info, llr = sg.gen_llr_square(23, 40, ebn0_db=2.0)
    fin, fout, fst = tmp_path / "llr.bin", tmp_path / "dec.bin", tmp_path / "st.bin"
    with open(fin, "wb") as f:
        f.write(llr.astype("<f8").tobytes())
        f.write(b"\0" * 100)  # a trailing partial codeword is ignored
    r = subprocess.run([sg.CLI_PATH, str(fin), str(fout), "--dim", "2", "--stats", str(fst)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    bits, st, _ = oracle.decode_squares(llr)
    assert open(fout, "rb").read() == bits.tobytes()
    lines = r.stdout.splitlines()
    assert lines[0] == "Decoding complete. 40 block(s) decoded."
    assert lines[3] == f"Average NG per info bit: {st['NG'].sum() / (40 * 64):.2f}"
    assert np.fromfile(fst, sg.CW_STATS_DTYPE).tolist() == st.tolist()
