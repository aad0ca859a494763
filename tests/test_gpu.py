"""Parity of the gfx950 decoder (through the C ABI) with the reference and the oracle.

The bar, on every line and every codeword tested:
- decoded bits, half-iteration counts, guess counts NG and NG_p and the
  convergence flag: bit-exact;
- soft outputs (component L_APP / L_E and final L_APP): within 1e-9 relative
  (1e-9 absolute floor, SOFT_RTOL in conftest.py).
The decoder's exp and log give the C library's results bit for bit (checked on
the CPU by tests/log_check.cpp), and it forms every other double with the
reference's operations in the reference's order, so nothing here is exempted:
not even the codewords on which the reference's own FMA build disagrees with
its flow build.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, golden, soft_close, soft_err, stable_codewords

pytestmark = pytest.mark.gpu

SNRS = ["1.0", "1.5", "2.0", "2.5", "3.0"]


def _sg():
    import sogrand_amd as sg

    return sg


def _manifest():
    return json.load(open(os.path.join(GOLDEN, "manifest.json")))


def _say(*a):
    print(*a, flush=True)


def test_lines_match_reference(decoder):
    g = golden("lines.npz")
    _say("lines:", g["llr"].shape)
    la, le, ng = decoder.bitso(g["llr"])
    assert (ng == g["N_guess"]).all(), np.flatnonzero(ng != g["N_guess"])[:10]
    assert soft_close(la, g["L_APP"]), soft_err(la, g["L_APP"])
    assert soft_close(le, g["L_E"]), soft_err(le, g["L_E"])
    _say("  lines bit-identical to the reference:",
         int((la == g["L_APP"]).all(axis=1).sum()), "of", len(la))


def test_lines_random_match_oracle(decoder, oracle):
    rng = np.random.default_rng(11)
    x = rng.normal(0.8, 1.5, (8192, 16)) * rng.choice([1.0, -1.0], (8192, 16))
    x[:500] = np.round(x[:500])  # ties
    la, le, ng = decoder.bitso(x)
    for i in range(len(x)):
        la2, le2, tr = oracle.siso(x[i])
        assert ng[i] == tr.T, i
        assert soft_close(la[i], la2), (i, soft_err(la[i], la2))
        assert soft_close(le[i], le2), (i, soft_err(le[i], le2))


@pytest.mark.parametrize("backend", [0, 1, 2], ids=["plain", "persistent", "lanegroup"])
@pytest.mark.parametrize("snr", SNRS)
def test_cubes_match_reference_binary(decoder, oracle, backend, snr):
    sg = _sg()
    m = _manifest()["cubes"][snr]
    g = golden(f"cubes_{snr}.npz")
    n = g["bits"].shape[0]  # the generator is counter-based: any prefix of the seed's stream
    _say("cubes", snr, "backend", backend, "n", n)
    decoder.set_params(backend=backend)
    llr = sg.gen_llr(m["seed"], n, ebn0_db=float(snr))
    r = decoder.decode(llr, soft=True)
    assert (r.bits == g["bits"]).all()
    for f in ("half_iters", "NG", "NG_p"):
        bad = np.flatnonzero(r.stats[f] != g[f])
        assert len(bad) == 0, (f, bad[:10])
    ref = np.zeros(n, sg.CW_STATS_DTYPE)
    fma = np.zeros(n, sg.CW_STATS_DTYPE)
    for f in ("half_iters", "NG", "NG_p"):
        ref[f] = g[f]
        fma[f] = g[f + "_fma"]
    unstable = ~stable_codewords(ref, fma, g["bits"], g["bits_fma"])
    _say("  exact on all", n, "codewords, including", int(unstable.sum()),
         "on which the reference's FMA build differs from its flow build")
    # final soft outputs: the reference's (recorded for the first codewords)
    ns = g["soft"].shape[0]
    assert soft_close(r.soft[:ns], g["soft"]), soft_err(r.soft[:ns], g["soft"])
    _say(f"  soft of {ns} codewords: within 1e-9 everywhere "
         f"(max {soft_err(r.soft[:ns], g['soft']):.2e})")


def test_lines_with_large_metrics_match_oracle(decoder, oracle):
    # magnitudes far above JacLog's cut at 50: the hard decision's weight is
    # exp(-0) = 1, so P_notGuess starts at 0 and goes negative (the search's
    # exact re-run), and hits have metrics above 512 (the library's
    # large-argument exp); plus lines mixing such magnitudes with ordinary ones
    rng = np.random.default_rng(77)
    n = 512
    mag = rng.uniform(60.0, 900.0, (n, 16))
    mag[n // 2:, :6] = rng.uniform(0.0, 4.0, (n // 2, 6))
    x = mag * rng.choice([1.0, -1.0], (n, 16))
    la, le, ng = decoder.bitso(x)
    for i in range(n):
        la2, le2, tr = oracle.siso(x[i])
        assert ng[i] == tr.T, i
        assert soft_close(la[i], la2), (i, soft_err(la[i], la2))
        assert soft_close(le[i], le2), (i, soft_err(le[i], le2))


@pytest.mark.parametrize("backend", [0, 1, 2], ids=["plain", "persistent", "lanegroup"])
def test_cubes_with_large_llrs_match_oracle(decoder, oracle, backend):
    # channel LLRs scaled far beyond the channel's range (metrics above 512,
    # negative P_notGuess chains), whole decodes
    sg = _sg()
    llr = sg.gen_llr(6060, 24, ebn0_db=1.0) * np.linspace(20.0, 400.0, 24)[:, None]
    decoder.set_params(backend=backend)
    r = decoder.decode(llr, soft=True)
    bits, st, soft = oracle.decode(llr, soft=True)
    assert (r.bits == bits).all()
    assert (r.stats == st).all()
    assert soft_close(r.soft, soft), soft_err(r.soft, soft)


@pytest.mark.parametrize("backend", [0, 1, 2], ids=["plain", "persistent", "lanegroup"])
def test_cubes_2000_match_oracle(decoder, oracle, backend):
    # the 2,000-codeword subset at 2 dB, checked against the oracle
    sg = _sg()
    llr = sg.gen_llr(4242, 2000, ebn0_db=2.0)
    decoder.set_params(backend=backend)
    r = decoder.decode(llr, soft=True)
    bits, st, soft = oracle.decode(llr, soft=True)
    assert (r.bits == bits).all()
    for f in ("half_iters", "NG", "NG_p", "converged"):
        assert (r.stats[f] == st[f]).all(), f
    assert soft_close(r.soft, soft), soft_err(r.soft, soft)


@pytest.mark.parametrize("backend", [0, 1, 2], ids=["plain", "persistent", "lanegroup"])
def test_cubes_with_ties_match_oracle(decoder, oracle, backend):
    # channel LLRs on a grid of 1/4: most lines start with equal magnitudes, so
    # the set-up's tie detection and the quicksort replay run in every kernel
    sg = _sg()
    llr = np.round(sg.gen_llr(2024, 64, ebn0_db=2.5) * 4.0) / 4.0
    lines = np.abs(llr.reshape(-1, 16))
    ties = np.array([len(np.unique(v)) < 16 for v in lines])
    assert ties.mean() > 0.9
    decoder.set_params(backend=backend)
    r = decoder.decode(llr, soft=True)
    bits, st, soft = oracle.decode(llr, soft=True)
    assert (r.bits == bits).all()
    assert (r.stats == st).all()
    assert soft_close(r.soft, soft), soft_err(r.soft, soft)


def test_backends_identical_and_batch_independent(decoder):
    sg = _sg()
    rng = np.random.default_rng(5)
    snr = rng.uniform(0.5, 3.5, 600)
    llr = sg.gen_llr(777, 600, ebn0_db=snr)
    out = []
    for backend in (0, 1, 2):
        decoder.set_params(backend=backend)
        out.append(decoder.decode(llr, soft=True))
    a = out[0]
    for b in out[1:]:
        assert (a.bits == b.bits).all() and (a.stats == b.stats).all()
        assert (a.soft == b.soft).all()
    # a codeword decodes the same alone, in a small batch and in the large one
    decoder.set_params(backend=1)
    for idx in ([17], [3, 400, 599], list(range(100, 164))):
        r = decoder.decode(llr[idx], soft=True)
        assert (r.bits == a.bits[idx]).all() and (r.stats == a.stats[idx]).all()
        assert (r.soft == a.soft[idx]).all()


@pytest.mark.parametrize("tmax", [1000, 100000])
def test_finite_tmax_matches_oracle(decoder, oracle, tmax):
    # the uneven configuration's budgets, per-codeword Eb/N0 over its whole
    # 0.5-3.5 dB range
    sg = _sg()
    ebn0 = np.random.default_rng(tmax).uniform(0.5, 3.5, 150)
    llr = sg.gen_llr(31337, 150, ebn0_db=ebn0)
    bits, st, soft = oracle.decode(llr, Tmax=tmax, soft=True)
    for backend in (0, 1, 2):
        decoder.set_params(backend=backend, Tmax=tmax)
        r = decoder.decode(llr, soft=True)
        assert (r.bits == bits).all()
        assert (r.stats == st).all()
        assert soft_close(r.soft, soft), soft_err(r.soft, soft)
    decoder.set_params(Tmax=sg.UINT64_MAX)
    if tmax >= 65536:  # at most 65536 guesses per line: a budget beyond it never binds
        r2 = decoder.decode(llr)
        assert (r2.bits == r.bits).all() and (r2.stats == r.stats).all()


@pytest.mark.parametrize("L", [1, 2])
def test_smaller_lists_match_oracle(decoder, oracle, L):
    # list sizes below the reference's L = 3 (the parameter is a runtime one):
    # the stopping rule fires at a list of L entries and the soft output sums L
    # entries; per-codeword Eb/N0 over 0.5-3.5 dB, both backends
    sg = _sg()
    ebn0 = np.random.default_rng(40 + L).uniform(0.5, 3.5, 96)
    llr = sg.gen_llr(9090 + L, 96, ebn0_db=ebn0)
    bits, st, soft = oracle.decode(llr, L=L, soft=True)
    try:
        for backend in (0, 1, 2):
            decoder.set_params(backend=backend, L=L)
            r = decoder.decode(llr, soft=True)
            assert (r.bits == bits).all(), backend
            assert (r.stats == st).all(), backend
            assert soft_close(r.soft, soft), soft_err(r.soft, soft)
    finally:
        decoder.set_params(L=3)


@pytest.mark.parametrize("shape", [
    {"wave_slots": 1, "lane_batch": 8},  # what the autotuner picks on the uneven mix
    {"wave_slots": 3, "lane_batch": 16, "step_len": 256},
    {"wave_slots": 4, "lane_batch": 1, "step_len": 16, "waves_per_cu": 4},
], ids=["tuned", "s3", "s4"])
def test_launch_shapes_match_oracle(decoder, oracle, shape):
    # the persistent kernel's shape parameters change only speed: per-codeword
    # Eb/N0 over the uneven configuration's 0.5-3.5 dB range
    sg = _sg()
    ebn0 = np.random.default_rng(8).uniform(0.5, 3.5, 160)
    llr = sg.gen_llr(8080, 160, ebn0_db=ebn0)
    bits, st, soft = oracle.decode(llr, soft=True)
    decoder.set_params(backend=1, **shape)
    try:
        r = decoder.decode(llr, soft=True)
    finally:
        decoder.set_params(wave_slots=0, lane_batch=0, step_len=0, waves_per_cu=0)
    assert (r.bits == bits).all()
    assert (r.stats == st).all()
    assert soft_close(r.soft, soft), soft_err(r.soft, soft)


def test_autotune_confirms_its_pick_on_the_whole_workload(decoder):
    # the tuner checks that every shape it tries decodes identically; given the
    # whole workload it times its pick against the default shape there and keeps
    # the faster, leaving the decoder on the shape it returns
    sg = _sg()
    from sogrand_amd import hip
    from sogrand_amd.autotune import DEFAULT, autotune

    n = 4096
    ebn0 = np.random.default_rng(12).uniform(0.5, 3.5, n)
    llr = sg.gen_llr(1212, n, ebn0_db=ebn0)
    d_llr = hip.DeviceArray.from_host(llr)
    d_sample = hip.DeviceArray.from_host(llr[:1024])
    backend = decoder.params.backend
    try:
        r = autotune(decoder, d_sample, 1024, rounds=1, repeats=1, confirm=(d_llr, n))
        got = {k: getattr(decoder.params, k) for k in DEFAULT}
    finally:
        decoder.set_params(backend=backend, wave_slots=0, lane_batch=0, step_len=0,
                           waves_per_cu=0)
        d_llr.free()
        d_sample.free()
    assert got == r["best"]
    if r["best"] != DEFAULT:
        assert r["confirm"]["n_cw"] == n
        assert r["confirm"]["tuned_ms"] < r["confirm"]["default_ms"]
    _say("autotune", r["best"], r.get("confirm"))


def test_generator_device_equals_host(decoder):
    sg = _sg()
    from sogrand_amd import hip

    sig = np.array([sg.sigma_for_ebn0(e) for e in np.linspace(1, 3, 37)])
    x = hip.DeviceArray((37, 4096))
    decoder.gen_llr_device(9, 1000, 37, sig, x)
    assert np.array_equal(x.to_host(), sg.gen_llr(9, 37, sigma=sig, cw_offset=1000))


def test_device_resident_buffers(decoder):
    sg = _sg()
    from sogrand_amd import hip

    n = 300
    decoder.set_params(backend=1)
    llr = sg.gen_llr(5, n, ebn0_db=2.0)
    d_llr = hip.DeviceArray.from_host(llr)
    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    decoder.decode(d_llr, out_bits=d_bits, out_stats=d_st)
    h = decoder.decode(llr)
    assert (d_bits.to_host() == h.bits).all() and (d_st.to_host() == h.stats).all()
    # at 2 dB the decoder recovers the transmitted information bits
    assert (h.bits == sg.gen_info_bits(5, n)).mean() > 0.999


def test_full_size_properties(decoder):
    # the plain configuration's size (100,000 codewords), checked through
    # properties that hold at any size: the two kernels decode identically, the
    # device generator matches the host generator (first codewords), every
    # codeword converges at 3 dB and the transmitted information comes back (no
    # errors for this seed; the measured BLER at 3 dB is 0 in 100,000)
    sg = _sg()
    from sogrand_amd import hip

    n = 100000
    seed = 424242
    d_llr = hip.DeviceArray((n, 4096))
    decoder.gen_llr_device(seed, 0, n, sg.sigma_for_ebn0(3.0), d_llr)
    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    _say("full size: persistent")
    decoder.set_params(backend=1)
    decoder.decode(d_llr, out_bits=d_bits, out_stats=d_st)
    bits, st = d_bits.to_host(), d_st.to_host()
    _say("full size: plain")
    decoder.set_params(backend=0)
    decoder.decode(d_llr, out_bits=d_bits, out_stats=d_st)
    assert (d_bits.to_host() == bits).all() and (d_st.to_host() == st).all()
    decoder.set_params(backend=1)
    assert (st["converged"] == 1).all()
    assert (bits == sg.gen_info_bits(seed, n)).all()
    head = hip.DeviceArray((3, 4096))
    hip.copy(head, d_llr, head.nbytes)
    assert np.array_equal(head.to_host(), sg.gen_llr(seed, 3, ebn0_db=3.0))
    d_llr.free()


def test_host_buffers_in_overlapped_chunks(decoder):
    # more than 65,536 codewords from host memory: three chunks alternate
    # between the two staging streams (one stage is reused), with soft outputs
    # and statistics; everything equals a single device-resident decode
    sg = _sg()
    from sogrand_amd import hip

    n = 66000
    decoder.set_params(backend=1)
    llr = sg.gen_llr(17, n, ebn0_db=3.0)
    _say("host chunks:", n, "codewords")
    h = decoder.decode(llr, soft=True)
    assert decoder.last_kernel_ms()[1] == 3
    d_llr = hip.DeviceArray.from_host(llr)
    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    d_soft = hip.DeviceArray((n, 4096))
    decoder.decode(d_llr, out_bits=d_bits, out_stats=d_st, out_soft=d_soft)
    assert (d_bits.to_host() == h.bits).all() and (d_st.to_host() == h.stats).all()
    assert (d_soft.to_host() == h.soft).all()
    d_llr.free()
    d_soft.free()


def _same_or_close(a, b) -> bool:
    """Equal where the reference's value is not finite (NaN to NaN, inf to the
    same inf), within SOFT_RTOL elsewhere."""
    a = np.asarray(a)
    b = np.asarray(b)
    fin = np.isfinite(b)
    same_nonfinite = np.array_equal(a[~fin], b[~fin], equal_nan=True)
    return same_nonfinite and soft_close(a[fin], b[fin])


def test_nonfinite_lines_match_reference(decoder):
    # lines with +-inf entries, with a NaN, all infinite (recorded from the
    # reference by tests/golden/make_golden_nonfinite.py)
    g = golden("nonfinite.npz")
    decoder.set_params(backend=1)
    la, le, ng = decoder.bitso(g["lines"])
    assert (ng == g["lines_N_guess"]).all(), np.flatnonzero(ng != g["lines_N_guess"])[:10]
    assert _same_or_close(la, g["lines_L_APP"])
    assert _same_or_close(le, g["lines_L_E"])


@pytest.mark.parametrize("backend", [0, 1, 2], ids=["plain", "persistent", "lanegroup"])
def test_nonfinite_cubes_match_reference(decoder, backend):
    # cubes with a few +-inf entries (which the reference's channel can emit)
    # and a NaN: the reference runs all 30 iterations on them, as the NaN of
    # inf - inf spreads; bits, counts and final L_APP as it computes them
    sys.path.insert(0, GOLDEN)
    from make_golden_nonfinite import cube_llr

    g = golden("nonfinite.npz")
    decoder.set_params(backend=backend)
    for snr in ("1.5", "2.0"):
        llr = cube_llr(snr)
        r = decoder.decode(llr, soft=True)
        assert (r.bits == g[f"bits_{snr}"]).all(), snr
        for f in ("half_iters", "NG", "NG_p"):
            assert (r.stats[f] == g[f"{f}_{snr}"]).all(), (snr, f)
        assert _same_or_close(r.soft, g[f"soft_{snr}"]), snr


def test_empty_batch(decoder):
    r = decoder.decode(np.zeros((0, 4096)))
    assert r.bits.shape == (0, 64)


def test_cli_drop_in(tmp_path, oracle):
    sg = _sg()
    llr = sg.gen_llr(8, 12, ebn0_db=2.0)
    fin, fout, fst = tmp_path / "llr.bin", tmp_path / "dec.bin", tmp_path / "st.bin"
    # a trailing partial codeword is ignored, as the reference does
    with open(fin, "wb") as f:
        f.write(llr.astype("<f8").tobytes())
        f.write(b"\0" * 1000)
    r = subprocess.run([sg.CLI_PATH, str(fin), str(fout), "--stats", str(fst)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    bits, st, _ = oracle.decode(llr)
    assert open(fout, "rb").read() == bits.tobytes()
    lines = r.stdout.splitlines()
    assert lines[0] == "Decoding complete. 12 block(s) decoded."
    assert lines[1] == f"Average iterations per block: {st['half_iters'].sum() / 2 / 12:.2f}"
    assert lines[2] == f"Average NG per block: {st['NG'].sum() / 12:.2f}"
    assert lines[3] == f"Average NG per info bit: {st['NG'].sum() / (12 * 512):.2f}"
    assert lines[4] == f"Average NG_p per block: {st['NG_p'].sum() / 12:.2f}"
    assert np.fromfile(fst, sg.CW_STATS_DTYPE).tolist() == st.tolist()


def test_cli_legacy_variant(tmp_path):
    # `--variant legacy` stands in for the reference's cubic_decoder.c binary:
    # its extra first summary line, and the fixture's bits and counts
    sg = _sg()
    g = golden("variant_legacy.npz")
    m = json.load(open(os.path.join(GOLDEN, "variant_legacy.json")))["cubes"]["3.0"]
    llr = sg.gen_llr(m["seed"], m["n"], ebn0_db=3.0)
    fin, fout = tmp_path / "llr.bin", tmp_path / "dec.bin"
    llr.astype("<f8").tofile(fin)
    r = subprocess.run([sg.CLI_PATH, str(fin), str(fout), "--variant", "legacy"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(np.fromfile(fout, np.uint8).reshape(-1, 64), g["bits_3.0"])
    n = m["n"]
    lines = r.stdout.splitlines()
    assert lines[0] == f"Decoding {fin} to {fout} using cubic code (n=16, k=8)..."
    assert lines[1] == f"Decoding complete. {n} block(s) decoded."
    assert lines[2] == f"Average iterations per block: {g['half_iters_3.0'].sum() / 2 / n:.2f}"
    assert lines[3] == f"Average NG per block: {g['NG_3.0'].sum() / n:.2f}"
    assert lines[5] == f"Average NG_p per block: {g['NG_p_3.0'].sum() / n:.2f}"


def test_cli_options_match_library(tmp_path):
    # every kernel and the FP32 variant through the drop-in's flags: bits,
    # statistics and soft outputs equal the library's for the same settings;
    # a setting the kernel does not support exits 1 with the library's message
    sg = _sg()
    n = 20
    llr = sg.gen_llr(33, n, ebn0_db=np.linspace(1.0, 3.0, n))
    fin = tmp_path / "llr.bin"
    llr.astype("<f8").tofile(fin)
    cases = [(["--backend", "plain"], dict(backend=0)),
             (["--backend", "persistent"], dict(backend=1)),
             (["--backend", "lanegroup"], dict(backend=2)),
             (["--precision", "fp32"], dict(backend=1, precision=1)),
             (["--backend", "lanegroup", "--tmax", "300"], dict(backend=2, Tmax=300)),
             ([], dict(backend=3))]
    for flags, params in cases:
        _say("cli", *flags)
        fout, fst, fsoft = tmp_path / "dec.bin", tmp_path / "st.bin", tmp_path / "soft.bin"
        r = subprocess.run([sg.CLI_PATH, str(fin), str(fout), "--stats", str(fst),
                            "--soft-out", str(fsoft), *flags],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (flags, r.stderr)
        with sg.Decoder(0, **params) as dec:  # the CLI's defaults otherwise
            ref = dec.decode(llr, soft=True)
        assert np.array_equal(np.fromfile(fout, np.uint8).reshape(n, 64), ref.bits), flags
        assert np.fromfile(fst, sg.CW_STATS_DTYPE).tolist() == ref.stats.tolist(), flags
        assert np.array_equal(np.fromfile(fsoft, "<f8").reshape(n, 4096), ref.soft), flags
    r = subprocess.run([sg.CLI_PATH, str(fin), str(tmp_path / "x.bin"), "--backend", "lanegroup",
                        "--precision", "fp32"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and r.stderr.startswith("Error: "), r.stderr


def test_auto_backend_matches_explicit_kernels(decoder):
    # the auto backend picks plain below one cube per resident persistent wave
    # (3,072 on MI355X), persistent with one slot per wave up to twice that and
    # two slots above: every pick decodes alike
    sg = _sg()
    from sogrand_amd import hip

    for n in (300, 4000, 7000):
        _say("auto", n)
        llr = sg.gen_llr(515 + n, n, ebn0_db=np.linspace(1.5, 3.0, n))
        d = hip.DeviceArray.from_host(llr)
        bits = hip.DeviceArray((n, 64), np.uint8)
        st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
        out = []
        for backend in (3, 0, 1):
            decoder.set_params(backend=backend)
            decoder.decode(d, out_bits=bits, out_stats=st)
            out.append((bits.to_host().copy(), st.to_host().copy()))
        for b, s in out[1:]:
            assert (b == out[0][0]).all() and (s == out[0][1]).all(), n
    decoder.set_params(backend=1)


@pytest.mark.parametrize("backend", [0, 1, 2], ids=["plain", "persistent", "lanegroup"])
def test_graph_replays_match_eager(decoder, backend):
    # a captured launch sequence replayed several times, with eager decodes on
    # another context and host synchronisations in between; every replay must
    # equal the eager result
    sg = _sg()
    from sogrand_amd import hip

    n = 300
    llr = sg.gen_llr(91, n, ebn0_db=np.linspace(1.0, 3.0, n))
    other = sg.gen_llr(92, 150, ebn0_db=2.0)
    decoder.set_params(backend=backend, use_graph=0)
    ref = decoder.decode(llr, soft=True)
    d_llr = hip.DeviceArray.from_host(llr)
    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    d_soft = hip.DeviceArray((n, 4096))
    with sg.Decoder(0, backend=backend) as eager:
        decoder.set_params(use_graph=1)
        for rep in range(4):
            d_bits.zero()
            d_st.zero()
            decoder.decode(d_llr, out_bits=d_bits, out_stats=d_st, out_soft=d_soft)
            assert decoder.graph_replays() == rep
            assert (d_bits.to_host() == ref.bits).all(), rep
            assert (d_st.to_host() == ref.stats).all(), rep
            assert (d_soft.to_host() == ref.soft).all(), rep
            eager.decode(other)
            hip.synchronize()
    decoder.set_params(use_graph=0)


def test_flow_script_end_to_end(tmp_path):
    # the reference pipeline's five steps with this decoder, at 3.5 dB
    env = dict(os.environ, SNR_DB="3.5", N_CW="24", SEED="5")
    r = subprocess.run(["bash", os.path.join(ROOT, "sogrand_amd", "cubic_flow.sh")], cwd=tmp_path,
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    assert "Decoding complete. 24 block(s) decoded." in out
    assert "SUCCESS: Files are identical." in out
    assert (tmp_path / "decoded_cubic.bin").stat().st_size == 24 * 64


@pytest.mark.parametrize("backend", [0, 1, 2], ids=["plain", "persistent", "lanegroup"])
def test_epsilon_variant_matches_reference_variant(decoder, oracle, backend):
    # the reference's DBL_EPSILON variant (cubic_decoder_updated.c): lines and
    # whole decodes to the same bar as the final decoder
    sg = _sg()
    from oracle import DBL_EPSILON

    g = golden("variant_eps.npz")
    meta = json.load(open(os.path.join(GOLDEN, "variant_eps.json")))
    decoder.set_params(backend=backend, epsilon=DBL_EPSILON)
    try:
        la, le, ng = decoder.bitso(g["llr"])
        assert (ng == g["N_guess"]).all()
        assert soft_close(la, g["L_APP"]), soft_err(la, g["L_APP"])
        assert soft_close(le, g["L_E"]), soft_err(le, g["L_E"])
        for tag, m in meta["cubes"].items():
            llr = sg.gen_llr(m["seed"], m["n"], ebn0_db=float(tag))
            r = decoder.decode(llr, soft=True)
            assert (r.bits == g[f"bits_{tag}"]).all()
            for f in ("half_iters", "NG", "NG_p"):
                assert (r.stats[f] == g[f"{f}_{tag}"]).all(), (tag, f)
            ref = g[f"soft_{tag}"]
            assert soft_close(r.soft[:ref.shape[0]], ref), soft_err(r.soft[:ref.shape[0]], ref)
    finally:
        decoder.set_params(epsilon=1e-40)


@pytest.mark.parametrize("backend", [0, 1], ids=["plain", "persistent"])
def test_legacy_variant_matches_reference_legacy(decoder, oracle, backend):
    # the reference's earlier decoder (cubic_decoder.c), recorded by
    # tests/golden/make_golden_variant.py legacy. It has no found-mass chain, so
    # every count is fixed by its arithmetic: guesses, bits, half-iterations, NG
    # and NG_p exact everywhere (its two builds agree on all of them). Soft
    # outputs are logs of sums of positive terms: a component decode within
    # 1e-12 absolute, a whole decode's final L_APP within 1e-10 relative
    # (1 absolute floor); both are about 1e-14 on MI355X.
    sg = _sg()
    g = golden("variant_legacy.npz")
    meta = json.load(open(os.path.join(GOLDEN, "variant_legacy.json")))
    decoder.set_params(backend=backend, variant=sg.VARIANT_LEGACY)
    try:
        la, le, ng = decoder.bitso(g["llr"])
        assert (ng == g["N_guess"]).all()
        assert np.abs(la - g["L_APP"]).max() <= 1e-12
        assert np.abs(le - g["L_E"]).max() <= 1e-12
        # a guess budget: the last guess's weight is the P_neg left behind
        decoder.set_params(Tmax=64)
        la, le, ng = decoder.bitso(g["llr"])
        assert (ng == g["N_guess_t64"]).all()
        assert np.abs(la - g["L_APP_t64"]).max() <= 1e-12
        assert np.abs(le - g["L_E_t64"]).max() <= 1e-12
        llr = sg.gen_llr(7001, 8, ebn0_db=3.0)
        r = decoder.decode(llr)
        oracle.set_variant(1)
        try:
            bits, st, _ = oracle.decode(llr, Tmax=64)
        finally:
            oracle.set_variant(0)
        assert (r.bits == bits).all()
        for f in ("half_iters", "NG", "NG_p"):
            assert (r.stats[f] == st[f]).all(), f
        decoder.set_params(Tmax=sg.UINT64_MAX)
        for tag, m in meta["cubes"].items():
            _say(f"legacy {tag} dB, backend {backend}")
            llr = sg.gen_llr(m["seed"], m["n"], ebn0_db=float(tag))
            r = decoder.decode(llr, soft=True)
            assert (r.bits == g[f"bits_{tag}"]).all()
            for f in ("half_iters", "NG", "NG_p"):
                assert (r.stats[f] == g[f"{f}_{tag}"]).all(), f
            ref = g[f"soft_{tag}"]
            ns = ref.shape[0]
            assert (np.abs(r.soft[:ns] - ref) <= 1e-10 * np.maximum(1.0, np.abs(ref))).all()
        with pytest.raises(Exception):
            decoder.set_params(precision=sg.FP32_MIXED)  # the earlier decoder is FP64 only
        with pytest.raises(Exception):
            decoder.set_params(backend=sg.BACKEND_LANEGROUP)  # final variant only
    finally:
        decoder.set_params(variant=sg.VARIANT_FINAL, Tmax=sg.UINT64_MAX, precision=sg.FP64)
