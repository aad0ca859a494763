"""The single-precision variant (SOGRAND_FP32_MIXED) against the FP64 path.

FP32: reliabilities sorted as floats, path-metric increments and their sums in
single precision, exp(-pm) of each guess in single precision. FP64: PM of the
hard decision, the 1 - found-mass chain (P_notGuess), list weights, confidence,
soft output. The variant is not bit-exact with the reference; the bar:
- both backends give identical results, independent of batching (same line
  decoder, deterministic);
- decoded info bits: at Eb/N0 >= 2 dB no codeword differs from the FP64 path
  on a 2,000-codeword set; guess counts within 0.1 % in total;
- component soft outputs: within 1e-4 relative (1e-4 floor) of the oracle on
  lines whose found mass has not cancelled (P_notGuess >= 1e-3 at the last
  list insertion); where it has cancelled, a single-precision metric cannot
  fix 1 - found-mass (each term carries ~1e-7 relative error), which is the
  variant's stated limitation, measured and reported by bench.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP32_SOFT_TOL = 1e-4


def _sg():
    import sogrand_amd as sg

    return sg


def _say(*a):
    print(*a, flush=True)


def test_fp32_backends_identical_and_batch_independent(decoder):
    sg = _sg()
    snr = np.random.default_rng(8).uniform(0.5, 3.5, 400)
    llr = sg.gen_llr(4321, 400, ebn0_db=snr)
    out = []
    for backend in (sg.BACKEND_PLAIN, sg.BACKEND_PERSISTENT):
        decoder.set_params(backend=backend, precision=sg.FP32_MIXED)
        out.append(decoder.decode(llr, soft=True))
    a, b = out
    assert (a.bits == b.bits).all() and (a.stats == b.stats).all() and (a.soft == b.soft).all()
    for idx in ([5], [0, 199, 399], list(range(40, 104))):
        r = decoder.decode(llr[idx], soft=True)
        assert (r.bits == b.bits[idx]).all() and (r.stats == b.stats[idx]).all()
        assert (r.soft == b.soft[idx]).all()
    decoder.set_params(precision=sg.FP64)


@pytest.mark.parametrize("snr", [2.0, 2.5])
def test_fp32_bits_match_fp64(decoder, snr):
    sg = _sg()
    llr = sg.gen_llr(77, 2000, ebn0_db=snr)
    decoder.set_params(backend=sg.BACKEND_PERSISTENT, precision=sg.FP64)
    r64 = decoder.decode(llr)
    decoder.set_params(precision=sg.FP32_MIXED)
    r32 = decoder.decode(llr)
    decoder.set_params(precision=sg.FP64)
    diff_cw = int((r32.bits != r64.bits).any(axis=1).sum())
    ng64 = r64.stats["NG"].astype(np.float64).sum()
    ng_rel = abs(r32.stats["NG"].astype(np.float64).sum() - ng64) / ng64
    _say(f"{snr} dB: codewords with different bits {diff_cw}, total NG relative difference "
         f"{ng_rel:.2e}, half-iteration counts equal on "
         f"{(r32.stats['half_iters'] == r64.stats['half_iters']).mean():.4f}")
    assert diff_cw == 0
    assert ng_rel < 1e-3


def test_fp32_lines_close_to_oracle(decoder, oracle):
    sg = _sg()
    rng = np.random.default_rng(12)
    x = rng.normal(0.8, 1.5, (4096, 16)) * rng.choice([1.0, -1.0], (4096, 16))
    decoder.set_params(precision=sg.FP32_MIXED)
    la, le, ng = decoder.bitso(x)
    decoder.set_params(precision=sg.FP64)
    same_t = 0
    checked = 0
    worst = 0.0
    for i in range(len(x)):
        la2, _, tr = oracle.siso(x[i])
        same_t += int(ng[i] == tr.T)
        if tr.n_list > 0 and 256.0 * abs(tr.pnl) >= 1e-3:
            err = np.max(np.abs(la[i] - la2) / np.maximum(np.abs(la2), 1.0))
            worst = max(worst, float(err))
            checked += 1
            assert err <= FP32_SOFT_TOL, (i, err)
    _say(f"fp32 lines: N_guess equal on {same_t / len(x):.4f}; soft checked on {checked} lines, "
         f"worst relative error {worst:.2e}")
    assert same_t >= 0.99 * len(x)
    assert checked >= 0.5 * len(x)
