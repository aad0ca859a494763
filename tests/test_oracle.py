"""The CPU oracle pinned against the reference's golden vectors and against closed forms.

Every function of oracle/sogrand_oracle.c is checked here against something it
could not agree with by accident: vectors recorded from the reference decoder
itself (tests/golden, made by make_golden.py), exhaustive enumeration, or a
textbook definition.
"""
import itertools
import json
import sys
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden

UINT64_MAX = (1 << 64) - 1

# H columns of the (16,8) CRC-0xeb code (bit j = row j), the known-answer table.
HCOL_KAT = [0x79, 0xF2, 0x33, 0x66, 0xCC, 0x4F, 0x9E, 0xEB, 1, 2, 4, 8, 16, 32, 64, 128]


def test_code_tables_match_reference(oracle):
    G, H, hcol = oracle.code_tables()
    g = golden("code_tables.npz")
    assert (G == g["G"]).all() and (H == g["H"]).all()
    assert list(hcol) == HCOL_KAT
    # G H^T = 0 over GF(2), G systematic, H = [P^T | I]
    assert not ((G.astype(int) @ H.T.astype(int)) % 2).any()
    assert (G[:, :8] == np.eye(8, dtype=np.uint8)).all()
    assert (H[:, 8:] == np.eye(8, dtype=np.uint8)).all()


def test_code_minimum_distance_and_parity():
    # brute force over all 255 nonzero messages: d_min = 5; the reference's
    # `even` flag (all generator rows even weight) is 0 for this code
    g = golden("code_tables.npz")["G"].astype(int)
    weights = []
    for m in range(1, 256):
        u = np.array([(m >> i) & 1 for i in range(8)])
        weights.append(int(((u @ g) % 2).sum()))
    assert min(weights) == 5
    assert any(int(r.sum()) % 2 for r in g)


def _key(ic, m):
    pos = [b for b in range(16) if (m >> b) & 1]
    w = len(pos)
    s = sum(p + 1 for p in pos)
    return (ic * w + s, w, tuple(pos))


@pytest.mark.parametrize("ic", [0, 1, 2, 5, 17, 54, 55, 60, 1000])
def test_schedule_is_the_defined_order(oracle, ic):
    # definition: logistic weight, then Hamming weight, then lexicographic
    # order of the ascending position tuple; every nonzero mask exactly once
    sched = oracle.schedule(ic)
    assert len(sched) == 65535
    assert len(set(sched.tolist())) == 65535 and 0 not in sched
    keys = [_key(ic, int(m)) for m in sched]
    assert keys == sorted(keys)


def test_schedule_matches_reference_probes(oracle):
    g = golden("schedule.npz")
    for ic in np.unique(g["ic"]):
        sel = g["ic"] == ic
        pos = {int(m): i for i, m in enumerate(oracle.schedule(int(ic)))}
        got = np.array([pos[int(m)] + 2 for m in g["mask"][sel]])  # T = 1 (hard decision) + 1-based position
        assert (got == g["T"][sel]).all(), f"IC {ic}"


def test_schedule_generic_beyond_ic55(oracle):
    base = oracle.schedule(55)
    for ic in (56, 80, 300, 40000):
        assert (oracle.schedule(ic) == base).all()
    assert not (oracle.schedule(54) == base).all()


def test_product_schedule_tables_equal_oracle(oracle, lib):
    import sogrand_amd as sg

    for ic in list(range(0, 58)) + [90, 12345]:
        t = sg.schedule_table(ic)
        assert t[65535] == 0
        assert (t[:65535] == oracle.schedule(ic)).all(), f"IC {ic}"


def test_wmin_formula_equals_integer_bounds():
    # The reference picks the first Hamming weight of a level with
    # w = max(1, ceil((t - sqrt(t^2 - 8 wt)) / 2)), t = 1 + 2(16 + IC)
    # (cubic_decoder_final.c:650-653). The oracle and the tables use the exact
    # integer rule: the smallest w whose largest logistic weight IC*w +
    # w(33-w)/2 reaches wt. Compare for every level the search can reach.
    def check(ic, wts):
        t = 1.0 + 2.0 * (16.0 + ic)
        w_fp = np.maximum(1, np.ceil((t - np.sqrt(t * t - 8.0 * wts)) / 2.0)).astype(np.int64)
        w = np.arange(1, 17)
        reach = ic * w[None, :] + (w * (33 - w) // 2)[None, :] >= wts[:, None]
        w_int = np.argmax(reach, axis=1) + 1
        assert (w_fp == w_int).all(), ic

    for ic in range(0, 2001):
        check(ic, np.arange(ic + 1, 16 * ic + 137, dtype=np.int64))
    rng = np.random.default_rng(3)
    for ic in rng.integers(2001, 1 << 24, 60):
        ic = int(ic)
        check(ic, np.unique(np.concatenate([np.arange(ic + 1, ic + 200),
                                            rng.integers(ic + 1, 16 * ic + 137, 20000)])))


def test_sort_ties_match_reference_quicksort(oracle):
    g = golden("quicksort.npz")
    for a, srt, perm in zip(g["a"], g["sorted"], g["perm"]):
        s, p = oracle.sort_with_perm(a)
        assert (s == srt).all() and (p == perm).all()


def test_sort_distinct_keys_is_argsort(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = rng.random(16)
        s, p = oracle.sort_with_perm(a)
        assert (p == np.argsort(a)).all() and (s == np.sort(a)).all()


def test_lines_match_reference_siso(oracle):
    g = golden("lines.npz")
    for i, x in enumerate(g["llr"]):
        la, le, tr = oracle.siso(x)
        assert tr.T == g["N_guess"][i], i
        # same libm, same operation order: equal to the last bit
        assert (la == g["L_APP"][i]).all() and (le == g["L_E"][i]).all(), i


def test_lines_match_reference_list_search(oracle):
    g = golden("lines.npz")
    for i, x in enumerate(g["llr"]):
        _, _, tr = oracle.siso(x)
        nl = int(g["curL"][i])
        assert tr.n_list == nl and tr.T == g["T"][i] and tr.pnl == g["pNL"][i], i
        for j in range(nl):
            assert tr.pm[j] == g["s_list"][i, j, 1]
            assert tr.t_hit[j] == g["s_list"][i, j, 2]
            assert tr.wt_hit[j] == g["s_list"][i, j, 0]
            cw = sum(int(b) << k for k, b in enumerate(g["chat_list"][i, j]))
            assert tr.cw[j] == cw


def test_lines_finite_tmax_match_reference(oracle):
    lines = golden("lines.npz")["llr"]
    recs = golden("lines_tmax.npz")["recs"]
    for r in recs:
        tmax, i, T, curL, pnl = int(r[0]), int(r[1]), int(r[2]), int(r[3]), r[4]
        _, _, tr = oracle.siso(lines[i], Tmax=tmax)
        assert (tr.T, tr.n_list) == (T, curL), (tmax, i)
        assert tr.pnl == pnl
        for j in range(curL):
            assert tr.pm[j] == r[5 + j] and tr.t_hit[j] == r[8 + j]


def test_early_termination_matches_reference(oracle):
    g = golden("early_term.npz")
    got = np.array([oracle.is_codeword(c.astype(np.float64)) for c in g["cubes"]], np.int8)
    assert (got == g["verdict"]).all()
    assert g["verdict"].any() and not g["verdict"].all()


def _manifest():
    p = os.path.join(GOLDEN, "manifest.json")
    if not os.path.exists(p):
        pytest.skip("manifest not recorded")
    return json.load(open(p))


@pytest.mark.parametrize("snr", ["1.0", "1.5", "2.0", "2.5", "3.0"])
def test_cubes_match_reference_binary(oracle, lib, snr):
    import hashlib

    import sogrand_amd as sg

    m = _manifest()["cubes"][snr]
    g = golden(f"cubes_{snr}.npz")
    n = 60 if snr == "1.0" else 100  # a bounded prefix keeps the CPU suite short
    llr = sg.gen_llr(m["seed"], m["n"], ebn0_db=float(snr))
    assert hashlib.sha256(llr.astype("<f8").tobytes()).hexdigest() == m["sha256_llr"]
    bits, st, soft = oracle.decode(llr[:n], soft=True)
    assert (bits == g["bits"][:n]).all()
    assert (st["half_iters"] == g["half_iters"][:n]).all()
    assert (st["NG"] == g["NG"][:n]).all()
    assert (st["NG_p"] == g["NG_p"][:n]).all()
    ns = g["soft"].shape[0]
    assert (soft[:ns] == g["soft"]).all()


def test_oracle_nonfinite_matches_reference(oracle):
    # +-inf and NaN inputs as the reference processes them (fixtures of
    # tests/golden/make_golden_nonfinite.py): lines and whole decodes bit for bit
    sys.path.insert(0, GOLDEN)
    from make_golden_nonfinite import CUBES, cube_llr

    g = golden("nonfinite.npz")
    for i, x in enumerate(g["lines"]):
        la, le, tr = oracle.siso(x)
        assert tr.T == g["lines_N_guess"][i], i
        assert np.array_equal(la, g["lines_L_APP"][i], equal_nan=True), i
        assert np.array_equal(le, g["lines_L_E"][i], equal_nan=True), i
    for snr in CUBES:
        bits, st, soft = oracle.decode(cube_llr(snr), soft=True)
        assert (bits == g[f"bits_{snr}"]).all()
        for f in ("half_iters", "NG", "NG_p"):
            assert (st[f] == g[f"{f}_{snr}"]).all(), f
        assert np.array_equal(soft, g[f"soft_{snr}"], equal_nan=True)


def test_oracle_epsilon_variant_matches_reference_variant(oracle):
    # the reference's DBL_EPSILON variant (cubic_decoder_updated.c), recorded by
    # tests/golden/make_golden_variant.py: lines and whole decodes, bit for bit
    from oracle import DBL_EPSILON

    g = golden("variant_eps.npz")
    try:
        oracle.set_epsilon(DBL_EPSILON)
        la = np.zeros_like(g["llr"])
        le = np.zeros_like(g["llr"])
        ng = np.zeros(len(la), np.int64)
        for i, x in enumerate(g["llr"]):
            la[i], le[i], tr = oracle.siso(x)
            ng[i] = tr.T
        assert (ng == g["N_guess"]).all()
        assert np.array_equal(la, g["L_APP"]) and np.array_equal(le, g["L_E"])
        import sogrand_amd as sg
        import json
        meta = json.load(open(os.path.join(GOLDEN, "variant_eps.json")))
        for tag, m in meta["cubes"].items():
            llr = sg.gen_llr(m["seed"], m["n"], ebn0_db=float(tag))
            bits, st, soft = oracle.decode(llr, soft=True)
            assert (bits == g[f"bits_{tag}"]).all()
            assert (st["half_iters"] == g[f"half_iters_{tag}"]).all()
            assert (st["NG"] == g[f"NG_{tag}"]).all() and (st["NG_p"] == g[f"NG_p_{tag}"]).all()
            ns = g[f"soft_{tag}"].shape[0]
            assert np.array_equal(soft[:ns], g[f"soft_{tag}"])
    finally:
        oracle.set_epsilon(1e-40)
    # the fixture tells the two EPSILON values apart
    la0 = np.array([oracle.siso(x)[0] for x in g["llr"][:200]])
    assert not np.array_equal(la0, g["L_APP"][:200])


def test_oracle_legacy_variant_matches_reference_legacy(oracle):
    # the reference's earlier decoder (cubic_decoder.c), recorded by
    # tests/golden/make_golden_variant.py legacy: lines and whole decodes, bit for bit
    import json

    import sogrand_amd as sg

    g = golden("variant_legacy.npz")
    try:
        oracle.set_variant(1)
        la = np.zeros_like(g["llr"])
        le = np.zeros_like(g["llr"])
        ng = np.zeros(len(la), np.int64)
        for i, x in enumerate(g["llr"]):
            la[i], le[i], tr = oracle.siso(x)
            ng[i] = tr.T
            assert tr.n_list == g["curL"][i]
        assert (ng == g["N_guess"]).all()
        assert np.array_equal(la, g["L_APP"]) and np.array_equal(le, g["L_E"])
        # under a guess budget the last guess's weight is the P_neg left behind
        for i, x in enumerate(g["llr"]):
            la[i], le[i], tr = oracle.siso(x, Tmax=64)
            ng[i] = tr.T
        assert (ng == g["N_guess_t64"]).all()
        assert np.array_equal(la, g["L_APP_t64"]) and np.array_equal(le, g["L_E_t64"])
        meta = json.load(open(os.path.join(GOLDEN, "variant_legacy.json")))
        for tag, m in meta["cubes"].items():
            llr = sg.gen_llr(m["seed"], m["n"], ebn0_db=float(tag))
            bits, st, soft = oracle.decode(llr, soft=True)
            assert (bits == g[f"bits_{tag}"]).all()
            assert (st["half_iters"] == g[f"half_iters_{tag}"]).all()
            assert (st["NG"] == g[f"NG_{tag}"]).all() and (st["NG_p"] == g[f"NG_p_{tag}"]).all()
            ns = g[f"soft_{tag}"].shape[0]
            assert np.array_equal(soft[:ns], g[f"soft_{tag}"])
    finally:
        oracle.set_variant(0)
    # the fixture tells the two decoders apart: the earlier one always fills its
    # list (P_notGuess stays 1, so the confidence never reaches 1 - 1e-5)
    assert (g["curL"] == 3).all()
    ng0 = np.array([oracle.siso(x)[2].T for x in g["llr"][:200]])
    assert (ng0 != g["N_guess"][:200]).any()
