"""Record the reference decoder's outputs on non-finite LLRs.

SOGRAND_C's channel can emit infinite LLRs (channel_sim.c:9-11 takes log(u1)
with u1 = 0), and its decoder (cubic_decoder_final.c) processes any double:
+-inf magnitudes give infinite path-metric increments, and the next phase
forms inf - inf = NaN, which then flows through the sort, the intercept and
the soft output. This script records what the reference computes for:
  - lines with +-inf entries, lines with a NaN entry, all-infinite lines
    (SOGRAND_bitSO and the list search, through the harness);
  - cubes of the seeded generator at 1.5 and 2.0 dB with a few entries set to
    +-inf (and two with a NaN), decoded by the unmodified reference binary one
    codeword per file (bits, n_iter, NG, NG_p) and by the harness (final L_APP).
Runs only where the reference tree exists (like make_golden.py).

    python tests/golden/make_golden_nonfinite.py

Writes tests/golden/nonfinite.npz and nonfinite.json.
"""
from __future__ import annotations

import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_golden import ROOT, run_ref_binary_one  # noqa: E402

sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import sogrand_amd as sg  # noqa: E402  (host generator only)
from oracle import Reference, build_reference  # noqa: E402

CUBES = {"1.5": (7101, 16), "2.0": (7102, 24)}


def make_lines(rng: np.random.Generator) -> np.ndarray:
    out = []
    for scale in (1.0, 3.0, 10.0):
        for k in (1, 2, 3):
            for _ in range(20):
                x = rng.normal(0.8, 1.3, 16) * scale * rng.choice([1.0, -1.0], 16)
                pos = rng.choice(16, k, replace=False)
                x[pos] = rng.choice([np.inf, -np.inf], k)
                out.append(x)
    for _ in range(20):  # one NaN
        x = rng.normal(0.8, 1.3, 16) * rng.choice([1.0, -1.0], 16)
        x[rng.integers(0, 16)] = np.nan
        out.append(x)
    for _ in range(10):  # a NaN and an infinity
        x = rng.normal(0.8, 1.3, 16) * 3.0
        p = rng.choice(16, 2, replace=False)
        x[p[0]] = np.nan
        x[p[1]] = rng.choice([np.inf, -np.inf])
        out.append(x)
    for _ in range(6):  # all infinite, and infinite beside huge finite values
        out.append(rng.choice([np.inf, -np.inf], 16))
        x = rng.choice([1.0, -1.0], 16) * rng.uniform(60.0, 900.0, 16)
        x[rng.integers(0, 16)] = np.inf
        out.append(x)
    return np.array(out)


def cube_llr(snr: str):
    """The seeded cubes at snr with some entries replaced by +-inf or NaN."""
    seed, n = CUBES[snr]
    llr = sg.gen_llr(seed, n, ebn0_db=float(snr))
    rng = np.random.default_rng(seed)
    for c in range(n):
        if c % 8 == 7:
            llr[c, rng.integers(0, 4096)] = np.nan
            continue
        k = (1, 2, 5)[c % 3]
        pos = rng.choice(4096, k, replace=False)
        llr[c, pos] = rng.choice([np.inf, -np.inf], k)
    return llr


def main() -> None:
    if not build_reference():
        sys.exit("reference tree not found: golden vectors can only be recorded where it exists")
    ref = Reference()
    H = ref.code_tables()[1].astype(np.uint8)
    rng = np.random.default_rng(20261204)
    lines = make_lines(rng)
    la = np.zeros_like(lines)
    le = np.zeros_like(lines)
    ng = np.zeros(len(lines), np.int64)
    T = np.zeros(len(lines), np.int64)
    curL = np.zeros(len(lines), np.int32)
    pnl = np.zeros(len(lines))
    for i, x in enumerate(lines):
        la[i], le[i], ng[i] = ref.siso(x)
        _, _, T[i], curL[i], pnl[i] = ref.main_logic(x, H, 8, IC=-1, L=3)
    out = {"lines": lines, "lines_L_APP": la, "lines_L_E": le, "lines_N_guess": ng,
           "lines_T": T, "lines_curL": curL, "lines_pNL": pnl}
    for snr in CUBES:
        llr = cube_llr(snr)
        with ThreadPoolExecutor(os.cpu_count() or 4) as ex:
            res = list(ex.map(run_ref_binary_one, llr))
        bits = np.array([np.frombuffer(r[0], np.uint8) for r in res])
        half = np.array([int(round(2 * r[1])) for r in res], np.int32)
        ngc = np.array([int(round(r[2])) for r in res], np.int64)
        ngpc = np.array([int(round(r[3])) for r in res], np.int64)
        soft = np.zeros_like(llr)
        for c in range(len(llr)):
            la_c, hi, ng_c, ngp_c, b = ref.decode_cube(llr[c])
            assert hi == half[c] and ng_c == ngc[c] and ngp_c == ngpc[c], c
            assert (np.packbits(b) == bits[c]).all(), c
            soft[c] = la_c
        out[f"bits_{snr}"] = bits
        out[f"half_iters_{snr}"] = half
        out[f"NG_{snr}"] = ngc
        out[f"NG_p_{snr}"] = ngpc
        out[f"soft_{snr}"] = soft
        print(f"{snr} dB: {len(llr)} cubes, half-iterations {half.tolist()}", flush=True)
    np.savez_compressed(os.path.join(HERE, "nonfinite.npz"), **out)
    with open(os.path.join(HERE, "nonfinite.json"), "w") as f:
        json.dump({"cubes": {k: {"seed": v[0], "n": v[1]} for k, v in CUBES.items()},
                   "lines": int(len(lines)),
                   "reference": "SOGRAND_C/cubic_decoder_final.c (flow build: gcc -lm)"},
                  f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
