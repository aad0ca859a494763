"""Record golden vectors from the reference's other decoder variants.

  eps     SOGRAND_C/cubic_decoder_updated.c: EPSILON = DBL_EPSILON, otherwise the
          semantics of _final. Writes variant_eps.npz / .json: 600 lines, and
          100 codewords per Eb/N0 in {1.5, 2.0} dB.
  legacy  SOGRAND_C/cubic_decoder.c, the earlier decoder (JacLog cut at +-30,
          IC guard, P_notGuess never decremented, clamps 1e-9 / 1e-30). Writes
          variant_legacy.npz / .json: 600 lines (also under Tmax = 64), and 64
          codewords per Eb/N0 in {2.5, 3.0} dB (it does not converge at 2 dB).

Runs only in a container that holds the reference tree (compiled where it lies by
`make -C oracle ref`). Lines go through the variant's SOGRAND_bitSO and list
search; codewords are decoded by the variant's flow build and its FMA build (one
codeword per file), with final L_APP for a few codewords from the harness.

    python tests/golden/make_golden_variant.py [eps|legacy]
"""
from __future__ import annotations

import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from make_golden import make_lines, run_ref_binary_one, sha256  # noqa: E402
from oracle import (REF_BIN_LEGACY, REF_BIN_LEGACY_NATIVE, REF_BIN_UPDATED,  # noqa: E402
                    REF_BIN_UPDATED_NATIVE, REF_SO_LEGACY, REF_SO_UPDATED, Reference,
                    build_reference)
import sogrand_amd as sg  # noqa: E402

VARIANTS = {
    # name: (harness, flow build, FMA build, source, line seed, cube seeds, SNRs, cw per SNR)
    "eps": (REF_SO_UPDATED, REF_BIN_UPDATED, REF_BIN_UPDATED_NATIVE,
            "SOGRAND_C/cubic_decoder_updated.c", 20261204, 6000, (1.5, 2.0), 100),
    "legacy": (REF_SO_LEGACY, REF_BIN_LEGACY, REF_BIN_LEGACY_NATIVE,
               "SOGRAND_C/cubic_decoder.c", 20261205, 7000, (2.5, 3.0), 64),
}


def main() -> None:
    name = sys.argv[1] if len(sys.argv) > 1 else "eps"
    so, flow_bin, fma_bin, source, line_seed, cube_seed, snrs, n_cw = VARIANTS[name]
    if not build_reference():
        sys.exit("reference tree not found")
    ref = Reference(so)
    rng = np.random.default_rng(line_seed)
    lines = make_lines(rng)[:600]
    la = np.zeros_like(lines)
    le = np.zeros_like(lines)
    ng = np.zeros(len(lines), np.int64)
    T = np.zeros(len(lines), np.int64)
    curL = np.zeros(len(lines), np.int32)
    pnl = np.zeros(len(lines))
    _, H = ref.code_tables()
    Hf = H.astype(np.uint8)
    for i, x in enumerate(lines):
        la[i], le[i], ng[i] = ref.siso(x)
        _, _, t, cl, p = ref.main_logic(x, Hf, 8, IC=-1, L=3)
        T[i], curL[i], pnl[i] = t, cl, p
    out = {"llr": lines, "L_APP": la, "L_E": le, "N_guess": ng, "T": T, "curL": curL, "pNL": pnl}
    if name == "legacy":
        # a guess budget (Tmax = 64): a search the budget stops keeps the last
        # guess's weight as P_neg (cubic_decoder.c:638)
        la64, le64, ng64 = np.zeros_like(lines), np.zeros_like(lines), np.zeros(len(lines), np.int64)
        for i, x in enumerate(lines):
            la64[i], le64[i], ng64[i] = ref.siso(x, Tmax=64)
        out.update({"L_APP_t64": la64, "L_E_t64": le64, "N_guess_t64": ng64})
    meta = {"source": source, "lines": len(lines), "cubes": {}}
    for k, snr in enumerate(snrs):
        seed, n, ns = cube_seed + k, n_cw, 3
        llr = sg.gen_llr(seed, n, ebn0_db=snr)

        def table(binary):
            with ThreadPoolExecutor(os.cpu_count() or 8) as ex:
                res = list(ex.map(lambda x: run_ref_binary_one(x, binary), llr))
            return (np.array([np.frombuffer(r[0], np.uint8) for r in res]),
                    np.array([int(round(2 * r[1])) for r in res], np.int32),
                    np.array([int(round(r[2])) for r in res], np.int64),
                    np.array([int(round(r[3])) for r in res], np.int64))

        bits, half, ngc, ngpc = table(flow_bin)
        bits_f, half_f, ng_f, ngp_f = table(fma_bin)
        soft = np.zeros((ns, 4096))
        for c in range(ns):
            la_c, hi, ng_c, ngp_c, b = ref.decode_cube(llr[c])
            assert hi == half[c] and ng_c == ngc[c] and ngp_c == ngpc[c]
            assert (np.packbits(b) == bits[c]).all()
            soft[c] = la_c
        tag = f"{snr:.1f}"
        out.update({f"bits_{tag}": bits, f"half_iters_{tag}": half, f"NG_{tag}": ngc,
                    f"NG_p_{tag}": ngpc, f"bits_fma_{tag}": bits_f, f"half_iters_fma_{tag}": half_f,
                    f"NG_fma_{tag}": ng_f, f"NG_p_fma_{tag}": ngp_f, f"soft_{tag}": soft})
        meta["cubes"][tag] = {"seed": seed, "n": n, "sha256_llr": sha256(llr.astype("<f8"))}
        print(f"variant {name} {snr} dB: mean half {half.mean():.2f}", flush=True)
    np.savez_compressed(os.path.join(HERE, f"variant_{name}.npz"), **out)
    with open(os.path.join(HERE, f"variant_{name}.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
