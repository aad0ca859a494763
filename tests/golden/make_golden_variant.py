"""Record golden vectors from the reference's EPSILON = DBL_EPSILON decoder variant
(SOGRAND_C/cubic_decoder_updated.c, otherwise the semantics of _final).

Runs only in a container that holds the reference tree (compiled where it lies by
`make -C oracle ref`). Writes tests/golden/variant_eps.npz: component decodes of
600 lines through the variant's SOGRAND_bitSO and list search, and 100 codewords
per Eb/N0 in {1.5, 2.0} dB decoded by the variant's flow build and its FMA build
(one codeword per file), with final L_APP for a few codewords.

    python tests/golden/make_golden_variant.py
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
from oracle import (REF_BIN_UPDATED, REF_BIN_UPDATED_NATIVE, REF_SO_UPDATED,  # noqa: E402
                    Reference, build_reference)
import sogrand_amd as sg  # noqa: E402

SNRS = (1.5, 2.0)


def main() -> None:
    if not build_reference():
        sys.exit("reference tree not found")
    ref = Reference(REF_SO_UPDATED)
    rng = np.random.default_rng(20261204)
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
    meta = {"source": "SOGRAND_C/cubic_decoder_updated.c", "lines": len(lines), "cubes": {}}
    for k, snr in enumerate(SNRS):
        seed, n, ns = 6000 + k, 100, 3
        llr = sg.gen_llr(seed, n, ebn0_db=snr)

        def table(binary):
            with ThreadPoolExecutor(os.cpu_count() or 8) as ex:
                res = list(ex.map(lambda x: run_ref_binary_one(x, binary), llr))
            return (np.array([np.frombuffer(r[0], np.uint8) for r in res]),
                    np.array([int(round(2 * r[1])) for r in res], np.int32),
                    np.array([int(round(r[2])) for r in res], np.int64),
                    np.array([int(round(r[3])) for r in res], np.int64))

        bits, half, ngc, ngpc = table(REF_BIN_UPDATED)
        bits_f, half_f, ng_f, ngp_f = table(REF_BIN_UPDATED_NATIVE)
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
        print(f"variant {snr} dB: mean half {half.mean():.2f}", flush=True)
    np.savez_compressed(os.path.join(HERE, "variant_eps.npz"), **out)
    with open(os.path.join(HERE, "variant_eps.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
