"""Record golden vectors from the reference decoder (SOGRAND_C/cubic_decoder_final.c).

Runs only in a container that holds the reference tree: it compiles the
reference's own sources with oracle/Makefile (`make -C oracle ref`, outputs in
oracle/_ref/) and calls them. Nothing from the reference is copied here; the
fixtures are inputs and the reference's outputs.

    python tests/golden/make_golden.py [--cubes-per-snr N]

Writes tests/golden/*.npz and manifest.json.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import sogrand_amd as sg  # noqa: E402  (host generator only)
from oracle import REF_BIN, REF_BIN_NATIVE, REF_SO_NATIVE, Reference, build_reference  # noqa: E402

UINT64_MAX = (1 << 64) - 1
SNRS = [1.0, 1.5, 2.0, 2.5, 3.0]


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_lines(rng: np.random.Generator) -> np.ndarray:
    """Random component inputs over the operating range plus edge lines."""
    out = []
    for eb in np.linspace(0.0, 4.0, 9):
        sigma = sg.sigma_for_ebn0(float(eb))
        var = sigma * sigma
        for scale in (1.0, 3.0, 8.0):  # later iterations carry larger |LLR|
            n = 60
            x = rng.choice([1.0, -1.0], (n, 16))
            y = x + sigma * rng.standard_normal((n, 16))
            out.append(scale * 2.0 * y / var)
    rnd = np.concatenate(out)
    edge = []
    # ties among magnitudes (quicksort tie order matters)
    for _ in range(120):
        v = rng.integers(1, 6, 16).astype(np.float64) * 0.5
        edge.append(v * rng.choice([1.0, -1.0], 16))
    # all magnitudes equal (beta = 0 -> IC from an out-of-range conversion)
    for m in (0.3, 1.0, 7.5):
        edge.append(np.full(16, m) * rng.choice([1.0, -1.0], 16))
    # zeros in the input (hard decision of +0.0 and -0.0 is 1)
    z = rng.normal(0, 1, 16)
    z[:3] = 0.0
    z[3] = -0.0
    edge.append(z)
    # |llr| > 50 (JacLog clamps) and large intercepts
    for _ in range(20):
        edge.append(rng.choice([1.0, -1.0], 16) * rng.uniform(40.0, 80.0, 16))
    for _ in range(20):
        base = rng.uniform(5.0, 30.0)
        edge.append(rng.choice([1.0, -1.0], 16) * (base + rng.uniform(0.0, 0.05 * base, 16)))
    # hard decision already a codeword (every line of an encoded cube)
    info = sg.gen_info_bits(77, 4)
    code = np.unpackbits(sg.encode_cubes(info), axis=1).reshape(4, 256, 16)
    for c in range(4):
        for l in range(0, 256, 32):
            s = 1.0 - 2.0 * code[c, l]
            edge.append(s * rng.uniform(0.2, 3.0, 16))
    return np.concatenate([rnd, np.array(edge)])


def kernel_h(v: np.ndarray) -> np.ndarray:
    """15x16 parity checks whose null space is {0, v} (v != 0)."""
    i0 = int(np.flatnonzero(v)[0])
    rows = []
    for j in range(16):
        if j == i0:
            continue
        r = np.zeros(16, np.uint8)
        r[j] = 1
        r[i0] = v[j]
        rows.append(r)
    return np.array(rows, np.uint8)


def schedule_probe(ref: Reference, rng: np.random.Generator):
    """Schedule position of chosen error patterns, read off the reference.

    With all inputs negative the hard decision c_HD is all ones; a parity-check
    matrix whose null space is {0, ~e} accepts exactly the error pattern e (and
    the all-ones pattern, the last one tested), so with L = 1 the reference
    stops at e and reports T = 1 + its schedule position. Magnitudes
    a_i = IC + 1 + i give perm = identity and intercept IC."""
    ics = [0, 1, 2, 3, 5, 9, 20, 54, 55, 56, 120]
    rec_ic, rec_mask, rec_T = [], [], []
    all_masks = np.arange(1, 65535, dtype=np.int64)
    w = np.array([bin(m).count("1") for m in range(65536)])
    for ic in ics:
        llr = -(ic + 1.0 + np.arange(16.0))
        low = all_masks[w[all_masks] <= 2]
        pick = np.concatenate([low, rng.choice(all_masks, 200, replace=False)])
        for m in pick:
            e = np.array([(m >> b) & 1 for b in range(16)], np.uint8)
            v = 1 - e  # c = c_HD ^ e = ~e must be the accepted codeword
            if not v.any():
                continue
            H = kernel_h(v)
            _, sl, T, curL, _ = ref.main_logic(llr, H, 15, IC=-1, L=1)
            assert curL == 1
            rec_ic.append(ic)
            rec_mask.append(int(m))
            rec_T.append(int(T))
    return np.array(rec_ic, np.int32), np.array(rec_mask, np.uint16), np.array(rec_T, np.int64)


def run_ref_binary_one(llr: np.ndarray, binary: str = REF_BIN) -> tuple[bytes, float, float, float]:
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        llr.astype("<f8").tofile(fi)
        r = subprocess.run([binary, fi, fo], capture_output=True, text=True, check=True)
        vals = {}
        for line in r.stdout.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
        n_iter = float(vals["Average iterations per block"])
        ng = float(vals["Average NG per block"])
        ngp = float(vals["Average NG_p per block"])
        return open(fo, "rb").read(), n_iter, ng, ngp


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--cubes-per-snr", type=int, default=2000)
    ap.add_argument("--soft-per-snr", type=int, default=20)
    ap.add_argument("--workers", type=int, default=os.cpu_count() or 8)
    args = ap.parse_args()
    if not build_reference():
        sys.exit("reference tree not found: golden vectors can only be recorded where it exists")
    ref = Reference()
    ref_fma = Reference(REF_SO_NATIVE)  # the same source built -O2 -march=native (FMA)
    rng = np.random.default_rng(20261203)
    manifest: dict = {"reference": "SOGRAND_C/cubic_decoder_final.c (flow build: gcc -lm)",
                      "reference_fma": "same source, gcc -O2 -march=native (fused multiply-add)"}
    src = open("/root/reference/SOGRAND_C/cubic_decoder_final.c", "rb").read()
    manifest["reference_sha256"] = hashlib.sha256(src).hexdigest()

    G, H = ref.code_tables()
    np.savez_compressed(os.path.join(HERE, "code_tables.npz"), G=G.astype(np.uint8),
                        H=H.astype(np.uint8))

    # line-level SISO (SOGRAND_bitSO)
    lines = make_lines(rng)
    la = np.zeros_like(lines)
    le = np.zeros_like(lines)
    ng = np.zeros(len(lines), np.int64)
    for i, x in enumerate(lines):
        la[i], le[i], ng[i] = ref.siso(x)
    # raw list search for the same lines: list contents and T at each hit
    T = np.zeros(len(lines), np.int64)
    curL = np.zeros(len(lines), np.int32)
    pnl = np.zeros(len(lines))
    slist = np.zeros((len(lines), 3, 4))
    chat = np.zeros((len(lines), 3, 16), np.uint8)
    Hf = H.astype(np.uint8)
    for i, x in enumerate(lines):
        c, sl, t, cl, p = ref.main_logic(x, Hf, 8, IC=-1, L=3)
        T[i], curL[i], pnl[i], slist[i] = t, cl, p, sl
        chat[i] = c.astype(np.uint8)
    la_f = np.array([ref_fma.siso(x)[0] for x in lines])
    np.savez_compressed(os.path.join(HERE, "lines.npz"), llr=lines, L_APP=la, L_E=le, N_guess=ng,
                        T=T, curL=curL, pNL=pnl, s_list=slist, chat_list=chat, L_APP_fma=la_f)
    manifest["lines"] = int(len(lines))

    # finite Tmax: raw search only (well defined even when the list stays empty)
    tm_lines = lines[:600]
    recs = []
    for tmax in (1, 2, 20, 200, 1000):
        for i, x in enumerate(tm_lines):
            c, sl, t, cl, p = ref.main_logic(x, Hf, 8, IC=-1, L=3, Tmax=tmax)
            recs.append((tmax, i, t, cl, p, *sl[:, 1], *sl[:, 2]))
    recs = np.array(recs, dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "lines_tmax.npz"), recs=recs)

    # schedule order
    ic, m, t = schedule_probe(ref, rng)
    np.savez_compressed(os.path.join(HERE, "schedule.npz"), ic=ic, mask=m, T=t)
    manifest["schedule_probes"] = int(len(t))

    # quicksort tie order
    qs_in = []
    for _ in range(400):
        k = int(rng.integers(1, 6))
        qs_in.append(rng.integers(0, k + 1, 16).astype(np.float64))
    qs_in = np.array(qs_in)
    qs_out = np.zeros_like(qs_in)
    qs_perm = np.zeros(qs_in.shape, np.int64)
    for i, a in enumerate(qs_in):
        qs_out[i], qs_perm[i] = ref.quicksort(a)
    np.savez_compressed(os.path.join(HERE, "quicksort.npz"), a=qs_in, sorted=qs_out, perm=qs_perm)

    # early termination verdicts on hard-decision cubes
    info = sg.gen_info_bits(99, 40)
    code = np.unpackbits(sg.encode_cubes(info), axis=1).astype(np.float64)
    cubes = []
    for c in range(40):
        sgn = 1.0 - 2.0 * code[c]
        v = sgn.copy()
        if c % 4 == 1:
            v[rng.integers(0, 4096)] *= -1
        elif c % 4 == 2:
            v[rng.integers(0, 4096, 3)] *= -1
        elif c % 4 == 3:
            v[:16] *= np.where(rng.random(16) < 0.3, -1.0, 1.0)  # scramble one line
        cubes.append(v)
    cubes = np.array(cubes).astype(np.int8)  # the verdict depends only on signs
    verdict = np.array([ref.early_term(c.astype(np.float64)) for c in cubes], np.int8)
    np.savez_compressed(os.path.join(HERE, "early_term.npz"), cubes=cubes, verdict=verdict)

    # cube level: the unmodified reference binary on one-codeword files
    seeds = {}
    for k, snr in enumerate(SNRS):
        seed = 5000 + k
        n = args.cubes_per_snr
        llr = sg.gen_llr(seed, n, ebn0_db=snr)
        seeds[str(snr)] = {"seed": seed, "n": n, "sha256_llr": sha256(llr.astype("<f8"))}
        def table(binary):
            with ThreadPoolExecutor(args.workers) as ex:
                res = list(ex.map(lambda x: run_ref_binary_one(x, binary), llr))
            return (np.array([np.frombuffer(r[0], np.uint8) for r in res]),
                    np.array([int(round(2 * r[1])) for r in res], np.int32),
                    np.array([int(round(r[2])) for r in res], np.int64),
                    np.array([int(round(r[3])) for r in res], np.int64))

        bits, half, ngc, ngpc = table(REF_BIN)
        bits_f, half_f, ng_f, ngp_f = table(REF_BIN_NATIVE)
        # final L_APP for a few codewords through the harness driving the
        # reference's component decoder and stopping rule; checked against the binary
        ns = args.soft_per_snr
        soft = np.zeros((ns, 4096))
        for c in range(ns):
            la_c, hi, ng_c, ngp_c, b = ref.decode_cube(llr[c])
            assert hi == half[c] and ng_c == ngc[c] and ngp_c == ngpc[c]
            assert (np.packbits(b) == bits[c]).all()
            soft[c] = la_c
        np.savez_compressed(os.path.join(HERE, f"cubes_{snr:.1f}.npz"), bits=bits, half_iters=half,
                            NG=ngc, NG_p=ngpc, soft=soft, bits_fma=bits_f, half_iters_fma=half_f,
                            NG_fma=ng_f, NG_p_fma=ngp_f)
        print(f"snr {snr}: {n} cw, mean half {half.mean():.2f} NG {ngc.mean():.0f}", flush=True)
    manifest["cubes"] = seeds
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
