"""File-level simulation flow around the drop-in decoder.

The reference runs its pipeline as separate programs over files
(SOGRAND_C/cubic_flow.sh): random bytes, the cubic encoder, the AWGN channel,
the decoder and a comparator. This module provides the same roles, with the same
file formats, so that `cubic_flow.sh` in this package can run the pipeline with
the MI355X decoder:

    python -m sogrand_amd.flow data     --n-cw N --seed S original_data.bin
    python -m sogrand_amd.flow encode   original_data.bin encoded_cubic.bin
    python -m sogrand_amd.flow channel  encoded_cubic.bin corrupted_llrs.bin EBN0_DB 3 --seed S
    sogrand_amd/cubic_decoder corrupted_llrs.bin decoded_cubic.bin
    python -m sogrand_amd.flow compare  original_data.bin decoded_cubic.bin

Formats: info data = 64 bytes per codeword (512 info bits, MSB first, bit
b = a0 + 8 a1 + 64 a2); encoded = 512 bytes per codeword (4096 code bits, index
a0 + 16 a1 + 256 a2, MSB first, SOGRAND_C/cubic_encoder.c:89-97); LLRs = 4096
little-endian float64 per codeword. The channel restates channel_sim.c:34-39,
61-89 (BPSK 0 -> +1, sigma from Eb/N0 at rate 1/8 for the cube or 1/4 for the
square code, LLR = 2y/sigma^2) with a seeded generator. The comparator prints
the reference comparator's report (comparator.c) and adds the block error rate
and 95 % Wilson intervals.
"""
from __future__ import annotations

import argparse
import math
import sys

import numpy as np


def _lib():
    import sogrand_amd as sg

    return sg


def wilson(k: int, n: int, z: float = 1.96) -> tuple[float, float]:
    """95 % Wilson score interval of a binomial rate k / n."""
    if n == 0:
        return (0.0, 1.0)
    p = k / n
    d = 1.0 + z * z / n
    c = (p + z * z / (2 * n)) / d
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / d
    return (max(0.0, c - h), min(1.0, c + h))


def cmd_data(a) -> int:
    sg = _lib()
    bits = sg.gen_info_bits(a.seed, a.n_cw)
    bits.tofile(a.out)
    print(f"Generating {bits.size} random bytes into {a.out}...")
    print(f"Successfully created binary file: {a.out}")
    return 0


def cmd_encode(a) -> int:
    sg = _lib()
    data = np.fromfile(a.inp, np.uint8)
    n = data.size // 64
    code = sg.encode_cubes(data[: n * 64].reshape(n, 64))
    code.tofile(a.out)
    print(f"Encoding {a.inp} to {a.out} using SYSTEMATIC (n=16, k=8) cubic code...")
    print(f"Encoding complete. {n} block(s) encoded.")
    return 0


def sigma_for(ebn0_db: float, dim: int) -> float:
    """channel_sim.c:34-39: R = (k/n)^dim, EsN0 = EbN0 + 10 log10(2R), sigma = 1/sqrt(EsN0)."""
    rate = (8.0 / 16.0) ** dim
    esn0 = ebn0_db + 10.0 * math.log10(2.0 * rate)
    return 1.0 / math.sqrt(10.0 ** (esn0 / 10.0))


def cmd_channel(a) -> int:
    if a.dim not in (2, 3):
        print("Error: Code dimension must be 2 or 3.", file=sys.stderr)
        return 1
    code = np.fromfile(a.inp, np.uint8)
    bits = np.unpackbits(code)  # MSB first
    sigma = sigma_for(a.ebn0, a.dim)
    print(f"Simulating AWGN channel for Eb/N0 = {a.ebn0:.2f} dB (sigma = {sigma:f})...")
    rng = np.random.default_rng(a.seed)
    x = 1.0 - 2.0 * bits.astype(np.float64)  # BPSK: 0 -> +1, 1 -> -1
    y = x + sigma * rng.standard_normal(x.size)
    llr = 2.0 * y / (sigma * sigma)
    llr.astype("<f8").tofile(a.out)
    print(f"Channel simulation complete. Output file '{a.out}' contains LLRs.")
    return 0


def compare(orig: bytes, dec: bytes) -> dict:
    """Bit and block errors of decoded data against the original (64 B blocks)."""
    o = np.frombuffer(orig, np.uint8)
    d = np.frombuffer(dec, np.uint8)
    n = min(o.size, d.size)
    missing = o.size - n
    err_bits = int(np.unpackbits(o[:n] ^ d[:n]).sum()) + 8 * max(missing, 0)
    total = 8 * (n + max(missing, 0))
    nb = o.size // 64
    blocks_bad = 0
    for b in range(nb):
        s = slice(64 * b, 64 * b + 64)
        if s.stop > n or bytes(o[s]) != bytes(d[s]):
            blocks_bad += 1
    return {"total_bits": total, "error_bits": err_bits, "blocks": nb, "block_errors": blocks_bad,
            "short": missing > 0}


def cmd_compare(a) -> int:
    orig = open(a.orig, "rb").read()
    dec = open(a.dec, "rb").read()
    r = compare(orig, dec)
    print(f"Comparing {a.orig} and {a.dec}...")
    if r["short"]:
        print("Error: Decoded file is shorter than original file.")
    print("\n--- Comparison Report ---")
    print("SUCCESS: Files are identical." if r["error_bits"] == 0 else "FAILURE: Files are different.")
    print(f"Total bits compared: {r['total_bits']}")
    print(f"Total bit errors:    {r['error_bits']}")
    if r["total_bits"] > 0:
        print(f"Bit Error Rate (BER): {r['error_bits'] / r['total_bits']:e}")
        lo, hi = wilson(r["error_bits"], r["total_bits"])
        print(f"BER 95% interval:     [{lo:e}, {hi:e}]")
    if r["blocks"] > 0:
        print(f"Block Error Rate (BLER): {r['block_errors'] / r['blocks']:e} "
              f"({r['block_errors']} of {r['blocks']} blocks)")
        lo, hi = wilson(r["block_errors"], r["blocks"])
        print(f"BLER 95% interval:       [{lo:e}, {hi:e}]")
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m sogrand_amd.flow")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("data", help="seeded random info data (64 B per codeword)")
    p.add_argument("--n-cw", type=int, default=1000)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("out")
    p.set_defaults(fn=cmd_data)
    p = sub.add_parser("encode", help="systematic (16,8)^3 encoding")
    p.add_argument("inp")
    p.add_argument("out")
    p.set_defaults(fn=cmd_encode)
    p = sub.add_parser("channel", help="BPSK over AWGN, LLRs as float64")
    p.add_argument("inp")
    p.add_argument("out")
    p.add_argument("ebn0", type=float)
    p.add_argument("dim", type=int)
    p.add_argument("--seed", type=int, default=1)
    p.set_defaults(fn=cmd_channel)
    p = sub.add_parser("compare", help="BER / BLER of decoded data")
    p.add_argument("orig")
    p.add_argument("dec")
    p.set_defaults(fn=cmd_compare)
    a = ap.parse_args(argv)
    return a.fn(a)


if __name__ == "__main__":
    sys.exit(main())
