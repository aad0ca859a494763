#!/usr/bin/env python3
"""Throughput and latency of the gfx950 SOGRAND decoder on the (16,8)^3 cubic code.

Default workload (one "step"): decode 100,000 codewords at Eb/N0 = 2.0 dB, L=3,
Imax=30, Tmax unlimited, FP64, LLRs resident in HBM before the timed region
(generated on the device by the library's seeded generator). `value` is the
persistent batched decoder's codewords/s; the plain drop-in (one workgroup per
codeword) is timed on the same data and both decodes must give identical bits
("BER-matched"). A short serving run at 70 % of the batched capacity adds the
p50/p99 frame latency. Prints one JSON line.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]

Other workloads, each printing its own JSON line:
    --workload uneven           1,000,000 codewords, per-codeword Eb/N0 ~ U[0.5, 3.5] dB,
                                Tmax in {unlimited, 1e5, 1e3}: cw/s of both kernels, BER/BLER
    --workload mixed-precision  1,000,000 codewords at 2.0 and 2.5 dB: FP32 variant vs FP64
    --workload serving          Poisson frames of U{1..256} codewords at U[1, 3] dB, offered
                                load 10-120 % of capacity: p50/p99 latency, sustained cw/s
    --workload square           100,000 squares of the (16,8)^2 code at 1, 2 and 3 dB
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the serving loop keeps 16 batches in flight; give each its own hardware queue
# (read by the HIP runtime when it starts, so set before the library loads). An
# environment that already sets fewer queues (4 is common) is raised to 16:
# with fewer queues than slots, a batch waits behind the slowest cube of the
# batch launched before it on the same queue.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

METRIC = "codewords/s @Eb/N0 2 dB (BER-matched), batched vs plain; p99 latency"
VALU_PEAK = 256 * 4 * 16 * 2.4e9  # lane-ops/s: CUs x SIMDs x lanes x clock
HBM_PEAK = 8.0e12  # bytes/s, MI355X HBM3E
C_GUESS = 36.0   # VALU lane-ops per guess (algorithmic budget)
C_LINE = 3500.0  # VALU lane-ops per component decode (set-up + soft output)
PUBLISHED_CW_S = 13.19  # the reference's published GPU throughput at 2 dB (README.md:137-143)
CHUNK = 16384


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


# The restatement's speed relative to the reference itself, measured in the
# build container on one core (same 40 codewords at 2 dB, identical bits):
# SOGRAND_C cubic_decoder_final.c built -O2 decodes REF_PER_CORE_CW_S codewords/s,
# the restatement ORACLE_PER_CORE_CW_S.
REF_PER_CORE_CW_S = 1.60
ORACLE_PER_CORE_CW_S = 22.81


def host_cores() -> int:
    """The host cores this process may run on (its CPU affinity)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(seed: int, ebn0: float, seconds: float) -> dict:
    """The C restatement of SOGRAND_C (oracle/liboracle.so, pinned bit for bit to
    the reference's own outputs) timed on every host core of this process, one
    codeword per thread, over a bounded sample of the same workload."""
    import sogrand_amd as sg

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle

    cores = host_cores()
    o = Oracle()
    # size the sample from a short probe of eight codewords per core (with one,
    # the slowest codeword sets the probe's time and the sample comes out short)
    probe = sg.gen_llr(seed + 7, 8 * cores, ebn0_db=ebn0)
    t0 = time.perf_counter()
    o.decode(probe, threads=cores)
    rate = 8 * cores / max(time.perf_counter() - t0, 1e-3)
    n = max(cores, int(seconds * rate))
    llr = sg.gen_llr(seed, n, ebn0_db=ebn0)
    t0 = time.perf_counter()
    o.decode(llr, threads=cores)
    dt = time.perf_counter() - t0
    out = {"value": round(n / dt, 2), "unit": "codewords/s", "cores": cores, "kind": "port",
           "sample": f"{n} codewords of the same workload (Eb/N0 {ebn0} dB), C restatement of "
                     f"SOGRAND_C built -O2 without FMA, {cores} threads, {dt:.1f} s on "
                     f"{cpu_model()}"}
    if REF_PER_CORE_CW_S and ORACLE_PER_CORE_CW_S:
        out["calibration"] = (f"per core, the restatement decodes {ORACLE_PER_CORE_CW_S} cw/s "
                              f"where SOGRAND_C itself (-O2) decodes {REF_PER_CORE_CW_S} "
                              f"(build container, 2 dB): the reference would reach about "
                              f"{round(n / dt * REF_PER_CORE_CW_S / ORACLE_PER_CORE_CW_S, 1)} "
                              f"cw/s on these cores")
    return out


class _View:
    """A device sub-range of a DeviceArray (pointer + shape) for the C ABI."""

    def __init__(self, base, byte_off: int, shape):
        self._p = base.data_ptr() + byte_off
        self.shape = shape

    def data_ptr(self) -> int:
        return self._p

    def is_contiguous(self) -> bool:
        return True

    def numel(self) -> int:
        return int(np.prod(self.shape))


def gen_device(dec, seed: int, n: int, sigma) -> object:
    """LLRs of codewords [0, n) on the device; sigma scalar or one per codeword."""
    from sogrand_amd import hip

    sigma = np.atleast_1d(np.asarray(sigma, dtype=np.float64))
    d = hip.DeviceArray((n, 4096))
    for off in range(0, n, CHUNK):
        m = min(CHUNK, n - off)
        sig = sigma if sigma.size == 1 else np.ascontiguousarray(sigma[off:off + m])
        dec.gen_llr_device(seed, off, m, sig, _View(d, off * 4096 * 8, (m, 4096)))
    hip.synchronize()
    return d


def timed_decode(dec, d_llr, d_bits, d_stats, d_soft=None) -> tuple[float, float]:
    """(wall seconds, kernel ms) of one decode of device-resident buffers."""
    from sogrand_amd import hip

    hip.synchronize()
    t0 = time.perf_counter()
    dec.decode(d_llr, out_bits=d_bits, out_stats=d_stats, out_soft=d_soft)
    hip.synchronize()
    wall = time.perf_counter() - t0
    return wall, dec.last_kernel_ms()[0]


def ber_bler(bits: np.ndarray, info: np.ndarray) -> tuple[int, int]:
    return int(np.unpackbits(bits ^ info).sum()), int((bits != info).any(axis=1).sum())


def wilson(k: int, n: int, z: float = 1.96) -> tuple[float, float]:
    """95 % Wilson interval of a binomial rate."""
    if n == 0:
        return (0.0, 1.0)
    p = k / n
    d = 1 + z * z / n
    c = (p + z * z / (2 * n)) / d
    h = z * math.sqrt(p * (1 - p) / n + z * z / (4 * n * n)) / d
    return (max(0.0, c - h), min(1.0, c + h))


# ---------------------------------------------------------------------------
# serving


def serve_load(dec, capacity: float, load: float, seconds: float, seed: int,
               max_inflight: int = 16, max_batch_cw: int = 0, plain_below_cw: int = 0,
               hold_above_cw: int = 0, lanegroup_max_cw: int = 0, small_batch_cw=None,
               backlog_cw: int = 0) -> dict:
    """Open-loop Poisson arrivals of frames (U{1..256} codewords, per-frame Eb/N0
    ~ U[1, 3] dB) at `load` x capacity, through the serving loop. Latency runs
    from a frame's scheduled arrival to its results being back in host memory."""
    import sogrand_amd as sg

    require_persistent(dec)
    rng = np.random.default_rng(seed)
    mean_frame = 128.5
    lam = load * capacity / mean_frame  # frames/s
    arrivals, t = [], 0.0
    while t < seconds:
        t += rng.exponential(1.0 / lam)
        arrivals.append(t)
    nf = len(arrivals)
    sizes = rng.integers(1, 257, nf)
    snr = rng.uniform(1.0, 3.0, nf)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    n = int(offs[-1])
    sig = np.repeat([sg.sigma_for_ebn0(float(e)) for e in snr], sizes)
    d_llr = gen_device(dec, seed, n, sig)
    bits = np.zeros((n, 64), np.uint8)
    views = [_View(d_llr, int(offs[k]) * 4096 * 8, (int(sizes[k]), 4096)) for k in range(nf)]
    done = []
    with sg.Server(dec, hbm_budget=8 << 30, max_inflight=max_inflight, max_batch_cw=max_batch_cw,
                   plain_below_cw=plain_below_cw, hold_above_cw=hold_above_cw,
                   lanegroup_max_cw=lanegroup_max_cw, small_batch_cw=small_batch_cw,
                   backlog_cw=backlog_cw) as srv:
        t0 = srv.clock() + 0.01
        lag = 0.0
        for k in range(nf):
            due = t0 + arrivals[k]
            while True:
                now = srv.clock()
                if now >= due:
                    break
                if due - now > 0.002:
                    time.sleep(due - now - 0.001)
                    done += srv.poll()
            lag = max(lag, srv.clock() - due)
            rc = srv.submit(k, views[k], bits[offs[k]:offs[k + 1]])
            if rc != sg.OK:
                raise RuntimeError(f"frame {k} refused ({rc})")
        srv.drain()
        done += srv.poll(max_n=nf + 1)
    assert len(done) == nf and all(c.status == sg.OK for c in done)
    lat = np.array([c.t_done - (t0 + arrivals[c.id]) for c in done])
    first = t0 + arrivals[0]
    last = max(c.t_done for c in done)
    # steady state: codewords completed inside the arrival window after its
    # first quarter (the whole-run figure also counts the drain after the
    # last arrival, which at overload takes seconds)
    w0, w1 = t0 + 0.25 * seconds, t0 + seconds
    steady = sum(int(sizes[c.id]) for c in done if w0 <= c.t_done < w1) / (w1 - w0)
    info = sg.gen_info_bits(seed, n)
    _, blk = ber_bler(bits, info)
    d_llr.free()
    return {"load": load, "offered_cw_s": round(load * capacity, 1), "frames": nf, "n_cw": n,
            "sustained_cw_s": round(n / (last - first), 1),
            "steady_cw_s": round(steady, 1),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "max_client_lag_ms": round(1e3 * lag, 3), "bler": blk / n}


def require_persistent(dec) -> None:
    """The serving leg measures the batched (persistent) kernel: refuse a decoder
    left on another backend (the server copies the context's parameters)."""
    import sogrand_amd as sg

    if dec.params.backend != sg.BACKEND_PERSISTENT:
        raise ValueError(f"serving runs on the persistent backend, not {dec.params.backend}")


def serving_capacity(dec, seed: int, n: int = 40000) -> float:
    """Batched codewords/s on the serving frame mix (per-codeword Eb/N0 ~ U[1, 3] dB)."""
    import sogrand_amd as sg
    from sogrand_amd import hip

    require_persistent(dec)

    snr = np.random.default_rng(seed).uniform(1.0, 3.0, n)
    d_llr = gen_device(dec, seed, n, [sg.sigma_for_ebn0(float(e)) for e in snr])
    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    wall, _ = timed_decode(dec, d_llr, d_bits, d_st)
    d_llr.free()
    return n / wall


def workload_serving(args) -> dict:
    import sogrand_amd as sg

    dec = sg.Decoder(0, backend=sg.BACKEND_PERSISTENT)
    if args.serve_capacity:
        cap = float(args.serve_capacity)  # a fixed reference, for runs compared with each other
    else:
        cap = serving_capacity(dec, args.seed)
    log(f"serving capacity on the frame mix: {cap:.0f} cw/s")
    loads = [float(x) for x in args.serve_loads.split(",")]
    pts = []
    for k, load in enumerate(loads):
        # a repeated load draws new frames (the seed advances per repeat)
        r = serve_load(dec, cap, load, args.serve_seconds,
                       args.seed + int(load * 100) + 1000 * loads[:k].count(load),
                       args.serve_slots, args.serve_batch, args.serve_plain_below,
                       args.serve_hold_above, args.serve_lanegroup_max, args.serve_small_batch,
                       args.serve_backlog)
        log(f"load {load:.1f}: {r}")
        pts.append(r)
    return {"metric": "serving: p50/p99 frame latency and sustained codewords/s vs offered load",
            "unit": "ms / codewords/s", "n_gpus": 1, "dtype": "f64",
            "data": "synthetic frames, Poisson arrivals, U{1..256} codewords, Eb/N0 ~ U[1, 3] dB",
            "config": {"workload": f"serving loop, {args.serve_slots} batch slots, 8 GiB budget",
                       "seconds_per_point": args.serve_seconds, "slots": args.serve_slots,
                       "max_batch_cw": args.serve_batch,
                       "plain_below_cw": args.serve_plain_below,
                       "hold_above_cw": args.serve_hold_above,
                       "lanegroup_max_cw": args.serve_lanegroup_max,
                       "small_batch_cw": args.serve_small_batch, "backlog_cw": args.serve_backlog},
            "capacity_cw_s": round(cap, 1), "points": pts}


# ---------------------------------------------------------------------------
# uneven work and mixed precision


def workload_uneven(args) -> dict:
    import sogrand_amd as sg
    from sogrand_amd import hip

    n = args.n_cw if args.n_cw_set else 1_000_000
    dec = sg.Decoder(0)
    snr = np.random.default_rng(args.seed).uniform(0.5, 3.5, n)
    d_llr = gen_device(dec, args.seed, n, [sg.sigma_for_ebn0(float(e)) for e in snr])
    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    info = sg.gen_info_bits(args.seed, n)
    # launch shape of the persistent kernel: autotuned on a sample of this workload,
    # the pick confirmed against the default shape on the whole of it
    from sogrand_amd.autotune import autotune

    n_tune = min(n, 32768)
    tuned = autotune(dec, _View(d_llr, 0, (n_tune, 4096)), n_tune, log=log,
                     confirm=(d_llr, n))
    log(f"autotuned shape {tuned['best']}: {tuned['speedup_vs_default']}x the default")
    rows = []
    ref_bits = None
    for tmax, label in ((sg.UINT64_MAX, "unlimited"), (100000, "1e5"), (1000, "1e3")):
        row = {"Tmax": label}
        pbits = None
        for backend, name in ((sg.BACKEND_PERSISTENT, "persistent"), (sg.BACKEND_PLAIN, "plain")):
            dec.set_params(backend=backend, Tmax=tmax)
            wall, _ = timed_decode(dec, d_llr, d_bits, d_st)
            bits = d_bits.to_host()
            row[f"{name}_cw_s"] = round(n / wall, 1)
            if backend == sg.BACKEND_PERSISTENT:
                st = d_st.to_host()
                be, bl = ber_bler(bits, info)
                row.update({"ber": be / (512.0 * n), "ber_ci95": wilson(be, 512 * n),
                            "bler": bl / n, "bler_ci95": wilson(bl, n),
                            "mean_half_iters": float(st["half_iters"].mean()),
                            "half_iters_range": [int(st["half_iters"].min()),
                                                 int(st["half_iters"].max())],
                            "mean_NG": float(st["NG"].astype(np.float64).mean())})
                pbits = bits
            else:
                row["plain_identical_bits"] = bool((bits == pbits).all())
            log(f"Tmax {label} {name}: {n / wall:.0f} cw/s")
        row["batched_over_plain"] = round(row["persistent_cw_s"] / row["plain_cw_s"], 3)
        if label == "unlimited":
            ref_bits = pbits
        elif label == "1e5":
            row["identical_to_unlimited"] = bool((pbits == ref_bits).all())
        rows.append(row)
    return {"metric": "codewords/s, uneven work (per-codeword Eb/N0 ~ U[0.5, 3.5] dB)",
            "unit": "codewords/s", "n_gpus": 1, "dtype": "f64",
            "data": "synthetic (seeded info bits, (16,8)^3 encoding, BPSK, AWGN)",
            "config": {"workload": f"{n} codewords, per-codeword Eb/N0 ~ U[0.5, 3.5] dB, L=3, "
                                   f"Imax=30", "n_cw": n,
                       "persistent_shape": tuned["best"]},
            "autotune": {k: v for k, v in tuned.items() if k != "trials"},
            "rows": rows}


def workload_mixed(args) -> dict:
    import sogrand_amd as sg
    from sogrand_amd import hip

    n = args.n_cw if args.n_cw_set else 1_000_000
    n_soft = min(n, 20000)
    dec = sg.Decoder(0, backend=sg.BACKEND_PERSISTENT)
    rows = []
    for snr in (2.0, 2.5):
        seed = args.seed + int(snr * 10)
        d_llr = gen_device(dec, seed, n, sg.sigma_for_ebn0(snr))
        d_bits = hip.DeviceArray((n, 64), np.uint8)
        d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
        d_soft = hip.DeviceArray((n_soft, 4096))
        info = sg.gen_info_bits(seed, n)
        res = {}
        for prec, name in ((sg.FP64, "fp64"), (sg.FP32_MIXED, "fp32")):
            dec.set_params(precision=prec)
            wall, _ = timed_decode(dec, d_llr, d_bits, d_st)
            bits = d_bits.to_host()
            st = d_st.to_host()
            timed_decode(dec, _View(d_llr, 0, (n_soft, 4096)), _View(d_bits, 0, (n_soft, 64)),
                         _View(d_st, 0, (n_soft,)), d_soft)
            res[name] = (n / wall, bits, st, d_soft.to_host())
            log(f"{snr} dB {name}: {n / wall:.0f} cw/s")
        c64, b64, s64, f64 = res["fp64"]
        c32, b32, s32, f32 = res["fp32"]
        be64, bl64 = ber_bler(b64, info)
        be32, bl32 = ber_bler(b32, info)
        hd_mis = int((np.signbit(f64) != np.signbit(f32)).sum())
        d = np.abs(f64 - f32)
        a = np.abs(f64)
        buckets = {}
        for lo, hi in ((0, 1), (1, 5), (5, 10), (10, 20), (20, np.inf)):
            m = (a >= lo) & (a < hi)
            if m.any():
                buckets[f"|L|in[{lo},{hi})"] = {"share": round(float(m.mean()), 4),
                                                "max": float(d[m].max()),
                                                "mean": float(d[m].mean())}
        ng64 = s64["NG"].astype(np.float64).sum()
        rows.append({
            "ebn0_db": snr, "fp64_cw_s": round(c64, 1), "fp32_cw_s": round(c32, 1),
            "speedup": round(c32 / c64, 3),
            "info_bit_mismatch_rate": float(np.unpackbits(b64 ^ b32).mean()),
            "codewords_differing": int((b64 != b32).any(axis=1).sum()),
            "hard_decision_mismatch_rate_all_positions": hd_mis / f64.size,
            "ber_fp64": be64 / (512.0 * n), "ber_fp32": be32 / (512.0 * n),
            "ber_delta": (be32 - be64) / (512.0 * n),
            "ber_ci95_fp64": wilson(be64, 512 * n), "ber_ci95_fp32": wilson(be32, 512 * n),
            "bler_fp64": bl64 / n, "bler_fp32": bl32 / n, "bler_delta": (bl32 - bl64) / n,
            "bler_ci95_fp64": wilson(bl64, n), "bler_ci95_fp32": wilson(bl32, n),
            "NG_rel_delta": float((s32["NG"].astype(np.float64).sum() - ng64) / ng64),
            "soft_abs_err_max": float(d.max()), "soft_abs_err_mean": float(d.mean()),
            "soft_err_by_fp64_magnitude": buckets, "soft_sample_cw": n_soft})
        for x in (d_llr, d_bits, d_st, d_soft):
            x.free()
    return {"metric": "FP32 variant vs FP64: mismatch, BER/BLER delta, soft-output error",
            "unit": "codewords/s", "n_gpus": 1, "dtype": "f32 metrics / f64 found mass",
            "data": "synthetic (seeded info bits, (16,8)^3 encoding, BPSK, AWGN)",
            "config": {"workload": f"{n} codewords at 2.0 and 2.5 dB, persistent kernel",
                       "n_cw": n}, "rows": rows}


def workload_snr_sweep(args) -> dict:
    """Both kernels at each Eb/N0 of the plain configuration's sweep, 100,000
    codewords per point by default: cw/s, error rates and work per codeword."""
    import sogrand_amd as sg
    from sogrand_amd import hip

    n = args.n_cw
    dec = sg.Decoder(0, backend=sg.BACKEND_PERSISTENT)
    rows = []
    for k, snr in enumerate((1.0, 1.5, 2.0, 2.5, 3.0)):
        seed = args.seed + 100 + k
        d_llr = gen_device(dec, seed, n, sg.sigma_for_ebn0(snr))
        d_bits = hip.DeviceArray((n, 64), np.uint8)
        d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
        info = sg.gen_info_bits(seed, n)
        row = {"ebn0_db": snr}
        ref_bits = None
        for backend, name in ((sg.BACKEND_PERSISTENT, "persistent"), (sg.BACKEND_PLAIN, "plain")):
            dec.set_params(backend=backend)
            wall, kms = timed_decode(dec, d_llr, d_bits, d_st)
            bits = d_bits.to_host()
            row[f"{name}_cw_s"] = round(n / wall, 1)
            if ref_bits is None:
                ref_bits = bits
                st = d_st.to_host()
                be, bl = ber_bler(bits, info)
                row.update({"ber": be / (512.0 * n), "ber_ci95": wilson(be, 512 * n),
                            "bler": bl / n, "bler_ci95": wilson(bl, n),
                            "mean_half_iters": float(st["half_iters"].mean()),
                            "converged": float(st["converged"].mean()),
                            "mean_NG": float(st["NG"].astype(np.float64).mean())})
            else:
                row["identical_bits"] = bool((bits == ref_bits).all())
            log(f"{snr} dB {name}: {n / wall:.0f} cw/s")
        row["batched_over_plain"] = round(row["persistent_cw_s"] / row["plain_cw_s"], 3)
        rows.append(row)
        for x in (d_llr, d_bits, d_st):
            x.free()
    return {"metric": "codewords/s per Eb/N0, batched and plain",
            "unit": "codewords/s", "n_gpus": 1, "dtype": "f64",
            "data": "synthetic (seeded info bits, (16,8)^3 encoding, BPSK, AWGN)",
            "config": {"workload": f"{n} codewords per Eb/N0 in {{1.0, 1.5, 2.0, 2.5, 3.0}} dB, "
                                   f"L=3, Imax=30, Tmax unlimited", "n_cw": n}, "rows": rows}


def workload_variants(args) -> dict:
    """The reference's three SOGRAND_C decoders on the same kernels: throughput,
    error rates and work of each, 100,000 codewords per Eb/N0 by default."""
    import sogrand_amd as sg
    from sogrand_amd import hip

    n = args.n_cw
    dec = sg.Decoder(0, backend=sg.BACKEND_PERSISTENT)
    variants = (("cubic_decoder_final.c", {"variant": sg.VARIANT_FINAL, "epsilon": 1e-40}),
                ("cubic_decoder_updated.c", {"variant": sg.VARIANT_FINAL,
                                             "epsilon": 2.220446049250313e-16}),
                ("cubic_decoder.c", {"variant": sg.VARIANT_LEGACY, "epsilon": 1e-40}))
    rows = []
    for snr in (2.0, 3.0):
        seed = args.seed + int(snr * 10)
        d_llr = gen_device(dec, seed, n, sg.sigma_for_ebn0(snr))
        d_bits = hip.DeviceArray((n, 64), np.uint8)
        d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
        info = sg.gen_info_bits(seed, n)
        for name, kw in variants:
            dec.set_params(**kw)
            timed_decode(dec, d_llr, d_bits, d_st)
            wall, kms = timed_decode(dec, d_llr, d_bits, d_st)
            bits, st = d_bits.to_host(), d_st.to_host()
            be, bl = ber_bler(bits, info)
            rows.append({"decoder": name, "ebn0_db": snr, "cw_s": round(n / wall, 1),
                         "kernel_ms": round(kms, 2), "ber": be / (512.0 * n),
                         "ber_ci95": wilson(be, 512 * n), "bler": bl / n,
                         "bler_ci95": wilson(bl, n),
                         "mean_half_iters": float(st["half_iters"].mean()),
                         "converged": float(st["converged"].mean()),
                         "mean_NG": float(st["NG"].astype(np.float64).mean())})
            log(f"{snr} dB {name}: {n / wall:.0f} cw/s, BLER {bl / n:.3e}")
        for x in (d_llr, d_bits, d_st):
            x.free()
    return {"metric": "codewords/s of the reference's decoder variants",
            "unit": "codewords/s", "n_gpus": 1, "dtype": "f64",
            "data": "synthetic (seeded info bits, (16,8)^3 encoding, BPSK, AWGN)",
            "config": {"workload": f"{n} codewords at 2.0 and 3.0 dB, persistent kernel, "
                                   f"L=3, Imax=30", "n_cw": n}, "rows": rows}


def workload_square(args) -> dict:
    """The (16,8)^2 square code: 100,000 squares per Eb/N0 (1, 2, 3 dB) in device
    memory, decoded by sogrand_decode_square; throughput and error rates."""
    import sogrand_amd as sg
    from sogrand_amd import hip

    n = args.n_cw
    dec = sg.Decoder(0)
    rows = []
    for snr in (1.0, 2.0, 3.0):
        info, llr = sg.gen_llr_square(args.seed + int(snr * 10), n, ebn0_db=snr)
        d_llr = hip.DeviceArray.from_host(llr)
        d_bits = hip.DeviceArray((n, 8), np.uint8)
        d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
        dec.decode_square(d_llr, out_bits=d_bits, out_stats=d_st)  # warm-up
        hip.synchronize()
        t0 = time.perf_counter()
        dec.decode_square(d_llr, out_bits=d_bits, out_stats=d_st)
        wall = time.perf_counter() - t0
        kms = dec.last_kernel_ms()[0]
        bits, st = d_bits.to_host(), d_st.to_host()
        be, bl = ber_bler(bits, info)
        rows.append({"ebn0_db": snr, "cw_s": round(n / wall, 1), "kernel_ms": round(kms, 2),
                     "ber": be / (64.0 * n), "ber_ci95": wilson(be, 64 * n), "bler": bl / n,
                     "bler_ci95": wilson(bl, n), "mean_half_iters": float(st["half_iters"].mean()),
                     "converged": float(st["converged"].mean()),
                     "mean_NG": float(st["NG"].astype(np.float64).mean())})
        log(f"square {snr} dB: {n / wall:.0f} cw/s, BLER {bl / n:.3e}")
        for x in (d_llr, d_bits, d_st):
            x.free()
    return {"metric": "codewords/s of the (16,8)^2 square code", "unit": "codewords/s",
            "n_gpus": 1, "dtype": "f64",
            "data": "synthetic (numpy-seeded info bits, (16,8)^2 encoding, BPSK, AWGN at rate 1/4)",
            "config": {"workload": f"{n} squares at 1.0, 2.0 and 3.0 dB, L=3, Imax=30",
                       "n_cw": n}, "rows": rows}


# ---------------------------------------------------------------------------
# HBM counter calibration


def workload_latency(args) -> dict:
    """Time to decode one small frame alone, per backend: frames of 1-4096
    codewords at 1 and 2 dB, device-resident, each decoded three times (the
    fastest kernel time, HIP events on the launch stream); bits and statistics
    checked identical across the backends."""
    import sogrand_amd as sg
    from sogrand_amd import hip

    dec = sg.Decoder(0)
    names = ((sg.BACKEND_PLAIN, "plain"), (sg.BACKEND_PERSISTENT, "persistent"),
             (sg.BACKEND_LANEGROUP, "lanegroup"))
    rows = []
    for ebn0 in (1.0, 2.0):
        for n in (1, 4, 16, 64, 256, 1024, 4096):
            llr = gen_device(dec, args.seed + n, n, sg.sigma_for_ebn0(ebn0))
            bits = hip.DeviceArray((n, 64), np.uint8)
            st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
            row = {"ebn0_db": ebn0, "n_cw": n}
            ref = None
            for b, name in names:
                dec.set_params(backend=b)
                best = None
                for _ in range(3):
                    dec.decode(llr, out_bits=bits, out_stats=st)
                    ms = dec.last_kernel_ms()[0]
                    best = ms if best is None else min(best, ms)
                got = (bits.to_host().copy(), st.to_host().copy())
                if ref is None:
                    ref = got
                elif not ((got[0] == ref[0]).all() and (got[1] == ref[1]).all()):
                    raise RuntimeError(f"backend {name} differs at {n} codewords, {ebn0} dB")
                row[name + "_ms"] = round(best, 3)
            row["NG_total"] = int(ref[1]["NG"].sum())  # guesses of the frame, every backend
            log(f"{ebn0} dB, {n} cw: {row}")
            rows.append(row)
    return {"metric": "frame decode time by backend (kernel, HIP events)", "unit": "ms",
            "n_gpus": 1, "dtype": "f64", "data": "synthetic (seeded AWGN frames)",
            "config": {"workload": "one frame alone, 1-4096 codewords at 1 and 2 dB, FP64"},
            "identical_across_backends": True, "rows": rows}


def workload_copy(args) -> dict:
    """Three device-to-device copies of 1 GiB: a kernel of known traffic (1 GiB
    read + 1 GiB written each) that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import sogrand_amd as sg  # noqa: F401  (loads the HIP runtime)
    from sogrand_amd import hip

    nbytes = 1 << 30
    a = hip.DeviceArray((nbytes,), np.uint8)
    b = hip.DeviceArray((nbytes,), np.uint8)
    a.zero()
    for _ in range(3):
        hip.copy(b, a, nbytes)
    hip.synchronize()
    return {"metric": "copy calibration", "bytes_read_per_copy": nbytes,
            "bytes_written_per_copy": nbytes, "copies": 3}


def kernel_source_hash() -> str:
    """SHA-256 over the decoder's device and host sources and build flags: the
    identity of the build a counter run measured."""
    import hashlib

    from sogrand_amd import build as b

    # flags with the tree's own paths taken out, so the hash names the same
    # build wherever the tree sits
    flags = [f.replace(ROOT, ".") for f in b.COMMON]
    h = hashlib.sha256(" ".join(flags).encode())
    for name in sorted(os.listdir(b.CSRC)):
        if name.endswith((".hip", ".h", ".cpp")):
            with open(os.path.join(b.CSRC, name), "rb") as f:
                h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()


def measured_traffic(n: int, ebn0: float):
    """HBM bytes per launch of the persistent kernel on this workload, from the
    committed counter runs (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel
    of known size) -> (bytes or None, note). None for another workload, or when
    the record was measured on other sources than the ones built here."""
    path = os.path.join(ROOT, "profiles", "persistent_100k_cw_2dB_traffic.json")
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None, "no counter record"
    if t.get("n_cw") != n or t.get("ebn0_db") != ebn0:
        return None, "no counter record for this workload"
    if t.get("source_sha256") != kernel_source_hash():
        return None, "counter record is stale (measured on other decoder sources)"
    return round(t["traffic_bytes_per_launch"]), f"profiles/{os.path.basename(path)}"


# ---------------------------------------------------------------------------
# headline


def workload_headline(args) -> dict:
    import sogrand_amd as sg
    from sogrand_amd import hip

    n = args.n_cw
    dec = sg.Decoder(0, backend=sg.BACKEND_PERSISTENT)
    d_llr = gen_device(dec, args.seed, n, sg.sigma_for_ebn0(args.ebn0))
    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_stats = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    log(f"generated {n} codewords at {args.ebn0} dB on the device")

    def run(backend: int) -> tuple[float, float]:
        dec.set_params(backend=backend)
        try:
            return timed_decode(dec, d_llr, d_bits, d_stats)
        finally:
            dec.set_params(backend=sg.BACKEND_PERSISTENT)

    for w in range(args.warmup):
        run(sg.BACKEND_PERSISTENT)
        log(f"warmup {w + 1}/{args.warmup} done")
    walls, kmss = [], []
    for s in range(args.steps):
        wall, kms = run(sg.BACKEND_PERSISTENT)
        walls.append(wall)
        kmss.append(kms)
        log(f"step {s + 1}/{args.steps}: {wall * 1e3:.1f} ms wall, {kms:.1f} ms kernel, "
            f"{n / wall:.0f} cw/s")
    bits_b = d_bits.to_host()
    st = d_stats.to_host()
    ms_per_step = 1e3 * sum(walls) / len(walls)
    value = n / (ms_per_step / 1e3)

    plain = None
    same_bits = None
    if not args.no_plain:
        wall_p, _ = run(sg.BACKEND_PLAIN)
        plain = n / wall_p
        same_bits = bool((d_bits.to_host() == bits_b).all())
        same_stats = bool((d_stats.to_host() == st).all())
        log(f"plain: {wall_p * 1e3:.1f} ms wall -> {plain:.0f} cw/s; identical bits {same_bits}, "
            f"identical stats {same_stats}")
    d_llr.free()

    info = sg.gen_info_bits(args.seed, n)
    bit_err, blk_err = ber_bler(bits_b, info)
    ng = float(st["NG"].astype(np.float64).sum())
    lines = 256.0 * float(st["half_iters"].astype(np.float64).sum())
    kernel_s = (sum(kmss) / len(kmss)) / 1e3
    achieved = (C_GUESS * ng + C_LINE * lines) / kernel_s
    traffic, traffic_note = measured_traffic(n, args.ebn0)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "codewords/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / PUBLISHED_CW_S, 1),
        "dtype": "f64",
        "data": "synthetic (seeded info bits, systematic (16,8)^3 encoding, BPSK, AWGN)",
        "config": {"workload": f"plain drop-in configuration: {n} codewords at Eb/N0 "
                               f"{args.ebn0} dB, L=3, Imax=30, Tmax=unlimited, FP64",
                   "n_cw": n, "ebn0_db": args.ebn0, "L": 3, "Imax": 30, "Tmax": "unlimited",
                   "backend": "persistent"},
        "plain_value": round(plain, 1) if plain else None,
        "batched_over_plain": round(value / plain, 3) if plain else None,
        "ber_matched": same_bits,
        "info_mbit_s": round(value * 512 / 1e6, 3),
        "guesses_per_s": round(ng / kernel_s, 1),
        "ber": bit_err / (512.0 * n),
        "bler": blk_err / float(n),
        "mean_half_iters": float(st["half_iters"].mean()),
        "mean_NG": ng / n,
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved / 1e12, 3),
            "peak": round(VALU_PEAK / 1e12, 2),
            "unit": "Tops/s (VALU lane-ops)",
            "frac": round(achieved / VALU_PEAK, 4),
            "traffic": traffic,
            "traffic_source": traffic_note,
            "per_unit": {"lane_ops_per_guess": C_GUESS, "lane_ops_per_line_decode": C_LINE},
            "kernel_ms": round(kernel_s * 1e3, 2),
            # the same launch against HBM, to show that bandwidth does not bound
            # it: the measured bytes (traffic) and the algorithmic stream (32 KB
            # of LLRs in, 64 B of bits out per codeword) over the launch time
            "hbm": {
                "measured_GB_s": round(traffic / kernel_s / 1e9, 1) if traffic else None,
                "algorithmic_GB_s": round(n * (4096 * 8 + 64) / kernel_s / 1e9, 1),
                "peak_GB_s": HBM_PEAK / 1e9,
                "frac": round(traffic / kernel_s / HBM_PEAK, 4) if traffic else None,
            },
        },
    }
    if not args.no_serve:
        cap = serving_capacity(dec, args.seed + 1)
        r = serve_load(dec, cap, 0.7, args.serve_seconds, args.seed + 2, args.serve_slots,
                       args.serve_batch, args.serve_plain_below, args.serve_hold_above,
                       args.serve_lanegroup_max, args.serve_small_batch, args.serve_backlog)
        log(f"serving at 70 % of {cap:.0f} cw/s: {r}")
        out["latency"] = {"p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
                          "workload": "serving loop, Poisson frames of U{1..256} codewords at "
                                      "Eb/N0 ~ U[1, 3] dB, offered 70 % of batched capacity",
                          "capacity_cw_s": round(cap, 1), "sustained_cw_s": r["sustained_cw_s"],
                          "frames": r["frames"],
                          "backend": "persistent" if dec.params.backend == sg.BACKEND_PERSISTENT
                          else str(dec.params.backend)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.seed, args.ebn0, args.cpu_seconds)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "snr-sweep", "uneven", "mixed-precision", "serving", "square",
                             "variants", "latency", "copy"])
    ap.add_argument("--n-cw", type=int, default=None)
    ap.add_argument("--ebn0", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-slots", type=int, default=16)
    ap.add_argument("--serve-batch", type=int, default=0)
    ap.add_argument("--serve-plain-below", type=int, default=0)
    ap.add_argument("--serve-hold-above", type=int, default=0)
    ap.add_argument("--serve-lanegroup-max", type=int, default=0,
                    help="a batch runs the lane-group kernel while it and the batches still "
                         "decoding hold at most this many codewords (0 = never)")
    ap.add_argument("--serve-capacity", type=float, default=0.0,
                    help="serving workload: offered loads relative to this rate in codewords/s "
                         "(0 = measure the batched capacity first)")
    ap.add_argument("--serve-loads", default="0.1,0.3,0.5,0.7,0.9,1.0,1.2",
                    help="serving workload: the offered loads, as fractions of the capacity")
    ap.add_argument("--serve-small-batch", type=int, default=None,
                    help="batch size while the waiting queue is short (library default 256, "
                         "0 = off)")
    ap.add_argument("--serve-backlog", type=int, default=0,
                    help="waiting codewords above which batches take everything that fits "
                         "(0 = library default, 8 x small batch)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the decoder shards codewords across GPUs with no "
              f"exchange step, and the bench runs on a single MI355X (asked for {args.gpus}).",
              file=sys.stderr)
        return 2
    args.n_cw_set = args.n_cw is not None
    if args.n_cw is None:
        args.n_cw = 100_000
    fn = {"headline": workload_headline, "uneven": workload_uneven,
          "mixed-precision": workload_mixed, "serving": workload_serving, "square": workload_square,
          "variants": workload_variants, "snr-sweep": workload_snr_sweep,
          "latency": workload_latency, "copy": workload_copy}[args.workload]
    print(json.dumps(fn(args)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
