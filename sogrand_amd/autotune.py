"""Launch-shape autotuner of the persistent batched decoder.

The persistent kernel has four shape parameters that change only speed, never
results (every shape decodes every codeword identically, which the tuner
checks): cube slots per wave (`wave_slots`), how many free lanes a wave
gathers before handing out new lines (`lane_batch`), how many guesses a lane
runs between scheduling points (`step_len`) and how many waves per compute
unit stay resident (`waves_per_cu`). The tuner times the kernel on a sample of
the caller's workload (device-resident) by coordinate descent over a grid of
those values and returns the fastest shape. A sample holds only a few cubes per
resident cube slot, so the end of the batch, where slots run dry, weighs more
in it than in a large workload; given the whole workload (`confirm`), the tuner
times its pick and the default shape on it once each and keeps the faster.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

SPACE = {
    "wave_slots": [1, 2, 3, 4],
    "lane_batch": [1, 2, 4, 8, 16, 32],
    "step_len": [16, 32, 64, 128, 256, 512],
    "waves_per_cu": [0, 8, 4],
}
DEFAULT = {"wave_slots": 2, "lane_batch": 4, "step_len": 64, "waves_per_cu": 0}


def _time(dec, d_llr, d_bits, d_st, shape, repeats):
    from sogrand_amd import hip

    dec.set_params(**shape)
    best = None
    for _ in range(repeats):
        hip.synchronize()
        dec.decode(d_llr, out_bits=d_bits, out_stats=d_st)
        hip.synchronize()
        ms = dec.last_kernel_ms()[0]
        best = ms if best is None else min(best, ms)
    return best


def autotune(dec, d_llr, n: int, rounds: int = 2, repeats: int = 2, log=None,
             confirm=None) -> dict:
    """Tune the persistent kernel's shape on n device-resident codewords.

    confirm: optional (d_llr_all, n_all), the whole workload the shape is for.
    Returns {"best": shape, "best_ms": t, "default_ms": t0, "trials": [...]}
    (plus "confirm": the two whole-workload times, if it ran) and leaves `dec`
    set to the best shape (backend persistent)."""
    import sogrand_amd as sg
    from sogrand_amd import hip

    d_bits = hip.DeviceArray((n, 64), np.uint8)
    d_st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
    dec.set_params(backend=sg.BACKEND_PERSISTENT, **DEFAULT)
    t_default = _time(dec, d_llr, d_bits, d_st, DEFAULT, repeats)
    ref_bits, ref_st = d_bits.to_host(), d_st.to_host()
    best, best_ms = dict(DEFAULT), t_default
    trials = [{"shape": dict(DEFAULT), "ms": round(t_default, 3)}]
    seen = {json.dumps(DEFAULT, sort_keys=True)}
    for _ in range(rounds):
        improved = False
        for key, values in SPACE.items():
            for v in values:
                cand = dict(best)
                cand[key] = v
                tag = json.dumps(cand, sort_keys=True)
                if tag in seen:
                    continue
                seen.add(tag)
                ms = _time(dec, d_llr, d_bits, d_st, cand, repeats)
                if not ((d_bits.to_host() == ref_bits).all() and (d_st.to_host() == ref_st).all()):
                    raise RuntimeError(f"shape {cand} changed the decoded results")
                trials.append({"shape": cand, "ms": round(ms, 3)})
                if log:
                    log(f"autotune {cand}: {ms:.2f} ms")
                if ms < best_ms * 0.995:
                    best, best_ms, improved = cand, ms, True
        if not improved:
            break
    d_bits.free()
    d_st.free()
    out = {"best": best, "best_ms": round(best_ms, 3), "default_ms": round(t_default, 3),
           "speedup_vs_default": round(t_default / best_ms, 4), "n_cw": n, "trials": trials}
    if confirm is not None and best != DEFAULT:
        d_all, n_all = confirm
        a_bits = hip.DeviceArray((n_all, 64), np.uint8)
        a_st = hip.DeviceArray((n_all,), sg.CW_STATS_DTYPE)
        ms_best = _time(dec, d_all, a_bits, a_st, best, 1)
        ms_def = _time(dec, d_all, a_bits, a_st, DEFAULT, 1)
        a_bits.free()
        a_st.free()
        out["confirm"] = {"n_cw": n_all, "tuned_ms": round(ms_best, 3),
                          "default_ms": round(ms_def, 3)}
        if log:
            log(f"autotune confirm on {n_all} codewords: {best} {ms_best:.1f} ms, "
                f"default {ms_def:.1f} ms")
        if ms_def <= ms_best:
            out["best"] = best = dict(DEFAULT)
    dec.set_params(**best)
    return out


def save(result: dict, path: str) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path: str) -> dict | None:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None
