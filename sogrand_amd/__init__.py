"""sogrand_amd - SOGRAND decoder for the (16,8)^3 cubic product code on AMD MI355X (gfx950).

Python view of the C ABI in include/sogrand.h (libsogrand.so, built in-tree by
``python -m sogrand_amd.build``). The decoder itself is HIP code; this module
only moves pointers. There is no CPU fallback: if the library is missing or no
gfx950 device is present, every entry point raises.

Mirrors the reference decoder's interface:
  Decoder.decode(llr)      ~ ``cubic_decoder <llr_file> <out_file>`` (SOGRAND_C/cubic_decoder_final.c:62-314)
  Decoder.bitso(lines)     ~ SOGRAND_bitSO (cubic_decoder_final.c:521)
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsogrand.so")
CLI_PATH = os.path.join(PKG, "cubic_decoder")

OK = 0
EINVAL = -1
ENOMEM = -2
EHIP = -3
ENONFINITE = -4
ENODEV = -5
EIO = -6
EBUSY = -7

FP64 = 0
FP32_MIXED = 1
VARIANT_FINAL = 0   # SOGRAND_C/cubic_decoder_final.c (and _updated.c via epsilon)
VARIANT_LEGACY = 1  # SOGRAND_C/cubic_decoder.c, the earlier decoder (FP64 only)
BACKEND_PLAIN = 0
BACKEND_PERSISTENT = 1
BACKEND_LANEGROUP = 2  # cube in LDS, four lanes per line (FP64, final variant)
BACKEND_AUTO = 3  # plain or persistent (one or two cube slots per wave) by batch size
UINT64_MAX = (1 << 64) - 1

CW_STATS_DTYPE = np.dtype(
    [("half_iters", "<u4"), ("converged", "<u4"), ("NG", "<u8"), ("NG_p", "<u8")]
)

# The C-ABI symbols include/sogrand.h declares.
EXPORTED = [
    "sogrand_params_default", "sogrand_strerror", "sogrand_ctx_create", "sogrand_ctx_set_params",
    "sogrand_ctx_destroy", "sogrand_decode_cubic", "sogrand_decode_cubic_noalloc",
    "sogrand_decode_square",
    "sogrand_bitso_lines", "sogrand_last_kernel_ms",
    "sogrand_graph_replays",
    "sogrand_sigma_for_ebn0", "sogrand_gen_info_bits", "sogrand_encode_cubes", "sogrand_gen_llr",
    "sogrand_gen_llr_device", "sogrand_schedule_table", "sogrand_server_cfg_default",
    "sogrand_server_create", "sogrand_server_batch_cw", "sogrand_server_submit",
    "sogrand_server_poll", "sogrand_server_drain", "sogrand_server_clock",
    "sogrand_server_destroy",
]


class SograndError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = _strerror(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


class Params(C.Structure):
    _fields_ = [
        ("L", C.c_uint32),
        ("Imax", C.c_uint32),
        ("Tmax", C.c_uint64),
        ("p_ET", C.c_double),
        ("alpha", C.c_double),
        ("precision", C.c_int32),
        ("backend", C.c_int32),
        ("wave_slots", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("lane_batch", C.c_int32),
        ("step_len", C.c_int32),
        ("use_graph", C.c_int32),
        ("epsilon", C.c_double),
        ("variant", C.c_int32),
    ]


class ServerCfg(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_size_t),
        ("max_inflight", C.c_uint32),
        ("max_batch_cw", C.c_uint32),
        ("max_pending_cw", C.c_uint32),
        ("with_soft", C.c_int32),
        ("plain_below_cw", C.c_uint32),
        ("hold_above_cw", C.c_uint32),
        ("lanegroup_max_cw", C.c_uint32),
        ("small_batch_cw", C.c_uint32),
        ("backlog_cw", C.c_uint32),
    ]


class Completion(C.Structure):
    _fields_ = [
        ("id", C.c_uint64),
        ("status", C.c_int32),
        ("n_cw", C.c_uint32),
        ("t_submit", C.c_double),
        ("t_start", C.c_double),
        ("t_done", C.c_double),
    ]


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libsogrand.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -m sogrand_amd.build` (hipcc, gfx950)"
        )
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.sogrand_params_default.argtypes = [C.POINTER(Params)]
    lib.sogrand_params_default.restype = None
    lib.sogrand_strerror.argtypes = [C.c_int]
    lib.sogrand_strerror.restype = C.c_char_p
    lib.sogrand_ctx_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(Params)]
    lib.sogrand_ctx_set_params.argtypes = [vp, C.POINTER(Params)]
    lib.sogrand_ctx_destroy.argtypes = [vp]
    lib.sogrand_ctx_destroy.restype = None
    lib.sogrand_decode_cubic.argtypes = [vp, vp, C.c_size_t, vp, vp, vp, vp]
    lib.sogrand_bitso_lines.argtypes = [vp, vp, C.c_size_t, vp, vp, vp, vp]
    lib.sogrand_decode_square.argtypes = [vp, vp, C.c_size_t, vp, vp, vp, vp]
    lib.sogrand_decode_cubic_noalloc.argtypes = [vp, vp, C.c_size_t, vp, vp, vp, vp]
    lib.sogrand_last_kernel_ms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
    lib.sogrand_graph_replays.argtypes = [vp, C.POINTER(C.c_uint32)]
    lib.sogrand_sigma_for_ebn0.argtypes = [C.c_double]
    lib.sogrand_sigma_for_ebn0.restype = C.c_double
    lib.sogrand_gen_info_bits.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, vp]
    lib.sogrand_encode_cubes.argtypes = [vp, C.c_size_t, vp]
    lib.sogrand_gen_llr.argtypes = [C.c_uint64, C.c_uint64, C.c_size_t, vp, C.c_size_t, vp]
    lib.sogrand_gen_llr_device.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_size_t, vp, C.c_size_t,
                                           vp, vp]
    lib.sogrand_schedule_table.argtypes = [C.c_int32, vp]
    lib.sogrand_server_cfg_default.argtypes = [C.POINTER(ServerCfg)]
    lib.sogrand_server_cfg_default.restype = None
    lib.sogrand_server_create.argtypes = [C.POINTER(vp), vp, C.POINTER(ServerCfg)]
    lib.sogrand_server_batch_cw.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.sogrand_server_submit.argtypes = [vp, C.c_uint64, vp, C.c_size_t, vp, vp, vp]
    lib.sogrand_server_poll.argtypes = [vp, C.POINTER(Completion), C.c_size_t,
                                        C.POINTER(C.c_size_t), C.c_int]
    lib.sogrand_server_drain.argtypes = [vp]
    lib.sogrand_server_clock.argtypes = [vp]
    lib.sogrand_server_clock.restype = C.c_double
    lib.sogrand_server_destroy.argtypes = [vp]
    lib.sogrand_server_destroy.restype = None
    _lib = lib
    return lib


def _strerror(code: int) -> str:
    try:
        return load_library().sogrand_strerror(code).decode()
    except ImportError:
        return "error"


def _check(rc: int, what: str) -> None:
    if rc != OK:
        raise SograndError(rc, what)


def default_params(**kw) -> Params:
    p = Params()
    load_library().sogrand_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _ptr(a) -> int | None:
    """Address of a numpy array or torch tensor (host or device), or None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return a.data_ptr()
    raise TypeError(type(a))


@dataclass
class DecodeResult:
    bits: object       # (n, 64) uint8, packed MSB-first
    stats: object      # (n,) CW_STATS_DTYPE, or None
    soft: object       # (n, 4096) float64 final L_APP, or None


class Decoder:
    """A decoder context on one gfx950 device (sogrand_ctx)."""

    def __init__(self, device: int = 0, **params):
        lib = load_library()
        self._lib = lib
        self.params = default_params(**params)
        h = C.c_void_p()
        _check(lib.sogrand_ctx_create(C.byref(h), device, C.byref(self.params)), "ctx_create")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sogrand_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, **kw) -> None:
        p = Params.from_buffer_copy(self.params)
        for k, v in kw.items():
            setattr(p, k, v)
        _check(self._lib.sogrand_ctx_set_params(self._h, C.byref(p)), "set_params")
        self.params = p

    def decode(self, llr, soft: bool = False, stats: bool = True, out_bits=None, out_stats=None,
               out_soft=None, stream=None) -> DecodeResult:
        """Decode cubes. llr: (n, 4096) float64, host (numpy) or device (torch) memory.
        Outputs are allocated like the input (numpy for numpy) unless given."""
        if isinstance(llr, np.ndarray):
            llr = np.ascontiguousarray(llr, dtype=np.float64).reshape(-1, 4096)
            n = llr.shape[0]
            bits = out_bits if out_bits is not None else np.zeros((n, 64), np.uint8)
            st = out_stats if out_stats is not None else (np.zeros(n, CW_STATS_DTYPE) if stats else None)
            sf = out_soft if out_soft is not None else (np.zeros((n, 4096)) if soft else None)
        else:
            n = llr.numel() // 4096
            bits, st, sf = out_bits, out_stats, out_soft
            if bits is None:
                raise ValueError("device decode needs out_bits (and optional out_stats/out_soft)")
        rc = self._lib.sogrand_decode_cubic(self._h, _ptr(llr), n, _ptr(bits), _ptr(sf), _ptr(st),
                                            stream)
        _check(rc, "decode_cubic")
        return DecodeResult(bits, st, sf)

    def decode_square(self, llr, soft: bool = False, stats: bool = True, out_bits=None,
                      out_stats=None, out_soft=None, stream=None) -> DecodeResult:
        """Decode squares of the (16,8)^2 code (sogrand_decode_square). llr: (n, 256)
        float64, host (numpy) or device (torch) memory; bits (n, 8)."""
        if isinstance(llr, np.ndarray):
            llr = np.ascontiguousarray(llr, dtype=np.float64).reshape(-1, 256)
            n = llr.shape[0]
            bits = out_bits if out_bits is not None else np.zeros((n, 8), np.uint8)
            st = out_stats if out_stats is not None else (np.zeros(n, CW_STATS_DTYPE) if stats else None)
            sf = out_soft if out_soft is not None else (np.zeros((n, 256)) if soft else None)
        else:
            n = llr.numel() // 256
            bits, st, sf = out_bits, out_stats, out_soft
            if bits is None:
                raise ValueError("device decode needs out_bits (and optional out_stats/out_soft)")
        rc = self._lib.sogrand_decode_square(self._h, _ptr(llr), n, _ptr(bits), _ptr(sf),
                                             _ptr(st), stream)
        _check(rc, "decode_square")
        return DecodeResult(bits, st, sf)

    def decode_noalloc(self, llr, out_bits, out_stats=None, out_soft=None, stream=None) -> None:
        """sogrand_decode_cubic_noalloc: device buffers only, nothing allocated."""
        n = llr.numel() // 4096
        _check(self._lib.sogrand_decode_cubic_noalloc(self._h, _ptr(llr), n, _ptr(out_bits),
                                                      _ptr(out_soft), _ptr(out_stats), stream),
               "decode_cubic_noalloc")

    def bitso(self, lines):
        """Component SISO on (n, 16) lines -> (L_APP, L_E, N_guess)."""
        lines = np.ascontiguousarray(lines, dtype=np.float64).reshape(-1, 16)
        n = lines.shape[0]
        la = np.zeros((n, 16))
        le = np.zeros((n, 16))
        ng = np.zeros(n, np.uint32)
        rc = self._lib.sogrand_bitso_lines(self._h, lines.ctypes.data, n, la.ctypes.data,
                                           le.ctypes.data, ng.ctypes.data, None)
        _check(rc, "bitso_lines")
        return la, le, ng

    def last_kernel_ms(self) -> tuple[float, int]:
        ms = C.c_double()
        k = C.c_uint32()
        _check(self._lib.sogrand_last_kernel_ms(self._h, C.byref(ms), C.byref(k)), "last_kernel_ms")
        return ms.value, k.value

    def graph_replays(self) -> int:
        k = C.c_uint32()
        _check(self._lib.sogrand_graph_replays(self._h, C.byref(k)), "graph_replays")
        return k.value

    def gen_llr_device(self, seed: int, cw_offset: int, n: int, sigma, out, stream=None) -> None:
        sig = np.ascontiguousarray(np.atleast_1d(sigma), dtype=np.float64) if not hasattr(
            sigma, "data_ptr") else sigma
        ns = sig.shape[0] if isinstance(sig, np.ndarray) else sig.numel()
        _check(self._lib.sogrand_gen_llr_device(self._h, seed, cw_offset, n, _ptr(sig), ns,
                                                _ptr(out), stream), "gen_llr_device")


class Server:
    """Continuous-batching serving loop on a Decoder (sogrand_server).

    submit() queues a frame of codewords (host or device LLRs) and returns at
    once; poll() returns completion records (id, status, n_cw and the server
    clock at submit, launch and completion). Output arrays passed to submit()
    are filled before the frame's completion is reported; the caller keeps the
    input and outputs alive until then."""

    def __init__(self, decoder: Decoder, hbm_budget: int | None = None, max_inflight: int = 16,
                 max_batch_cw: int = 0, max_pending_cw: int = 1 << 20, with_soft: bool = False,
                 plain_below_cw: int = 0, hold_above_cw: int = 0, lanegroup_max_cw: int = 0,
                 small_batch_cw: int | None = None, backlog_cw: int = 0):
        lib = decoder._lib
        self._lib = lib
        cfg = ServerCfg()
        lib.sogrand_server_cfg_default(C.byref(cfg))
        if hbm_budget is not None:
            cfg.hbm_budget = hbm_budget
        cfg.max_inflight = max_inflight
        cfg.max_batch_cw = max_batch_cw
        cfg.max_pending_cw = max_pending_cw
        cfg.with_soft = 1 if with_soft else 0
        cfg.plain_below_cw = plain_below_cw
        cfg.hold_above_cw = hold_above_cw
        cfg.lanegroup_max_cw = lanegroup_max_cw
        if small_batch_cw is not None:
            cfg.small_batch_cw = small_batch_cw
        cfg.backlog_cw = backlog_cw
        h = C.c_void_p()
        _check(lib.sogrand_server_create(C.byref(h), decoder._h, C.byref(cfg)), "server_create")
        self._h = h
        self._keep = {}
        b = C.c_size_t()
        _check(lib.sogrand_server_batch_cw(h, C.byref(b)), "server_batch_cw")
        self.batch_cw = b.value

    def submit(self, rid: int, llr, info_out, stats_out=None, soft_out=None) -> int:
        """Queue frame rid; returns 0, or EBUSY when the admission queue is full."""
        n = llr.shape[0] if isinstance(llr, np.ndarray) else llr.numel() // 4096
        rc = self._lib.sogrand_server_submit(self._h, rid, _ptr(llr), n, _ptr(info_out),
                                             _ptr(soft_out), _ptr(stats_out))
        if rc == EBUSY:
            return rc
        _check(rc, "server_submit")
        self._keep[rid] = (llr, info_out, stats_out, soft_out)
        return OK

    def poll(self, max_n: int = 1024, wait_ms: int = 0) -> list:
        buf = (Completion * max_n)()
        k = C.c_size_t()
        _check(self._lib.sogrand_server_poll(self._h, buf, max_n, C.byref(k), wait_ms),
               "server_poll")
        out = []
        for i in range(k.value):
            c = buf[i]
            self._keep.pop(c.id, None)
            out.append(c)
        return out

    def drain(self) -> None:
        _check(self._lib.sogrand_server_drain(self._h), "server_drain")

    def clock(self) -> float:
        return self._lib.sogrand_server_clock(self._h)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sogrand_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---- host-side data helpers (no GPU needed) ----

def schedule_table(ic: int) -> np.ndarray:
    """The 65535-pattern schedule for intercept ic (plus the 0 end marker)."""
    out = np.zeros(65536, np.uint16)
    _check(load_library().sogrand_schedule_table(ic, out.ctypes.data), "schedule_table")
    return out


def sigma_for_ebn0(ebn0_db: float, dim: int = 3) -> float:
    """Noise sigma of SOGRAND_C/channel_sim.c:32-36 for the cubic (dim 3) or
    square (dim 2) code: rate (8/16)^dim, Es/N0 = Eb/N0 + 10 log10(2 R)."""
    if dim == 3:
        return load_library().sogrand_sigma_for_ebn0(ebn0_db)
    if dim != 2:
        raise ValueError("dim must be 2 or 3")
    r = (8.0 / 16.0) ** 2
    esn0 = ebn0_db + 10.0 * math.log10(2.0 * r)
    return 1.0 / math.sqrt(10.0 ** (esn0 / 10.0))


def gen_info_bits(seed: int, n: int, cw_offset: int = 0) -> np.ndarray:
    out = np.zeros((n, 64), np.uint8)
    _check(load_library().sogrand_gen_info_bits(seed, cw_offset, n, out.ctypes.data), "gen_info")
    return out


def encode_cubes(info_bits: np.ndarray) -> np.ndarray:
    info_bits = np.ascontiguousarray(info_bits, dtype=np.uint8).reshape(-1, 64)
    out = np.zeros((info_bits.shape[0], 512), np.uint8)
    _check(load_library().sogrand_encode_cubes(info_bits.ctypes.data, info_bits.shape[0],
                                               out.ctypes.data), "encode")
    return out


def encode_squares(info_bits: np.ndarray) -> np.ndarray:
    """Codewords of the (16,8)^2 square code: (n, 8) info bytes (bit b = a0 + 8 a1,
    MSB first) -> (n, 256) code bits by position a0 + 16 a1. A square codeword is
    the plane a2 = 0 of the cubic codeword whose other info planes are zero (the
    product code's encoding leaves an information plane a square codeword)."""
    info_bits = np.ascontiguousarray(info_bits, dtype=np.uint8).reshape(-1, 8)
    cube_info = np.zeros((info_bits.shape[0], 64), np.uint8)
    cube_info[:, :8] = info_bits
    code = np.unpackbits(encode_cubes(cube_info), axis=1)  # bit index a0 + 16 a1 + 256 a2
    return code[:, :256].copy()


def gen_llr_square(seed: int, n: int, ebn0_db=None, sigma=None):
    """(info (n, 8), LLRs (n, 256)) of n random squares through BPSK and AWGN
    (numpy's generator seeded with seed; LLR = 2 y / sigma^2 as in
    channel_sim.c). Give ebn0_db (scalar or per-square array) or sigma."""
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 256, (n, 8), dtype=np.uint8)
    if sigma is None:
        e = np.broadcast_to(np.asarray(ebn0_db, dtype=np.float64), (n,))
        sigma = np.array([sigma_for_ebn0(float(x), dim=2) for x in e])
    sig = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n,))[:, None]
    x = 1.0 - 2.0 * encode_squares(info).astype(np.float64)
    y = x + sig * rng.standard_normal((n, 256))
    return info, 2.0 * y / (sig * sig)


def gen_llr(seed: int, n: int, ebn0_db=None, sigma=None, cw_offset: int = 0) -> np.ndarray:
    """LLRs of n codewords (host). Give ebn0_db (scalar or per-codeword array) or sigma."""
    if sigma is None:
        e = np.atleast_1d(np.asarray(ebn0_db, dtype=np.float64))
        sigma = np.array([sigma_for_ebn0(float(x)) for x in e])
    sig = np.ascontiguousarray(np.atleast_1d(sigma), dtype=np.float64)
    out = np.zeros((n, 4096))
    _check(load_library().sogrand_gen_llr(seed, cw_offset, n, sig.ctypes.data, sig.shape[0],
                                          out.ctypes.data), "gen_llr")
    return out
