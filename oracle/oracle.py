"""ctypes bindings for the CPU oracle (liboracle.so) and, where it was built, the
reference harness (_ref/libsogrand_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product never imports this.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
ORACLE_FMA_SO = os.path.join(HERE, "liboracle_fma.so")
REF_SO = os.path.join(HERE, "_ref", "libsogrand_ref.so")
REF_SO_NATIVE = os.path.join(HERE, "_ref", "libsogrand_ref_native.so")
# the reference's EPSILON = DBL_EPSILON variant (cubic_decoder_updated.c)
REF_SO_UPDATED = os.path.join(HERE, "_ref", "libsogrand_ref_updated.so")
REF_BIN_UPDATED = os.path.join(HERE, "_ref", "cubic_decoder_updated")
REF_BIN_UPDATED_NATIVE = os.path.join(HERE, "_ref", "cubic_decoder_updated_native")
DBL_EPSILON = 2.220446049250313e-16
# the reference's earlier decoder (cubic_decoder.c)
REF_SO_LEGACY = os.path.join(HERE, "_ref", "libsogrand_ref_legacy.so")
REF_BIN_LEGACY = os.path.join(HERE, "_ref", "cubic_decoder_legacy")
REF_BIN_LEGACY_NATIVE = os.path.join(HERE, "_ref", "cubic_decoder_legacy_native")
REF_BIN_NATIVE = os.path.join(HERE, "_ref", "cubic_decoder_final_native")
REF_BIN = os.path.join(HERE, "_ref", "cubic_decoder_final")
REF_BIN_O2 = os.path.join(HERE, "_ref", "cubic_decoder_final_O2")
REF_ENCODER = os.path.join(HERE, "_ref", "cubic_encoder")

UINT64_MAX = (1 << 64) - 1
MAXL = 8

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class LineTrace(C.Structure):
    _fields_ = [
        ("ic", C.c_int32),
        ("n_list", C.c_int32),
        ("T", C.c_uint32),
        ("pad0", C.c_uint32),
        ("pnl", C.c_double),
        ("pm", C.c_double * MAXL),
        ("conf", C.c_double * MAXL),
        ("t_hit", C.c_uint32 * MAXL),
        ("wt_hit", C.c_int32 * MAXL),
        ("cw", C.c_uint16 * MAXL),
        ("perm", C.c_uint8 * 16),
    ]


CW_STATS_DTYPE = np.dtype(
    [("half_iters", "<u4"), ("converged", "<u4"), ("NG", "<u8"), ("NG_p", "<u8")]
)


def build_oracle() -> None:
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so", "liboracle_fma.so"], check=True)


def build_reference() -> bool:
    """Compile the reference's own sources into _ref/ (only where they exist)."""
    import subprocess

    if not os.path.isdir("/root/reference/SOGRAND_C"):
        return False
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
    return True


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        lib = C.CDLL(path)
        lib.og_code_tables.argtypes = [_u8p, _u8p, _u8p]
        lib.og_code_tables.restype = None
        lib.og_sort_with_perm.argtypes = [_dp, _u8p, C.c_int]
        lib.og_sort_with_perm.restype = None
        lib.og_schedule_prefix.argtypes = [C.c_int32, _u16p, C.c_int]
        lib.og_schedule_prefix.restype = C.c_int
        lib.og_siso.argtypes = [_dp, C.c_int, C.c_uint64, C.c_double, C.c_void_p, C.c_void_p,
                                C.POINTER(LineTrace)]
        lib.og_siso.restype = C.c_int
        lib.og_cube_is_codeword.argtypes = [_dp]
        lib.og_cube_is_codeword.restype = C.c_int
        lib.og_decode_cubes.argtypes = [_dp, C.c_size_t, C.c_int, C.c_int, C.c_uint64, C.c_double,
                                        C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        lib.og_decode_cubes.restype = C.c_int
        lib.og_decode_cubes_cond.argtypes = [_dp, C.c_size_t, C.c_int, C.c_int, C.c_uint64,
                                             C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_int]
        lib.og_decode_cubes_cond.restype = C.c_int
        lib.og_square_is_codeword.argtypes = [_dp]
        lib.og_square_is_codeword.restype = C.c_int
        lib.og_decode_squares.argtypes = [_dp, C.c_size_t, C.c_int, C.c_int, C.c_uint64,
                                          C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_int]
        lib.og_decode_squares.restype = C.c_int
        lib.og_set_epsilon.argtypes = [C.c_double]
        lib.og_set_epsilon.restype = None
        lib.og_set_variant.argtypes = [C.c_int]
        lib.og_set_variant.restype = None
        self.lib = lib

    def set_epsilon(self, eps: float) -> None:
        """EPSILON of the soft output and confidence test, for every later call."""
        self.lib.og_set_epsilon(eps)

    def set_variant(self, variant: int) -> None:
        """0: cubic_decoder_final.c semantics; 1: the earlier cubic_decoder.c."""
        self.lib.og_set_variant(variant)

    def code_tables(self):
        G = np.zeros((8, 16), np.uint8)
        H = np.zeros((8, 16), np.uint8)
        hcol = np.zeros(16, np.uint8)
        self.lib.og_code_tables(G, H, hcol)
        return G, H, hcol

    def sort_with_perm(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        perm = np.arange(len(a), dtype=np.uint8)
        self.lib.og_sort_with_perm(a, perm, len(a))
        return a, perm

    def schedule(self, ic: int, count: int = 65535):
        out = np.zeros(count, np.uint16)
        n = self.lib.og_schedule_prefix(ic, out, count)
        return out[:n]

    def siso(self, llr, L=3, Tmax=UINT64_MAX, thres=1.0 - 1e-5):
        llr = np.ascontiguousarray(llr, dtype=np.float64)
        la = np.zeros(16)
        le = np.zeros(16)
        tr = LineTrace()
        rc = self.lib.og_siso(llr, L, Tmax, thres, la.ctypes.data, le.ctypes.data, C.byref(tr))
        if rc:
            raise ValueError("non-finite LLR")
        return la, le, tr

    @staticmethod
    def chain_cond(tr) -> float:
        """Relative rounding-error bound of the reference's P_notGuess for a line
        (same definition as og_chain_cond)."""
        if tr.n_list == 0:
            return 0.0
        P = 256.0 * tr.pnl
        return (tr.T + 64.0) * 2.0 ** -50 / abs(P) if P != 0 else float("inf")

    def is_codeword(self, cube) -> bool:
        return bool(self.lib.og_cube_is_codeword(np.ascontiguousarray(cube, dtype=np.float64)))

    def is_square_codeword(self, square) -> bool:
        return bool(self.lib.og_square_is_codeword(np.ascontiguousarray(square, dtype=np.float64)))

    def decode_squares(self, llr, L=3, Imax=30, Tmax=UINT64_MAX, p_et=1e-5, alpha=0.7,
                       soft=False, threads=0):
        """Decode squares of the (16,8)^2 code -> (bits (n, 8), stats, soft or None)."""
        llr = np.ascontiguousarray(llr, dtype=np.float64).reshape(-1, 256)
        n = llr.shape[0]
        bits = np.zeros((n, 8), np.uint8)
        stats = np.zeros(n, CW_STATS_DTYPE)
        sarr = np.zeros((n, 256)) if soft else None
        self.lib.og_decode_squares(llr, n, L, Imax, Tmax, p_et, alpha, bits.ctypes.data,
                                   sarr.ctypes.data if soft else None, stats.ctypes.data,
                                   threads)
        return bits, stats, sarr

    def decode(self, llr, L=3, Imax=30, Tmax=UINT64_MAX, p_et=1e-5, alpha=0.7, soft=False,
               threads=0, cond=False):
        """Decode cubes -> (bits, stats, soft or None) [+ cond if cond=True]."""
        llr = np.ascontiguousarray(llr, dtype=np.float64).reshape(-1, 4096)
        n = llr.shape[0]
        bits = np.zeros((n, 64), np.uint8)
        stats = np.zeros(n, CW_STATS_DTYPE)
        sarr = np.zeros((n, 4096)) if soft else None
        carr = np.zeros((n, 4096)) if cond else None
        rc = self.lib.og_decode_cubes_cond(llr, n, L, Imax, Tmax, p_et, alpha, bits.ctypes.data,
                                           sarr.ctypes.data if soft else None,
                                           carr.ctypes.data if cond else None, stats.ctypes.data,
                                           threads)
        if rc:
            raise ValueError("non-finite LLR")
        if cond:
            return bits, stats, sarr, carr
        return bits, stats, sarr


class Reference:
    """The reference's own functions (SOGRAND_C/cubic_decoder_final.c) via the harness."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        lib.refh_code_tables.argtypes = [_i32p, _i32p]
        lib.refh_siso.argtypes = [_dp, _dp, _dp, C.POINTER(C.c_int), C.c_uint64, C.c_double]
        lib.refh_main_logic.argtypes = [_dp, _u8p, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                                        C.c_double, _dp, _dp, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.refh_quicksort.argtypes = [_dp, _u64p, C.c_uint64]
        lib.refh_early_term.argtypes = [_dp]
        lib.refh_early_term.restype = C.c_int
        lib.refh_decode_cube.argtypes = [_dp, C.c_int, C.c_uint64, C.c_double, _dp,
                                         C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), _u8p]
        self.lib = lib

    def code_tables(self):
        G = np.zeros(128, np.int32)
        H = np.zeros(128, np.int32)
        self.lib.refh_code_tables(G, H)
        return G.reshape(8, 16), H.reshape(8, 16)

    def siso(self, llr, Tmax=UINT64_MAX, thres=1.0 - 1e-5):
        llr = np.ascontiguousarray(llr, dtype=np.float64)
        la = np.zeros(16)
        le = np.zeros(16)
        ng = C.c_int(0)
        self.lib.refh_siso(llr, la, le, C.byref(ng), Tmax, thres)
        return la, le, ng.value

    def main_logic(self, llr, H_flat, s, IC=-1, L=3, Tmax=UINT64_MAX, thres=1.0 - 1e-5):
        llr = np.ascontiguousarray(llr, dtype=np.float64)
        chat = np.zeros(16 * L)
        sl = np.zeros(4 * L)
        T = C.c_double(0)
        cl = C.c_double(0)
        pnl = C.c_double(0)
        self.lib.refh_main_logic(llr, np.ascontiguousarray(H_flat, dtype=np.uint8), s, IC, L,
                                 Tmax, thres, chat, sl, C.byref(T), C.byref(cl), C.byref(pnl))
        return chat.reshape(L, 16), sl.reshape(L, 4), T.value, int(cl.value), pnl.value

    def quicksort(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        perm = np.arange(len(a), dtype=np.uint64)
        self.lib.refh_quicksort(a, perm, len(a))
        return a, perm

    def early_term(self, cube) -> bool:
        return bool(self.lib.refh_early_term(np.ascontiguousarray(cube, dtype=np.float64)))

    def decode_cube(self, llr, Imax=30, Tmax=UINT64_MAX, alpha=0.7):
        llr = np.ascontiguousarray(llr, dtype=np.float64)
        la = np.zeros(4096)
        hi = C.c_int(0)
        ng = C.c_int64(0)
        ngp = C.c_int64(0)
        bits = np.zeros(512, np.uint8)
        self.lib.refh_decode_cube(llr, Imax, Tmax, alpha, la, C.byref(hi), C.byref(ng),
                                  C.byref(ngp), bits)
        return la, hi.value, ng.value, ngp.value, bits
