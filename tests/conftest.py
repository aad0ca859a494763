import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

# Soft outputs: relative 1e-9 with an absolute floor of 1e-9. An L_APP value is
# a difference of two logarithms, so its rounding error is absolute near zero.
SOFT_RTOL = 1e-9


def soft_close(a, b) -> bool:
    a = np.asarray(a)
    b = np.asarray(b)
    return bool(np.all(np.abs(a - b) <= SOFT_RTOL * np.maximum(np.abs(b), 1.0)))


def soft_err(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)))


def chain_cond(T, pnl, n_list):
    """Relative rounding-error bound of the reference's 1 - found-mass term at a
    line's last list insertion: T ordered subtractions from 1.0 of increments
    whose own error is a few ulps, divided by the value (og_chain_cond)."""
    T = np.asarray(T, dtype=np.float64)
    P = 256.0 * np.abs(np.asarray(pnl, dtype=np.float64))
    with np.errstate(divide="ignore"):
        k = (T + 64.0) * 2.0 ** -50 / P
    return np.where(np.asarray(n_list) > 0, k, 0.0)


def within_chain_bound(a, b, kappa) -> np.ndarray:
    """|a - b| <= 1e-9 * max(|b|, 1) + 4 * kappa, per value (kappa broadcast per line)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b) <= SOFT_RTOL * np.maximum(np.abs(b), 1.0) + 4.0 * kappa


def stable_codewords(st_a, st_b, bits_a, bits_b) -> np.ndarray:
    """Codewords on which two valid evaluations of the reference arithmetic agree."""
    same = (bits_a == bits_b).all(axis=1)
    for f in ("half_iters", "NG", "NG_p"):
        same &= st_a[f] == st_b[f]
    return same


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name: str):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} not recorded")
    return np.load(path, allow_pickle=False)


@pytest.fixture(scope="session")
def oracle():
    from oracle import ORACLE_SO, Oracle

    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"],
                       check=True)
    return Oracle()


@pytest.fixture(scope="session")
def oracle_fma():
    """The oracle built with fused multiply-adds: a second valid evaluation."""
    from oracle import ORACLE_FMA_SO, Oracle

    if not os.path.exists(ORACLE_FMA_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_fma.so"],
                       check=True)
    return Oracle(ORACLE_FMA_SO)


@pytest.fixture(scope="session")
def lib():
    import sogrand_amd

    return sogrand_amd.load_library()


@pytest.fixture(scope="module")
def decoder():
    import sogrand_amd as sg

    d = sg.Decoder(0)
    yield d
    d.close()
