"""Build the gfx950 library (libsogrand.so) and the drop-in CLI (cubic_decoder) in-tree.

    python -m sogrand_amd.build

Everything is compiled with hipcc for gfx950 only, with -ffp-contract=off (no
fused multiply-add anywhere: the decoder reproduces the reference's separately
rounded multiplies and adds) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libsogrand.so")
CLI = os.path.join(PKG, "cubic_decoder")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE,
    "-I" + CSRC,
    "-Wall",
]

LIB_SOURCES = [
    "sg_kernels.hip", "sg_persistent.hip", "sg_lanegroup.hip", "sg_api.cpp", "sg_schedule.cpp",
    "sg_serve.cpp", "sg_noalloc.hip",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature: drop that noise
    noise = ("packed-fp32-ops", "warning: unknown target feature")
    err = "\n".join(l for l in r.stderr.splitlines() if not any(n in l for n in noise))
    if r.returncode != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))


def build(verbose: bool = False) -> str:
    tmpdir = os.path.join(PKG, "_obj")
    os.makedirs(tmpdir, exist_ok=True)
    cmds, objs = [], []
    for src in LIB_SOURCES:
        obj = os.path.join(tmpdir, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC, *COMMON, "-pthread", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
        if verbose:
            print(" ".join(cmds[-1]))
    with ThreadPoolExecutor(max_workers=min(len(cmds), 6)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    cli_cmd = [HIPCC, *COMMON, "-pthread", os.path.join(CSRC, "cubic_decoder_main.cpp"), "-o",
               CLI + ".tmp",
               "-L" + PKG, "-lsogrand", "-Wl,-rpath,$ORIGIN"]
    _run(cli_cmd)
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
