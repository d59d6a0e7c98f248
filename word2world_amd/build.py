"""Builds the HIP rasterizer library ``libw2w_raster.so`` for gfx950.

One hipcc invocation compiles ``csrc/w2w_capi.hip`` (which includes every
kernel header) into a shared library exposing the C ABI of
``include/w2w_raster.h``.  Flags:

* ``-ffp-contract=off``: no implicit fused multiply-adds; the kernels spell
  out every fma, so tile counts, keys and ranges are reproducible bit for bit.
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE division and sqrt.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: keep the compiler
  from emitting packed FP32 (v_pk_*_f32) instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_NAME = "libw2w_raster.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["csrc/w2w_capi.hip"]


def deps() -> list[str]:
    """Every file the library is built from: the sources, every kernel header
    under csrc/ and the public header."""
    csrc = os.path.join(HERE, "csrc")
    names = sorted(f for f in os.listdir(csrc) if f.endswith((".hip", ".hpp", ".h")))
    return [os.path.join("csrc", f) for f in names] + [os.path.join("..", "include", "w2w_raster.h")]


ARCH = os.environ.get("W2W_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found; the rasterizer needs ROCm's hipcc to build")


def command(out: str = LIB_PATH) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-Wno-unused-result",
            "-o", out] + [os.path.join(HERE, s) for s in SOURCES]


def is_stale() -> bool:
    if not os.path.exists(LIB_PATH):
        return True
    t = os.path.getmtime(LIB_PATH)
    return any(os.path.getmtime(os.path.join(HERE, d)) > t for d in deps())


def build(force: bool = True, verbose: bool = False) -> str:
    """Compile the library in-tree (always when ``force``)."""
    if not force and not is_stale():
        return LIB_PATH
    tmp = LIB_PATH + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed with status {r.returncode}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(force=True, verbose=True))
