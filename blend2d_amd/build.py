"""Builds libblend2d_amd.so in-tree with hipcc for gfx950.

    python -m blend2d_amd.build

The library is the product: HIP kernels (csrc/kernels.hip) and the host
frontend (csrc/context.cpp, csrc/style.cpp) behind the C ABI declared in
include/blend2d_amd.h.  libb2amd_player.so (csrc/scene_player.cpp) is a
native client of it that plays recorded scenes through that C ABI.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libblend2d_amd.so")
PLAYER = os.path.join(HERE, "libb2amd_player.so")
OBJ_DIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("BLEND2D_AMD_ARCH", "gfx950")

# -ffp-contract=off: the reference never fuses multiply-adds (Math::madd is
# x*y+a, support/math_p.h:134-135); bit-exact edges and gradients need the
# same rounding.  fp32 division and square root stay correctly rounded.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-math-errno", "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=" + ARCH, "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out:
        print(out)


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    for src, device in (("kernels.hip", True), ("cost_model.hip", True), ("context.cpp", False), ("style.cpp", False)):
        obj = os.path.join(OBJ_DIR, src + ".o")
        flags = COMMON + (DEVICE if device else [])
        _run([hipcc, *flags, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    # the scene player: a native client of the library (no device code)
    tmp = PLAYER + ".tmp"
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared",
          os.path.join(CSRC, "scene_player.cpp"), "-o", tmp, "-L" + HERE, "-lblend2d_amd",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, PLAYER)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
