"""Build the gfx950 shared library ``glide_amd/libglide_hip.so`` with hipcc.

The library is built in-tree so that it travels with the source tree; it is
rebuilt on every call of :func:`build`.  Packed FP32 instructions are disabled
for all device code (``-packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libglide_hip.so"
ARCH = os.environ.get("GLIDE_OFFLOAD_ARCH", "gfx950")

DEVICE_SOURCES = ["ssa_kernels.hip", "mc_kernels.hip", "transfer_kernels.hip", "enthalpy_kernels.hip",
                  "noalloc.hip"]
HOST_SOURCES = ["solver.cpp", "enthalpy_abi.cpp", "timing.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall"]
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-ffp-contract=fast"]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the gfx950 library cannot be built")
    return exe


def build(verbose: bool = False) -> Path:
    """Compile every HIP/C++ source and link ``libglide_hip.so``."""
    hipcc = _hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in DEVICE_SOURCES:
        obj = objdir / (src + ".o")
        cmd = [hipcc, *COMMON, *DEVICE_FLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
        _run(cmd, verbose)
        objs.append(str(obj))
    for src in HOST_SOURCES:
        obj = objdir / (src + ".o")
        cmd = [hipcc, *COMMON, f"--offload-arch={ARCH}", "-c", str(CSRC / src), "-o", str(obj)]
        _run(cmd, verbose)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose and (res.stdout or res.stderr):
        print(res.stdout, res.stderr, flush=True)


if __name__ == "__main__":
    print(build(verbose=True))
