"""Builds the native library ``mach_amd/_lib/libmach_hip.so`` with hipcc for gfx950.

Run as ``python mach_amd/build.py`` (a plain script: importing the package would first load the
library being rebuilt). The build always recompiles. Flags:
  * ``-ffp-contract=off``: products and sums round separately, so the float64 geometry makes the same
    discrete decisions (aperture, bounds, tap index) as a float64 evaluation of the same formulas;
    the kernels spell out every fused multiply-add they want with ``fmaf``.
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-FP32 instructions.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
SOURCES = ["das_kernel.hip", "das_band.hip", "das_batch.hip", "mach_capi.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libmach_hip.so")
    tmp = out + ".tmp"
    cmd = [
        hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
        "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-o", tmp,
    ] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
