"""Builds the native library ``mach_amd/_lib/libmach_hip.so`` with hipcc for gfx950.

Run as ``python mach_amd/build.py`` (a plain script: importing the package would first load the
library being rebuilt). The build always recompiles: every source to an object file, in parallel
processes (``MACH_AMD_BUILD_JOBS``, default: the CPU count, at most 16), then one link. Flags:
  * ``-ffp-contract=off``: products and sums round separately, so the float64 geometry makes the same
    discrete decisions (aperture, bounds, tap index) as a float64 evaluation of the same formulas;
    the kernels spell out every fused multiply-add they want with ``fmaf``.
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-FP32 instructions.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
# largest first, so the parallel build is not left waiting on one long compile at the end
SOURCES = ["das_batch_tcn.hip", "das_batch_tc1.hip", "das_kernel_c64.hip", "das_kernel_c32h.hip",
           "das_kernel_f32.hip", "das_band.hip", "das_batch.hip", "das_kernel.hip", "das_post.hip", "das_hilbert.hip",
           "mach_capi.hip"]
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}: {' '.join(cmd)}")


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libmach_hip.so")
    tmp = out + ".tmp"
    cc = hipcc()
    jobs = int(os.environ.get("MACH_AMD_BUILD_JOBS", "0")) or min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory(prefix="mach_amd_build_") as objdir:
        objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        cmds = [[cc] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
            for f in [pool.submit(_run, c, verbose) for c in cmds]:
                f.result()
        _run([cc] + FLAGS + ["-shared", "-o", tmp] + objs, verbose)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
