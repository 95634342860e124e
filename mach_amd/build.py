"""Builds the native library ``mach_amd/_lib/libmach_hip.so`` for gfx950 with CMake
(``mach_amd/CMakeLists.txt``: the HIP language, ROCm's clang, Ninja).

Run as ``python mach_amd/build.py`` (a plain script: importing the package would first load the
library being rebuilt). The build always recompiles: every call configures a fresh build directory
in a temporary folder (nothing but the library lands in the tree), compiles the translation units in
parallel (``MACH_AMD_BUILD_JOBS``, default: the CPU count, at most 16) and then replaces the library
atomically, so a failed build leaves the previous one in place. ``debug=True`` builds
``libmach_hip_debug.so`` instead, with the kernels' bounds checks (``MACH_DEBUG``); the package loads
it when ``MACH_AMD_DEBUG=1``. Compiler flags (set in CMakeLists.txt):
  * ``-ffp-contract=off``: products and sums round separately, so the float64 geometry makes the same
    discrete decisions (aperture, bounds, tap index) as a float64 evaluation of the same formulas; the
    kernels spell out every fused multiply-add they want with ``fmaf``.
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-FP32 instructions.
  * ``-mllvm -structurizecfg-skip-uniform-regions=1`` (complex64 and lateral-group kernels): branches on
    wave-uniform values stay scalar branches, so the lateral kernels' accumulators update in place
    instead of being copied at the end of every element iteration (control flow only).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "_lib")
ARCH = "gfx950"


def rocm_path() -> str:
    for cand in (os.environ.get("ROCM_PATH"), "/opt/rocm"):
        if cand and os.path.exists(os.path.join(cand, "llvm", "bin", "clang++")):
            return cand
    raise RuntimeError("ROCm (llvm/bin/clang++) not found; set ROCM_PATH")


def _tool(name: str) -> str:
    path = shutil.which(name)
    if path is None:
        raise RuntimeError(f"{name} not found (the build uses CMake with Ninja)")
    return path


def _run(cmd, verbose, cwd=None):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd)
    if verbose and res.stdout:
        print(res.stdout[-4000:], flush=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout[-20000:] + res.stderr[-20000:])
        raise RuntimeError(f"build step failed with exit code {res.returncode}: {' '.join(cmd)}")


def build(verbose: bool = False, debug: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    name = "libmach_hip_debug.so" if debug else "libmach_hip.so"
    out = os.path.join(LIB_DIR, name)
    jobs = int(os.environ.get("MACH_AMD_BUILD_JOBS", "0")) or min(16, os.cpu_count() or 1)
    cmake, ninja = _tool("cmake"), _tool("ninja")
    clang = os.path.join(rocm_path(), "llvm", "bin", "clang++")
    with tempfile.TemporaryDirectory(prefix="mach_amd_build_") as bdir:
        stage = os.path.join(bdir, "lib")
        _run([cmake, "-S", HERE, "-B", bdir, "-G", "Ninja", f"-DCMAKE_MAKE_PROGRAM={ninja}",
              "-DCMAKE_BUILD_TYPE=Release", f"-DCMAKE_HIP_COMPILER={clang}", f"-DCMAKE_HIP_ARCHITECTURES={ARCH}",
              f"-DMACH_OUTPUT_DIR={stage}", f"-DMACH_DEBUG={'ON' if debug else 'OFF'}"], verbose)
        _run([cmake, "--build", bdir, "-j", str(jobs)] + (["-v"] if verbose else []), verbose)
        tmp = out + ".tmp"
        shutil.copyfile(os.path.join(stage, "libmach_hip.so"), tmp)
        os.chmod(tmp, 0o755)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True, debug="--debug" in sys.argv))
