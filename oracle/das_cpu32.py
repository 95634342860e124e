"""ctypes front end of the FP32 CPU delay-and-sum (``oracle/das_cpu32.c``).

CPU BASELINE ONLY: timed by bench.py's cpu_baseline leg and checked against the FP64 oracle by
``tests/test_cpu_baseline.py``; the mach_amd package never loads it. Same call convention as
``das_c.das_c`` (float32 scalars, as the reference's binding passes them), returning complex64 (IQ) or
float32 (RF) [n_vox, n_frames].

``load(native=True)`` compiles the library for the host it runs on (``-march=native``, into
``oracle/_native/``, keyed by the CPU model) and falls back to the portable x86-64-v3 build
(``oracle/libdas_cpu32.so``, made by ``__graft_entry__.build()``) when no compiler is available.
"""

from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PORTABLE = os.path.join(_HERE, "libdas_cpu32.so")


class _Args(ctypes.Structure):
    _fields_ = [
        ("channel_data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("interp", ctypes.c_int32),
        ("rx_coords_m", ctypes.c_void_p),
        ("scan_coords_m", ctypes.c_void_p),
        ("tx_arrivals_s", ctypes.c_void_p),
        ("n_tx", ctypes.c_int64),
        ("n_rx", ctypes.c_int64),
        ("n_samples", ctypes.c_int64),
        ("n_frames", ctypes.c_int64),
        ("n_vox", ctypes.c_int64),
        ("f_number", ctypes.c_float),
        ("rx_start_s", ctypes.c_float),
        ("sampling_freq_hz", ctypes.c_float),
        ("sound_speed_m_s", ctypes.c_float),
        ("modulation_freq_hz", ctypes.c_float),
        ("tukey_alpha", ctypes.c_float),
        ("out", ctypes.c_void_p),
    ]


_loaded = {}


def cpu_model() -> str:
    """The host CPU's model name (``lscpu``'s "Model name"), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def build(native: bool = False) -> str:
    """Compile the library with the committed Makefile: portable (x86-64-v3), or for this host."""
    if not native:
        subprocess.run(["make", "-s", "-C", _HERE, "libdas_cpu32.so"], check=True)
        return _PORTABLE
    key = hashlib.sha1(cpu_model().encode()).hexdigest()[:12]
    out = os.path.join(_HERE, "_native", f"libdas_cpu32_{key}.so")
    if not os.path.exists(out):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.run(["make", "-s", "-B", "-C", _HERE, f"CPU32_OUT={out}", "CPU32_ARCH=-march=native", out],
                       check=True, capture_output=True)
    return out


def load(native: bool = True):
    """(library, build description); native builds fall back to the portable one."""
    if native in _loaded:
        return _loaded[native]
    path, desc = None, None
    if native:
        try:
            path, desc = build(native=True), "-O3 -march=native -fopenmp"
        except (OSError, subprocess.CalledProcessError):
            path = None
    if path is None:
        if not os.path.exists(_PORTABLE):
            build(native=False)
        path, desc = _PORTABLE, "-O3 -march=x86-64-v3 -fopenmp"
    lib = ctypes.CDLL(path)
    lib.das_cpu32_run.argtypes = [ctypes.POINTER(_Args), ctypes.c_int]
    lib.das_cpu32_run.restype = ctypes.c_int
    lib.das_cpu32_max_threads.restype = ctypes.c_int
    _loaded[native] = (lib, desc)
    return _loaded[native]


def das_cpu32(channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, *, f_number, rx_start_s,
              sampling_freq_hz, sound_speed_m_s, modulation_freq_hz=0.0, tukey_alpha=0.5, interp_type=1,
              n_threads=0, native=True):
    """FP32 delay-and-sum of [T, E, S, F] (or [E, S, F]) channel data: complex64, float32, or float16
    [..., 2] FP16 IQ pairs. Returns [V, F] complex64 (IQ) or float32 (RF)."""
    lib, _ = load(native)
    D = np.asarray(channel_data)
    arr = np.asarray(tx_wave_arrivals_s, dtype=np.float32)
    fp16 = D.dtype == np.float16
    if (D.ndim == 4 and fp16) or (D.ndim == 3 and not fp16):
        D, arr = D[None], arr[None]
    if fp16:
        dtype, core = 2, D.shape[:4]
    elif np.iscomplexobj(D):
        dtype, core, D = 1, D.shape, D.astype(np.complex64, copy=False)
    else:
        dtype, core, D = 0, D.shape, D.astype(np.float32, copy=False)
    D = np.ascontiguousarray(D)
    T, E, S, F = core
    rx = np.ascontiguousarray(rx_coords_m, dtype=np.float32)
    vox = np.ascontiguousarray(scan_coords_m, dtype=np.float32)
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    V = vox.shape[0]
    out = np.zeros((V, F), dtype=np.float32 if dtype == 0 else np.complex64)
    a = _Args(D.ctypes.data, dtype, int(interp_type), rx.ctypes.data, vox.ctypes.data, arr.ctypes.data,
              T, E, S, F, V, f_number, rx_start_s, sampling_freq_hz, sound_speed_m_s,
              modulation_freq_hz or 0.0, tukey_alpha, out.ctypes.data)
    st = lib.das_cpu32_run(ctypes.byref(a), int(n_threads))
    if st != 0:
        raise RuntimeError(f"das_cpu32_run failed with status {st}")
    return out


def max_threads(native=True) -> int:
    return int(load(native)[0].das_cpu32_max_threads())
