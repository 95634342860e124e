"""ctypes front end of the C oracle (``oracle/das_oracle.c``).

TEST INFRASTRUCTURE ONLY (see ``oracle/das64.py``). Same call signature and return convention as
``das64.das64``; scalars are rounded to float32 first, as the reference's binding does
(``src/mach/kernel.cu:1039-1044``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .das64 import f32

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_das.so")


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
        ("f_number", ctypes.c_double),
        ("rx_start_s", ctypes.c_double),
        ("sampling_freq_hz", ctypes.c_double),
        ("sound_speed_m_s", ctypes.c_double),
        ("modulation_freq_hz", ctypes.c_double),
        ("tukey_alpha", ctypes.c_double),
        ("out", ctypes.c_void_p),
    ]


_lib = None


def build() -> str:
    """Compile ``liboracle_das.so`` with the committed Makefile; returns its path."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle_das.so"], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.das_oracle_run.argtypes = [ctypes.POINTER(_Args), ctypes.c_int]
        lib.das_oracle_run.restype = ctypes.c_int
        lib.das_oracle_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def max_threads() -> int:
    return int(_load().das_oracle_max_threads())


def das_c(
    channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, *, f_number, rx_start_s,
    sampling_freq_hz, sound_speed_m_s, modulation_freq_hz=0.0, tukey_alpha=0.5, interp_type=1,
    out=None, n_threads=0,
):
    """FP64 delay-and-sum through the C oracle. Shapes as ``das64.das64``."""
    lib = _load()
    cd = np.asarray(channel_data)
    if cd.dtype == np.float16:
        dtype = 2
        core_ndim = 4
    elif np.iscomplexobj(cd):
        cd = cd.astype(np.complex64)
        dtype = 1
        core_ndim = 3
    else:
        cd = cd.astype(np.float32)
        dtype = 0
        core_ndim = 3
    arr = np.asarray(tx_wave_arrivals_s, dtype=np.float32)
    if cd.ndim == core_ndim:
        cd = cd[None]
        arr = arr[None]
    cd = np.ascontiguousarray(cd)
    arr = np.ascontiguousarray(arr)
    T, E, S, F = cd.shape[:4]
    rx = np.ascontiguousarray(np.asarray(rx_coords_m, dtype=np.float32))
    vox = np.ascontiguousarray(np.asarray(scan_coords_m, dtype=np.float32))
    V = vox.shape[0]
    is_complex = dtype != 0
    res = np.zeros((V, F), dtype=np.complex128 if is_complex else np.float64)
    args = _Args(
        cd.ctypes.data, dtype, int(interp_type), rx.ctypes.data, vox.ctypes.data, arr.ctypes.data,
        T, E, S, F, V,
        f32(f_number), f32(rx_start_s), f32(sampling_freq_hz), f32(sound_speed_m_s),
        f32(modulation_freq_hz), f32(tukey_alpha), res.ctypes.data,
    )
    rc = lib.das_oracle_run(ctypes.byref(args), int(n_threads))
    if rc != 0:
        raise RuntimeError(f"das_oracle_run failed with status {rc}")
    if out is not None:
        res = res + np.asarray(out)
    return res
