"""Loader for the native library ``mach_amd/_lib/libmach_hip.so`` (C ABI in ``include/mach_hip.h``).

The library is built in-tree by ``mach_amd.build`` (hipcc, gfx950). Loading fails loudly when it is
missing: there is no CPU fallback for the beamform path.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "_lib")
RELEASE_PATH = os.path.join(LIB_DIR, "libmach_hip.so")
DEBUG_PATH = os.path.join(LIB_DIR, "libmach_hip_debug.so")
# MACH_AMD_DEBUG=1 loads the debug build: the kernels count failed bounds checks (instead of trapping)
# and every call waits for its launches and fails with MACH_ERR_DEBUG_CHECK if any check failed.
DEBUG = os.environ.get("MACH_AMD_DEBUG", "0") not in ("", "0")
LIB_PATH = DEBUG_PATH if DEBUG else RELEASE_PATH

MACH_OK = 0
STATUS_NAMES = {
    0: "MACH_OK",
    1: "MACH_ERR_INVALID_ARGUMENT",
    2: "MACH_ERR_DIMENSION",
    3: "MACH_ERR_HIP",
    4: "MACH_ERR_NO_DEVICE",
    5: "MACH_ERR_UNSUPPORTED",
    6: "MACH_ERR_DEBUG_CHECK",
}

ERR_INVALID_ARGUMENT, ERR_DIMENSION, ERR_HIP, ERR_NO_DEVICE, ERR_UNSUPPORTED, ERR_DEBUG_CHECK = 1, 2, 3, 4, 5, 6

# Data-type codes (mach_data_type)
DATA_F32, DATA_C64, DATA_C32H = 0, 1, 2


class BeamformArgs(ctypes.Structure):
    """Mirror of ``mach_beamform_args`` (include/mach_hip.h)."""

    _fields_ = [
        ("channel_data", ctypes.c_void_p),
        ("rx_coords_m", ctypes.c_void_p),
        ("scan_coords_m", ctypes.c_void_p),
        ("tx_wave_arrivals_s", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("n_tx", ctypes.c_uint64),
        ("n_rx", ctypes.c_uint64),
        ("n_samples", ctypes.c_uint64),
        ("n_frames", ctypes.c_uint64),
        ("n_vox", ctypes.c_uint64),
        ("data_type", ctypes.c_int32),
        ("interp_type", ctypes.c_int32),
        ("f_number", ctypes.c_float),
        ("rx_start_s", ctypes.c_float),
        ("sampling_freq_hz", ctypes.c_float),
        ("sound_speed_m_s", ctypes.c_float),
        ("modulation_freq_hz", ctypes.c_float),
        ("tukey_alpha", ctypes.c_float),
        ("accumulate", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("voxel_order", ctypes.c_void_p),
        ("output", ctypes.c_int32),
        ("peak", ctypes.c_void_p),
        ("tx_waves", ctypes.c_void_p),
    ]


_SINGLE_ARGTYPES = [
    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
    ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
    ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
    ctypes.c_int32, ctypes.c_void_p,
]

# Every symbol include/mach_hip.h declares, with its ctypes signature (restype, argtypes).
SIGNATURES = {
    "mach_beamform": (ctypes.c_int, [ctypes.POINTER(BeamformArgs), ctypes.c_void_p]),
    "mach_beamform_c64": (ctypes.c_int, _SINGLE_ARGTYPES),
    "mach_beamform_f32": (ctypes.c_int, _SINGLE_ARGTYPES),
    "mach_batch_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "mach_batch_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BeamformArgs), ctypes.c_uint32,
                                         ctypes.c_void_p]),
    "mach_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "mach_batch_submit_noalloc": (ctypes.c_int, [ctypes.POINTER(BeamformArgs), ctypes.c_uint32, ctypes.c_void_p]),
    "mach_to_db": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int32,
                                  ctypes.c_void_p]),
    "mach_analytic_signal": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                            ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "mach_tx_arrivals": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "mach_last_error": (ctypes.c_char_p, []),
    "mach_last_timing": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "mach_abi_version": (ctypes.c_int32, []),
    "mach_hip_runtime_version": (ctypes.c_char_p, []),
    "mach_device_arch": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
}

_lib = None


class NativeLibraryMissing(ImportError):
    pass


def load():
    """Load and type the native library (cached). Raises ``NativeLibraryMissing`` if not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"mach_amd native library not found at {LIB_PATH}; build it with "
            f"`python mach_amd/build.py{' --debug' if DEBUG else ''}` (CMake + ROCm clang, gfx950). "
            "There is no CPU fallback."
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().mach_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "mach_beamform") -> None:
    """Raise RuntimeError for a non-zero status, with the library's message."""
    if status != MACH_OK:
        msg = last_error()
        raise RuntimeError(msg if msg else f"{what} failed with {STATUS_NAMES.get(status, status)}")
