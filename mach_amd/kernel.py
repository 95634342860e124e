"""Python drop-in for ``mach.kernel`` (reference: ``src/mach/kernel.py:1-292``).

``beamform`` keeps the reference's signature, keyword names, defaults, casts, error types and
messages, and ``out`` semantics; ``nb_beamform`` is the strict native entry (the reference's nanobind
function) and ``InterpolationType`` the enum. Arrays may be NumPy arrays, PyTorch tensors (CPU or
ROCm) or any DLPack producer.
"""

from __future__ import annotations

import numpy as np

from . import _xp
from ._check import ensure_contiguous, is_contiguous, is_writeable
from ._hip_impl import InterpolationType, __hip_version__  # noqa: F401
from ._hip_impl import beamform as nb_beamform

__all__ = ["InterpolationType", "beamform", "nb_beamform"]


def _torch():
    import torch

    return torch


def _dtype_class(a):
    """'bool' | 'int' | 'float' | 'complex' | None for NumPy arrays and PyTorch tensors."""
    dt = getattr(a, "dtype", None)
    if dt is None:
        return None
    if _xp.is_torch(a):
        torch = _torch()
        if dt == torch.bool:
            return "bool"
        if dt.is_complex:
            return "complex"
        if dt.is_floating_point:
            return "float"
        return "int"
    try:
        dt = np.dtype(dt)
    except TypeError:
        return None
    if dt.kind == "b":
        return "bool"
    if dt.kind in "iu":
        return "int"
    if dt.kind == "f":
        return "float"
    if dt.kind == "c":
        return "complex"
    return None


def _is_num(a) -> bool:
    return hasattr(a, "shape") and _dtype_class(a) in ("int", "float", "complex")


def _is_real(a) -> bool:
    return hasattr(a, "shape") and _dtype_class(a) in ("int", "float")


def _astype(a, kind: str):
    """Cast to float32 / complex64 without copying when already of that dtype (``astype(copy=False)``)."""
    if _xp.is_torch(a):
        torch = _torch()
        return a.to({"f32": torch.float32, "c64": torch.complex64}[kind])
    if isinstance(a, np.ndarray):
        return a.astype({"f32": np.float32, "c64": np.complex64}[kind], copy=False)
    torch = _torch()
    return _astype(torch.from_dlpack(a), kind)


def _is_fp16_iq(a, ndim: int) -> bool:
    """FP16 IQ storage: float16 [..., n_rx, n_samples, n_frames, 2] with ``ndim`` axes in total, or a
    complex32 tensor."""
    if _xp.is_torch(a):
        torch = _torch()
        return (a.dtype == torch.float16 and a.dim() == ndim and a.shape[-1] == 2) or a.dtype == torch.complex32
    return isinstance(a, np.ndarray) and a.dtype == np.float16 and a.ndim == ndim and a.shape[-1] == 2


def _zeros_like_data(data, shape, kind: str):
    if _xp.is_torch(data):
        torch = _torch()
        dt = torch.complex64 if kind == "c64" else torch.float32
        return torch.zeros(shape, dtype=dt, device=data.device)
    return np.zeros(shape, dtype=np.complex64 if kind == "c64" else np.float32)


def _dtype_matches(out, kind: str) -> bool:
    if _xp.is_torch(out):
        torch = _torch()
        return out.dtype == (torch.complex64 if kind == "c64" else torch.float32)
    return np.dtype(out.dtype) == (np.complex64 if kind == "c64" else np.float32)


def prepare_inputs(channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, modulation_freq_hz,
                   *, n_leading: int = 0):
    """Type checks, casts, contiguity and ``out`` allocation of ``kernel.py:180-275``.

    ``n_leading`` is the number of leading axes before [n_rx, n_samples, n_frames] (1 for the
    transmit axis of the compounding entry).
    """
    if not _is_num(channel_data):
        raise TypeError(
            f"channel_data must be array with dtype=numeric, got type={type(channel_data)}, "
            f"dtype={getattr(channel_data, 'dtype', None)}"
        )
    for arr, name in ((rx_coords_m, "rx_coords_m"), (scan_coords_m, "scan_coords_m"),
                      (tx_wave_arrivals_s, "tx_wave_arrivals_s")):
        if not _is_real(arr):
            raise TypeError(
                f"{name} must be array with dtype=Real, got type={type(arr)}, dtype={getattr(arr, 'dtype', None)}"
            )
    if out is not None and not _is_num(out):
        raise TypeError(f"out must be array with dtype=numeric, got type={type(out)}, dtype={getattr(out, 'dtype', None)}")
    for arr, name in ((channel_data, "channel_data"), (rx_coords_m, "rx_coords_m"),
                      (scan_coords_m, "scan_coords_m"), (tx_wave_arrivals_s, "tx_wave_arrivals_s")):
        if not hasattr(arr, "__dlpack_device__"):
            raise TypeError(f"Array '{name}' does not support DLPack protocol")
    if out is not None and not hasattr(out, "__dlpack_device__"):
        raise TypeError("Array 'out' does not support DLPack protocol")

    rx_coords_m = _astype(rx_coords_m, "f32")
    scan_coords_m = _astype(scan_coords_m, "f32")
    tx_wave_arrivals_s = _astype(tx_wave_arrivals_s, "f32")

    if _is_fp16_iq(channel_data, 4 + n_leading):
        out_kind = "c64"
        if modulation_freq_hz is None:
            raise ValueError(
                "modulation_freq_hz is required for complex phase-correction. set it to 0 if no demodulation was used."
            )
        if isinstance(channel_data, np.ndarray):
            channel_data = _torch().from_numpy(np.ascontiguousarray(channel_data))
    elif _dtype_class(channel_data) == "complex":
        channel_data = _astype(channel_data, "c64")
        out_kind = "c64"
        if modulation_freq_hz is None:
            raise ValueError(
                "modulation_freq_hz is required for complex phase-correction. set it to 0 if no demodulation was used."
            )
    else:
        channel_data = _astype(channel_data, "f32")
        out_kind = "f32"
        if modulation_freq_hz is None:
            modulation_freq_hz = 0.0

    channel_data = ensure_contiguous(channel_data)
    rx_coords_m = ensure_contiguous(rx_coords_m)
    scan_coords_m = ensure_contiguous(scan_coords_m)
    tx_wave_arrivals_s = ensure_contiguous(tx_wave_arrivals_s)

    frames_axis = n_leading + 2
    n_frames = channel_data.shape[frames_axis]
    n_scan = scan_coords_m.shape[0]
    if out is None:
        out = _zeros_like_data(channel_data, (n_scan, n_frames), out_kind)
    elif not _dtype_matches(out, out_kind):
        expected = "complex64" if out_kind == "c64" else "float32"
        raise ValueError(f"Output array dtype {out.dtype} doesn't match expected {expected}")
    if not is_writeable(out):
        raise ValueError("Output array `out` is not writable. Try using cupy or pytorch for the output array.")
    if not is_contiguous(out):
        raise ValueError("Output array `out` must be contiguous.")
    return channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, modulation_freq_hz


def beamform(
    channel_data,
    rx_coords_m,
    scan_coords_m,
    tx_wave_arrivals_s,
    out=None,
    *,
    rx_start_s: float,
    sampling_freq_hz: float,
    f_number: float,
    sound_speed_m_s: float,
    modulation_freq_hz: float | None = None,
    tukey_alpha: float = 0.5,
    interp_type: InterpolationType = InterpolationType.Linear,
):
    """Delay-and-sum beamforming on the GPU (drop-in for ``mach.kernel.beamform``).

    Args:
        channel_data: [n_rx, n_samples, n_frames]; complex (IQ, cast to complex64) or real (RF, cast to
            float32). FP16 IQ storage is float16 [n_rx, n_samples, n_frames, 2] (or a complex32
            tensor) and yields a complex64 output.
        rx_coords_m: [n_rx, 3] element positions (m).
        scan_coords_m: [n_scan, 3] voxel positions (m).
        tx_wave_arrivals_s: [n_scan] transmit arrival time at each voxel (s), e.g.
            ``wavefront.plane(...) / sound_speed_m_s``.
        out: optional [n_scan, n_frames] output. A GPU ``out`` is accumulated into; a CPU ``out`` is
            overwritten (as in the reference).
        rx_start_s: time of the first sample relative to transmit (t0).
        sampling_freq_hz, f_number, sound_speed_m_s: acquisition and aperture parameters.
        modulation_freq_hz: demodulation frequency for IQ data (required for complex data; 0 disables
            the phase rotation); ignored for RF.
        tukey_alpha: Tukey apodization taper in [0, 1] (0 = rectangular).
        interp_type: NearestNeighbor, Linear (default) or Quadratic.

    Returns:
        ``out`` (allocated as zeros in channel_data's namespace and device if not given).
    """
    channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, modulation_freq_hz = prepare_inputs(
        channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, modulation_freq_hz
    )
    nb_beamform(
        channel_data,
        rx_coords_m,
        scan_coords_m,
        tx_wave_arrivals_s,
        out,
        f_number,
        rx_start_s,
        sampling_freq_hz,
        sound_speed_m_s,
        modulation_freq_hz,
        tukey_alpha,
        interp_type,
    )
    return out
