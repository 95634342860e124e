"""Fused post-processing: power Doppler and B-mode images straight from channel data.

The reference returns the beamformed signal y[v, f] and its users post-process it on the host:

* power Doppler ``np.square(np.abs(y)).sum(axis=-1)`` (``examples/doppler.py:252``,
  ``tests/compare/test_pymust.py:317-318``);
* B-mode ``db_zero(y)`` = 20 log10 |y| normalised to 0 dB at the maximum, with exact zeros replaced
  by the float64 machine epsilon (``src/mach/_vis.py:17-40``).

Here the delay-and-sum kernel computes these in its epilogue (``mach_beamform_args.output``): the
per-frame signal stays in registers, and only the [n_vox] power image or the float32 envelope reaches
HBM. For the 256-frame Doppler ensemble that removes the 134 MB complex64 output write. The
logarithmic compression is a separate element-wise kernel (``mach_to_db``) normalised by the peak
the epilogue records with an order-independent atomic maximum.

Channel data is [n_rx, n_samples, n_frames] (one transmit, arrivals [n_vox]) or
[n_tx, n_rx, n_samples, n_frames] (coherently compounded, arrivals [n_tx, n_vox]); complex64 IQ,
float32 RF, or FP16 IQ (float16 [..., 2] or complex32). Inputs on the CPU are staged to the GPU with
the reference's latency warning; results are device tensors.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._hip_impl import (OUTPUT_ENVELOPE, OUTPUT_POWER, InterpolationType, _as_tensor, _core_shape, _data_kind,
                        _device_for, _require, _shape_str, check_dimensions, groups_per_block, make_args, order_on,
                        voxel_order)

__all__ = ["power_doppler", "envelope", "bmode_db", "power_doppler_db", "to_db"]


def _prepare(channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, modulation_freq_hz, name):
    cd = _as_tensor(channel_data, "channel_data")
    kind = _data_kind(cd)
    if kind != _lib.DATA_F32 and modulation_freq_hz is None:
        raise ValueError(f"{name}(): modulation_freq_hz is required for complex channel data")
    fmod = 0.0 if modulation_freq_hz is None else float(modulation_freq_hz)
    pairs = kind == _lib.DATA_C32H and cd.dtype == torch.float16
    core_dims = cd.dim() - (1 if pairs else 0)
    if core_dims not in (3, 4) or not cd.is_contiguous():
        raise TypeError(f"{name}(): incompatible channel_data shape {_shape_str(cd)} (C-contiguous "
                        "[n_rx, n_samples, n_frames] or [n_tx, n_rx, n_samples, n_frames] expected)")
    rx = _as_tensor(rx_coords_m, "rx_coords_m")
    scan = _as_tensor(scan_coords_m, "scan_coords_m")
    arr = _as_tensor(tx_wave_arrivals_s, "tx_wave_arrivals_s")
    _require(rx, "rx_coords_m", torch.float32, 2, 3)
    _require(scan, "scan_coords_m", torch.float32, 2, 3)
    _require(arr, "tx_wave_arrivals_s", torch.float32, core_dims - 2)
    core = _core_shape(cd, kind)
    n_tx = core[0] if core_dims == 4 else 1
    if core_dims == 4 and arr.shape[0] != n_tx:
        raise RuntimeError(f"Dimension mismatch in transmits: channel_data.shape {list(core)} vs "
                           f"tx_wave_arrivals_s.shape {_shape_str(arr)}")
    arrays = [cd, rx, scan, arr]
    device = _device_for(arrays)
    if any(not t.is_cuda for t in arrays):
        import warnings

        n_cpu = sum(1 for t in arrays if not t.is_cuda)
        warnings.warn(
            f"Found {n_cpu} input array(s) on CPU. This will add latency due to CPU<->GPU memory "
            "transfers. For optimal performance with HIP beamforming, move arrays to the GPU "
            "(e.g. PyTorch tensors on a ROCm device).",
            UserWarning,
            stacklevel=3,
        )
    cd, rx, scan, arr = (t if t.is_cuda else t.to(device) for t in arrays)
    if any(t.device != device for t in (cd, rx, scan, arr)):
        raise RuntimeError("All GPU arrays must be on the same device")
    return cd, rx, scan, arr, kind, n_tx, core, fmod, device


def _launch(cd, rx, scan, arr, out, peak, *, kind, n_tx, output, accumulate, f_number, rx_start_s,
            sampling_freq_hz, sound_speed_m_s, fmod, tukey_alpha, interp_type, device):
    if peak is not None and not (isinstance(peak, torch.Tensor) and peak.device == device
                                 and peak.dtype == torch.float32 and peak.numel() >= 1):
        raise TypeError(f"peak must be a float32 tensor of at least one element on {device}")
    from . import autotune

    core = _core_shape(cd, kind)
    layout = autotune.layout_for(kind, int(interp_type), n_tx, core[-3], core[-1], scan.shape[0], cd.data_ptr(),
                                 out.data_ptr(), output=output)
    args = make_args(cd, rx, scan, arr, out, n_tx=n_tx, kind=kind, f_number=f_number, rx_start_s=rx_start_s,
                     sampling_freq_hz=sampling_freq_hz, sound_speed_m_s=sound_speed_m_s,
                     modulation_freq_hz=fmod, tukey_alpha=tukey_alpha, interp_type=interp_type,
                     accumulate=accumulate, layout=layout,
                     order=order_on(voxel_order(scan, layout, groups_per_block(n_tx, core[-1])),
                                    torch.cuda.current_stream(device)), output=output,
                     peak=peak)
    lib = _lib.load()
    with torch.cuda.device(device):
        stream = torch.cuda.current_stream(device).cuda_stream
        _lib.check(lib.mach_beamform(ctypes.byref(args), ctypes.c_void_p(stream)))


def _out_tensor(out, shape, device, name):
    if out is None:
        return torch.empty(shape, dtype=torch.float32, device=device)
    o = _as_tensor(out, "out")
    if not o.is_cuda or o.device != device:
        raise ValueError(f"{name}(): out must be a float32 tensor on {device}")
    _require(o, "out", torch.float32, len(shape))
    if tuple(o.shape) != tuple(shape):
        raise RuntimeError(f"Dimension mismatch in output: expected {list(shape)}, got out.shape {_shape_str(o)}")
    return o


def power_doppler(channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out=None, *, rx_start_s: float,
                  sampling_freq_hz: float, f_number: float, sound_speed_m_s: float,
                  modulation_freq_hz: float | None = None, tukey_alpha: float = 0.5,
                  interp_type: InterpolationType = InterpolationType.Linear, accumulate: bool = False,
                  peak=None) -> torch.Tensor:
    """Power Doppler image sum_f |y[v, f]|^2, float32 [n_vox], in one kernel launch.

    ``accumulate=True`` adds to ``out`` (e.g. summing power over several ensembles). ``peak`` (a
    float32 device tensor of one element) receives the maximum of the written image."""
    cd, rx, scan, arr, kind, n_tx, core, fmod, device = _prepare(
        channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, modulation_freq_hz, "power_doppler")
    o = _out_tensor(out, (scan.shape[0],), device, "power_doppler")
    _check_voxels(core, rx, scan, arr)
    _launch(cd, rx, scan, arr, o, peak, kind=kind, n_tx=n_tx, output=OUTPUT_POWER, accumulate=accumulate,
            f_number=f_number, rx_start_s=rx_start_s, sampling_freq_hz=sampling_freq_hz,
            sound_speed_m_s=sound_speed_m_s, fmod=fmod, tukey_alpha=tukey_alpha, interp_type=interp_type,
            device=device)
    return o


def envelope(channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out=None, *, rx_start_s: float,
             sampling_freq_hz: float, f_number: float, sound_speed_m_s: float,
             modulation_freq_hz: float | None = None, tukey_alpha: float = 0.5,
             interp_type: InterpolationType = InterpolationType.Linear, peak=None) -> torch.Tensor:
    """Envelope |y[v, f]| (complex magnitude for IQ, absolute value for RF), float32 [n_vox, n_frames].
    ``peak`` (float32 device tensor of one element) receives its maximum."""
    cd, rx, scan, arr, kind, n_tx, core, fmod, device = _prepare(
        channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, modulation_freq_hz, "envelope")
    o = _out_tensor(out, (scan.shape[0], core[-1]), device, "envelope")
    check_dimensions(core, rx, scan, arr, o, rx.shape[0], scan.shape[0], core[-1])
    _launch(cd, rx, scan, arr, o, peak, kind=kind, n_tx=n_tx, output=OUTPUT_ENVELOPE, accumulate=False,
            f_number=f_number, rx_start_s=rx_start_s, sampling_freq_hz=sampling_freq_hz,
            sound_speed_m_s=sound_speed_m_s, fmod=fmod, tukey_alpha=tukey_alpha, interp_type=interp_type,
            device=device)
    return o


def to_db(x: torch.Tensor, peak: torch.Tensor, power_mode: bool = False) -> torch.Tensor:
    """In place: x = k log10(a(x)) - k log10(a(peak)), a(v) = |v| (float64 epsilon for v == 0),
    k = 10 for power images and 20 for amplitudes (``db_zero``, ``src/mach/_vis.py:17-40``)."""
    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32 and x.is_contiguous()):
        raise TypeError("to_db(): x must be a contiguous float32 device tensor")
    if not (isinstance(peak, torch.Tensor) and peak.device == x.device and peak.dtype == torch.float32
            and peak.numel() == 1):
        raise TypeError("to_db(): peak must be a one-element float32 tensor on the device of x")
    lib = _lib.load()
    with torch.cuda.device(x.device):
        stream = torch.cuda.current_stream(x.device).cuda_stream
        _lib.check(lib.mach_to_db(ctypes.c_void_p(x.data_ptr()), x.numel(), ctypes.c_void_p(peak.data_ptr()),
                                  1 if power_mode else 0, ctypes.c_void_p(stream)))
    return x


def bmode_db(channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out=None, **kwargs) -> torch.Tensor:
    """B-mode image in dB, 0 dB at its maximum: ``db_zero(y)`` of ``src/mach/_vis.py:30-40`` over the
    envelope, float32 [n_vox, n_frames]. Keyword arguments as ``envelope``."""
    kwargs.pop("peak", None)
    cd = _as_tensor(channel_data, "channel_data")
    device = _device_for([cd])
    peak = torch.empty(1, dtype=torch.float32, device=device)
    env = envelope(cd, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, peak=peak, **kwargs)
    return to_db(env, peak, power_mode=False)


def power_doppler_db(channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out=None,
                     **kwargs) -> torch.Tensor:
    """Power Doppler image in dB (10 log10, 0 dB at its maximum), float32 [n_vox]. Keyword arguments
    as ``power_doppler`` (without ``accumulate``)."""
    kwargs.pop("peak", None)
    cd = _as_tensor(channel_data, "channel_data")
    device = _device_for([cd])
    peak = torch.empty(1, dtype=torch.float32, device=device)
    pd = power_doppler(cd, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, peak=peak, **kwargs)
    return to_db(pd, peak, power_mode=True)


def _check_voxels(core, rx, scan, arr):
    # element and voxel counts as check_dimensions (kernel.cu:914-951); frames are summed away
    if core[-3] != rx.shape[0]:
        raise RuntimeError(
            "Dimension mismatch in receive elements:\n"
            f"  channel_data.shape: {list(core)}\n"
            f"  rx_coords_m.shape: {_shape_str(rx)}\n"
            "→ channel_data.shape[0] and rx_coords_m.shape[0] must both equal n_receive_elements"
        )
    if arr.shape[-1] != scan.shape[0]:
        raise RuntimeError(
            "Dimension mismatch in output voxels:\n"
            f"  scan_coords_m.shape: {_shape_str(scan)}\n"
            f"  tx_wave_arrivals_s.shape: {_shape_str(arr)}\n"
            "→ scan_coords_m.shape[0] and tx_wave_arrivals_s.shape[-1] must both equal n_output_voxels"
        )

