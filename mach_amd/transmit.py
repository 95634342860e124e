"""Transmit wavefronts evaluated on the device.

The reference's callers compute the transmit arrival of every voxel on the host --
``wavefront.plane(origin, points, direction) / c`` (``src/mach/wavefront.py:8-50``,
``tests/compare/test_pymust.py:69-95``), ``wavefront.spherical(origin, points, focus) / c``
(``src/mach/wavefront.py:53-116``), plus a per-transmit sequence delay in the UFF loader
(``src/mach/io/uff.py:37-67, 156-158``) -- and hand the kernel a [n_tx, n_vox] float32 array.

Here a transmit is a small descriptor (``PlaneWave`` / ``SphericalWave``, packed as ``mach_wave``).
``beamform`` lets the kernel evaluate each voxel's arrival itself (no arrivals array is built, copied
or read), and ``arrivals`` materialises the same float32 values on the device for callers that need
the array. The value is float32(distance / c + time_offset), evaluated in float64 from the float32
voxel coordinates.
"""

from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib
from ._hip_impl import (InterpolationType, _as_tensor, _core_shape, _data_kind, _require, _shape_str,
                        check_dimensions, groups_per_block, make_args, order_on, voxel_order)

__all__ = ["PlaneWave", "SphericalWave", "pack", "arrivals", "beamform"]

# Layout of ``mach_wave`` (include/mach_hip.h): int32 kind, int32 reserved, 8 float64.
WAVE_DTYPE = np.dtype([("kind", "<i4"), ("reserved", "<i4"), ("origin", "<f8", (3,)), ("vector", "<f8", (3,)),
                       ("sound_speed", "<f8"), ("time_offset", "<f8")])
assert WAVE_DTYPE.itemsize == 72


@dataclasses.dataclass(frozen=True)
class PlaneWave:
    """Plane wave through ``origin_m`` travelling along the unit vector ``direction``
    (``wavefront.plane``)."""

    direction: tuple
    sound_speed_m_s: float
    origin_m: tuple = (0.0, 0.0, 0.0)
    time_offset_s: float = 0.0

    def __post_init__(self):
        n = float(np.linalg.norm(np.asarray(self.direction, dtype=np.float64)))
        if not abs(n - 1.0) < 1e-6:  # the reference's unit-vector check (wavefront.py:36-43)
            raise ValueError("direction must be a unit vector")


@dataclasses.dataclass(frozen=True)
class SphericalWave:
    """Focused (focus in front of the origin) or diverging (focus behind) spherical wave timed through
    ``origin_m`` (``wavefront.spherical``)."""

    focus_m: tuple
    sound_speed_m_s: float
    origin_m: tuple = (0.0, 0.0, 0.0)
    time_offset_s: float = 0.0


def pack(waves, device=None) -> torch.Tensor:
    """Wavefronts as a device array of ``mach_wave`` records (returned as a uint8 tensor)."""
    rec = np.zeros(len(waves), dtype=WAVE_DTYPE)
    for i, w in enumerate(waves):
        if isinstance(w, PlaneWave):
            rec[i]["kind"], vec = 0, w.direction
        elif isinstance(w, SphericalWave):
            rec[i]["kind"], vec = 1, w.focus_m
        else:
            raise TypeError(f"not a wavefront: {w!r}")
        if not w.sound_speed_m_s > 0:
            raise ValueError("sound_speed_m_s must be positive")
        rec[i]["origin"] = np.asarray(w.origin_m, dtype=np.float64)
        rec[i]["vector"] = np.asarray(vec, dtype=np.float64)
        rec[i]["sound_speed"] = float(w.sound_speed_m_s)
        rec[i]["time_offset"] = float(w.time_offset_s)
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return torch.from_numpy(rec.view(np.uint8).copy()).to(device)


def _device_scan(scan_coords_m):
    scan = _as_tensor(scan_coords_m, "scan_coords_m")
    _require(scan, "scan_coords_m", torch.float32, 2, 3)
    if not scan.is_cuda:
        raise RuntimeError("scan_coords_m must be on a ROCm device")
    return scan


def arrivals(waves, scan_coords_m) -> torch.Tensor:
    """Transmit arrivals [n_tx, n_vox] float32 on the device of ``scan_coords_m`` -- the values the
    beamform kernel computes from the same wavefronts."""
    scan = _device_scan(scan_coords_m)
    packed = pack(waves, scan.device)
    out = torch.empty((len(waves), scan.shape[0]), dtype=torch.float32, device=scan.device)
    lib = _lib.load()
    with torch.cuda.device(scan.device):
        stream = torch.cuda.current_stream(scan.device).cuda_stream
        _lib.check(lib.mach_tx_arrivals(ctypes.c_void_p(packed.data_ptr()), len(waves), ctypes.c_void_p(scan.data_ptr()),
                                        scan.shape[0], ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
    return out


def beamform(channel_data, rx_coords_m, scan_coords_m, waves, out=None, *, rx_start_s: float,
             sampling_freq_hz: float, f_number: float, modulation_freq_hz: float | None = None,
             tukey_alpha: float = 0.5, interp_type: InterpolationType = InterpolationType.Linear,
             accumulate: bool = False, sound_speed_m_s: float | None = None):
    """Beamform and coherently compound the transmits described by ``waves`` (one per transmit), with
    the transmit arrivals evaluated inside the kernel.

    Args:
        channel_data: [n_rx, n_samples, n_frames] (one wave) or [n_tx, n_rx, n_samples, n_frames];
            complex64 IQ, float32 RF, or FP16 IQ (float16 [..., 2] / complex32); on the device.
        waves: sequence of ``PlaneWave`` / ``SphericalWave``, one per transmit.
        sound_speed_m_s: receive-path sound speed; defaults to the first wave's.
        accumulate: add into ``out`` instead of overwriting it.
        Other arguments as ``mach_amd.kernel.beamform``.

    Returns:
        [n_vox, n_frames] image (``out`` if given).
    """
    cd = _as_tensor(channel_data, "channel_data")
    kind = _data_kind(cd)
    if kind != _lib.DATA_F32 and modulation_freq_hz is None:
        raise ValueError("beamform(): modulation_freq_hz is required for complex channel data")
    fmod = 0.0 if modulation_freq_hz is None else float(modulation_freq_hz)
    pairs = kind == _lib.DATA_C32H and cd.dtype == torch.float16
    core_dims = cd.dim() - (1 if pairs else 0)
    if core_dims not in (3, 4) or not cd.is_contiguous() or not cd.is_cuda:
        raise TypeError(f"beamform(): incompatible channel_data shape {_shape_str(cd)} (C-contiguous device array "
                        "[n_rx, n_samples, n_frames] or [n_tx, n_rx, n_samples, n_frames] expected)")
    core = _core_shape(cd, kind)
    n_tx = core[0] if core_dims == 4 else 1
    waves = list(waves)
    if len(waves) != n_tx:
        raise RuntimeError(f"Dimension mismatch in transmits: channel_data.shape {list(core)} vs {len(waves)} waves")
    rx = _as_tensor(rx_coords_m, "rx_coords_m")
    scan = _device_scan(scan_coords_m)
    _require(rx, "rx_coords_m", torch.float32, 2, 3)
    if rx.device != cd.device or scan.device != cd.device:
        raise RuntimeError("All GPU arrays must be on the same device")
    out_dtype = torch.float32 if kind == _lib.DATA_F32 else torch.complex64
    if out is None:
        out = torch.zeros((scan.shape[0], core[-1]), dtype=out_dtype, device=cd.device)
    o = _as_tensor(out, "out")
    _require(o, "out", out_dtype, 2)
    if o.device != cd.device:
        raise RuntimeError("All GPU arrays must be on the same device")
    shape_only = torch.empty((n_tx, scan.shape[0]), device="meta")  # arrivals are computed in the kernel
    check_dimensions(core, rx, scan, shape_only, o, rx.shape[0], scan.shape[0], core[-1])
    c = float(waves[0].sound_speed_m_s) if sound_speed_m_s is None else float(sound_speed_m_s)
    packed = pack(waves, cd.device)
    from . import autotune

    layout = autotune.layout_for(kind, int(interp_type), n_tx, core[-3], core[-1], scan.shape[0], cd.data_ptr(),
                                 o.data_ptr())
    args = make_args(cd, rx, scan, None, o, n_tx=n_tx, kind=kind, f_number=f_number, rx_start_s=rx_start_s,
                     sampling_freq_hz=sampling_freq_hz, sound_speed_m_s=c, modulation_freq_hz=fmod,
                     tukey_alpha=tukey_alpha, interp_type=interp_type, accumulate=accumulate, layout=layout,
                     order=order_on(voxel_order(scan, layout, groups_per_block(n_tx, core[-1])),
                                    torch.cuda.current_stream(cd.device)),
                     tx_waves=packed)
    lib = _lib.load()
    with torch.cuda.device(cd.device):
        stream = torch.cuda.current_stream(cd.device).cuda_stream
        _lib.check(lib.mach_beamform(ctypes.byref(args), ctypes.c_void_p(stream)))
    return out
