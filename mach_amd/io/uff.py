"""Beamforming setup from UFF (ultrasound file format) objects (reference: ``src/mach/io/uff.py``).

The reference builds the beamformer's inputs from the ``ChannelData`` and ``Scan`` objects that
``pyuff_ustb`` reads from a UFF file; its callers read the files (the PICMUS datasets are downloads).
Those objects are plain attribute containers, so the functions here take any object with the same
attributes:

* ``channel_data.data``: RF [n_samples, n_elements, n_waves] or [..., n_frames] (float32), or
  demodulated IQ of that shape (complex) when ``modulation_frequency`` is not 0;
* ``channel_data.sound_speed``, ``.modulation_frequency``, ``.sampling_frequency``, ``.initial_time``;
* ``channel_data.probe.xyz``: [n_elements, 3] element positions (m);
* ``channel_data.sequence[i].source.azimuth`` / ``.elevation`` (rad) and ``.delay`` (s): the plane
  waves;
* ``scan.xyz``: [1, n_points, 3] (m).

``pyuff_ustb`` is not installed here; nothing in this module reads a file.

Where this differs from the reference:

* RF data becomes the analytic signal on the GPU (``mach_amd.preprocess.analytic_signal``), already in
  the beamformer's [n_waves, n_elements, n_samples, n_frames] layout, instead of
  ``scipy.signal.hilbert`` on the host followed by a rearrangement (``uff.py:70-103``).
* IQ data (``modulation_frequency`` not 0) is rearranged to that layout. The reference leaves its
  result unassigned in that case (``uff.py:85-88`` assigns it only for RF data) and fails.
* Arrays come back as torch tensors on ``device`` (the first GPU by default).
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from .. import _xp
from ..geometry import ultrasound_angles_to_cartesian
from ..preprocess import analytic_signal
from ..wavefront import plane

__all__ = ["extract_wave_directions", "compute_tx_wave_arrivals_s", "preprocess_signal", "extract_sequence_delays",
           "create_beamforming_setup", "create_single_transmit_beamforming_setup"]


def extract_wave_directions(sequence: list[Any], xp=np):
    """Unit propagation direction of each transmit, [n_transmits, 3] (``uff.py:19-34``): the
    sequence's (azimuth, elevation) through ``ultrasound_angles_to_cartesian``.

    ``xp`` is numpy or torch (the array type of the result)."""
    azimuth = np.asarray([float(w.source.azimuth) for w in sequence], dtype=np.float64)
    elevation = np.asarray([float(w.source.elevation) for w in sequence], dtype=np.float64)
    if xp is torch:
        azimuth, elevation = torch.from_numpy(azimuth), torch.from_numpy(elevation)
    return ultrasound_angles_to_cartesian(azimuth, elevation)


def compute_tx_wave_arrivals_s(directions, scan_coords_m, speed_of_sound: float, origin=None):
    """Plane-wave transmit arrival times [n_transmits, n_points] (``uff.py:37-67``): for each
    direction d, ((p - origin) . d) / c, with the origin at 0 unless given.

    Works on numpy arrays and torch tensors (CPU or GPU); the result has the type of ``scan_coords_m``
    and float64 precision (the beamform call rounds arrivals to float32)."""
    if _xp.is_torch(scan_coords_m):
        scan = scan_coords_m.to(torch.float64)
        dirs = torch.as_tensor(directions, dtype=torch.float64, device=scan.device)
        org = torch.zeros(3, dtype=torch.float64, device=scan.device) if origin is None else \
            torch.as_tensor(origin, dtype=torch.float64, device=scan.device)
        return torch.stack([plane(org, scan, d) / speed_of_sound for d in dirs], dim=0)
    scan = np.asarray(scan_coords_m, dtype=np.float64)
    dirs = np.asarray(directions, dtype=np.float64)
    org = np.zeros(3) if origin is None else np.asarray(origin, dtype=np.float64)
    return np.stack([plane(org, scan, d) / speed_of_sound for d in dirs], axis=0)


def preprocess_signal(signal_data, modulation_frequency: float, device=None, storage: str = "c64") -> torch.Tensor:
    """Channel data in the beamformer's layout [n_waves, n_elements, n_samples, n_frames] from the
    UFF order [n_samples, n_elements, n_waves(, n_frames)] (``uff.py:70-103``).

    RF data (``modulation_frequency == 0``) becomes its analytic signal on the GPU
    (``mach_amd.preprocess.analytic_signal``, complex64 or FP16 pairs); IQ data is rearranged.

    Raises:
        ValueError: if the data has neither 3 nor 4 dimensions.
        NotImplementedError: for RF records whose length the device FFT does not support.
    """
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    x = signal_data if isinstance(signal_data, torch.Tensor) else torch.as_tensor(np.asarray(signal_data))
    if x.dim() not in (3, 4):
        raise ValueError(f"Expected signal data to have 3 or 4 dimensions, got {x.dim()}")
    if x.dim() == 3:
        x = x.unsqueeze(-1)
    if modulation_frequency == 0:
        rf = x.real if x.is_complex() else x
        return analytic_signal(rf.to(device=dev, dtype=torch.float32).contiguous(), storage)
    iq = x.to(device=dev, dtype=torch.complex64).permute(2, 1, 0, 3).contiguous()
    if storage == "fp16":
        return torch.view_as_real(iq).to(torch.float16).contiguous()
    return iq


def extract_sequence_delays(sequence: list[Any], xp=np):
    """Each transmit's delay in the sequence (s), [n_transmits] (``uff.py:106-121``)."""
    delays = np.asarray([float(w.delay) for w in sequence], dtype=np.float64)
    return torch.from_numpy(delays) if xp is torch else delays


def create_beamforming_setup(channel_data, scan, f_number: float = 1.7, device=None,
                             storage: str = "c64") -> dict[str, Any]:
    """Keyword arguments of ``mach_amd.experimental.beamform`` for all transmits of a UFF acquisition
    (``uff.py:124-175``): channel data in the beamformer's layout, element and scan coordinates,
    plane-wave arrivals plus each transmit's sequence delay, the sampling and modulation frequencies,
    the sound speed, and ``rx_start_s`` = the channel data's ``initial_time``.

    Arrays are torch tensors on ``device`` (the current GPU by default): coordinates and arrivals in
    float32, channel data in complex64 (or FP16 pairs with ``storage="fp16"``)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    speed_of_sound = float(channel_data.sound_speed)
    modulation_frequency = float(channel_data.modulation_frequency)
    sampling_freq_hz = float(channel_data.sampling_frequency)
    signal = preprocess_signal(channel_data.data, modulation_frequency, dev, storage)
    xyz = np.asarray(scan.xyz)
    if xyz.shape[0] != 1:
        raise ValueError(f"Expected scan.xyz to be one spatial grid, but got {xyz.shape}")
    scan_coords_m = torch.as_tensor(xyz[0], dtype=torch.float64, device=dev)
    rx_coords_m = torch.as_tensor(np.asarray(channel_data.probe.xyz), dtype=torch.float32, device=dev).contiguous()
    directions = extract_wave_directions(channel_data.sequence)
    arrivals = compute_tx_wave_arrivals_s(directions, scan_coords_m, speed_of_sound)
    delays = torch.as_tensor(extract_sequence_delays(channel_data.sequence), dtype=torch.float64, device=dev)
    arrivals = arrivals + delays[:, None]  # each transmit fires `delay` after the sequence's time zero
    return {
        "channel_data": signal,
        "rx_coords_m": rx_coords_m,
        "scan_coords_m": scan_coords_m.to(torch.float32).contiguous(),
        "tx_wave_arrivals_s": arrivals.to(torch.float32).contiguous(),
        "out": None,
        "f_number": f_number,
        "sampling_freq_hz": sampling_freq_hz,
        "sound_speed_m_s": speed_of_sound,
        "modulation_freq_hz": modulation_frequency,
        "rx_start_s": float(channel_data.initial_time),
    }


def create_single_transmit_beamforming_setup(channel_data, scan, wave_index: int = 0, f_number: float = 1.7,
                                             device=None, storage: str = "c64") -> dict[str, Any]:
    """The setup of one transmit, for the single-transmit ``mach_amd.kernel.beamform``
    (``uff.py:178-202``): the multi-transmit setup with ``channel_data`` and ``tx_wave_arrivals_s``
    indexed at ``wave_index``."""
    setup = create_beamforming_setup(channel_data, scan, f_number, device, storage)
    single = dict(setup)
    single["channel_data"] = setup["channel_data"][wave_index]
    single["tx_wave_arrivals_s"] = setup["tx_wave_arrivals_s"][wave_index]
    return single
