"""Coherent compounding entry (reference: ``src/mach/experimental.py:23-106``).

The reference loops ``kernel.beamform`` over transmits, each call accumulating into the same device
``out`` (``experimental.py:80-97``). Here the transmit sum runs inside one kernel launch
(``mach_beamform`` with ``n_tx`` > 1): the same result, one read-modify-write of ``out`` instead of
``n_tx``. As in the reference, the interpolation is linear unless ``interp_type`` is given (the
reference does not forward it, ``experimental.py:84-97``), a GPU ``out`` is accumulated into and a
CPU ``out`` is overwritten with the compounded image (``experimental.py:66-106``).
"""

from __future__ import annotations

from ._hip_impl import InterpolationType, beamform_compound
from .kernel import prepare_inputs


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
    """Beamform and coherently compound all transmits.

    Args:
        channel_data: [n_transmits, n_rx, n_samples, n_frames] (FP16 IQ: float16 [..., 2]).
        tx_wave_arrivals_s: [n_transmits, n_points].
        Other arguments as ``mach_amd.kernel.beamform``.

    Returns:
        [n_points, n_frames] compounded image (``out`` if given).
    """
    channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, modulation_freq_hz = prepare_inputs(
        channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, modulation_freq_hz, n_leading=1
    )
    beamform_compound(
        channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, out, f_number, rx_start_s,
        sampling_freq_hz, sound_speed_m_s, modulation_freq_hz, tukey_alpha, interp_type,
    )
    return out
