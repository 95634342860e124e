"""UFF beamforming setup end to end on a ROCm GPU (mach_amd.io.uff; reference src/mach/io/uff.py): a
synthetic plane-wave RF acquisition with pyuff_ustb's attribute names (pyuff_ustb is not installed; the
PICMUS files are downloads), three steered waves fired with sequence delays. The setup's analytic signal
matches the oracle's (scipy.signal.hilbert's definition) and the compounded beamform matches the FP64
oracle on it (<= 1e-5 rel-L2); the image peaks at the scatterer, which only holds when each wave's
sequence delay is added to its arrivals."""

from types import SimpleNamespace as NS

import numpy as np
import pytest
import torch

from mach_amd import experimental
from mach_amd.io import uff
from oracle.das64 import analytic_signal64, das64, rel_l2

pytestmark = pytest.mark.gpu

C, FS, F0, T0 = 1540.0, 20.8e6, 5.2e6, 2e-6


def rf_acquisition(scatterer, angles_deg=(-8.0, 0.0, 8.0), delays=(0.0, 1e-6, 2e-6), n_e=64, n_s=1024):
    ex = (np.arange(n_e) - (n_e - 1) / 2) * 0.3e-3
    t = T0 + np.arange(n_s) / FS
    rf = np.zeros((n_s, n_e, len(angles_deg)), np.float64)
    sx, _, sz = scatterer
    for w, (a, d) in enumerate(zip(np.deg2rad(angles_deg), delays)):
        tx = (sx * np.sin(a) + sz * np.cos(a)) / C + d
        tau = tx + np.hypot(ex - sx, sz) / C  # [n_e]
        u = t[:, None] - tau[None, :]
        rf[:, :, w] = np.cos(2 * np.pi * F0 * u) * np.exp(-(u / 0.25e-6) ** 2)
    seq = [NS(source=NS(azimuth=a, elevation=0.0), delay=d) for a, d in zip(np.deg2rad(angles_deg), delays)]
    cd = NS(data=rf.astype(np.float32), sound_speed=C, modulation_frequency=0.0, sampling_frequency=FS,
            initial_time=T0, probe=NS(xyz=np.stack([ex, np.zeros(n_e), np.zeros(n_e)], -1)), sequence=seq)
    x, z = np.meshgrid(np.linspace(-3e-3, 5e-3, 81), np.linspace(8e-3, 16e-3, 81), indexing="ij")
    scan = NS(xyz=np.stack([x.ravel(), np.zeros(x.size), z.ravel()], -1)[None])
    return cd, scan


def test_uff_setup_beamforms_like_the_oracle_and_focuses(device):
    scatterer = (1e-3, 0.0, 12e-3)
    cd, scan = rf_acquisition(scatterer)
    s = uff.create_beamforming_setup(cd, scan, device=device)
    iq = s["channel_data"]
    assert iq.shape == (3, 64, 1024, 1) and iq.dtype == torch.complex64
    iq64 = analytic_signal64(cd.data)
    assert rel_l2(iq.cpu().numpy(), iq64) <= 1e-5
    y = experimental.beamform(**s)
    torch.cuda.synchronize()
    y = y.cpu().numpy()
    ref = das64(iq.cpu().numpy(), s["rx_coords_m"].cpu().numpy(), s["scan_coords_m"].cpu().numpy(),
                s["tx_wave_arrivals_s"].cpu().numpy(), f_number=1.7, rx_start_s=T0, sampling_freq_hz=FS,
                sound_speed_m_s=C, modulation_freq_hz=0.0, tukey_alpha=0.5, interp_type=1)
    assert np.abs(ref).max() > 0
    assert rel_l2(y, ref) <= 1e-5
    peak = s["scan_coords_m"].cpu().numpy()[int(np.argmax(np.abs(y[:, 0])))]
    assert abs(peak[0] - scatterer[0]) <= 0.15e-3 and abs(peak[2] - scatterer[2]) <= 0.15e-3, peak
    # the single-transmit setup of the delayed wave focuses there too
    one = uff.create_single_transmit_beamforming_setup(cd, scan, wave_index=2, device=device)
    from mach_amd.kernel import beamform
    y2 = beamform(**{k: v for k, v in one.items() if k != "out"})
    torch.cuda.synchronize()
    peak2 = one["scan_coords_m"].cpu().numpy()[int(np.argmax(np.abs(y2.cpu().numpy()[:, 0])))]
    assert abs(peak2[0] - scatterer[0]) <= 0.3e-3 and abs(peak2[2] - scatterer[2]) <= 0.3e-3, peak2
