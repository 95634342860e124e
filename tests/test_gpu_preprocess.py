"""RF -> IQ analytic signal on the device (``mach_amd.preprocess``, ``mach_analytic_signal``) against
the oracle's restatement of scipy.signal.hilbert + the UFF rearrangement (runs on a ROCm GPU).

Bars: <= 1e-5 relative L2 (float32 FFT against the float64 oracle), per element <= 1e-4 of the
column's RMS; FP16 output within binary16 rounding (<= 1e-3 relative L2); an RF beamform of the
converted data with modulation frequency 0 matches the oracle chain within 1e-5.
"""

import numpy as np
import pytest
import torch

from mach_amd import experimental, preprocess
from mach_amd.kernel import InterpolationType
from oracle.das64 import analytic_signal64, das64, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 3, 5, 8, 12, 45, 64, 96, 100, 250, 1000, 1024, 1536, 2048, 3000, 4096])
def test_analytic_signal_matches_oracle(device, n):
    rng = np.random.default_rng(n)
    frames = 3 if n <= 2048 else 2
    rf = rng.normal(size=(n, 4, 2, frames)).astype(np.float32)
    got = preprocess.analytic_signal(torch.as_tensor(rf).to(device))
    torch.cuda.synchronize()
    ref = analytic_signal64(rf)
    g = got.cpu().numpy()
    assert g.shape == ref.shape == (2, 4, n, frames)
    err = rel_l2(g, ref)
    rms = np.sqrt(np.mean(np.abs(ref) ** 2))
    print(f"n={n}: rel_l2={err:.2e}", flush=True)
    assert err <= 1e-5
    assert np.abs(g - ref).max() <= 1e-4 * rms


def test_three_dimensional_input_and_fp16_storage(device):
    rng = np.random.default_rng(7)
    rf = rng.normal(size=(1536, 8, 3)).astype(np.float32)
    x = torch.as_tensor(rf).to(device)
    c64 = preprocess.preprocess_signal(x, 0.0)
    h16 = preprocess.analytic_signal(x, storage="fp16")
    torch.cuda.synchronize()
    ref = analytic_signal64(rf)
    assert c64.shape == (3, 8, 1536, 1) and h16.shape == (3, 8, 1536, 1, 2)
    assert rel_l2(c64.cpu().numpy(), ref) <= 1e-5
    h = h16.float().cpu().numpy()
    assert rel_l2(h[..., 0] + 1j * h[..., 1], ref) <= 1e-3
    with pytest.raises(NotImplementedError):
        preprocess.preprocess_signal(x, 5e6)
    with pytest.raises(NotImplementedError, match="product of 2, 3 and 5"):
        preprocess.analytic_signal(torch.zeros((7 * 64, 2, 1), device=device))


def test_rf_chain_beamforms_like_the_oracle(device):
    """RF -> analytic signal (device) -> beamform with f_mod = 0, against the same chain in float64."""
    rng = np.random.default_rng(9)
    S, E, T, F, V = 512, 16, 2, 4, 300
    t = np.arange(S) / 20e6
    rf = (np.cos(2 * np.pi * 5e6 * t)[:, None, None, None] * rng.normal(size=(1, E, T, F))
          + 0.1 * rng.normal(size=(S, E, T, F))).astype(np.float32)
    rx = np.stack([np.linspace(-3e-3, 3e-3, E), np.zeros(E), np.zeros(E)], -1).astype(np.float32)
    vox = (rng.uniform(-1, 1, size=(V, 3)) * [4e-3, 0, 4e-3] + [0, 0, 8e-3]).astype(np.float32)
    arr = (vox[None, :, 2] / 1540.0 + np.array([0.0, 1e-7])[:, None]).astype(np.float32)
    iq = preprocess.analytic_signal(torch.as_tensor(rf).to(device))
    y = experimental.beamform(iq, torch.as_tensor(rx).to(device), torch.as_tensor(vox).to(device),
                              torch.as_tensor(arr).to(device), f_number=1.0, rx_start_s=0.0, sampling_freq_hz=20e6,
                              sound_speed_m_s=1540.0, modulation_freq_hz=0.0, tukey_alpha=0.5,
                              interp_type=InterpolationType.Linear)
    torch.cuda.synchronize()
    iq64 = analytic_signal64(rf).astype(np.complex64)  # the API's complex64 cast of the host result
    ref = das64(iq64, rx, vox, arr, f_number=1.0, rx_start_s=0.0, sampling_freq_hz=20e6, sound_speed_m_s=1540.0,
                modulation_freq_hz=0.0, tukey_alpha=0.5, interp_type=1)
    assert np.abs(ref).max() > 0
    assert rel_l2(y.cpu().numpy(), ref) <= 1e-5
