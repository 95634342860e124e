"""Channel blocks of 4 GiB or more (n_samples * n_frames * element bytes >= 2^32), which the kernel
addresses with 64-bit offsets (the reference refuses tensors of 2^32 elements or more,
kernel.cu:738-743; this build lifts that cap).

Two channels of 2^27 + 4096 samples hold data only in a window of 256 rows that starts at sample
2^27 (8 s at fs = 2^24 Hz); every other row is zero. Arrivals of 8 s + 4 * 2^-20 s put every tap in
that window. The window, cut out and beamformed with arrivals 8 s earlier through the ordinary
32-bit path, must give the same image: the sample position (tau - t0) * fs shifts by exactly 2^27
(times 2^24 is exact), the phase by 5 MHz * 8 s = 4e7 whole cycles, and the geometry is unchanged.
The only difference is the float64 rounding of tau = T + d / c near 8 s (about 3e-8 samples), so the
bar is 1e-6 relative L2. The cut-out image is checked against the oracle (1e-5, FP16 storage against
the oracle on the same FP16 data).
"""

import sys

import numpy as np
import pytest
import torch

import mach_amd as mach
from mach_amd.kernel import InterpolationType
from oracle.das64 import das64, rel_l2

pytestmark = pytest.mark.gpu

TOL_SHIFT = 1e-6
TOL_ORACLE = 1e-5

FS = float(2**24)  # Hz; a power of two, so shifting the sample position by 2^27 is exact
LO = 2**27  # first sample of the window (8 s)
W = 256  # window rows
S = LO + 4096  # samples per channel
E = 2  # channels
V = 64  # voxels
T_SHIFT = 8.0  # s, = LO / FS
T_WIN = 4 * 2.0**-20  # s: the window run's arrivals (float32-exact, as is T_SHIFT + T_WIN)

# storage -> frames per row (each channel block is >= 4 GiB)
FRAMES = {"iq": 5, "rf": 9, "iq16": 9}


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def geometry(rng):
    rx = np.array([[-1e-3, 0.0, 0.0], [1e-3, 0.0, 0.0]], dtype=np.float32)
    vox = np.stack([rng.uniform(-3e-3, 3e-3, V), np.zeros(V), rng.uniform(3e-3, 10e-3, V)], axis=1)
    return rx, vox.astype(np.float32)


def window_data(rng, kind, F):
    shape = (E, W, F)
    if kind == "rf":
        return rng.normal(size=shape).astype(np.float32)
    z = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    if kind == "iq16":
        return np.stack([z.real, z.imag], axis=-1).astype(np.float16)
    return z


@pytest.mark.parametrize("kind", ["iq", "rf", "iq16"])
def test_channel_blocks_over_4_gib(device, kind):
    F = FRAMES[kind]
    rng = np.random.default_rng({"iq": 1, "rf": 2, "iq16": 3}[kind])
    rx, vox = geometry(rng)
    win = window_data(rng, kind, F)
    elem = {"iq": 8, "rf": 4, "iq16": 4}[kind]
    assert S * F * elem >= 2**32  # the 64-bit path
    fmod = None if kind == "rf" else 5e6
    arr_wide = np.full(V, T_SHIFT + T_WIN, dtype=np.float32)
    arr_win = np.full(V, T_WIN, dtype=np.float32)
    assert float(arr_wide[0]) - float(arr_win[0]) == T_SHIFT and LO / FS == T_SHIFT

    r = torch.as_tensor(rx, device=device)
    v = torch.as_tensor(vox, device=device)
    big_shape = (E, S, F) + ((2,) if kind == "iq16" else ())
    big = torch.zeros(big_shape, dtype=torch.from_numpy(win).dtype, device=device)
    big[:, LO:LO + W] = torch.as_tensor(win, device=device)
    small = torch.as_tensor(win, device=device).contiguous()

    for interp in (0, 1, 2):
        k = dict(f_number=1.0, rx_start_s=0.0, sampling_freq_hz=FS, sound_speed_m_s=1540.0,
                 modulation_freq_hz=fmod, tukey_alpha=0.5, interp_type=InterpolationType(interp))
        log(f"wide kind={kind} interp={interp} frames={F} samples={S} ({S * F * elem} bytes per channel)")
        y_wide = mach.beamform(big, r, v, torch.as_tensor(arr_wide, device=device), **k)
        y_win = mach.beamform(small, r, v, torch.as_tensor(arr_win, device=device), **k)
        torch.cuda.synchronize()
        y_wide, y_win = y_wide.cpu().numpy(), y_win.cpu().numpy()
        assert np.abs(y_win).max() > 0
        assert rel_l2(y_wide, y_win) <= TOL_SHIFT, (kind, interp)
        ref = das64(win, rx, vox, arr_win, f_number=1.0, rx_start_s=0.0, sampling_freq_hz=FS,
                    sound_speed_m_s=1540.0, modulation_freq_hz=fmod or 0.0, tukey_alpha=0.5, interp_type=interp)
        assert rel_l2(y_win, ref) <= TOL_ORACLE, (kind, interp)
    del big, small
    torch.cuda.empty_cache()
