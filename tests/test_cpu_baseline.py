"""The FP32 CPU baseline (oracle/das_cpu32.c), which bench.py times beside the GPU: it computes the
same delay-and-sum as the FP64 oracle, to float32 rounding, for every interpolation, storage type,
apodization and transmit count, and its result does not depend on the thread count."""

import numpy as np
import pytest

from oracle.das64 import das64, rel_l2
from oracle.das_cpu32 import cpu_model, das_cpu32


def case(seed, T, E, S, F, V, kind):
    """Band-limited channel data (a few slow complex tones along the samples, as demodulated echoes
    are): float32 and float64 may round an interpolation centre (or an aperture edge) differently, and
    only for smooth data is the result then close. White noise would make quadratic interpolation
    jump by O(1) where its centre tap switches."""
    rng = np.random.default_rng(seed)
    k = np.arange(S)[:, None]
    D = np.zeros((T, E, S, F), dtype=np.complex128)
    for _ in range(4):
        w = rng.uniform(-0.3, 0.3, size=(T, E, 1, F))
        amp = rng.normal(size=(T, E, 1, F)) + 1j * rng.normal(size=(T, E, 1, F))
        D += amp * np.exp(1j * w * k)
    D = D.astype(np.complex64) if kind != "rf" else D.real.astype(np.float32)
    if kind == "iq16":
        D = np.stack([D.real, D.imag], axis=-1).astype(np.float16)
    rx = (rng.normal(size=(E, 3)) * 1e-3).astype(np.float32)
    rx[:, 2] = 0.0
    vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 3e-3]).astype(np.float32)
    arr = (vox[None, :, 2] / 1540.0 + rng.normal(size=(T, V)) * 1e-8).astype(np.float32)
    return D, rx, vox, arr


def params(kind, interp, alpha):
    return dict(f_number=0.8, rx_start_s=-2e-7, sampling_freq_hz=20.8e6, sound_speed_m_s=1540.0,
                modulation_freq_hz=0.0 if kind == "rf" else 5.2e6, tukey_alpha=alpha, interp_type=interp)


@pytest.mark.parametrize("kind", ["rf", "iq", "iq16"])
@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("alpha", [0.0, 0.5])
def test_cpu_baseline_matches_oracle(kind, interp, alpha):
    D, rx, vox, arr = case(11 + interp, 3, 16, 160, 5, 300, kind)
    p = params(kind, interp, alpha)
    y = das_cpu32(D, rx, vox, arr, **p, native=False)
    ref = das64(D, rx, vox, arr, **p)
    assert np.abs(ref).max() > 0
    # float32 delays and phases (the reference's arithmetic): a few float32 ulps of the ~100 rad phase
    assert rel_l2(y, ref) <= 1e-4, rel_l2(y, ref)


def test_cpu_baseline_is_thread_count_independent():
    D, rx, vox, arr = case(5, 2, 12, 128, 8, 500, "iq")
    p = params("iq", 2, 0.5)
    a = das_cpu32(D, rx, vox, arr, **p, n_threads=1, native=False)
    b = das_cpu32(D, rx, vox, arr, **p, n_threads=4, native=False)
    assert np.array_equal(a, b)


def test_native_build_agrees_with_portable_build():
    D, rx, vox, arr = case(6, 1, 10, 100, 4, 200, "iq")
    p = params("iq", 1, 0.5)
    a = das_cpu32(D, rx, vox, arr, **p, native=False)
    b = das_cpu32(D, rx, vox, arr, **p, native=True)
    assert rel_l2(b, a) <= 1e-6
    assert cpu_model() != ""
