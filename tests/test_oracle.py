"""Pins the delay-and-sum oracle (oracle/das64.py, oracle/das_oracle.c) before it is trusted.

Each check is chosen so a plausible mistake fails it: closed-form geometries built from exactly
representable Pythagorean distances (delay formula, transmit term, t0 sign, sampling), aperture
straddles at and around R = z / (2F) (skip only when h^2 > R^2), Tukey weights against scipy's
window and the Hann closed form, exact reproduction of affine/quadratic signals by linear/quadratic
interpolation, the phase-rotation sign, a physical single-scatterer echo, linearity, frame
permutation, compounding = sum of single transmits, accumulation into ``out``, a scalar brute-force
loop, and agreement of the two independent restatements (NumPy and C).
"""

import math

import numpy as np
import pytest
import scipy.signal

from oracle.das64 import das64, das_bruteforce, interp_taps, rel_l2, tukey
from oracle.das_c import das_c

U = 1.0 / 1024.0  # exactly representable length unit (m); with c = 1 m/s and fs = 1024 Hz, s = 1024 * tau


def kw(**over):
    base = dict(f_number=0.5, rx_start_s=0.0, sampling_freq_hz=1024.0, sound_speed_m_s=1.0,
                modulation_freq_hz=0.0, tukey_alpha=0.0, interp_type=1)
    base.update(over)
    return base


def run_both(*args, **k):
    a = das64(*args, **k)
    b = das_c(*args, **k)
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13)
    return a


def pythagoras_setup(n_samples=32, n_frames=1):
    """Voxel at (0, 0, 4U), 0 deg plane wave (arrival = z = 4U), elements at x = -3U, 0, +3U.

    Receive distances are 5U, 4U, 5U exactly, so tau = 9U, 8U, 9U and s = 9, 8, 9 with fs = 1024, c = 1.
    """
    rx = np.array([[-3 * U, 0, 0], [0, 0, 0], [3 * U, 0, 0]], dtype=np.float32)
    vox = np.array([[0, 0, 4 * U]], dtype=np.float32)
    arr = np.array([4 * U], dtype=np.float32)
    D = np.zeros((3, n_samples, n_frames), dtype=np.float32)
    return rx, vox, arr, D


@pytest.mark.parametrize("interp", [0, 1, 2])
def test_pythagorean_delays_select_exact_samples(interp):
    rx, vox, arr, D = pythagoras_setup()
    D[0, 9, 0] = 1.0
    D[1, 8, 0] = 10.0
    D[2, 9, 0] = 100.0
    y = run_both(D, rx, vox, arr, **kw(interp_type=interp))
    assert y.shape == (1, 1)
    assert y[0, 0] == pytest.approx(111.0, abs=1e-12)
    # transmit arrival matters: shifting it by one sample period moves every tap by one sample
    y2 = run_both(D, rx, vox, arr + np.float32(U), **kw(interp_type=interp))
    assert y2[0, 0] == pytest.approx(0.0, abs=1e-12)


def test_rx_start_shifts_sample_index():
    rx, vox, arr, D = pythagoras_setup()
    D[:, 7, 0] = [1.0, 0.0, 1.0]  # s = (9U - t0) * 1024 = 7 when t0 = 2U
    D[1, 6, 0] = 5.0
    y = run_both(D, rx, vox, arr, **kw(rx_start_s=2 * U, interp_type=0))
    assert y[0, 0] == pytest.approx(7.0, abs=1e-12)


def test_aperture_boundary_is_inclusive():
    # R = z / (2F) = 4U with F = 0.5. Elements at h = 3U, 4U (on the boundary) and 5U.
    rx = np.array([[3 * U, 0, 0], [4 * U, 0, 0], [5 * U, 0, 0]], dtype=np.float32)
    vox = np.array([[0, 0, 4 * U]], dtype=np.float32)
    arr = np.array([0.0], dtype=np.float32)
    D = np.ones((3, 64, 1), dtype=np.float32)
    D[1] *= 10.0
    D[2] *= 100.0
    y = run_both(D, rx, vox, arr, **kw(interp_type=0))
    assert y[0, 0] == pytest.approx(11.0)  # h = 5U > R excluded, h = R included
    y = run_both(D, rx, vox, arr, **kw(interp_type=0, f_number=0.375))  # R = 16U/3 > 5U
    assert y[0, 0] == pytest.approx(111.0)
    # f_number reaches the kernel as float32: f32(0.4) > 0.4 puts R just below 5U, excluding h = 5U
    y = run_both(D, rx, vox, arr, **kw(interp_type=0, f_number=0.4))
    assert y[0, 0] == pytest.approx(11.0)
    # the aperture uses only lateral distance: the same elements moved in y behave identically
    rx_y = rx[:, [1, 0, 2]]
    assert run_both(D, rx_y, vox, arr, **kw(interp_type=0))[0, 0] == pytest.approx(11.0)


def test_tukey_matches_scipy_and_hann():
    # scipy's symmetric Tukey window of odd length M: sample n at normalised distance r = |n - c| / c
    M = 2001
    c = (M - 1) // 2
    r = np.abs(np.arange(M) - c) / c
    for alpha in (0.25, 0.5, 0.75, 1.0):
        np.testing.assert_allclose(tukey(r, alpha), scipy.signal.windows.tukey(M, alpha), atol=1e-12)
    rr = np.linspace(0, 1, 101)
    np.testing.assert_allclose(tukey(rr, 1.0), 0.5 * (1 + np.cos(np.pi * rr)), atol=1e-12)  # Hann
    np.testing.assert_array_equal(tukey(rr, 0.0), np.ones_like(rr))  # rectangular
    assert tukey(np.array([1.0]), 0.5)[0] == 0.0 and tukey(np.array([0.5]), 0.5)[0] == 1.0
    assert tukey(np.array([-0.1, 1.1]), 0.5).tolist() == [0.0, 0.0]


def test_tukey_weight_applied_through_das():
    # element at h = R/2 with alpha = 1 -> w = 0.5 + 0.5 cos(pi/2) = 0.5; at h = 0 -> 1
    rx = np.array([[0, 0, 0], [2 * U, 0, 0]], dtype=np.float32)
    vox = np.array([[0, 0, 4 * U]], dtype=np.float32)
    D = np.ones((2, 64, 1), dtype=np.float32)
    D[1] *= 10.0
    y = run_both(D, rx, vox, np.zeros(1, np.float32), **kw(interp_type=0, tukey_alpha=1.0))
    assert y[0, 0] == pytest.approx(1.0 + 10.0 * 0.5)


@pytest.mark.parametrize("frac", [0.0, 0.25, 0.5, 0.75])
def test_linear_reproduces_affine_signal(frac):
    rx, vox, arr, _ = pythagoras_setup(n_samples=32, n_frames=2)
    k = np.arange(32, dtype=np.float64)
    D = np.stack([np.stack([3.0 + 2.0 * k, -1.0 + 0.5 * k], -1)] * 3).astype(np.float32)
    t0 = -frac * U  # s = 1024 * (tau - t0) = {9, 8, 9} + frac
    y = run_both(D, rx, vox, arr, **kw(rx_start_s=t0))
    s = np.array([9, 8, 9]) + frac
    np.testing.assert_allclose(y[0], [np.sum(3 + 2 * s), np.sum(-1 + 0.5 * s)], rtol=1e-12)


@pytest.mark.parametrize("frac", [0.0, 0.25, 0.375, 0.5, 0.625])  # exact binary fractions
def test_quadratic_reproduces_quadratic_signal(frac):
    rx, vox, arr, _ = pythagoras_setup(n_samples=32)
    k = np.arange(32, dtype=np.float64)
    D = np.stack([(1.0 - 0.5 * k + 0.25 * k * k)[:, None]] * 3).astype(np.float32)
    t0 = -frac * U
    y = run_both(D, rx, vox, arr, **kw(rx_start_s=t0, interp_type=2))
    s = np.array([9, 8, 9]) + frac
    assert y[0, 0] == pytest.approx(np.sum(1 - 0.5 * s + 0.25 * s * s), rel=1e-12)
    # linear interpolation is *not* exact on a quadratic away from integer samples
    if frac not in (0.0,):
        y_lin = run_both(D, rx, vox, arr, **kw(rx_start_s=t0, interp_type=1))
        assert abs(y_lin[0, 0] - y[0, 0]) > 1e-3


def test_nearest_rounds_half_to_even_and_clamps():
    idx, w, valid = interp_taps(np.array([2.5, 3.5, -0.5, -0.51, 7.5, 7.49, 6.5]), 8, 0)
    assert idx[:, 0].tolist() == [2, 4, 0, 0, 7, 7, 6]
    assert valid.tolist() == [True, True, True, False, True, True, True]
    # linear bounds [0, n-1], quadratic [1, n-2]
    assert interp_taps(np.array([0.0, 7.0, 7.01, -0.01]), 8, 1)[2].tolist() == [True, True, False, False]
    assert interp_taps(np.array([1.0, 6.0, 0.99, 6.01]), 8, 2)[2].tolist() == [True, True, False, False]


def test_phase_rotation_sign_and_magnitude():
    rx, vox, arr, _ = pythagoras_setup()
    D = np.ones((3, 32, 1), dtype=np.complex64)
    f_mod = 32.0  # cycles = 32 * tau; tau = 9U, 8U, 9U
    y = run_both(D, rx, vox, arr, **kw(modulation_freq_hz=f_mod))
    tau = np.array([9, 8, 9]) * U
    expected = np.sum(np.exp(2j * np.pi * f_mod * tau))
    assert abs(y[0, 0] - expected) < 1e-12
    assert abs(y[0, 0] - np.conj(expected)) > 0.1  # the sign is +j (kernel.cu:653-661)
    # f_mod = 0 disables rotation; RF data is never rotated
    assert run_both(D, rx, vox, arr, **kw(modulation_freq_hz=0.0))[0, 0] == pytest.approx(3.0)
    assert run_both(np.ones((3, 32, 1), np.float32), rx, vox, arr, **kw(modulation_freq_hz=f_mod))[0, 0] == pytest.approx(3.0)


def test_negative_tau_and_negative_depth_are_skipped_or_kept_as_reference():
    rx, vox, arr, D = pythagoras_setup()
    D[:] = 1.0
    # arrival so negative that tau < 0 for every element: skipped (kernel.cu:634)
    assert run_both(D, rx, vox, np.array([-20 * U], np.float32), **kw(interp_type=0))[0, 0] == 0.0
    # voxel at negative z: R < 0, R^2 > 0; alpha = 0 keeps in-aperture elements; alpha > 0 gives r = h/R < 0
    # -> w = 0, except the on-axis element where r = 0/R = -0.0, which is not < 0 -> w = 1 (kernel.cu:336-341)
    vneg = np.array([[0, 0, -4 * U]], np.float32)
    arr_n = np.array([12 * U], np.float32)
    y0 = run_both(D, rx, vneg, arr_n, **kw(interp_type=0, tukey_alpha=0.0))
    y1 = run_both(D, rx, vneg, arr_n, **kw(interp_type=0, tukey_alpha=0.5))
    assert y0[0, 0] == pytest.approx(3.0) and y1[0, 0] == pytest.approx(1.0)


def test_single_scatterer_echo_focuses():
    from mach_amd import phantoms, probes

    rx = probes.linear_probe_positions(64, 0.3e-3)
    p = np.array([[[1e-3, 0.0, 20e-3]]])
    D = phantoms.synthesize_echoes(p, np.array([1.0 + 0j]), rx, lambda q, t: q[..., 2] / 1540.0, 1, 2048).numpy()[0]
    vox = p[0].astype(np.float32)
    arr = (vox[:, 2] / 1540.0).astype(np.float32)
    k = dict(f_number=1.0, rx_start_s=0.0, sampling_freq_hz=20.8e6, sound_speed_m_s=1540.0,
             modulation_freq_hz=5.2e6, tukey_alpha=0.0, interp_type=2)
    y = run_both(D, rx, vox, arr, **k)
    n_ap = int(np.sum(np.abs(rx[:, 0] - 1e-3) <= 20e-3 / 2))
    assert abs(y[0, 0]) == pytest.approx(n_ap, rel=0.01)  # every aperture element adds ~g(0) = 1 in phase
    assert abs(np.angle(y[0, 0])) < 0.01
    off = run_both(D, rx, vox + np.float32([0, 0, 1e-3]), arr + np.float32(1e-3 / 1540), **k)
    assert abs(off[0, 0]) < 0.2 * abs(y[0, 0])


def random_case(rng, E=5, S=40, F=3, V=9, T=1, cplx=True):
    D = rng.normal(size=(T, E, S, F)).astype(np.float32)
    if cplx:
        D = (D + 1j * rng.normal(size=D.shape)).astype(np.complex64)
    rx = (rng.normal(size=(E, 3)) * 1e-3).astype(np.float32)
    vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 2e-3]).astype(np.float32)  # some negative z
    arr = (rng.normal(size=(T, V)) * 1e-6 + 0.5e-6).astype(np.float32)  # some tau < 0
    return D, rx, vox, arr


RK = dict(f_number=0.8, rx_start_s=-1e-7, sampling_freq_hz=20e6, sound_speed_m_s=1540.0)


@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("alpha", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("cplx", [False, True])
def test_bruteforce_numpy_and_c_agree(interp, alpha, cplx):
    rng = np.random.default_rng(interp * 10 + int(alpha * 4) + cplx)
    D, rx, vox, arr = random_case(rng, T=2, cplx=cplx)
    k = dict(RK, modulation_freq_hz=5e6, tukey_alpha=alpha, interp_type=interp)
    a = das_bruteforce(D, rx, vox, arr, **k)
    b = das64(D, rx, vox, arr, **k)
    c = das_c(D, rx, vox, arr, **k)
    assert np.abs(a).max() > 0
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(c, a, rtol=1e-12, atol=1e-12)


def test_invariants_linearity_frames_compounding_accumulation():
    rng = np.random.default_rng(7)
    D1, rx, vox, arr = random_case(rng, T=3, F=4)
    D2 = (rng.normal(size=D1.shape) + 1j * rng.normal(size=D1.shape)).astype(np.complex64)
    k = dict(RK, modulation_freq_hz=5e6, tukey_alpha=0.5, interp_type=2)
    y1, y2 = das64(D1, rx, vox, arr, **k), das64(D2, rx, vox, arr, **k)
    a, b = np.complex64(0.5 - 2j), np.complex64(3.0)
    y12 = das64((a * D1.astype(np.complex128) + b * D2.astype(np.complex128)), rx, vox, arr, **k)
    np.testing.assert_allclose(y12, a * y1 + b * y2, rtol=1e-6, atol=1e-6)  # inputs re-rounded to complex64
    perm = rng.permutation(4)
    np.testing.assert_allclose(das64(D1[..., perm], rx, vox, arr, **k), y1[:, perm], rtol=1e-13, atol=1e-13)
    single = sum(das64(D1[t], rx, vox, arr[t], **k) for t in range(3))
    np.testing.assert_allclose(y1, single, rtol=1e-12, atol=1e-12)
    y0 = (rng.normal(size=y1.shape) + 1j * rng.normal(size=y1.shape))
    np.testing.assert_allclose(das64(D1, rx, vox, arr, out=y0, **k), y0 + y1, rtol=1e-13)


def test_fp16_iq_storage_is_exact_upcast():
    rng = np.random.default_rng(3)
    D, rx, vox, arr = random_case(rng)
    h = np.stack([D.real, D.imag], -1).astype(np.float16)
    k = dict(RK, modulation_freq_hz=5e6, tukey_alpha=0.5, interp_type=1)
    ref = das64((h[..., 0].astype(np.float64) + 1j * h[..., 1].astype(np.float64)).astype(np.complex64), rx, vox, arr, **k)
    np.testing.assert_allclose(das64(h, rx, vox, arr, **k), ref, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(das_c(h, rx, vox, arr, **k), ref, rtol=1e-13, atol=1e-13)


def test_rel_l2_metric():
    y = np.array([3 + 4j, 0])
    assert rel_l2(y, y) == 0.0
    assert rel_l2(y * 1.1, y) == pytest.approx(0.1)
    assert math.isfinite(rel_l2(np.zeros(2), np.zeros(2)))


def test_wave_arrivals_match_wavefront_functions_and_closed_forms():
    """wave_arrivals64 against the reference-behaviour wavefront module (itself pinned by the reference's
    KATs) and closed forms: a 0 deg plane wave arrives at z / c; a diverging wave from a source d behind
    the origin reaches depth z on axis at (z + d - d) / c = z / c; the offset adds after the division."""
    from mach_amd import wavefront
    from oracle.das64 import wave_arrivals64

    rng = np.random.default_rng(11)
    scan = (rng.normal(size=(500, 3)) * 1e-2 + [0, 0, 3e-2]).astype(np.float32)
    d = np.array([math.sin(0.2), 0.0, math.cos(0.2)])
    src = np.array([1e-3, -2e-3, -10e-3])
    waves = [(0, (0.0, 0.0, 0.0), d, 1540.0, 0.0), (1, (0.0, 0.0, 0.0), src, 1480.0, 2e-6),
             (0, (1e-3, 0.0, 0.0), (0.0, 0.0, 1.0), 1540.0, -1e-6)]
    t = wave_arrivals64(waves, scan)
    s64 = scan.astype(np.float64)
    ref0 = wavefront.plane(np.zeros(3), s64, d) / 1540.0
    ref1 = wavefront.spherical(np.zeros(3), s64, src) / 1480.0 + 2e-6
    np.testing.assert_allclose(t[0], ref0.astype(np.float32), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(t[1], ref1.astype(np.float32), rtol=1e-6, atol=1e-12)
    np.testing.assert_array_equal(t[2], (s64[:, 2] / 1540.0 - 1e-6).astype(np.float32))
    on_axis = np.array([[0.0, 0.0, 0.02]], dtype=np.float32)
    t_axis = wave_arrivals64([(1, (0.0, 0.0, 0.0), (0.0, 0.0, -0.01), 1500.0, 0.0)], on_axis)[0, 0]
    assert t_axis == np.float32(np.float64(np.float32(0.02)) / 1500.0)
    assert t.dtype == np.float32 and t.shape == (3, 500)
