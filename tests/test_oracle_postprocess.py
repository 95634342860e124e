"""Pins the oracle's post-processing restatements (power Doppler, envelope, dB) with closed forms.

References: power Doppler ``np.square(np.abs(y)).sum(axis=-1)`` (examples/doppler.py:252);
``db`` / ``db_zero`` (src/mach/_vis.py:17-40).
"""

import numpy as np
import pytest

from oracle.das64 import db64, db_zero64, envelope64, power_doppler64


def test_power_of_a_pure_tone_is_amplitude_squared_times_frames():
    # Parseval: |a e^{j w f}|^2 = |a|^2 for every frame, whatever the Doppler frequency
    n = 256
    f = np.arange(n)
    a = np.array([3 + 4j, 0.5j, -2.0])
    y = a[:, None] * np.exp(1j * 2 * np.pi * np.array([0.1, 0.37, 0.0])[:, None] * f[None, :])
    np.testing.assert_allclose(power_doppler64(y), np.abs(a) ** 2 * n, rtol=1e-12)


def test_power_of_rf_is_sum_of_squares_and_ignores_sign():
    y = np.array([[1.0, -2.0, 3.0], [0.0, 0.0, 0.0]], dtype=np.float32)
    np.testing.assert_array_equal(power_doppler64(y), [14.0, 0.0])


def test_power_drops_a_wrong_axis_or_missing_square():
    # a plausible slip (summing over voxels, or |y| instead of |y|^2) fails this case
    y = np.array([[3 + 4j, 0j], [1j, 1 + 0j]])
    np.testing.assert_array_equal(power_doppler64(y), [25.0, 2.0])


def test_envelope_is_complex_magnitude_and_rf_absolute_value():
    np.testing.assert_array_equal(envelope64(np.array([3 + 4j, -5 - 12j, 0j])), [5.0, 13.0, 0.0])
    np.testing.assert_array_equal(envelope64(np.array([-2.5, 0.0, 7.0], dtype=np.float32)), [2.5, 0.0, 7.0])


def test_db_decades_and_power_mode():
    np.testing.assert_allclose(db64([1.0, 10.0, 0.1, 1000.0]), [0.0, 20.0, -20.0, 60.0], atol=1e-12)
    np.testing.assert_allclose(db64([1.0, 10.0, 0.01], power_mode=True), [0.0, 10.0, -20.0], atol=1e-12)


def test_db_replaces_exact_zero_by_float64_epsilon():
    eps = np.finfo(float).eps
    assert db64([0.0])[0] == pytest.approx(20 * np.log10(eps), rel=1e-15)
    assert db64([0.0], power_mode=True)[0] == pytest.approx(10 * np.log10(eps), rel=1e-15)
    # a tiny non-zero value is kept, not floored
    assert db64([1e-20])[0] == pytest.approx(-400.0, rel=1e-12)


def test_db_zero_normalises_to_the_maximum():
    d = db_zero64(np.array([[2.0, 20.0], [0.2, 0.0]]))
    np.testing.assert_allclose(d[:, :3].ravel()[:3], [-20.0, 0.0, -40.0], atol=1e-12)
    assert d.max() == 0.0
    # all-zero image: every value is epsilon, so the image is 0 dB everywhere
    np.testing.assert_array_equal(db_zero64(np.zeros(4)), np.zeros(4))


def test_analytic_signal_matches_scipy_hilbert_and_rearranges():
    """analytic_signal64 = scipy.signal.hilbert along samples (the reference's UFF preprocessing,
    src/mach/io/uff.py:86-88), then (S, E, T, F) -> (T, E, S, F)."""
    import scipy.signal

    from oracle.das64 import analytic_signal64

    rng = np.random.default_rng(3)
    for n in (9, 10, 64, 96):
        rf = rng.normal(size=(n, 3, 2, 4))
        got = analytic_signal64(rf)
        ref = scipy.signal.hilbert(rf, axis=0).transpose(2, 1, 0, 3)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
        assert got.shape == (2, 3, n, 4)
        np.testing.assert_allclose(got.real, rf.transpose(2, 1, 0, 3), atol=1e-12)  # real part is the RF


def test_analytic_signal_of_a_cosine_is_the_complex_exponential():
    # a bin-centred cosine has the exact analytic signal exp(j w t); a wrong sign or a dropped doubling fails
    from oracle.das64 import analytic_signal64

    n, k = 64, 5
    t = np.arange(n)
    rf = np.cos(2 * np.pi * k * t / n)[:, None, None]
    np.testing.assert_allclose(analytic_signal64(rf)[0, 0, :, 0], np.exp(1j * 2 * np.pi * k * t / n), atol=1e-12)


def test_bmode_db_diff_closed_forms():
    """bmode_db_diff (the FP16 B-mode dB figure): each case fails a plausible mistake. Exact decades
    give exact dB; a factor 2 is 20 log10 2 = 6.0206 dB (10 log10 would give 3.01); each image is
    normalised by its own maximum (a global scale changes nothing); the -60 dB mask is taken on the
    reference image and is strict (a voxel at exactly -60 dB, or below it, is ignored however wrong)."""
    from oracle.das64 import bmode_db_diff

    ref = np.array([1.0, 1e-1, 1e-2, 1e-3, 1e-4])
    assert bmode_db_diff(ref, ref) == 0.0
    assert bmode_db_diff(7.5 * ref, ref) == pytest.approx(0.0, abs=1e-12)  # own-maximum normalisation
    y = ref.copy()
    y[1] = 2e-1
    assert bmode_db_diff(y, ref) == pytest.approx(20 * np.log10(2.0), rel=1e-12)
    y = ref.copy()
    y[3:] = [0.5, 0.9]  # only voxels at or below -60 dB in the reference differ
    assert bmode_db_diff(y, ref) == pytest.approx(0.0, abs=1e-12)
    # the mask comes from the reference: a voxel that collapses in y still counts
    assert bmode_db_diff(np.array([1.0, 1e-4]), np.array([1.0, 0.5])) == pytest.approx(80 - 20 * np.log10(2.0),
                                                                                      rel=1e-12)
    # magnitudes of complex values; a phase change alone is no difference
    assert bmode_db_diff(np.array([1j, -0.1]), np.array([1.0, 0.1 + 0j])) == pytest.approx(0.0, abs=1e-12)
    # normalised by each image's own maximum: [2, 1] vs [1, 1] differ by 6.02 dB at the second voxel
    assert bmode_db_diff(np.array([2.0, 1.0]), np.array([1.0, 1.0])) == pytest.approx(20 * np.log10(2.0), rel=1e-12)
    # a floor other than -60 dB moves the mask
    assert bmode_db_diff(y, ref, floor_db=-90.0) > 50
