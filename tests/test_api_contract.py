"""Host-side contract of the Python drop-in (reference: ``tests/test_nanobind.py`` and
``src/mach/kernel.py:180-275``): names, type/dtype errors, required modulation frequency, ``out``
validation and the reference's dimension-mismatch messages. Everything here is decided before any
device work, so it runs without a GPU; the numeric behaviour is covered by the GPU parity tests."""

import numpy as np
import pytest
import torch

import mach_amd as mach
from mach_amd import experimental, kernel
from mach_amd.kernel import InterpolationType, beamform


def make(n_rx=4, n_s=100, n_f=2, n_v=8, seed=0):
    rng = np.random.default_rng(seed)
    rf = rng.normal(size=(n_rx, n_s, n_f)).astype(np.float32)
    iq = (rf + 1j * rng.normal(size=rf.shape)).astype(np.complex64)
    coords = rng.normal(size=(n_rx, 3)).astype(np.float32)
    grid = rng.normal(size=(n_v, 3)).astype(np.float32)
    arr = np.zeros(n_v, np.float32)
    return rf, iq, coords, grid, arr


KW = dict(f_number=2.0, rx_start_s=0.0, sampling_freq_hz=1e6, sound_speed_m_s=1500.0)


def test_public_names():
    assert mach.beamform is kernel.beamform
    assert [m.name for m in InterpolationType] == ["NearestNeighbor", "Linear", "Quadratic"]
    assert [int(m) for m in InterpolationType] == [0, 1, 2]
    assert hasattr(kernel, "nb_beamform") and hasattr(kernel, "__hip_version__")
    assert callable(experimental.beamform)
    from mach_amd import _hip_impl

    assert _hip_impl.Linear == InterpolationType.Linear


def test_bool_channel_data_is_a_type_error():
    _, iq, coords, grid, arr = make()
    with pytest.raises(TypeError, match="channel_data must be array with dtype=numeric"):
        beamform(iq.astype(bool), coords, grid, arr, modulation_freq_hz=5e6, **KW)
    with pytest.raises(TypeError, match="rx_coords_m must be array with dtype=Real"):
        beamform(iq, coords.astype(np.complex64), grid, arr, modulation_freq_hz=5e6, **KW)
    with pytest.raises(TypeError, match="scan_coords_m must be array with dtype=Real"):
        beamform(iq, coords, [[0, 0, 1]], arr, modulation_freq_hz=5e6, **KW)


def test_complex_requires_modulation_frequency():
    _, iq, coords, grid, arr = make()
    with pytest.raises(ValueError, match="modulation_freq_hz is required"):
        beamform(iq, coords, grid, arr, **KW)
    with pytest.raises(ValueError, match="modulation_freq_hz is required"):
        beamform(torch.view_as_real(torch.from_numpy(iq)).half(), coords, grid, arr, **KW)


def test_out_validation():
    rf, iq, coords, grid, arr = make()
    with pytest.raises(ValueError, match="doesn't match expected"):
        beamform(iq, coords, grid, arr, out=np.zeros((8, 2), np.float32), modulation_freq_hz=5e6, **KW)
    with pytest.raises(ValueError, match="doesn't match expected"):
        beamform(rf, coords, grid, arr, out=np.zeros((8, 2), np.complex64), **KW)
    ro = np.zeros((8, 2), np.complex64)
    ro.flags.writeable = False
    with pytest.raises(ValueError, match="not writable"):
        beamform(iq, coords, grid, arr, out=ro, modulation_freq_hz=5e6, **KW)
    with pytest.raises(ValueError, match="must be contiguous"):
        beamform(iq, coords, grid, arr, out=np.zeros((2, 8), np.complex64).T, modulation_freq_hz=5e6, **KW)


def test_dimension_mismatch_messages():
    _, iq, coords, grid, arr = make()
    with pytest.raises(RuntimeError, match="Dimension mismatch in receive elements"):
        beamform(iq[:2], coords, grid, arr, modulation_freq_hz=5e6, **KW)
    with pytest.raises(RuntimeError, match="Dimension mismatch in output voxels"):
        beamform(iq, coords, grid, arr[:5], modulation_freq_hz=5e6, **KW)
    with pytest.raises(RuntimeError, match="Dimension mismatch in frames"):
        beamform(iq, coords, grid, arr, out=np.zeros((8, 3), np.complex64), modulation_freq_hz=5e6, **KW)
    try:
        beamform(iq[:2], coords, grid, arr, modulation_freq_hz=5e6, **KW)
    except RuntimeError as e:
        assert "channel_data.shape: [2, 100, 2]" in str(e) and "rx_coords_m.shape: [4, 3]" in str(e)


def test_raw_entry_is_strict():
    rf, iq, coords, grid, arr = make()
    out = np.zeros((8, 2), np.complex64)
    with pytest.raises(TypeError):
        kernel.nb_beamform(iq.astype(np.complex128), coords, grid, arr, out, 2.0, 0.0, 1e6, 1500.0, 5e6)
    with pytest.raises(TypeError):
        kernel.nb_beamform(iq, coords.astype(np.float64), grid, arr, out, 2.0, 0.0, 1e6, 1500.0, 5e6)
    with pytest.raises(TypeError, match="modulation_freq_hz is required"):
        kernel.nb_beamform(iq, coords, grid, arr, out, 2.0, 0.0, 1e6, 1500.0)


def test_non_contiguous_input_warns():
    _, iq, coords, grid, arr = make()
    nc = np.asfortranarray(iq)
    with pytest.warns(UserWarning, match="not contiguous"):
        try:
            beamform(nc, coords, grid, arr, modulation_freq_hz=5e6, **KW)
        except RuntimeError as e:  # no GPU in the CPU tier: the call then fails loudly
            assert "No HIP device" in str(e)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure mode")
def test_no_gpu_fails_loudly():
    _, iq, coords, grid, arr = make()
    with pytest.raises(RuntimeError, match="No HIP device"):
        beamform(iq, coords, grid, arr, modulation_freq_hz=5e6, **KW)
    with pytest.raises(RuntimeError, match="No HIP device"):
        experimental.beamform(iq[None], coords, grid, arr[None], modulation_freq_hz=5e6, **KW)


def test_reference_import_name():
    """``import mach`` (the reference's package name) gives this package: the same module objects under
    the reference's dotted names, so code written for the reference imports unchanged."""
    import importlib

    import mach
    import mach_amd

    assert mach.beamform is mach_amd.beamform
    assert mach.InterpolationType is mach_amd.InterpolationType
    for name in ("kernel", "experimental", "wavefront", "geometry", "io", "io.must", "io.uff"):
        mod = importlib.import_module(f"mach.{name}")
        assert mod is importlib.import_module(f"mach_amd.{name}"), name
    from mach.io.uff import create_beamforming_setup  # the reference's UFF test imports it so
    from mach.kernel import nb_beamform  # the raw entry the reference's benchmarks call

    assert create_beamforming_setup is mach_amd.io.uff.create_beamforming_setup

    assert nb_beamform is mach_amd.kernel.nb_beamform
    assert set(["beamform", "geometry", "kernel", "wavefront", "__version__", "Array"]) <= set(mach.__all__)
