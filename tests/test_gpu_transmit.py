"""On-device transmit arrivals (``mach_amd.transmit``, ``mach_beamform_args.tx_waves``,
``mach_tx_arrivals``) against the oracle (runs on a ROCm GPU).

Bars: the materialised arrivals equal the oracle's ``wave_arrivals64`` bit for bit (float32); a
beamform with in-kernel arrivals equals, bit for bit, the same beamform given those arrivals as an
array, and is within 1e-5 relative L2 of the FP64 oracle.
"""

import ctypes
import math

import numpy as np
import pytest
import torch

from mach_amd import _hip_impl, _lib, experimental, phantoms, transmit
from mach_amd.kernel import InterpolationType
from oracle.das64 import das64, rel_l2, wave_arrivals64

pytestmark = pytest.mark.gpu

KW = dict(f_number=1.2, rx_start_s=-2e-7, sampling_freq_hz=20e6)


def waves_and_specs():
    waves, specs = [], []
    for ang in (-0.2, 0.0, 0.13):
        d = (math.sin(ang), 0.0, math.cos(ang))
        waves.append(transmit.PlaneWave(direction=d, sound_speed_m_s=1540.0, time_offset_s=1e-7 * ang))
        specs.append((0, (0.0, 0.0, 0.0), d, 1540.0, 1e-7 * ang))
    for src, off in (((0.0, 0.0, -8e-3), 0.0), ((2e-3, -1e-3, -10e-3), 3e-7), ((0.0, 0.0, 12e-3), -1e-7)):
        waves.append(transmit.SphericalWave(focus_m=src, sound_speed_m_s=1540.0, origin_m=(0.0, 0.0, 0.0),
                                            time_offset_s=off))
        specs.append((1, (0.0, 0.0, 0.0), src, 1540.0, off))
    return waves, specs


def test_arrivals_match_oracle_bitwise(device):
    rng = np.random.default_rng(31)
    scan = (rng.normal(size=(5000, 3)) * [8e-3, 8e-3, 1e-2] + [0, 0, 2.5e-2]).astype(np.float32)
    scan[:10, 2] = 12e-3  # voxels at the depth of the focused wave's focus (sign(0) branch)
    waves, specs = waves_and_specs()
    got = transmit.arrivals(waves, torch.as_tensor(scan, device=device))
    torch.cuda.synchronize()
    ref = wave_arrivals64(specs, scan)
    np.testing.assert_array_equal(got.cpu().numpy(), ref)


@pytest.mark.parametrize("kind", ["iq", "rf", "fp16"])
@pytest.mark.parametrize("interp", [0, 1, 2])
def test_in_kernel_arrivals_equal_array_path_and_oracle(device, kind, interp):
    rng = np.random.default_rng(32 + interp)
    waves, specs = waves_and_specs()
    T, E, S, F, V = len(waves), 9, 64, 4, 400
    D = rng.normal(size=(T, E, S, F)).astype(np.float32)
    if kind != "rf":
        D = (D + 1j * rng.normal(size=D.shape)).astype(np.complex64)
    rx = np.stack([np.linspace(-4e-3, 4e-3, E), np.zeros(E), np.zeros(E)], -1).astype(np.float32)
    vox = (rng.uniform(-1, 1, size=(V, 3)) * [6e-3, 0, 1e-3] + [0, 0, 1.5e-3]).astype(np.float32)
    arr = wave_arrivals64(specs, vox)
    if kind == "fp16":
        D16 = np.stack([D.real, D.imag], -1).astype(np.float16)
        D = (D16[..., 0].astype(np.float32) + 1j * D16[..., 1].astype(np.float32)).astype(np.complex64)
        d = torch.as_tensor(D16).to(device)
    else:
        d = torch.as_tensor(D).to(device)
    r, v, a = (torch.as_tensor(x).to(device) for x in (rx, vox, arr))
    fmod = None if kind == "rf" else 5e6
    y_waves = transmit.beamform(d, r, v, waves, modulation_freq_hz=fmod, tukey_alpha=0.5,
                                interp_type=InterpolationType(interp), **KW)
    y_array = experimental.beamform(d, r, v, a, modulation_freq_hz=fmod, tukey_alpha=0.5, sound_speed_m_s=1540.0,
                                    interp_type=InterpolationType(interp), **KW)
    torch.cuda.synchronize()
    assert torch.equal(y_waves, y_array)
    ref = das64(D, rx, vox, arr, **KW, sound_speed_m_s=1540.0, modulation_freq_hz=fmod or 0.0, tukey_alpha=0.5,
                interp_type=interp)
    assert np.abs(ref).max() > 0
    assert rel_l2(y_waves.cpu().numpy(), ref) <= 1e-5


def test_volumetric_workload_with_diverging_waves(device):
    """The volumetric workload's five diverging waves as descriptors instead of its arrival array."""
    acq = phantoms.make("volumetric", device=device, scale=0.25)
    p = dict(acq.params)
    waves = [transmit.SphericalWave(focus_m=tuple(s), sound_speed_m_s=p["sound_speed_m_s"])
             for s in phantoms._volumetric_sources()]
    cd = acq.fp16_channel_data()
    kw = {k: p[k] for k in ("f_number", "rx_start_s", "sampling_freq_hz", "modulation_freq_hz", "tukey_alpha")}
    y_waves = transmit.beamform(cd, acq.rx_coords_m, acq.scan_coords_m, waves,
                                interp_type=InterpolationType(p["interp_type"]), **kw)
    arr = transmit.arrivals(waves, acq.scan_coords_m)
    y_array = experimental.beamform(cd, acq.rx_coords_m, acq.scan_coords_m, arr, sound_speed_m_s=p["sound_speed_m_s"],
                                    interp_type=InterpolationType(p["interp_type"]), **kw)
    # the workload's own arrivals (host float64 from float64 grid points) differ only in rounding
    y_host = experimental.beamform(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                                   sound_speed_m_s=p["sound_speed_m_s"],
                                   interp_type=InterpolationType(p["interp_type"]), **kw)
    torch.cuda.synchronize()
    assert torch.equal(y_waves, y_array)
    err = rel_l2(y_waves.cpu().numpy(), y_host.cpu().numpy())
    n_diff = int((arr != acq.tx_wave_arrivals_s).sum())
    print(f"in-kernel vs host arrivals: rel_l2={err:.2e}, arrivals differing: {n_diff}", flush=True)
    assert torch.allclose(arr, acq.tx_wave_arrivals_s, rtol=1e-6, atol=0.0)
    assert err <= 5e-4


def test_waves_in_a_batch_and_errors(device):
    lib = _lib.load()
    rng = np.random.default_rng(33)
    waves, specs = waves_and_specs()
    T, E, S, F, V = len(waves), 5, 48, 2, 300
    D = (rng.normal(size=(T, E, S, F)) + 1j * rng.normal(size=(T, E, S, F))).astype(np.complex64)
    rx = np.stack([np.linspace(-3e-3, 3e-3, E), np.zeros(E), np.zeros(E)], -1).astype(np.float32)
    vox = (rng.uniform(-1, 1, size=(V, 3)) * [5e-3, 0, 1e-3] + [0, 0, 1.5e-3]).astype(np.float32)
    d, r, v = (torch.as_tensor(x).to(device) for x in (D, rx, vox))
    packed = transmit.pack(waves, device)
    a = torch.as_tensor(wave_arrivals64(specs, vox)).to(device)
    base = dict(n_tx=T, kind=_lib.DATA_C64, **KW, sound_speed_m_s=1540.0, modulation_freq_hz=5e6, tukey_alpha=0.5,
                interp_type=1, accumulate=False)
    out_plain = torch.empty((V, F), dtype=torch.complex64, device=device)
    out_batch = torch.empty_like(out_plain)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = _hip_impl.make_args(d, r, v, a, out_plain, **base)
    assert lib.mach_beamform(ctypes.byref(args), stream) == 0, _lib.last_error()
    bargs = _hip_impl.make_args(d, r, v, None, out_batch, tx_waves=packed, **base)
    bp = ctypes.c_void_p()
    assert lib.mach_batch_create(2, 1, ctypes.byref(bp)) == 0
    try:
        assert lib.mach_batch_submit(bp, ctypes.byref(bargs), 1, stream) == 0, _lib.last_error()
        torch.cuda.synchronize()
    finally:
        lib.mach_batch_destroy(bp)
    assert torch.equal(out_plain, out_batch)
    both = _hip_impl.make_args(d, r, v, a, out_plain, tx_waves=packed, **base)
    assert lib.mach_beamform(ctypes.byref(both), stream) == _lib.ERR_INVALID_ARGUMENT
    assert "not both" in _lib.last_error()
    with pytest.raises(ValueError, match="unit vector"):
        transmit.PlaneWave(direction=(0.0, 0.0, 2.0), sound_speed_m_s=1540.0)


@pytest.mark.parametrize("wave", [0, 4])
def test_lateral_layouts_with_in_kernel_arrivals(device, wave):
    """The lateral-group kernel evaluates the transmit arrival per voxel too: with ``tx_waves`` every
    lateral layout equals, bit for bit, the general kernel given the same arrivals as an array; and
    accumulation into ``out`` adds exactly the same values."""
    from mach_amd import autotune

    waves, specs = waves_and_specs()
    waves, specs = [waves[wave]], [specs[wave]]
    acq = phantoms.doppler(device=device, scale=0.25, n_frames=64)
    scan = acq.scan_coords_m
    arr = torch.as_tensor(wave_arrivals64(specs, scan.cpu().numpy()), device=device)
    packed = transmit.pack(waves, device=device)
    order = _hip_impl.voxel_order(scan)
    p = acq.params
    lib = _lib.load()

    def run(out, layout, use_waves, accumulate=False):
        args = _hip_impl.make_args(acq.channel_data, acq.rx_coords_m, scan, None if use_waves else arr, out, n_tx=1,
                                   kind=_hip_impl._data_kind(acq.channel_data), f_number=p["f_number"],
                                   rx_start_s=p["rx_start_s"], sampling_freq_hz=p["sampling_freq_hz"],
                                   sound_speed_m_s=p["sound_speed_m_s"], modulation_freq_hz=p["modulation_freq_hz"],
                                   tukey_alpha=p["tukey_alpha"], interp_type=p["interp_type"], accumulate=accumulate,
                                   layout=layout, order=order, tx_waves=packed if use_waves else None)
        _lib.check(lib.mach_beamform(ctypes.byref(args), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    ref = torch.zeros((acq.n_vox, 64), dtype=torch.complex64, device=device)
    run(ref, autotune.encode(16, 4), False)
    for lay in [x for x in autotune.candidates(64, band=False, lateral=True) if x & autotune.LATERAL_FLAG]:
        got = torch.zeros_like(ref)
        run(got, lay, True)
        twice = got.clone()
        run(twice, lay, True, accumulate=True)
        torch.cuda.synchronize()
        assert torch.abs(ref).max() > 0
        assert torch.equal(got, ref), hex(lay)
        assert torch.equal(twice, ref + ref), hex(lay)
