"""UFF beamforming setup (mach_amd.io.uff; reference src/mach/io/uff.py) on CPU: wave directions and
plane-wave arrivals against their closed forms, the sequence delays, the layout of IQ channel data, and
the setup dictionary the reference's callers pass to beamform. The UFF objects are small stand-ins with
pyuff_ustb's attribute names (pyuff_ustb is not installed; the PICMUS files are downloads)."""

from types import SimpleNamespace as NS

import numpy as np
import pytest
import torch

from mach_amd.io import uff

C = 1540.0


def acquisition(n_s=64, n_e=8, angles_deg=(-10.0, 0.0, 10.0), delays=(0.0, 2e-6, 4e-6), frames=None, iq=True,
                seed=0):
    rng = np.random.default_rng(seed)
    shape = (n_s, n_e, len(angles_deg)) + ((frames,) if frames else ())
    data = rng.normal(size=shape).astype(np.float32)
    if iq:
        data = (data + 1j * rng.normal(size=shape)).astype(np.complex64)
    probe = NS(xyz=np.stack([np.linspace(-2e-3, 2e-3, n_e), np.zeros(n_e), np.zeros(n_e)], -1))
    seq = [NS(source=NS(azimuth=np.deg2rad(a), elevation=0.0), delay=d) for a, d in zip(angles_deg, delays)]
    cd = NS(data=data, sound_speed=C, modulation_frequency=5.2e6 if iq else 0.0, sampling_frequency=20.8e6,
            initial_time=1.5e-6, probe=probe, sequence=seq)
    x, z = np.meshgrid(np.linspace(-3e-3, 3e-3, 5), np.linspace(5e-3, 9e-3, 4), indexing="ij")
    scan = NS(xyz=np.stack([x.ravel(), np.zeros(x.size), z.ravel()], -1)[None])
    return cd, scan


def test_wave_directions_closed_form():
    az, el = np.deg2rad([-12.0, 0.0, 7.5, 0.0]), np.deg2rad([0.0, 0.0, 0.0, 9.0])
    seq = [NS(source=NS(azimuth=a, elevation=e), delay=0.0) for a, e in zip(az, el)]
    d = uff.extract_wave_directions(seq)
    want = np.stack([np.sin(az), np.sin(el) * np.cos(az), np.cos(el) * np.cos(az)], -1)
    assert d.shape == (4, 3) and np.allclose(d, want, atol=1e-15)
    assert np.allclose(np.linalg.norm(d, axis=-1), 1.0)
    dt = uff.extract_wave_directions(seq, xp=torch)
    assert isinstance(dt, torch.Tensor) and np.allclose(dt.numpy(), want, atol=1e-15)


def test_plane_wave_arrivals_closed_form():
    th = np.deg2rad([-16.0, 0.0, 5.0])
    dirs = np.stack([np.sin(th), np.zeros(3), np.cos(th)], -1)
    pts = np.array([[0.0, 0.0, 10e-3], [2e-3, 0.0, 20e-3], [-4e-3, 1e-3, 5e-3]])
    t = uff.compute_tx_wave_arrivals_s(dirs, pts, C)
    want = (pts[None, :, 0] * np.sin(th)[:, None] + pts[None, :, 2] * np.cos(th)[:, None]) / C
    assert t.shape == (3, 3) and np.allclose(t, want, rtol=1e-14, atol=0)
    # a shifted origin: distances from it; torch input gives a torch result with the same values
    org = np.array([1e-3, 0.0, -2e-3])
    t2 = uff.compute_tx_wave_arrivals_s(dirs, pts, C, origin=org)
    assert np.allclose(t2, ((pts - org) @ dirs.T).T / C, rtol=1e-14)
    tt = uff.compute_tx_wave_arrivals_s(dirs, torch.as_tensor(pts), C)
    assert isinstance(tt, torch.Tensor) and np.allclose(tt.numpy(), want, rtol=1e-14)


def test_sequence_delays():
    seq = [NS(source=NS(azimuth=0.0, elevation=0.0), delay=d) for d in (0.0, 1e-6, 3.5e-6)]
    assert np.array_equal(uff.extract_sequence_delays(seq), [0.0, 1e-6, 3.5e-6])
    assert torch.equal(uff.extract_sequence_delays(seq, xp=torch), torch.tensor([0.0, 1e-6, 3.5e-6], dtype=torch.float64))


@pytest.mark.parametrize("frames", [None, 3])
def test_setup_of_iq_acquisition(frames):
    cd, scan = acquisition(frames=frames)
    s = uff.create_beamforming_setup(cd, scan, device="cpu")
    # the keys the reference returns (uff.py:164-175), so that beamform(**setup) works
    assert set(s) == {"channel_data", "rx_coords_m", "scan_coords_m", "tx_wave_arrivals_s", "out", "f_number",
                      "sampling_freq_hz", "sound_speed_m_s", "modulation_freq_hz", "rx_start_s"}
    assert s["out"] is None and s["f_number"] == 1.7 and s["rx_start_s"] == 1.5e-6
    assert s["sampling_freq_hz"] == 20.8e6 and s["sound_speed_m_s"] == C and s["modulation_freq_hz"] == 5.2e6
    # [samples, elements, waves(, frames)] -> [waves, elements, samples, frames]
    d = s["channel_data"]
    n_f = frames or 1
    assert d.shape == (3, 8, 64, n_f) and d.dtype == torch.complex64 and d.is_contiguous()
    src = cd.data if frames else cd.data[..., None]
    for w, e, k, f in [(0, 0, 0, 0), (2, 5, 17, n_f - 1), (1, 7, 63, 0)]:
        assert d[w, e, k, f].item() == complex(src[k, e, w, f])
    # arrivals: plane wave from the origin over c, plus each transmit's sequence delay, as float32
    pts = scan.xyz[0]
    th = np.deg2rad([-10.0, 0.0, 10.0])
    want = (pts[None, :, 0] * np.sin(th)[:, None] + pts[None, :, 2] * np.cos(th)[:, None]) / C
    want = want + np.array([0.0, 2e-6, 4e-6])[:, None]
    arr = s["tx_wave_arrivals_s"]
    assert arr.dtype == torch.float32 and arr.shape == (3, 20)
    assert np.array_equal(arr.numpy(), want.astype(np.float32))
    assert s["scan_coords_m"].dtype == torch.float32 and np.array_equal(s["scan_coords_m"].numpy(), pts.astype(np.float32))
    assert np.array_equal(s["rx_coords_m"].numpy(), cd.probe.xyz.astype(np.float32))


def test_single_transmit_setup_and_fp16_storage():
    cd, scan = acquisition()
    multi = uff.create_beamforming_setup(cd, scan, device="cpu")
    one = uff.create_single_transmit_beamforming_setup(cd, scan, wave_index=2, f_number=1.2, device="cpu")
    assert one["f_number"] == 1.2
    assert torch.equal(one["channel_data"], multi["channel_data"][2])
    assert torch.equal(one["tx_wave_arrivals_s"], multi["tx_wave_arrivals_s"][2])
    h = uff.create_beamforming_setup(cd, scan, device="cpu", storage="fp16")["channel_data"]
    assert h.dtype == torch.float16 and h.shape == (3, 8, 64, 1, 2)
    assert torch.equal(h[..., 0], multi["channel_data"].real.to(torch.float16))


def test_setup_rejects_what_the_reference_rejects():
    cd, scan = acquisition()
    with pytest.raises(ValueError, match="one spatial grid"):
        uff.create_beamforming_setup(cd, NS(xyz=np.concatenate([scan.xyz, scan.xyz])), device="cpu")
    cd.data = cd.data[:, :, 0, None, None, None]
    with pytest.raises(ValueError, match="3 or 4 dimensions"):
        uff.create_beamforming_setup(cd, scan, device="cpu")


def test_sequence_delays_are_added_to_the_arrivals_physics():
    """On a synthetic RF acquisition whose waves fire `delay` after time zero, the oracle's compounded
    image peaks at the scatterer with the setup's arrivals (plane wave + delay) and not with the delays
    subtracted: the convention of uff.py:160 against the physics, on CPU (oracle only)."""
    import importlib.util
    import os

    from oracle.das64 import analytic_signal64, das64

    spec = importlib.util.spec_from_file_location("gpu_uff", os.path.join(os.path.dirname(__file__), "test_gpu_uff.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    cd, scan = m.rf_acquisition((1e-3, 0.0, 12e-3), n_e=32, n_s=1024)
    pts = scan.xyz[0].reshape(81, 81, 3)[::4, ::4].reshape(-1, 3)
    iq = analytic_signal64(cd.data).astype(np.complex64)
    base = uff.compute_tx_wave_arrivals_s(uff.extract_wave_directions(cd.sequence), pts, m.C)
    dl = uff.extract_sequence_delays(cd.sequence)[:, None]
    peaks = []
    for arr in (base + dl, base - dl):
        y = das64(iq, cd.probe.xyz.astype(np.float32), pts.astype(np.float32), arr.astype(np.float32), f_number=1.7,
                  rx_start_s=m.T0, sampling_freq_hz=m.FS, sound_speed_m_s=m.C, modulation_freq_hz=0.0,
                  tukey_alpha=0.5, interp_type=1)
        peaks.append((pts[int(np.argmax(np.abs(y[:, 0])))], np.abs(y).max()))
    (p_add, a_add), (p_sub, a_sub) = peaks
    assert np.allclose(p_add, [1e-3, 0.0, 12e-3], atol=1e-9) and a_add > 2 * a_sub
    assert not np.allclose(p_sub, [1e-3, 0.0, 12e-3], atol=0.2e-3)
