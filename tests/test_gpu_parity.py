"""Parity of the HIP beamformer with the FP64 oracle (runs on a ROCm GPU).

Bar (stated per test): FP32 path <= 1e-5 relative L2 against the FP64 oracle evaluated on the same
float32/complex64 inputs; FP16-IQ storage <= 1e-5 against the oracle on the same FP16-rounded data
and <= 1e-3 end to end against the FP32 data. Every call goes through the Python drop-in and the C
ABI (``mach_beamform*`` in include/mach_hip.h).
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

import mach_amd as mach
from mach_amd import _lib, experimental, phantoms
from mach_amd.kernel import InterpolationType
from oracle.das64 import bmode_db_diff, das64, rel_l2
from oracle.das_c import das_c

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5
TOL_FP16_SAME_DATA = 1e-5
TOL_FP16_END_TO_END = 1e-3
TOL_BMODE_DB = 0.1  # B-mode (envelope, dB) difference of FP16 storage against FP32 data, over -60 dB


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_case(rng, T=1, E=5, S=40, F=3, V=9, kind="iq"):
    D = rng.normal(size=(T, E, S, F)).astype(np.float32)
    if kind != "rf":
        D = (D + 1j * rng.normal(size=D.shape)).astype(np.complex64)
    rx = (rng.normal(size=(E, 3)) * 1e-3).astype(np.float32)
    vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 2e-3]).astype(np.float32)
    arr = (rng.normal(size=(T, V)) * 1e-6 + 0.5e-6).astype(np.float32)
    return D, rx, vox, arr


KW = dict(f_number=0.8, rx_start_s=-1e-7, sampling_freq_hz=20e6, sound_speed_m_s=1540.0)


def to_dev(*arrays, device):
    return [torch.as_tensor(a).to(device) for a in arrays]


def gpu_single(D, rx, vox, arr, device, **k):
    d, r, v, a = to_dev(D, rx, vox, arr, device=device)
    if D.dtype == np.float16:
        d = d.contiguous()
    y = mach.beamform(d, r, v, a, **k)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("kind", ["rf", "iq"])
@pytest.mark.parametrize("alpha", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("frames", [1, 2, 3, 8, 33, 64, 256])
def test_random_cases_match_oracle(device, interp, kind, alpha, frames):
    log(f"case interp={interp} kind={kind} alpha={alpha} frames={frames}")
    rng = np.random.default_rng(1000 * interp + 10 * frames + int(alpha * 4) + (kind == "iq"))
    V = 70 if frames < 64 else 13
    D, rx, vox, arr = random_case(rng, E=7, S=48, F=frames, V=V, kind=kind)
    k = dict(KW, modulation_freq_hz=5e6 if kind == "iq" else None, tukey_alpha=alpha,
             interp_type=InterpolationType(interp))
    y = gpu_single(D[0], rx, vox, arr[0], device, **k)
    k_or = dict(k, modulation_freq_hz=k["modulation_freq_hz"] or 0.0, interp_type=interp)
    ref = das64(D[0], rx, vox, arr[0], **k_or)
    assert np.abs(ref).max() > 0
    err = rel_l2(y, ref)
    log(f"  rel_l2={err:.3e}")
    assert err <= TOL_FP32


@pytest.mark.parametrize("interp", [0, 1, 2])
def test_golden_vectors(device, interp):
    import os

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "das_small.npz"), allow_pickle=False)
    names = sorted({key.split("__")[0] for key in z.files})
    for name in names:
        case = {key.split("__")[1]: z[key] for key in z.files if key.startswith(name + "__")}
        p = case["params"]
        k = dict(f_number=p[0], rx_start_s=p[1], sampling_freq_hz=p[2], sound_speed_m_s=p[3],
                 modulation_freq_hz=p[4], tukey_alpha=p[5], interp_type=int(p[6]))
        if k["interp_type"] != interp:
            continue
        log(f"golden {name}")
        D = case["D"]
        d, r, v, a = to_dev(D, case["rx"], case["vox"], case["arr"], device=device)
        out = torch.zeros((v.shape[0], D.shape[3]), dtype=torch.float32 if name.startswith("rf") else torch.complex64,
                          device=device)
        experimental.beamform(d, r, v, a, out=out, f_number=k["f_number"], rx_start_s=k["rx_start_s"],
                              sampling_freq_hz=k["sampling_freq_hz"], sound_speed_m_s=k["sound_speed_m_s"],
                              modulation_freq_hz=k["modulation_freq_hz"], tukey_alpha=k["tukey_alpha"],
                              interp_type=InterpolationType(interp))
        torch.cuda.synchronize()
        assert rel_l2(out.cpu().numpy(), case["y"]) <= TOL_FP32


def test_c_abi_entry_points_directly(device):
    """mach_beamform_c64 / mach_beamform_f32 called through ctypes with raw device pointers."""
    lib = _lib.load()
    rng = np.random.default_rng(5)
    for kind, fn in (("iq", lib.mach_beamform_c64), ("rf", lib.mach_beamform_f32)):
        D, rx, vox, arr = random_case(rng, F=4, V=50, kind=kind)
        d, r, v, a = to_dev(D[0], rx, vox, arr[0], device=device)
        out = torch.zeros((50, 4), dtype=torch.complex64 if kind == "iq" else torch.float32, device=device)
        out += 1.0  # the C entry accumulates into out
        st = fn(d.data_ptr(), r.data_ptr(), v.data_ptr(), a.data_ptr(), out.data_ptr(), 5, 40, 4, 50,
                0.8, -1e-7, 20e6, 1540.0, 5e6, 0.5, 2, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert st == 0, _lib.last_error()
        torch.cuda.synchronize()
        ref = das64(D[0], rx, vox, arr[0], **KW, modulation_freq_hz=5e6 if kind == "iq" else 0.0,
                    tukey_alpha=0.5, interp_type=2) + 1.0
        assert rel_l2(out.cpu().numpy(), ref) <= TOL_FP32


def test_compounding_equals_sum_of_single_transmits(device):
    rng = np.random.default_rng(11)
    D, rx, vox, arr = random_case(rng, T=6, F=2, V=200)
    k = dict(KW, modulation_freq_hz=5e6, tukey_alpha=0.5)
    d, r, v, a = to_dev(D, rx, vox, arr, device=device)
    y = experimental.beamform(d, r, v, a, **k)
    singles = torch.zeros_like(y)
    for t in range(6):
        mach.beamform(d[t], r, v, a[t], out=singles, **k)
    torch.cuda.synchronize()
    ref = das64(D, rx, vox, arr, **k, interp_type=1)
    assert rel_l2(y.cpu().numpy(), ref) <= TOL_FP32
    assert rel_l2(singles.cpu().numpy(), ref) <= TOL_FP32
    assert rel_l2(y.cpu().numpy(), singles.cpu().numpy()) <= TOL_FP32


def test_out_semantics_device_accumulates_host_overwrites(device):
    rng = np.random.default_rng(12)
    D, rx, vox, arr = random_case(rng, F=3, V=40)
    k = dict(KW, modulation_freq_hz=5e6, tukey_alpha=0.5)
    ref = das64(D[0], rx, vox, arr[0], **k, interp_type=1)
    d, r, v, a = to_dev(D[0], rx, vox, arr[0], device=device)
    base = torch.full((40, 3), 2.0 - 1.0j, dtype=torch.complex64, device=device)
    out = base.clone()
    mach.beamform(d, r, v, a, out=out, **k)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref + (2.0 - 1.0j)) <= TOL_FP32
    host = np.full((40, 3), 7.0, np.complex64)
    with pytest.warns(UserWarning, match="latency"):
        mach.beamform(d, r, v, a, out=host, **k)
    assert rel_l2(host, ref) <= TOL_FP32  # a host out is overwritten (reference kernel.cu:1135-1139)
    with pytest.warns(UserWarning, match="latency"):
        y = mach.beamform(D[0], rx, vox, arr[0], **k)  # all-NumPy inputs are staged through the GPU
    assert isinstance(y, np.ndarray) and rel_l2(y, ref) <= TOL_FP32


def test_fp16_iq_storage(device):
    rng = np.random.default_rng(13)
    D, rx, vox, arr = random_case(rng, F=8, V=120)
    h = np.stack([D.real, D.imag], -1).astype(np.float16)
    k = dict(KW, modulation_freq_hz=5e6, tukey_alpha=0.5)
    y16 = gpu_single(h[0], rx, vox, arr[0], device, **k)
    same = das64(h[0], rx, vox, arr[0], **k, interp_type=1)
    full = das64(D[0], rx, vox, arr[0], **k, interp_type=1)
    e_same, e_full = rel_l2(y16, same), rel_l2(y16, full)
    log(f"fp16 storage: rel_l2 vs oracle(same fp16 data)={e_same:.2e}, vs oracle(fp32 data)={e_full:.2e}")
    assert e_same <= TOL_FP16_SAME_DATA and e_full <= TOL_FP16_END_TO_END


def test_edge_cases(device):
    rng = np.random.default_rng(14)
    k = dict(KW, modulation_freq_hz=5e6, tukey_alpha=0.5)
    # no voxels, no elements, one sample
    D, rx, vox, arr = random_case(rng, F=2, V=5)
    d, r, v, a = to_dev(D[0], rx, vox, arr[0], device=device)
    empty = mach.beamform(d, r, v[:0], a[:0], **k)
    assert empty.shape == (0, 2)
    y0 = mach.beamform(d[:0], r[:0], v, a, **k)
    torch.cuda.synchronize()
    assert torch.count_nonzero(y0).item() == 0
    one = np.ones((5, 1, 2), np.complex64)
    for interp in (0, 1, 2):
        y = gpu_single(one, rx, vox, arr[0], device, **dict(k, interp_type=InterpolationType(interp)))
        ref = das64(one, rx, vox, arr[0], **k, interp_type=interp)
        assert np.allclose(y, ref, atol=1e-6)
    # sample indices exactly on the interpolation bounds (s = 0, n-1, n-0.5 and 1, n-2)
    U = 1.0 / 1024.0
    rxe = np.array([[0, 0, 0]], np.float32)
    voxe = np.array([[0, 0, 8 * U]], np.float32)
    S = 16
    Dm = (np.arange(S, dtype=np.float32)[None, :, None] + 1j).astype(np.complex64)
    kk = dict(f_number=1.0, sampling_freq_hz=1024.0, sound_speed_m_s=1.0, modulation_freq_hz=0.0, tukey_alpha=0.0)
    for s_target in (0.0, 1.0, S - 2.0, S - 1.0, S - 0.5, -0.5, S - 0.75):
        t0 = 8 * U - s_target * U  # tau = 8U (rx) + arrival 0; s = (tau - t0) * 1024
        for interp in (0, 1, 2):
            y = gpu_single(Dm, rxe, voxe, np.zeros(1, np.float32), device,
                           **dict(kk, rx_start_s=t0, interp_type=InterpolationType(interp)))
            ref = das64(Dm, rxe, voxe, np.zeros(1, np.float32), **kk, rx_start_s=t0, interp_type=interp)
            assert np.allclose(y, ref, atol=1e-6), (s_target, interp, y, ref)


def test_repeatable_bitwise(device):
    acq = phantoms.plane_wave(device=device, scale=0.25)
    p = dict(acq.params, interp_type=InterpolationType(acq.params["interp_type"]))
    y1 = experimental.beamform(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, **p)
    y2 = experimental.beamform(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, **p)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)  # one owner per output element, fixed summation order, no atomics


def workload_parity(acq, device, voxel_sample=None, fp16=False, seed=0):
    """Full-size kernel run; oracle on all voxels or on a seeded voxel sample (voxels are independent)."""
    p = dict(acq.params)
    cd = acq.fp16_channel_data() if fp16 else acq.channel_data
    y = experimental.beamform(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                              **dict(p, interp_type=InterpolationType(p["interp_type"])))
    torch.cuda.synchronize()
    y = y.cpu().numpy()
    V = acq.n_vox
    idx = np.arange(V) if voxel_sample is None or voxel_sample >= V else np.sort(
        np.random.default_rng(seed).choice(V, voxel_sample, replace=False))
    cd_np = cd.cpu().numpy()
    ref = das_c(cd_np, acq.rx_coords_m.cpu().numpy(), acq.scan_coords_m.cpu().numpy()[idx],
                acq.tx_wave_arrivals_s.cpu().numpy()[:, idx], **p)
    return y[idx], ref, cd_np


@pytest.mark.parametrize("name,scale,sample", [
    ("plane_wave", 1.0, None),
    ("compound", 0.3, 4000),
    ("doppler", 0.5, 600),
    ("volumetric", 0.25, 3000),
    ("pymust_disk", 1.0, 2000),
])
def test_workloads_fp32(device, name, scale, sample):
    log(f"workload {name} scale={scale}")
    acq = phantoms.make(name, device=device, scale=scale)
    y, ref, _ = workload_parity(acq, device, sample)
    err = rel_l2(y, ref)
    log(f"  {name}: rel_l2={err:.3e} dB-diff={bmode_db_diff(y.mean(-1), ref.mean(-1)):.2e}")
    assert err <= TOL_FP32


def test_doppler_full_size_voxel_sample(device):
    """The headline workload at full size (128 ch x 1024 samples x 256 frames, 256x256 voxels); the
    oracle checks a seeded sample of 256 voxels with all frames."""
    acq = phantoms.doppler(device=device)
    y, ref, _ = workload_parity(acq, device, 256, seed=2)
    err = rel_l2(y, ref)
    log(f"  doppler full: rel_l2={err:.3e}")
    assert err <= TOL_FP32


@pytest.mark.parametrize("name,scale,sample", [("volumetric", 0.25, 3000), ("doppler", 0.5, 600),
                                               ("compound", 0.3, 4000)])
def test_workloads_fp16_storage(device, name, scale, sample):
    acq = phantoms.make(name, device=device, scale=scale)
    y16, ref16, _ = workload_parity(acq, device, sample, fp16=True)
    _, ref32, _ = workload_parity(acq, device, sample, fp16=False)
    e_same, e_e2e = rel_l2(y16, ref16), rel_l2(y16, ref32)
    db = bmode_db_diff(np.abs(y16).mean(-1), np.abs(ref32).mean(-1))
    log(f"  {name} fp16: rel_l2 same-data={e_same:.2e} end-to-end={e_e2e:.2e} B-mode dB diff={db:.3f}")
    assert e_same <= TOL_FP16_SAME_DATA and e_e2e <= TOL_FP16_END_TO_END
    assert db <= TOL_BMODE_DB


@pytest.mark.parametrize("frames", [1, 8, 12, 256])
def test_every_layout_gives_identical_results(device, frames):
    """The autotuner's layouts differ only in work assignment: outputs must be bitwise identical."""
    from mach_amd import _hip_impl, autotune

    rng = np.random.default_rng(frames)
    D, rx, vox, arr = random_case(rng, T=2, E=9, S=48, F=frames, V=150)
    d, r, v, a = to_dev(D, rx, vox, arr, device=device)
    outs = {}
    for lay in autotune.candidates(frames, band=False):
        out = torch.zeros((150, frames), dtype=torch.complex64, device=device)
        _hip_impl.beamform_compound(d, r, v, a, out, 0.8, -1e-7, 20e6, 1540.0, 5e6, 0.5, 2, layout=lay)
        outs[lay] = out
    torch.cuda.synchronize()
    ref = das64(D, rx, vox, arr, **KW, modulation_freq_hz=5e6, tukey_alpha=0.5, interp_type=2)
    first = next(iter(outs.values()))
    assert rel_l2(first.cpu().numpy(), ref) <= TOL_FP32
    for lay, o in outs.items():
        assert torch.equal(o, first), hex(lay)


@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("kind", ["iq", "iq16", "rf"])
@pytest.mark.parametrize("frames", [64, 128, 256])
def test_band_kernel_matches_plain_and_oracle(device, interp, kind, frames):
    """LDS-staged band kernel (selected by a band layout code) vs the general kernel with a forced
    layout: bitwise equal; and vs the oracle on a voxel sample. A regular depth-major
    grid keeps windows in LDS; the random cloud forces the global fallback for wide windows."""
    from mach_amd import _hip_impl, autotune

    log(f"band interp={interp} kind={kind} frames={frames}")
    acq = phantoms.doppler(device=device, scale=0.25, n_frames=frames)
    rng = np.random.default_rng(frames + interp)
    cloud = torch.as_tensor((rng.normal(size=(4200, 3)) * [8e-3, 0, 10e-3] + [0, 0, 25e-3]).astype(np.float32),
                            device=device)
    cloud_arr = (cloud[:, 2] / 1540.0)[None].contiguous()
    for scan, arr in ((acq.scan_coords_m, acq.tx_wave_arrivals_s), (cloud, cloud_arr)):
        cd = acq.channel_data
        if kind == "iq16":
            cd = acq.fp16_channel_data()
        elif kind == "rf":
            cd = acq.channel_data.real.contiguous()
        otype = torch.float32 if kind == "rf" else torch.complex64
        fmod = 0.0 if kind == "rf" else acq.params["modulation_freq_hz"]
        args = (acq.params["f_number"], 0.0, acq.params["sampling_freq_hz"], acq.params["sound_speed_m_s"], fmod, 0.5,
                interp)
        band = torch.zeros((scan.shape[0], frames), dtype=otype, device=device)
        plain = torch.zeros_like(band)
        for k in (4, 2, 1):  # widest band layout this dtype / frame count accepts
            try:
                _hip_impl.beamform_compound(cd, acq.rx_coords_m, scan, arr, band, *args, layout=autotune.encode(0, k))
                break
            except RuntimeError:
                continue
        else:
            raise AssertionError("no band layout accepted")
        _hip_impl.beamform_compound(cd, acq.rx_coords_m, scan, arr, plain, *args, layout=autotune.encode(64, 1))
        torch.cuda.synchronize()
        assert torch.equal(band, plain)
        idx = np.sort(rng.choice(scan.shape[0], 300, replace=False))
        ref = das_c(cd.cpu().numpy(), acq.rx_coords_m.cpu().numpy(), scan.cpu().numpy()[idx], arr.cpu().numpy()[:, idx],
                    f_number=args[0], rx_start_s=0.0, sampling_freq_hz=args[2], sound_speed_m_s=args[3],
                    modulation_freq_hz=fmod, tukey_alpha=0.5, interp_type=interp)
        assert np.abs(ref).max() > 0
        assert rel_l2(band.cpu().numpy()[idx], ref) <= TOL_FP32


def _beamform_raw(cd, rx, scan, arr, out, args, layout, output=0, order=None):
    """mach_beamform with a forced layout and output kind (0 signal, 1 power, 2 envelope)."""
    from mach_amd import _hip_impl

    kind = _hip_impl._data_kind(cd)
    a = _hip_impl.make_args(cd, rx, scan, arr, out, n_tx=arr.shape[0], kind=kind, f_number=args[0],
                            rx_start_s=args[1], sampling_freq_hz=args[2], sound_speed_m_s=args[3],
                            modulation_freq_hz=args[4], tukey_alpha=args[5], interp_type=args[6], accumulate=False,
                            layout=layout, order=order, output=output)
    _lib.check(_lib.load().mach_beamform(ctypes.byref(a), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))


@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("kind", ["iq", "iq16", "rf"])
@pytest.mark.parametrize("frames", [64, 256])
def test_lateral_layouts_match_general_and_oracle(device, interp, kind, frames):
    """Lateral-group layouts (2 or 4 voxels per lane reading shared sample windows) against the general
    kernel: bitwise equal for the signal and envelope outputs and, against the 64-lane general layout
    with the same frames per lane (same per-lane sums and butterfly), for the power image; and against
    the oracle on a voxel sample. The regular grid in depth-major order takes the window path; a
    random cloud (ordered and in caller order) also exercises the voxel-by-voxel fallback."""
    from mach_amd import _hip_impl, autotune

    log(f"lateral interp={interp} kind={kind} frames={frames}")
    acq = phantoms.doppler(device=device, scale=0.25, n_frames=frames)
    rng = np.random.default_rng(frames + 7 * interp)
    cloud = torch.as_tensor((rng.normal(size=(4100, 3)) * [8e-3, 0, 10e-3] + [0, 0, 25e-3]).astype(np.float32),
                            device=device)
    cloud_arr = (cloud[:, 2] / 1540.0)[None].contiguous()
    cd = acq.channel_data
    if kind == "iq16":
        cd = acq.fp16_channel_data()
    elif kind == "rf":
        cd = acq.channel_data.real.contiguous()
    otype = torch.float32 if kind == "rf" else torch.complex64
    fmod = 0.0 if kind == "rf" else acq.params["modulation_freq_hz"]
    args = (acq.params["f_number"], 0.0, acq.params["sampling_freq_hz"], acq.params["sound_speed_m_s"], fmod, 0.5,
            interp)
    lateral = [lay for lay in autotune.candidates(frames, band=False, lateral=True) if lay & autotune.LATERAL_FLAG]
    assert lateral
    cases = ((acq.scan_coords_m, acq.tx_wave_arrivals_s, True), (cloud, cloud_arr, True), (cloud, cloud_arr, False))
    for scan, arr, ordered in cases:
        order = _hip_impl.voxel_order(scan) if ordered else None
        n = scan.shape[0]
        ref_sig = torch.zeros((n, frames), dtype=otype, device=device)
        _beamform_raw(cd, acq.rx_coords_m, scan, arr, ref_sig, args, autotune.encode(32, 2), order=order)
        ref_env = torch.zeros((n, frames), dtype=torch.float32, device=device)
        _beamform_raw(cd, acq.rx_coords_m, scan, arr, ref_env, args, autotune.encode(32, 2), 2, order)
        for lay in lateral:
            k = lay & 0xF
            sig = torch.zeros_like(ref_sig)
            _beamform_raw(cd, acq.rx_coords_m, scan, arr, sig, args, lay, order=order)
            env = torch.zeros_like(ref_env)
            _beamform_raw(cd, acq.rx_coords_m, scan, arr, env, args, lay, 2, order)
            pw = torch.zeros((n,), dtype=torch.float32, device=device)
            _beamform_raw(cd, acq.rx_coords_m, scan, arr, pw, args, lay, 1, order)
            pw_ref = torch.zeros_like(pw)
            _beamform_raw(cd, acq.rx_coords_m, scan, arr, pw_ref, args, autotune.encode(64, k), 1, order)
            torch.cuda.synchronize()
            assert torch.equal(sig, ref_sig), hex(lay)
            assert torch.equal(env, ref_env), hex(lay)
            assert torch.equal(pw, pw_ref), hex(lay)
        idx = np.sort(rng.choice(n, 300, replace=False))
        ref = das_c(cd.cpu().numpy(), acq.rx_coords_m.cpu().numpy(), scan.cpu().numpy()[idx], arr.cpu().numpy()[:, idx],
                    f_number=args[0], rx_start_s=0.0, sampling_freq_hz=args[2], sound_speed_m_s=args[3],
                    modulation_freq_hz=fmod, tukey_alpha=0.5, interp_type=interp)
        assert np.abs(ref).max() > 0
        assert rel_l2(sig.cpu().numpy()[idx], ref) <= TOL_FP32


def test_lateral_layout_rejected_for_several_transmits(device):
    """Lateral groups sum a single transmit; with two transmits the layout is refused, not approximated."""
    from mach_amd import _hip_impl, autotune

    rng = np.random.default_rng(3)
    D, rx, vox, arr = random_case(rng, T=2, E=9, S=48, F=64, V=150)
    d, r, v, a = to_dev(D, rx, vox, arr, device=device)
    out = torch.zeros((150, 64), dtype=torch.complex64, device=device)
    with pytest.raises(RuntimeError, match="lateral"):
        _hip_impl.beamform_compound(d, r, v, a, out, 0.8, -1e-7, 20e6, 1540.0, 5e6, 0.5, 2,
                                    layout=autotune.encode_lateral(2, 2))


@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("kind", ["iq", "rf"])
@pytest.mark.parametrize("S", [1, 2, 3, 4, 6])
def test_lateral_layouts_short_records(device, interp, kind, S):
    """Lateral groups on records shorter than their sample window (the window is never used, or only
    where it fits) and a voxel count that is no multiple of the group: bitwise equal to a general
    layout and within 1e-5 of the oracle."""
    from mach_amd import _hip_impl, autotune

    rng = np.random.default_rng(100 * S + 10 * interp + (kind == "rf"))
    E, F, V = 7, 64, 37
    D, rx, vox, arr = random_case(rng, T=1, E=E, S=S, F=F, V=V, kind=kind)
    tau = arr[0][:, None] + np.linalg.norm(vox[:, None, :].astype(np.float64) - rx[None], axis=-1) / 1540.0
    t0 = float(np.median(tau) - 0.5 * S / 20e6)  # centres the taps on the short record
    d, r, v, a = to_dev(D, rx, vox, arr, device=device)
    fmod = 0.0 if kind == "rf" else 5e6
    args = (2.0, t0, 20e6, 1540.0, fmod, 0.5, interp)
    ref_gpu = torch.zeros((V, F), dtype=torch.float32 if kind == "rf" else torch.complex64, device=device)
    _hip_impl.beamform_compound(d, r, v, a, ref_gpu, *args, layout=autotune.encode(32, 2))
    for lay in [x for x in autotune.candidates(F, band=False, lateral=True) if x & autotune.LATERAL_FLAG]:
        got = torch.zeros_like(ref_gpu)
        _hip_impl.beamform_compound(d, r, v, a, got, *args, layout=lay)
        torch.cuda.synchronize()
        assert torch.equal(got, ref_gpu), hex(lay)
    ref = das64(D, rx, vox, arr, f_number=2.0, rx_start_s=t0, sampling_freq_hz=20e6, sound_speed_m_s=1540.0,
                modulation_freq_hz=fmod, tukey_alpha=0.5, interp_type=interp)
    if np.abs(ref).max() > 0:
        assert rel_l2(ref_gpu.cpu().numpy(), ref) <= TOL_FP32
    else:
        assert not torch.any(ref_gpu != 0)
