"""Fused post-processing epilogues against the FP64 oracle (runs on a ROCm GPU).

Power Doppler sum_f |y|^2, envelope |y| and their dB images (``mach_amd.postprocess``, C ABI
``mach_beamform_args.output`` and ``mach_to_db``) are compared with ``oracle.das64``'s restatement of
examples/doppler.py:252 and src/mach/_vis.py:17-40 applied to the FP64 beamformed signal.

Bars: power and envelope <= 1e-5 relative L2 against the oracle (the power image also per voxel
<= 1e-4 relative where it is above 1e-6 of its maximum); dB images within 1e-3 dB of the oracle
over the -60 dB mask and 0 dB exactly at the peak.
"""

import ctypes
import sys

import numpy as np
import pytest
import torch

from mach_amd import _lib, phantoms, postprocess
from mach_amd.kernel import InterpolationType
from oracle.das64 import das64, db_zero64, envelope64, power_doppler64, rel_l2
from oracle.das_c import das_c

pytestmark = pytest.mark.gpu

TOL = 1e-5
TOL_DB = 1e-3

KW = dict(f_number=0.8, rx_start_s=-1e-7, sampling_freq_hz=20e6, sound_speed_m_s=1540.0)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_case(rng, T=1, E=7, S=48, F=3, V=70, kind="iq"):
    D = rng.normal(size=(T, E, S, F)).astype(np.float32)
    if kind != "rf":
        D = (D + 1j * rng.normal(size=D.shape)).astype(np.complex64)
    rx = (rng.normal(size=(E, 3)) * 1e-3).astype(np.float32)
    vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 2e-3]).astype(np.float32)
    arr = (rng.normal(size=(T, V)) * 1e-6 + 0.5e-6).astype(np.float32)
    return D, rx, vox, arr


def dev(*arrays, device):
    return [torch.as_tensor(a).to(device) for a in arrays]


def oracle_signal(D, rx, vox, arr, kind, interp, alpha):
    return das64(D, rx, vox, arr, **KW, modulation_freq_hz=5e6 if kind == "iq" else 0.0, tukey_alpha=alpha,
                 interp_type=interp)


def kwargs(kind, interp, alpha):
    return dict(KW, modulation_freq_hz=5e6 if kind == "iq" else None, tukey_alpha=alpha,
                interp_type=InterpolationType(interp))


@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("kind", ["rf", "iq"])
@pytest.mark.parametrize("n_tx", [1, 3])
@pytest.mark.parametrize("frames", [1, 3, 8, 64, 256])
def test_power_doppler_matches_oracle(device, interp, kind, n_tx, frames):
    log(f"power interp={interp} kind={kind} n_tx={n_tx} frames={frames}")
    rng = np.random.default_rng(7000 + 100 * interp + 10 * n_tx + frames + (kind == "iq"))
    D, rx, vox, arr = random_case(rng, T=n_tx, F=frames, V=70 if frames < 64 else 21, kind=kind)
    d, r, v, a = dev(D if n_tx > 1 else D[0], rx, vox, arr if n_tx > 1 else arr[0], device=device)
    peak = torch.empty(1, dtype=torch.float32, device=device)
    pd = postprocess.power_doppler(d, r, v, a, peak=peak, **kwargs(kind, interp, 0.5))
    torch.cuda.synchronize()
    ref = power_doppler64(oracle_signal(D, rx, vox, arr, kind, interp, 0.5))
    got = pd.cpu().numpy()
    assert ref.max() > 0
    err = rel_l2(got, ref)
    big = ref > 1e-6 * ref.max()
    worst = float(np.max(np.abs(got[big] - ref[big]) / ref[big]))
    log(f"  rel_l2={err:.3e} worst={worst:.3e}")
    assert err <= TOL and worst <= 1e-4
    assert peak.item() == got.max()


@pytest.mark.parametrize("kind", ["rf", "iq", "fp16"])
@pytest.mark.parametrize("frames", [1, 4, 33])
def test_envelope_matches_oracle(device, kind, frames):
    log(f"envelope kind={kind} frames={frames}")
    rng = np.random.default_rng(8000 + frames)
    D, rx, vox, arr = random_case(rng, T=2, F=frames, kind="rf" if kind == "rf" else "iq")
    if kind == "fp16":  # oracle on the same FP16-rounded data
        D16 = np.stack([D.real, D.imag], axis=-1).astype(np.float16)
        D = (D16[..., 0].astype(np.float32) + 1j * D16[..., 1].astype(np.float32)).astype(np.complex64)
        d = torch.as_tensor(D16).to(device)
    else:
        d = torch.as_tensor(D).to(device)
    r, v, a = dev(rx, vox, arr, device=device)
    peak = torch.empty(1, dtype=torch.float32, device=device)
    env = postprocess.envelope(d, r, v, a, peak=peak, **kwargs("rf" if kind == "rf" else "iq", 1, 0.5))
    torch.cuda.synchronize()
    ref = envelope64(oracle_signal(D, rx, vox, arr, "rf" if kind == "rf" else "iq", 1, 0.5))
    got = env.cpu().numpy()
    err = rel_l2(got, ref)
    log(f"  rel_l2={err:.3e}")
    assert err <= TOL
    assert (got >= 0).all()
    assert peak.item() == got.max()


def test_bmode_and_power_db_match_oracle_db_zero(device):
    rng = np.random.default_rng(81)
    D, rx, vox, arr = random_case(rng, T=1, F=16, V=200, kind="iq")
    d, r, v, a = dev(D[0], rx, vox, arr[0], device=device)
    y64 = oracle_signal(D, rx, vox, arr, "iq", 2, 0.5)
    k = kwargs("iq", 2, 0.5)
    bm = postprocess.bmode_db(d, r, v, a, **k).cpu().numpy()
    pdb = postprocess.power_doppler_db(d, r, v, a, **k).cpu().numpy()
    for got, ref in ((bm, db_zero64(envelope64(y64))), (pdb, db_zero64(power_doppler64(y64), power_mode=True))):
        mask = ref > -60.0
        diff = float(np.max(np.abs(got[mask] - ref[mask])))
        log(f"  dB diff={diff:.2e} over {mask.sum()} values")
        assert diff <= TOL_DB
        assert got.max() == 0.0


def test_to_db_zero_and_epsilon(device):
    x = torch.tensor([0.0, 1.0, 10.0, 0.1], device=device)
    peak = torch.tensor([10.0], device=device)
    postprocess.to_db(x, peak)
    eps_db = 20 * np.log10(np.finfo(float).eps)
    np.testing.assert_allclose(x.cpu().numpy(), [eps_db - 20.0, -20.0, 0.0, -40.0], atol=1e-4)
    z = torch.zeros(3, device=device)
    postprocess.to_db(z, torch.zeros(1, device=device), power_mode=True)
    np.testing.assert_array_equal(z.cpu().numpy(), np.zeros(3))


def test_power_accumulates_and_zero_frames(device):
    rng = np.random.default_rng(82)
    D, rx, vox, arr = random_case(rng, T=1, F=8, kind="iq")
    d, r, v, a = dev(D[0], rx, vox, arr[0], device=device)
    k = kwargs("iq", 1, 0.0)
    pd = postprocess.power_doppler(d, r, v, a, **k)
    base = torch.full_like(pd, 2.0)
    postprocess.power_doppler(d, r, v, a, out=base, accumulate=True, **k)
    torch.cuda.synchronize()
    ref = power_doppler64(oracle_signal(D, rx, vox, arr, "iq", 1, 0.0))
    assert rel_l2(base.cpu().numpy() - 2.0, ref) <= TOL
    # a sum over no frames is 0 (overwrite mode), like np.sum over an empty axis
    d0 = torch.zeros((7, 48, 0), dtype=torch.complex64, device=device)
    out = torch.full((70,), 5.0, device=device)
    postprocess.power_doppler(d0, r, v, a, out=out, **k)
    torch.cuda.synchronize()
    assert (out == 0).all()


def test_power_on_doppler_workload_matches_oracle_on_voxel_sample(device):
    """Full-size 256-frame Doppler ensemble (128 ch x 1024 samples, 65536 voxels, quadratic): the power
    image of a seeded voxel sample against the oracle (voxels are independent)."""
    acq = phantoms.make("doppler", device=device)
    p = acq.params
    pd = postprocess.power_doppler(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                                   f_number=p["f_number"], rx_start_s=p["rx_start_s"],
                                   sampling_freq_hz=p["sampling_freq_hz"], sound_speed_m_s=p["sound_speed_m_s"],
                                   modulation_freq_hz=p["modulation_freq_hz"], tukey_alpha=p["tukey_alpha"],
                                   interp_type=InterpolationType(p["interp_type"]))
    torch.cuda.synchronize()
    idx = np.sort(np.random.default_rng(3).choice(acq.n_vox, 256, replace=False))
    y = das_c(acq.channel_data.cpu().numpy(), acq.rx_coords_m.cpu().numpy(), acq.scan_coords_m.cpu().numpy()[idx],
              acq.tx_wave_arrivals_s.cpu().numpy()[:, idx], **p)
    err = rel_l2(pd.cpu().numpy()[idx], power_doppler64(y))
    log(f"doppler power rel_l2={err:.3e}")
    assert err <= TOL


def test_epilogue_errors(device):
    lib = _lib.load()
    rng = np.random.default_rng(83)
    D, rx, vox, arr = random_case(rng, T=1, F=4, kind="iq")
    d, r, v, a = dev(D[0], rx, vox, arr[0], device=device)
    out = torch.empty((70, 4), dtype=torch.float32, device=device)
    from mach_amd._hip_impl import OUTPUT_ENVELOPE, OUTPUT_SIGNAL, make_args

    base = dict(n_tx=1, kind=_lib.DATA_C64, **KW, modulation_freq_hz=5e6, tukey_alpha=0.5, interp_type=1)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = make_args(d, r, v, a, out, accumulate=True, output=OUTPUT_ENVELOPE, **base)
    assert lib.mach_beamform(ctypes.byref(args), stream) == _lib.ERR_INVALID_ARGUMENT
    assert "envelope" in _lib.last_error()
    peak = torch.empty(1, device=device)
    args = make_args(d, r, v, a, out, accumulate=False, output=OUTPUT_SIGNAL, peak=peak, **base)
    assert lib.mach_beamform(ctypes.byref(args), stream) == _lib.ERR_INVALID_ARGUMENT
    args = make_args(d, r, v, a, out, accumulate=False, output=7, **base)
    assert lib.mach_beamform(ctypes.byref(args), stream) == _lib.ERR_INVALID_ARGUMENT
    # batches run the signal output only
    bp = ctypes.c_void_p()
    assert lib.mach_batch_create(4, 0, ctypes.byref(bp)) == 0
    try:
        args = make_args(d, r, v, a, out, accumulate=False, output=OUTPUT_ENVELOPE, **base)
        assert lib.mach_batch_submit(bp, ctypes.byref(args), 1, stream) == _lib.ERR_UNSUPPORTED
    finally:
        lib.mach_batch_destroy(bp)


def test_power_is_identical_for_every_frame_group_count(device):
    """Frame groups per lane (layout bits 12+) change only how often the pair setup is recomputed: the
    power image (and the signal) must be bitwise identical for 1, 2 and 4 groups."""
    from mach_amd._hip_impl import OUTPUT_POWER, OUTPUT_SIGNAL, make_args
    from mach_amd.autotune import encode

    lib = _lib.load()
    rng = np.random.default_rng(84)
    D, rx, vox, arr = random_case(rng, T=1, F=256, V=300, kind="iq")
    d, r, v, a = dev(D[0], rx, vox, arr[0], device=device)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    base = dict(n_tx=1, kind=_lib.DATA_C64, **KW, modulation_freq_hz=5e6, tukey_alpha=0.5, interp_type=2,
                accumulate=False)
    res = {}
    for lf, k in ((16, 4), (32, 2), (64, 1), (8, 4)):
        for ch in (1, 2, 4):
            pd = torch.empty(300, dtype=torch.float32, device=device)
            sig = torch.empty((300, 256), dtype=torch.complex64, device=device)
            for out, kind in ((pd, OUTPUT_POWER), (sig, OUTPUT_SIGNAL)):
                args = make_args(d, r, v, a, out, layout=encode(lf, k, ch), output=kind, **base)
                assert lib.mach_beamform(ctypes.byref(args), stream) == 0, _lib.last_error()
            res[(lf, k, ch)] = (pd, sig)
    torch.cuda.synchronize()
    ref = oracle_signal(D, rx, vox, arr, "iq", 2, 0.5)
    for (lf, k, ch), (pd, sig) in res.items():
        assert torch.equal(sig, res[(16, 4, 1)][1]), (lf, k, ch)
        assert torch.equal(pd, res[(lf, k, 1)][0]), (lf, k, ch)
        assert rel_l2(pd.cpu().numpy(), power_doppler64(ref)) <= TOL
