"""The debug build (libmach_hip_debug.so, loaded with MACH_AMD_DEBUG=1): its kernels check every tap
index against the record and every voxel index from the processing order against the grid, count
failures on the device instead of trapping, and each call fails with MACH_ERR_DEBUG_CHECK if any check
failed. On valid requests every kernel family must run clean and match the FP64 oracle. The library is
loaded once per process, so the cases run in one child process with MACH_AMD_DEBUG=1."""

import os
import subprocess
import sys
import textwrap

import pytest

from mach_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = textwrap.dedent("""
    import sys
    import numpy as np
    import torch
    from mach_amd import _hip_impl, _lib, autotune, batch, phantoms
    from oracle.das64 import rel_l2
    from oracle.das_c import das_c

    assert _lib.DEBUG and _lib.load()._name == _lib.DEBUG_PATH, _lib.load()._name
    dev = torch.device("cuda", 0)

    def check(name, acq, layout, cd=None):
        p = acq.params
        cd = acq.channel_data if cd is None else cd
        out = torch.zeros((acq.n_vox, acq.n_frames), dtype=torch.complex64, device=dev)
        _hip_impl.beamform_compound(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, out,
                                    p["f_number"], p["rx_start_s"], p["sampling_freq_hz"], p["sound_speed_m_s"],
                                    p["modulation_freq_hz"], p["tukey_alpha"], p["interp_type"], layout=layout)
        torch.cuda.synchronize()
        idx = np.sort(np.random.default_rng(0).choice(acq.n_vox, min(acq.n_vox, 300), replace=False))
        ref = das_c(cd.cpu().numpy(), acq.rx_coords_m.cpu().numpy(), acq.scan_coords_m.cpu().numpy()[idx],
                    acq.tx_wave_arrivals_s.cpu().numpy()[:, idx], **p)
        err = rel_l2(out.cpu().numpy()[idx], ref)
        print(f"{name} layout={hex(layout)}: rel_l2={err:.2e}", flush=True)
        assert err <= 1e-5, (name, err)
        return out

    dop = phantoms.doppler(device=dev, scale=0.25, n_frames=64)
    check("doppler general 16x4", dop, autotune.encode(16, 4))
    check("doppler lateral 4x2", dop, autotune.encode_lateral(4, 2))
    check("doppler band", dop, autotune.encode(0, 1))
    vol = phantoms.volumetric(device=dev, scale=0.25)
    check("volume fp16 1x4x2 (per-transmit setup)", vol, autotune.encode(1, 4, 2), cd=vol.fp16_channel_data())
    check("volume c64 2x4", vol, autotune.encode(2, 4))
    comp = phantoms.compound(device=dev, scale=0.25)
    check("compound 1x1 (transmit chunk loads)", comp, autotune.encode(1, 1))

    reqs = []
    for acq in (dop, vol, comp, phantoms.plane_wave(device=dev, scale=0.25)):
        p = acq.params
        reqs.append(batch.Request(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                                  torch.zeros((acq.n_vox, acq.n_frames), dtype=torch.complex64, device=dev),
                                  f_number=p["f_number"], rx_start_s=p["rx_start_s"],
                                  sampling_freq_hz=p["sampling_freq_hz"], sound_speed_m_s=p["sound_speed_m_s"],
                                  modulation_freq_hz=p["modulation_freq_hz"], tukey_alpha=p["tukey_alpha"],
                                  interp_type=p["interp_type"], voxel_order=_hip_impl.voxel_order(acq.scan_coords_m)))
    plain = []
    for r in reqs:
        batch.run_plain(r)
        plain.append(r.out.clone())
    with batch.BatchRunner(max_requests=8) as runner:
        runner.submit(reqs)
        torch.cuda.synchronize()
    for r, want in zip(reqs, plain):
        assert torch.equal(r.out, want)
    batch.submit_noalloc(reqs)
    torch.cuda.synchronize()
    for r, want in zip(reqs, plain):
        assert torch.equal(r.out, want)
    print("batch and allocation-free batch: clean, equal to plain", flush=True)
    print("DEBUG-BUILD-OK", flush=True)
""")


def test_debug_build_runs_every_kernel_family_clean(device):
    if not os.path.exists(_lib.DEBUG_PATH):
        pytest.fail(f"{_lib.DEBUG_PATH} is not built (python mach_amd/build.py --debug)")
    env = dict(os.environ, MACH_AMD_DEBUG="1", PYTHONPATH=ROOT)
    print("debug build: child process with MACH_AMD_DEBUG=1", flush=True)
    res = subprocess.run([sys.executable, "-u", "-c", CASES], env=env, cwd=ROOT, capture_output=True, text=True,
                         timeout=240)
    print(res.stdout[-4000:], res.stderr[-4000:], flush=True)
    assert res.returncode == 0 and "DEBUG-BUILD-OK" in res.stdout
