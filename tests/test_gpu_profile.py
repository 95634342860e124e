"""Opt-in timing (MACH_AMD_PROFILE=1, read when the library loads, so the case runs in a child
process): each native call prints its kernel and call times to stderr and keeps them for
mach_last_timing; staging CPU inputs and copying the result back are timed by the Python layer."""

import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASE = textwrap.dedent("""
    import torch
    import mach_amd as mach
    from mach_amd import _hip_impl, phantoms
    acq = phantoms.doppler(device="cuda", scale=0.25, n_frames=64)
    p = dict(acq.params, interp_type=mach.InterpolationType(acq.params["interp_type"]))
    y = mach.beamform(acq.channel_data[0], acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s[0], **p)
    k, c = _hip_impl.last_timing()
    print(f"TIMING {k} {c}", flush=True)
    assert 0 < k <= c, (k, c)
    # CPU inputs and output: the staging and the copy back are timed too
    mach.beamform(acq.channel_data[0].cpu(), acq.rx_coords_m.cpu(), acq.scan_coords_m.cpu(),
                  acq.tx_wave_arrivals_s[0].cpu(), **p)
    print("PROFILE-OK", flush=True)
""")


def test_profile_mode_times_each_call(device):
    env = dict(os.environ, MACH_AMD_PROFILE="1", PYTHONPATH=ROOT)
    print("profile mode: child process with MACH_AMD_PROFILE=1", flush=True)
    res = subprocess.run([sys.executable, "-u", "-W", "ignore", "-c", CASE], env=env, cwd=ROOT, capture_output=True,
                         text=True, timeout=240)
    print(res.stdout[-3000:], res.stderr[-3000:], flush=True)
    assert res.returncode == 0 and "PROFILE-OK" in res.stdout
    err = res.stderr
    assert "mach_amd profile: mach_beamform kernel" in err
    assert "lateral kernel" in err or "general kernel" in err
    assert "stage CPU inputs to the device" in err and "copy the result to the host" in err
