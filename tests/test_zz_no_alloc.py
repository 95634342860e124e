"""Allocation-free batch entry (mach_batch_submit_noalloc) called through the C ABI: its outputs equal
mach_batch_submit's bit for bit, on two calls in a row (the second with fewer, smaller requests, over
a table that still holds the first call's descriptors), and one request agrees with the FP64 oracle."""

import ctypes
import sys

import numpy as np
import pytest
import torch

from mach_amd import _lib, batch
from oracle.das64 import das64, rel_l2

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rand_request(rng, device, T, E, S, F, V, kind, interp, alpha):
    D = rng.normal(size=(T, E, S, F)).astype(np.float32)
    if kind != "rf":
        D = (D + 1j * rng.normal(size=D.shape)).astype(np.complex64)
    rx = (rng.normal(size=(E, 3)) * 1e-3).astype(np.float32)
    vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 2e-3]).astype(np.float32)
    arr = (rng.normal(size=(T, V)) * 1e-6 + 0.5e-6).astype(np.float32)
    cd = torch.as_tensor(D, device=device)
    if kind == "iq16":
        cd = torch.view_as_real(cd).to(torch.float16).contiguous()
    out_dtype = torch.float32 if kind == "rf" else torch.complex64
    req = batch.Request(cd, torch.as_tensor(rx, device=device), torch.as_tensor(vox, device=device),
                        torch.as_tensor(arr, device=device), torch.zeros((V, F), dtype=out_dtype, device=device),
                        f_number=0.8, rx_start_s=-1e-7, sampling_freq_hz=20e6, sound_speed_m_s=1540.0,
                        modulation_freq_hz=0.0 if kind == "rf" else 5e6, tukey_alpha=alpha, interp_type=interp)
    req.check()
    return req, (D, rx, vox, arr)


def make_requests(device, seed, shapes):
    rng = np.random.default_rng(seed)
    kinds = ["iq", "rf", "iq16"]
    return [rand_request(rng, device, *shape, kind=kinds[i % 3], interp=i % 3, alpha=(0.0, 0.5, 1.0)[i % 3])
            for i, shape in enumerate(shapes)]


def noalloc_submit(reqs):
    """mach_batch_submit_noalloc through the C ABI on the current stream."""
    arr = (_lib.BeamformArgs * len(reqs))()
    for i, r in enumerate(reqs):
        arr[i] = r.args()
    lib = _lib.load()
    st = lib.mach_batch_submit_noalloc(arr, len(reqs), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, _lib.last_error()


def reference_outputs(reqs):
    """The same requests through the existing entry point, mach_batch_submit."""
    with batch.BatchRunner(max_requests=64) as runner:
        runner.submit(reqs)
        torch.cuda.synchronize()
    return [r.out.clone() for r in reqs]


# (T, E, S, F, V): 20 requests (two table writes of 14 + 6 descriptors), both sections, 1-256 frames
LARGE = [(1, 7, 48, 1, 300), (2, 5, 40, 3, 200), (1, 6, 64, 8, 150), (3, 4, 32, 33, 90), (1, 8, 48, 256, 20),
         (1, 9, 40, 64, 40), (2, 7, 56, 5, 120), (1, 5, 30, 2, 500), (9, 6, 40, 1, 260), (1, 16, 64, 4, 400),
         (1, 3, 24, 16, 64), (4, 8, 48, 2, 100), (1, 12, 80, 128, 12), (1, 4, 20, 1, 1000), (2, 6, 36, 7, 77),
         (1, 10, 50, 32, 33), (1, 5, 40, 3, 210), (12, 4, 32, 1, 150), (1, 7, 44, 64, 45), (1, 6, 60, 12, 99)]
# a smaller second call: fewer requests, fewer voxels, frames and transmits
SMALL = [(1, 5, 32, 2, 60), (3, 4, 24, 1, 40), (1, 6, 40, 8, 25)]


def check_call(device, seed, shapes):
    made = make_requests(device, seed, shapes)
    reqs = [r for r, _ in made]
    ref = reference_outputs(reqs)
    for r in reqs:
        r.out.fill_(float("nan"))  # overwrite mode: every output value must be written
    noalloc_submit(reqs)
    torch.cuda.synchronize()
    for i, (r, want) in enumerate(zip(reqs, ref)):
        log(f"seed {seed} request {i}: shape={tuple(r.channel_data.shape)} kind={r.kind()} interp={r.interp_type}")
        assert torch.equal(r.out, want), i
    return made


def test_noalloc_matches_batch_submit_on_two_calls(device):
    log("no-alloc call 1: 20 requests")
    made = check_call(device, 7, LARGE)
    log("no-alloc call 2: 3 smaller requests")
    check_call(device, 8, SMALL)
    # one request of the first call against the FP64 oracle
    r, (D, rx, vox, arr) = made[1]
    y = das64(D, rx, vox, arr, f_number=r.f_number, rx_start_s=r.rx_start_s, sampling_freq_hz=r.sampling_freq_hz,
              sound_speed_m_s=r.sound_speed_m_s, modulation_freq_hz=r.modulation_freq_hz,
              tukey_alpha=r.tukey_alpha, interp_type=r.interp_type)
    err = rel_l2(r.out.cpu().numpy(), y)
    log(f"request 1 vs oracle: rel_l2 {err:.2e}")
    assert err <= 1e-5


def test_noalloc_rejects_more_than_its_table_holds(device):
    made = make_requests(device, 9, [(1, 4, 20, 1, 10)] * 65)
    arr = (_lib.BeamformArgs * 65)()
    for i, (r, _) in enumerate(made):
        arr[i] = r.args()
    lib = _lib.load()
    assert lib.mach_batch_submit_noalloc(arr, 65, None) == _lib.ERR_INVALID_ARGUMENT
    assert "allocation-free maximum of 64" in _lib.last_error()
