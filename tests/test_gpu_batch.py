"""Batched persistent kernel (mach_batch_*): a mix of requests in one launch gives exactly the plain
per-request results, on every graph replay, and agrees with the FP64 oracle."""

import sys

import numpy as np
import pytest
import torch

from mach_amd import batch, phantoms
from mach_amd._hip_impl import voxel_order
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
    return req


def phantom_request(name, device, scale, **kw):
    acq = phantoms.make(name, device=device, scale=scale, **kw)
    p = acq.params
    return batch.Request(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                         torch.zeros((acq.n_vox, acq.n_frames), dtype=torch.complex64, device=device),
                         f_number=p["f_number"], rx_start_s=p["rx_start_s"], sampling_freq_hz=p["sampling_freq_hz"],
                         sound_speed_m_s=p["sound_speed_m_s"], modulation_freq_hz=p["modulation_freq_hz"],
                         tukey_alpha=p["tukey_alpha"], interp_type=p["interp_type"],
                         voxel_order=voxel_order(acq.scan_coords_m))


def mixed_requests(device):
    rng = np.random.default_rng(99)
    reqs = []
    shapes = [(1, 7, 48, 1, 300), (2, 5, 40, 3, 200), (1, 6, 64, 8, 150), (3, 4, 32, 33, 90), (1, 8, 48, 256, 20),
              (1, 9, 40, 64, 40), (2, 7, 56, 5, 120), (1, 5, 30, 2, 500)]
    kinds = ["iq", "rf", "iq16", "iq", "iq", "rf", "iq16", "iq"]
    for i, ((T, E, S, F, V), kind) in enumerate(zip(shapes, kinds)):
        reqs.append(rand_request(rng, device, T, E, S, F, V, kind, interp=i % 3, alpha=(0.0, 0.5, 1.0)[i % 3]))
    reqs.append(phantom_request("plane_wave", device, 0.25))
    reqs.append(phantom_request("doppler", device, 0.25, n_frames=64))
    return reqs


def test_batch_matches_plain_bitwise_and_oracle(device):
    reqs = mixed_requests(device)
    plain = []
    for r in reqs:
        batch.run_plain(r)
        plain.append(r.out.clone())
        r.out.zero_()
    torch.cuda.synchronize()
    with batch.BatchRunner(max_requests=32) as runner:
        runner.submit(reqs)
        torch.cuda.synchronize()
        for i, (r, ref) in enumerate(zip(reqs, plain)):
            log(f"request {i}: shape={tuple(r.channel_data.shape)} kind={r.kind()} interp={r.interp_type}")
            assert torch.equal(r.out, ref), i
    # one request against the oracle (the batch adds no numerical path of its own)
    r = reqs[1]
    D = r.channel_data.cpu().numpy()
    ref = das64(D, r.rx_coords_m.cpu().numpy(), r.scan_coords_m.cpu().numpy(), r.tx_wave_arrivals_s.cpu().numpy(),
                f_number=0.8, rx_start_s=-1e-7, sampling_freq_hz=20e6, sound_speed_m_s=1540.0,
                modulation_freq_hz=0.0, tukey_alpha=r.tukey_alpha, interp_type=r.interp_type)
    assert rel_l2(r.out.cpu().numpy(), ref) <= 1e-5


@pytest.mark.parametrize("use_graph", [True, False])
def test_graph_replays_are_all_correct(device, use_graph):
    reqs = mixed_requests(device)[:6]
    expected = []
    for r in reqs:
        batch.run_plain(r)
        expected.append(r.out.clone())
    torch.cuda.synchronize()
    with batch.BatchRunner(max_requests=8, use_graph=use_graph) as runner:
        for rep in range(5):
            for r in reqs:
                r.out.fill_(float("nan"))  # eager work between replays
            torch.cuda.synchronize()
            runner.submit(reqs)
            torch.cuda.synchronize()
            for i, (r, e) in enumerate(zip(reqs, expected)):
                assert torch.equal(r.out, e), (rep, i)


def test_two_submissions_in_flight_and_slot_reuse(device):
    reqs = mixed_requests(device)
    a, b = reqs[:5], reqs[5:]
    expected = []
    for r in reqs:
        batch.run_plain(r)
        expected.append(r.out.clone())
        r.out.zero_()
    torch.cuda.synchronize()
    with batch.BatchRunner(max_requests=16) as runner:
        runner.submit(a)
        runner.submit(b)  # second slot while the first may still run
        runner.submit(a)  # reuses the first slot after waiting for it
        torch.cuda.synchronize()
    for r, e in zip(reqs, expected):
        assert torch.equal(r.out, e)


def test_capacity_and_validation_errors(device):
    reqs = mixed_requests(device)[:3]
    with batch.BatchRunner(max_requests=2) as runner:
        with pytest.raises(ValueError, match="capacity"):
            runner.submit(reqs)
        bad = reqs[0]
        bad.tukey_alpha = 3.0
        with pytest.raises(RuntimeError, match="request 0: Error: tukey_alpha"):
            runner.submit([bad])
        runner.submit([])  # empty batch is a no-op
