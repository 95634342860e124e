"""Single-frame requests with 32 or more transmits: the workgroup's 4 waves each sum a contiguous
quarter of the transmits and the first adds the others' partial sums in transmit order (das_block's
transmit split, chosen from the request alone). Every layout and the batched kernel must give the same
bits, the oracle must agree at the FP32 bar, and the split must not change what accumulate mode and the
epilogues write.
"""

import sys

import numpy as np
import pytest
import torch

from mach_amd import _hip_impl, autotune, batch
from oracle.das64 import das64, rel_l2

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5
KW = dict(f_number=0.8, rx_start_s=-1e-7, sampling_freq_hz=20e6, sound_speed_m_s=1540.0)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def case(rng, T, E=9, S=48, V=300, kind="iq"):
    """V = 300 voxels: 4 full 64-voxel groups and a partial one (whose blocks' waves hold no voxel for
    some lanes and must still pass the partial-sum barriers)."""
    D = rng.normal(size=(T, E, S, 1)).astype(np.float32)
    if kind != "rf":
        D = (D + 1j * rng.normal(size=D.shape)).astype(np.complex64)
    rx = (rng.normal(size=(E, 3)) * 1e-3).astype(np.float32)
    vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 2e-3]).astype(np.float32)
    arr = (rng.normal(size=(T, V)) * 1e-6 + 0.5e-6).astype(np.float32)
    return D, rx, vox, arr


def device_inputs(D, rx, vox, arr, kind, device):
    d = torch.as_tensor(D).to(device)
    if kind == "iq16":
        d = torch.view_as_real(d).to(torch.float16).contiguous()
    return d, torch.as_tensor(rx).to(device), torch.as_tensor(vox).to(device), torch.as_tensor(arr).to(device)


def oracle_data(D, kind):
    """The channel data the kernel sees: FP16 IQ is compared with the oracle on the same rounded data."""
    if kind == "iq16":
        return (D.real.astype(np.float16).astype(np.float32) + 1j * D.imag.astype(np.float16).astype(np.float32)
                ).astype(np.complex64)
    return D


def run(d, r, v, a, out, kind, interp, layout=0, accumulate=False):
    fmod = 0.0 if kind == "rf" else 5e6
    _hip_impl.beamform_compound(d, r, v, a, out, 0.8, -1e-7, 20e6, 1540.0, fmod, 0.5, interp, layout=layout,
                                accumulate=accumulate)


@pytest.mark.parametrize("kind", ["iq", "rf", "iq16"])
@pytest.mark.parametrize("interp", [0, 1, 2])
def test_split_every_layout_identical_and_oracle(device, kind, interp):
    rng = np.random.default_rng(100 + interp)
    D, rx, vox, arr = case(rng, T=40, kind=kind)
    d, r, v, a = device_inputs(D, rx, vox, arr, kind, device)
    dtype = torch.float32 if kind == "rf" else torch.complex64
    outs = {}
    for lay in [0] + autotune.candidates(1, band=False):
        out = torch.zeros((vox.shape[0], 1), dtype=dtype, device=device)
        run(d, r, v, a, out, kind, interp, layout=lay)
        outs[lay] = out
    torch.cuda.synchronize()
    ref = das64(oracle_data(D, kind), rx, vox, arr, **KW, modulation_freq_hz=0.0 if kind == "rf" else 5e6,
                tukey_alpha=0.5, interp_type=interp)
    first = outs[0]
    err = rel_l2(first.cpu().numpy(), ref)
    log(f"split kind={kind} interp={interp}: rel_l2 {err:.2e} over {len(outs)} layouts")
    assert np.abs(ref).max() > 0 and err <= TOL_FP32
    for lay, o in outs.items():
        assert torch.equal(o, first), hex(lay)


@pytest.mark.parametrize("n_tx", [31, 32, 33, 75])
def test_split_boundary_against_oracle(device, n_tx):
    """31 transmits run unsplit, 32 and more split (uneven quarters for 33 and 75); all at the FP32 bar."""
    rng = np.random.default_rng(n_tx)
    D, rx, vox, arr = case(rng, T=n_tx)
    d, r, v, a = device_inputs(D, rx, vox, arr, "iq", device)
    out = torch.zeros((vox.shape[0], 1), dtype=torch.complex64, device=device)
    run(d, r, v, a, out, "iq", 1)
    torch.cuda.synchronize()
    ref = das64(D, rx, vox, arr, **KW, modulation_freq_hz=5e6, tukey_alpha=0.5, interp_type=1)
    err = rel_l2(out.cpu().numpy(), ref)
    log(f"n_tx={n_tx}: rel_l2 {err:.2e}")
    assert err <= TOL_FP32


def test_split_accumulate_adds_to_out(device):
    """Accumulate mode adds the split sum once (only the group's first wave writes)."""
    rng = np.random.default_rng(7)
    D, rx, vox, arr = case(rng, T=40)
    d, r, v, a = device_inputs(D, rx, vox, arr, "iq", device)
    y = torch.zeros((vox.shape[0], 1), dtype=torch.complex64, device=device)
    run(d, r, v, a, y, "iq", 2)
    prev = torch.as_tensor((rng.normal(size=(vox.shape[0], 1)) + 1j * rng.normal(size=(vox.shape[0], 1)))
                           .astype(np.complex64)).to(device)
    acc = prev.clone()
    run(d, r, v, a, acc, "iq", 2, accumulate=True)
    torch.cuda.synchronize()
    assert torch.equal(acc, prev + y)


def test_split_batch_matches_plain(device):
    """Split requests inside the persistent batched kernel (which runs them as 1 x 1) beside unsplit
    multi-transmit and single-transmit requests: each bitwise equal to its plain launch, on every
    graph replay."""
    rng = np.random.default_rng(11)
    reqs = []
    for T, F, V, interp, lay in ((40, 1, 300, 1, 0), (75, 1, 200, 2, autotune.encode(4, 1)), (31, 1, 150, 0, 0),
                                 (3, 2, 120, 1, 0), (1, 4, 90, 2, 0)):
        D = (rng.normal(size=(T, 9, 48, F)) + 1j * rng.normal(size=(T, 9, 48, F))).astype(np.complex64)
        rx = (rng.normal(size=(9, 3)) * 1e-3).astype(np.float32)
        vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 2e-3]).astype(np.float32)
        arr = (rng.normal(size=(T, V)) * 1e-6 + 0.5e-6).astype(np.float32)
        req = batch.Request(torch.as_tensor(D, device=device), torch.as_tensor(rx, device=device),
                            torch.as_tensor(vox, device=device), torch.as_tensor(arr, device=device),
                            torch.zeros((V, F), dtype=torch.complex64, device=device), modulation_freq_hz=5e6,
                            interp_type=interp, layout=lay, **KW)
        req.check()
        reqs.append(req)
    plain = []
    for req in reqs:
        batch.run_plain(req)
        plain.append(req.out.clone())
    torch.cuda.synchronize()
    with batch.BatchRunner(max_requests=8) as runner:
        for rep in range(3):
            for req in reqs:
                req.out.fill_(float("nan"))
            torch.cuda.synchronize()
            runner.submit(reqs)
            torch.cuda.synchronize()
            for i, (req, ref) in enumerate(zip(reqs, plain)):
                assert torch.equal(req.out, ref), (rep, i)
