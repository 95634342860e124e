"""Size-independent properties of the HIP beamformer at the full benchmark configurations (runs on a
ROCm GPU). The oracle checks voxel samples at these sizes (test_gpu_parity.py); these properties
cover every voxel and frame:

* frame permutation: each frame is beamformed on its own, with a fixed per-frame summation order,
  so permuting the channel data's frames permutes the output columns bit for bit (Doppler ensemble);
* voxel permutation: voxels are independent, so permuting the scan coordinates permutes the output
  rows bit for bit (Doppler ensemble);
* exact scaling: multiplying the channel data by 2 is exact in float16 and float32, and every step
  after it is linear, so the output doubles bit for bit (volumetric, FP16 storage);
* linearity: beamform(a + 2 b) = beamform(a) + 2 beamform(b) to float32 rounding (75-angle compound);
* compounding: one launch over 75 transmits equals the sum of 75 single-transmit beamforms
  (src/mach/experimental.py:80-97) to float32 rounding.
"""

import numpy as np
import pytest
import torch

from mach_amd import experimental, kernel, phantoms
from mach_amd.kernel import InterpolationType
from oracle.das64 import rel_l2

pytestmark = pytest.mark.gpu


def _kw(acq):
    p = dict(acq.params)
    return dict(p, interp_type=InterpolationType(p["interp_type"]))


@pytest.fixture(scope="module")
def doppler(device):
    return phantoms.make("doppler", device=device)


def test_frame_permutation_full_doppler(device, doppler):
    acq = doppler
    y = experimental.beamform(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                              **_kw(acq))
    perm = torch.randperm(acq.n_frames, generator=torch.Generator().manual_seed(0)).to(device)
    yp = experimental.beamform(acq.channel_data[..., perm].contiguous(), acq.rx_coords_m, acq.scan_coords_m,
                               acq.tx_wave_arrivals_s, **_kw(acq))
    torch.cuda.synchronize()
    assert torch.equal(yp, y[:, perm])


def test_voxel_permutation_full_doppler(device, doppler):
    acq = doppler
    y = experimental.beamform(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                              **_kw(acq))
    perm = torch.randperm(acq.n_vox, generator=torch.Generator().manual_seed(1)).to(device)
    yp = experimental.beamform(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m[perm].contiguous(),
                               acq.tx_wave_arrivals_s[:, perm].contiguous(), **_kw(acq))
    torch.cuda.synchronize()
    assert torch.equal(yp, y[perm])


def test_exact_scaling_full_volumetric_fp16(device):
    acq = phantoms.make("volumetric", device=device)
    cd = acq.fp16_channel_data()
    y = experimental.beamform(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, **_kw(acq))
    y2 = experimental.beamform((cd * 2).contiguous(), acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                               **_kw(acq))
    torch.cuda.synchronize()
    assert torch.isfinite(y).all() and torch.abs(y).max() > 0
    assert torch.equal(y2, y * 2)


@pytest.fixture(scope="module")
def compound(device):
    return phantoms.make("compound", device=device)


def test_linearity_full_compound(device, compound):
    acq = compound
    a = acq.channel_data
    b = torch.flip(acq.channel_data, dims=[2]).contiguous()  # a second, different data set of the same shape
    run = lambda d: experimental.beamform(d, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,  # noqa: E731
                                          **_kw(acq))
    lhs = run((a + 2 * b).contiguous())
    rhs = run(a) + 2 * run(b)
    torch.cuda.synchronize()
    err = rel_l2(lhs.cpu().numpy(), rhs.cpu().numpy())
    print(f"linearity rel_l2={err:.2e}", flush=True)
    assert err <= 1e-5


def test_compounding_equals_sum_of_single_transmits_full_compound(device, compound):
    acq = compound
    kw = _kw(acq)
    y = experimental.beamform(acq.channel_data, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, **kw)
    acc = torch.zeros_like(y)
    for t in range(acq.n_tx):  # the reference's per-transmit loop (experimental.py:80-97)
        kernel.beamform(acq.channel_data[t], acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s[t], out=acc,
                        **kw)
    torch.cuda.synchronize()
    err = rel_l2(y.cpu().numpy(), acc.cpu().numpy())
    print(f"compound vs sum of transmits rel_l2={err:.2e}", flush=True)
    assert err <= 1e-5
    assert np.isfinite(err)

