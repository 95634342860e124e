"""XCD wedges (layout code bits 16-19 = 15): launches in the hardware block order (workgroup bx of every
block row on XCD bx % 8, rows padded to a multiple of 8 workgroups) over a voxel order that gives each
XCD one angular sector of every depth band of a volume. The block order and the voxel order change only
which workgroup computes a voxel, so every output must be bitwise the same as under the other block
orders, and match the FP64 oracle (<= 1e-5 relative L2 on the same stored data) on a voxel sample.
Runs on a ROCm GPU through the C ABI (mach_beamform)."""

import ctypes
import sys

import numpy as np
import pytest
import torch

from mach_amd import _hip_impl, _lib, autotune, phantoms
from oracle.das64 import rel_l2
from oracle.das_c import das_c

pytestmark = pytest.mark.gpu

TOL = 1e-5


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def beamform(acq, cd, scan, arr, layout, frames):
    p = acq.params
    kind = _hip_impl._data_kind(cd)
    n = scan.shape[0]
    out = torch.zeros((n, frames), dtype=torch.complex64, device=cd.device)
    order = _hip_impl.voxel_order(scan, layout, _hip_impl.groups_per_block(arr.shape[0], frames))
    a = _hip_impl.make_args(cd, acq.rx_coords_m, scan, arr, out, n_tx=arr.shape[0], kind=kind,
                            f_number=p["f_number"], rx_start_s=p["rx_start_s"], sampling_freq_hz=p["sampling_freq_hz"],
                            sound_speed_m_s=p["sound_speed_m_s"], modulation_freq_hz=p["modulation_freq_hz"],
                            tukey_alpha=p["tukey_alpha"], interp_type=p["interp_type"], accumulate=False,
                            layout=layout, order=order)
    _lib.check(_lib.load().mach_beamform(ctypes.byref(a), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return out


def oracle_sample(acq, cd, scan, arr, n_sample, seed):
    idx = np.sort(np.random.default_rng(seed).choice(scan.shape[0], n_sample, replace=False))
    ref = das_c(cd.cpu().numpy(), acq.rx_coords_m.cpu().numpy(), scan.cpu().numpy()[idx], arr.cpu().numpy()[:, idx],
                **acq.params)
    assert np.abs(ref).max() > 0
    return idx, ref


@pytest.mark.parametrize("storage", ["c64", "fp16"])
@pytest.mark.parametrize("n_grid,frames", [(32, 8), (30, 32)])
def test_wedge_block_order_matches_other_orders_and_oracle(device, storage, n_grid, frames):
    """The volume (5 diverging waves, matrix probe) on a 32^3 grid (full wave tiles) and a 30^3 grid
    (partial tiles at the edges; 32 frames make 4 block rows of 106 workgroups for 1 x 4 x 2, padded to
    112), under general layouts of 64 and 32 voxels per wave."""
    acq = phantoms.volumetric(device=device, scale=n_grid / 128, n_frames=frames)
    assert acq.n_vox == n_grid ** 3
    cd = acq.fp16_channel_data() if storage == "fp16" else acq.channel_data
    scan, arr = acq.scan_coords_m, acq.tx_wave_arrivals_s
    log(f"wedges: {storage} grid {n_grid}^3 frames {frames}")
    idx, ref = oracle_sample(acq, cd, scan, arr, 400, n_grid + frames)
    for base in (autotune.encode(1, 4, 2), autotune.encode(2, 4)):
        outs = {c: beamform(acq, cd, scan, arr, autotune.with_xcd(base, c), frames) for c in (0, 2, 15)}
        for c, o in outs.items():
            assert torch.equal(o, outs[0]), (hex(base), c)
        err = rel_l2(outs[15].cpu().numpy()[idx], ref)
        log(f"  layout {hex(autotune.with_xcd(base, 15))}: rel_l2 vs oracle {err:.2e}")
        assert err <= TOL


def test_wedge_block_order_pads_lateral_rows(device):
    """A lateral-group layout (no wave tiles: depth-major rows) under the hardware block order: 62x62
    voxels in groups of 4 are 241 workgroups per row, 256 frames make 2 rows, padded to 248."""
    acq = phantoms.doppler(device=device, scale=62 / 256, n_frames=256)
    assert acq.n_vox == 62 * 62
    cd, scan, arr = acq.channel_data, acq.scan_coords_m, acq.tx_wave_arrivals_s
    lat = autotune.encode_lateral(4, 2)
    ref_gpu = beamform(acq, cd, scan, arr, autotune.encode(32, 2), 256)
    for c in (0, 2, 15):
        assert torch.equal(beamform(acq, cd, scan, arr, autotune.with_xcd(lat, c), 256), ref_gpu), c
    idx, ref = oracle_sample(acq, cd, scan, arr, 200, 9)
    assert rel_l2(ref_gpu.cpu().numpy()[idx], ref) <= TOL
