"""The benchmark configurations at their full stated sizes against the FP64 oracle.

Voxels are independent, so the kernel runs on the whole problem and the oracle checks a seeded voxel
sample (all frames, transmits and elements of each sampled voxel):
  * 75-angle compounding, 75 tx x 128 ch x 2048 samples, 387 x 609 voxels (FP32 and FP16 storage);
  * the volume, 5 diverging waves x 1024 ch x 1536 samples x 8 frames, 128^3 voxels (FP32 and FP16);
  * the Doppler ensemble with FP16 storage, 128 ch x 1024 samples x 256 frames, 256 x 256 voxels;
  * the served mixed stream at its stated shapes: full-size templates of the four workloads, grids
    x0.5-1.0, 1-256 frames; every request checked against the oracle, and alone == in the mix.
FP32 bar: 1e-5 relative L2. FP16 storage: 1e-5 against the oracle on the same FP16 data, 1e-3 end to
end against the FP32 data, and the B-mode dB difference (max over the -60 dB mask of the
frame-averaged envelope) within 0.1 dB."""

import sys

import numpy as np
import pytest
import torch

from mach_amd import experimental, phantoms, serving
from mach_amd.kernel import InterpolationType
from oracle.das64 import bmode_db_diff, rel_l2
from oracle.das_c import das_c

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5
TOL_FP16_SAME_DATA = 1e-5
TOL_FP16_END_TO_END = 1e-3
TOL_BMODE_DB = 0.1


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def acquisitions(device):
    """Full-size acquisitions, synthesized once for the module."""
    cache = {}

    def get(name):
        if name not in cache:
            log(f"synthesizing full-size {name}")
            cache[name] = phantoms.make(name, device=device)
            torch.cuda.synchronize()
        return cache[name]

    yield get
    cache.clear()


def beamform_sample(acq, fp16, idx):
    """Kernel on the whole problem; returns (kernel output at idx, channel data as numpy)."""
    p = dict(acq.params)
    cd = acq.fp16_channel_data() if fp16 else acq.channel_data
    y = experimental.beamform(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                              **dict(p, interp_type=InterpolationType(p["interp_type"])))
    torch.cuda.synchronize()
    return y.cpu().numpy()[idx], cd.cpu().numpy()


def oracle_sample(acq, cd_np, idx):
    return das_c(cd_np, acq.rx_coords_m.cpu().numpy(), acq.scan_coords_m.cpu().numpy()[idx],
                 acq.tx_wave_arrivals_s.cpu().numpy()[:, idx], **acq.params)


def sample(acq, n, seed):
    return np.sort(np.random.default_rng(seed).choice(acq.n_vox, n, replace=False))


@pytest.mark.parametrize("name,n", [("compound", 400), ("volumetric", 300)])
def test_full_size_fp32(acquisitions, name, n):
    acq = acquisitions(name)
    idx = sample(acq, n, 10)
    log(f"{name} full size fp32: shape={tuple(acq.channel_data.shape)} voxels={acq.n_vox} sample={n}")
    y, cd = beamform_sample(acq, False, idx)
    ref = oracle_sample(acq, cd, idx)
    err = rel_l2(y, ref)
    log(f"  rel_l2={err:.3e}")
    assert np.abs(ref).max() > 0 and err <= TOL_FP32


@pytest.mark.parametrize("name,n", [("compound", 400), ("volumetric", 300), ("doppler", 200)])
def test_full_size_fp16_storage(acquisitions, name, n):
    acq = acquisitions(name)
    idx = sample(acq, n, 11)
    log(f"{name} full size fp16: shape={tuple(acq.channel_data.shape)} voxels={acq.n_vox} sample={n}")
    y16, cd16 = beamform_sample(acq, True, idx)
    ref16 = oracle_sample(acq, cd16, idx)
    ref32 = oracle_sample(acq, acq.channel_data.cpu().numpy(), idx)
    e_same, e_e2e = rel_l2(y16, ref16), rel_l2(y16, ref32)
    db = bmode_db_diff(np.abs(y16).mean(-1), np.abs(ref32).mean(-1))
    log(f"  rel_l2 same-data={e_same:.2e} end-to-end={e_e2e:.2e} B-mode dB diff={db:.4f}")
    assert e_same <= TOL_FP16_SAME_DATA
    assert e_e2e <= TOL_FP16_END_TO_END
    assert db <= TOL_BMODE_DB


def test_served_stream_at_stated_shapes(device):
    """c4 as stated: full-size templates, grids x0.5-1.0, frames up to 256, arrivals fast enough that
    requests batch, fuse and run as direct launches. Every request agrees with the oracle on a voxel
    sample, and its output is bitwise the one it gets alone (plain kernel, one request at a time)."""
    log("served stream at stated shapes: synthesizing full-size templates")
    reqs = serving.make_stream(device, n_requests=20, qps=3000.0, seed=14, template_scale=1.0)
    kinds = sorted({r.kind for r in reqs})
    log(f"  {len(reqs)} requests, kinds {kinds}, frames {sorted({r.source.shape[3] for r in reqs})}")
    serving.serve_plain(reqs, device, keep_outputs=True)
    alone = {r.rid: r.out for r in reqs}
    for r in reqs:
        r.done_s = r.admit_s = float("nan")
        r.out = None
    rep = serving.serve_batched(reqs, device, keep_outputs=True)
    log(f"  served: batches={rep.n_batches} fused={rep.n_fused} direct={rep.n_direct} "
        f"p50/p99={rep.p50_ms:.1f}/{rep.p99_ms:.1f} ms")
    assert rep.n_requests == len(reqs) and rep.n_batches < len(reqs)
    host = {}
    for r in reqs:
        assert torch.equal(r.out, alone[r.rid]), r.rid
        key = id(r.source)
        if key not in host:
            host[key] = r.source.cpu().numpy()
        g = r.geometry
        V = g.scan_coords_m.shape[0]
        idx = np.sort(np.random.default_rng(r.rid).choice(V, min(V, 48), replace=False))
        ref = das_c(host[key], g.rx_coords_m.cpu().numpy(), g.scan_coords_m.cpu().numpy()[idx],
                    g.tx_wave_arrivals_s.cpu().numpy()[:, idx], **g.params)
        err = rel_l2(r.out.cpu().numpy()[idx], ref)
        log(f"  request {r.rid} {r.kind} data={tuple(r.source.shape)} voxels={V}: rel_l2={err:.2e}")
        assert np.abs(ref).max() > 0 and err <= TOL_FP32, r.rid
