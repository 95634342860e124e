"""Voxel processing orders (host logic): depth-major rows and wave-sized tiles are permutations of the
voxels, and in the tiled order every wave-aligned run of voxels is one compact tile of the grid."""

import numpy as np
import pytest
import torch

from mach_amd import probes
from mach_amd._hip_impl import _depth_major, _tiled, wave_tile


def grid(nx, ny, nz):
    x = np.linspace(-10e-3, 10e-3, nx)
    y = np.linspace(-5e-3, 5e-3, ny) if ny > 1 else np.array([0.0])
    z = np.linspace(5e-3, 45e-3, nz)
    return torch.as_tensor(probes.scan_grid(x, y, z).astype(np.float32))


def is_perm(idx, n):
    return torch.equal(torch.sort(idx).values, torch.arange(n))


def test_wave_tile_of_layout_codes():
    assert wave_tile(0) is None                  # automatic layout: chosen by the library later
    assert wave_tile(0x11) == 64                 # 1 lane per voxel x 1 frame
    assert wave_tile(0x24) == 32                 # 2 lanes per voxel
    assert wave_tile(0x41) == 16                 # 4 lanes per voxel
    assert wave_tile(0x104) is None              # 16 lanes per voxel: 4 voxels per wave
    assert wave_tile(0x9402) is None             # lateral groups keep depth-major rows
    assert wave_tile(0x004) is None              # LDS-staged band kernel


@pytest.mark.parametrize("shape,per_wave", [((387, 1, 609), 64), ((256, 1, 256), 16), ((40, 40, 40), 32),
                                            ((33, 17, 9), 16), ((10, 1, 7), 64)])
def test_tiles_are_compact_and_aligned(shape, per_wave):
    scan = grid(*shape)
    n = scan.shape[0]
    order = _tiled(scan, per_wave)
    assert is_perm(order, n)
    pts = scan[order].numpy()
    xs, ys, zs = (np.unique(pts[:, a]) for a in range(3))
    ranks = np.stack([np.searchsorted(xs, pts[:, 0]), np.searchsorted(ys, pts[:, 1]),
                      np.searchsorted(zs, pts[:, 2])], -1)
    # the full tiles of the grid lead the order, one per wave-aligned run, each exactly the tile shape
    if shape[1] == 1:
        tz, tl = {64: (8, 8), 32: (4, 8), 16: (4, 4)}[per_wave]
        full, want = (shape[2] // tz) * (shape[0] // tl), (tl, 1, tz)
    else:
        tz, tx, ty = {64: (4, 4, 4), 32: (2, 4, 4), 16: (2, 2, 4)}[per_wave]
        full, want = (shape[2] // tz) * (shape[0] // tx) * (shape[1] // ty), (tx, ty, tz)
    for k in range(full):
        r = ranks[k * per_wave:(k + 1) * per_wave]
        assert tuple(r.max(0) - r.min(0) + 1) == want, k
    n_tiles = full
    # inside a tile, consecutive voxels run along the last lateral axis at one depth
    if n_tiles:
        r = ranks[:per_wave]
        assert r[0, 2] == r[1, 2]


def test_scattered_cloud_falls_back_to_depth_major():
    rng = np.random.default_rng(0)
    scan = torch.as_tensor((rng.normal(size=(5000, 3)) * 1e-3).astype(np.float32))
    order = _tiled(scan, 32)
    assert is_perm(order, 5000)
    assert torch.equal(order, _depth_major(scan))  # no tile of a random cloud is full
