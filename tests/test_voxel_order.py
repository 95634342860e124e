"""Voxel processing orders (host logic): depth-major rows and wave-sized tiles are permutations of the
voxels, and in the tiled order every wave-aligned run of voxels is one compact tile of the grid."""

import numpy as np
import pytest
import torch

from mach_amd import probes
from mach_amd._hip_impl import _depth_major, _tiled, _xcd_wedges, wave_tile


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


def _ranks(scan, order):
    pts = scan[order].numpy()
    return np.stack([np.searchsorted(np.unique(pts[:, a]), pts[:, a]) for a in range(3)], -1)


@pytest.mark.parametrize("shape,per_wave,groups", [((64, 64, 64), 64, 4), ((80, 80, 80), 64, 4),
                                                   ((64, 64, 32), 32, 4), ((48, 48, 16), 16, 1),
                                                   ((30, 26, 22), 64, 4)])
def test_xcd_wedges_give_each_xcd_one_sector(shape, per_wave, groups):
    """Workgroup h runs on XCD h % 8 in the hardware block order: over the full tiles, XCD x's voxels lie
    in one angular sector of the lateral plane, every sector holds the same number of tiles of every
    depth band, an XCD walks the bands in order, and the grid's partial tiles follow."""
    scan = grid(*shape)
    n = scan.shape[0]
    order = _xcd_wedges(scan, per_wave, groups)
    assert is_perm(order, n)
    tz, tx, ty = {64: (4, 4, 4), 32: (2, 4, 4), 16: (2, 2, 4)}[per_wave]
    nxt, nyt, nzt = shape[0] // tx, shape[1] // ty, shape[2] // tz
    n_full = nxt * nyt * nzt
    r = _ranks(scan, order)
    wg = per_wave * groups
    cx, cy = (shape[0] - 1) / 2.0, (shape[1] - 1) / 2.0
    tiles = [[] for _ in range(8)]
    for k in range(n_full):
        t = r[k * per_wave:(k + 1) * per_wave]
        assert tuple(t.max(0) - t.min(0) + 1) == (tx, ty, tz), k  # still one whole wave tile
        tiles[(k * per_wave // wg) % 8].append((int(t[0, 2]) // tz, float(np.arctan2(t[:, 1].mean() - cy,
                                                                                    t[:, 0].mean() - cx))))
    equal = nxt * nyt % 8 == 0 and nxt % 2 == 0 and nyt % 2 == 0
    for x in range(8):
        bands = [b for b, _ in tiles[x]]
        if equal:
            assert bands == sorted(bands)  # band after band
            assert np.bincount(bands).tolist() == [nxt * nyt // 8] * nzt  # equal tiles per band
            a = np.array([ang for _, ang in tiles[x]])
            lo, hi = -np.pi + x * np.pi / 4, -np.pi + (x + 1) * np.pi / 4
            assert np.all((a >= lo - 1e-9) & (a <= hi + 1e-9)), x  # one sector of pi/4 (diagonals shared)
    # partial tiles (grid edges) follow the full ones in depth-major order
    rest = order[n_full * per_wave:]
    if len(rest):
        assert torch.equal(rest, torch.tensor([i for i in _depth_major(scan).tolist() if i in set(rest.tolist())]))


def test_xcd_wedges_sectors_are_balanced_on_a_symmetric_grid():
    """On a square grid centred on the aperture the eight sectors are images of each other under the
    square's symmetries (diagonal tiles alternate between neighbours), so any work that depends on the
    distance from the centre is the same per XCD: here the summed distance of each XCD's tiles."""
    scan = grid(128, 128, 8)
    order = _xcd_wedges(scan, 64, 4)
    r = _ranks(scan, order)
    per_tile = r.reshape(-1, 64, 3).mean(1)
    xcd = (np.arange(per_tile.shape[0]) // 4) % 8
    d = np.hypot(per_tile[:, 0] - 63.5, per_tile[:, 1] - 63.5)
    sums = np.array([d[xcd == x].sum() for x in range(8)])
    assert np.ptp(sums) / sums.mean() < 2e-3, sums


def test_xcd_wedge_order_needs_two_lateral_axes():
    assert _xcd_wedges(grid(64, 1, 64), 64, 4) is None  # planar grids keep their tiles


def test_large_scattered_cloud_is_not_tiled():
    """A scattered cloud has about n distinct coordinates per axis: its rank grid (n^3 cells) is never
    materialised. The tiled order is then the depth-major one, and there are no XCD wedges."""
    rng = np.random.default_rng(1)
    scan = torch.as_tensor((rng.normal(size=(200_000, 3)) * 1e-3).astype(np.float32))
    assert torch.equal(_tiled(scan, 64), _depth_major(scan))
    assert _xcd_wedges(scan, 64, 4) is None
