"""Native entry points of the HIP beamformer: the drop-in for the reference's ``mach._cuda_impl``.

The reference exposes a nanobind module ``_cuda_impl`` with ``beamform`` (two overloads, complex64
and float32, ``src/mach/kernel.cu:1229-1255``; stub ``src/mach/_cuda_impl.pyi:24-53``),
``InterpolationType`` (``kernel.cu:1222-1226``) and ``__nvcc_version__`` (``kernel.cu:1215``). This
module exposes the same names over the C ABI of ``include/mach_hip.h`` (ctypes), with
``__hip_version__`` in place of ``__nvcc_version__``.

Arrays are taken through DLPack (PyTorch tensors directly, NumPy arrays via ``torch.from_numpy``,
anything else via ``torch.from_dlpack``); ROCm arrays (DLPack device type 10) are used in place.
CPU arrays are staged to the GPU with the reference's latency warning (``kernel.cu:1092-1100``); a
CPU ``out`` is overwritten, a device ``out`` is accumulated into (``kernel.cu:1135-1139, 1203-1205,
674-680``). Without a HIP device this module raises: there is no CPU implementation.
"""

from __future__ import annotations

import collections
import ctypes
import enum
import os
import sys
import time
import warnings
import weakref

import numpy as np
import torch

from . import _lib

_LIB = _lib.load()

__hip_version__: str = _LIB.mach_hip_runtime_version().decode()


class InterpolationType(enum.IntEnum):
    """Interpolation selector with the reference's values (``src/mach/kernel.cu:25-29``)."""

    NearestNeighbor = 0
    Linear = 1
    Quadratic = 2


NearestNeighbor = InterpolationType.NearestNeighbor
Linear = InterpolationType.Linear
Quadratic = InterpolationType.Quadratic


def _shape_str(t) -> str:
    return "[" + ", ".join(str(int(d)) for d in t.shape) + "]"


def _as_tensor(x, name: str) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x
    if isinstance(x, np.ndarray):
        with warnings.catch_warnings():  # read-only NumPy inputs are only read
            warnings.simplefilter("ignore")
            return torch.from_numpy(x)
    if hasattr(x, "__dlpack__"):
        return torch.from_dlpack(x)
    raise TypeError(f"Array '{name}' does not support DLPack protocol")


def _data_kind(cd: torch.Tensor) -> int:
    if cd.dtype == torch.complex64:
        return _lib.DATA_C64
    if cd.dtype == torch.float32:
        return _lib.DATA_F32
    if cd.dtype == torch.complex32:
        return _lib.DATA_C32H
    if cd.dtype == torch.float16 and cd.shape[-1] == 2:
        return _lib.DATA_C32H
    raise TypeError(
        "beamform(): incompatible channel_data dtype "
        f"{cd.dtype}; expected complex64 (IQ), float32 (RF), or float16 [..., 2] / complex32 (FP16 IQ)"
    )


def _core_shape(cd: torch.Tensor, kind: int):
    """Logical [.., n_rx, n_samples, n_frames] shape (drops the trailing FP16 IQ axis)."""
    if kind == _lib.DATA_C32H and cd.dtype == torch.float16:
        return tuple(cd.shape[:-1])
    return tuple(cd.shape)


def check_dimensions(channel_shape, rx, scan, arrivals, out, n_rx, n_vox, n_frames):
    """Shape checks with the reference's messages (``src/mach/kernel.cu:914-951``)."""

    def s(shape):
        return "[" + ", ".join(str(int(d)) for d in shape) + "]"

    if n_rx != channel_shape[-3] or n_rx != rx.shape[0]:
        raise RuntimeError(
            "Dimension mismatch in receive elements:\n"
            f"  Expected n_receive_elements: {n_rx}\n"
            f"  channel_data.shape: {s(channel_shape)}\n"
            f"  rx_coords_m.shape: {_shape_str(rx)}\n"
            "→ channel_data.shape[0] and rx_coords_m.shape[0] must both equal n_receive_elements"
        )
    if n_vox != arrivals.shape[-1] or n_vox != scan.shape[0] or n_vox != out.shape[0]:
        raise RuntimeError(
            "Dimension mismatch in output voxels:\n"
            f"  Expected n_output_voxels: {n_vox}\n"
            f"  scan_coords_m.shape: {_shape_str(scan)}\n"
            f"  tx_wave_arrivals_s.shape: {_shape_str(arrivals)}\n"
            f"  out.shape: {_shape_str(out)}\n"
            "→ scan_coords_m.shape[0], tx_wave_arrivals_s.shape[0], and out.shape[0] must all equal n_output_voxels"
        )
    if n_frames != channel_shape[-1] or n_frames != out.shape[1]:
        raise RuntimeError(
            "Dimension mismatch in frames:\n"
            f"  Expected n_frames: {n_frames}\n"
            f"  channel_data.shape: {s(channel_shape)}\n"
            f"  out.shape: {_shape_str(out)}\n"
            "→ channel_data.shape[2] and out.shape[1] must both equal n_frames"
        )


def _require(t: torch.Tensor, name: str, dtype, ndim: int, last=None):
    if t.dtype != dtype:
        raise TypeError(f"beamform(): incompatible {name} dtype {t.dtype}; expected {dtype}")
    if t.dim() != ndim or (last is not None and t.shape[-1] != last):
        raise TypeError(f"beamform(): incompatible {name} shape {_shape_str(t)}")
    if not t.is_contiguous():
        raise TypeError(f"beamform(): {name} must be C-contiguous")


def _check_writable(out):
    if isinstance(out, np.ndarray) and not out.flags.writeable:
        raise ValueError("Output array `out` is not writable.")


# MACH_AMD_SYNC=1 restores the reference's synchronise-after-every-call behaviour (for timing a call the
# way the reference's benchmarks do); by default calls are ordered on the current stream only.
_SYNC_EACH_CALL = os.environ.get("MACH_AMD_SYNC", "0") == "1"

# Voxels below this count are processed in the caller's order (no reordering benefit).
_ORDER_MIN_VOXELS = 4096
_ORDER_CACHE: "collections.OrderedDict" = collections.OrderedDict()
_ORDER_CACHE_SIZE = 16


# Voxel tiles (depth, then the varying lateral axes) of one wave of a general-kernel layout, by voxels
# per wave; see wave_tile.
_PLANAR_TILES = {64: (8, 8), 32: (4, 8), 16: (4, 4)}
_VOLUME_TILES = {64: (4, 4, 4), 32: (2, 4, 4), 16: (2, 2, 4)}


def wave_tile(layout: int):
    """Voxels per wave of a general-kernel layout code when the wave holds 16 or more voxels (lanes per
    voxel 1, 2 or 4), else None: lateral groups, the LDS-staged band kernel and narrow waves are tuned
    for depth-major rows, and an automatic layout (0) is chosen by the library later."""
    if layout <= 0 or layout & 0x8000:
        return None
    lf = (layout >> 4) & 0xFF
    if lf == 0 or 64 // lf < 16:
        return None
    return 64 // lf


def _depth_major(scan: torch.Tensor) -> torch.Tensor:
    idx = torch.argsort(scan[:, 1], stable=True)
    idx = idx[torch.argsort(scan[idx, 0], stable=True)]
    return idx[torch.argsort(scan[idx, 2], stable=True)]


def _is_grid(dims, n: int) -> bool:
    """Whether voxels with these per-axis coordinate counts fill most of their rank grid (a grid or a
    grid with holes), so that tiling it pays; a scattered cloud has about n distinct values per axis,
    and its rank grid (n^3 cells) is never materialised."""
    return dims[0] * dims[1] * dims[2] <= 2 * n


def _tiled(scan: torch.Tensor, per_wave: int) -> torch.Tensor:
    """Wave-sized voxel tiles. Each axis's coordinates are ranked (exact for grids); voxels are grouped
    into tiles of (depth x lateral) ranks, `per_wave` voxels each, ordered depth-major at the tile
    level and inside a tile. Tiles holding exactly `per_wave` voxels come first, so every wave of the
    kernel gets one whole tile (a compact blob whose voxels share nearly the same in-aperture elements
    and sample windows); the remaining voxels (partial tiles at the grid's edges, or a scattered cloud)
    follow in depth-major order."""
    n = scan.shape[0]
    ranks = [torch.unique(scan[:, a], return_inverse=True)[1] for a in range(3)]
    dims = [int(r.max()) + 1 for r in ranks]
    if not _is_grid(dims, n):
        return _depth_major(scan)
    lateral = [a for a in (0, 1) if dims[a] > 1]
    sizes = {2: 1, 0: 1, 1: 1}
    if len(lateral) == 2:
        sizes[2], sizes[0], sizes[1] = _VOLUME_TILES[per_wave]
    else:
        sizes[2], lat = _PLANAR_TILES[per_wave]
        for a in lateral:
            sizes[a] = lat
    tid = torch.zeros(n, dtype=torch.int64, device=scan.device)
    local = torch.zeros(n, dtype=torch.int64, device=scan.device)
    for a in (2, 0, 1):  # depth, then x, then y (y fastest)
        nt = (dims[a] + sizes[a] - 1) // sizes[a]
        tid = tid * nt + ranks[a] // sizes[a]
        local = local * sizes[a] + ranks[a] % sizes[a]
    _, inv, count = torch.unique(tid, return_inverse=True, return_counts=True)
    full = count[inv] == per_wave
    dm_rank = torch.empty(n, dtype=torch.int64, device=scan.device)
    dm_rank[_depth_major(scan)] = torch.arange(n, device=scan.device)
    base = (int(tid.max()) + 1) * per_wave
    key = torch.where(full, tid * per_wave + local, base + dm_rank)
    return torch.argsort(key, stable=True)


# Bits 16-19 of a layout code that select the hardware block order with XCD wedges (kXcdWedgeCode).
XCD_WEDGE_CODE = 15
_N_XCD = 8


def groups_per_block(n_tx: int, n_frames: int) -> int:
    """Voxel groups (waves' worth of voxels) per workgroup of a general-kernel launch: 4, or 1 when the
    workgroup's 4 waves split one group's transmits (one frame and at least 32 transmits;
    tx_split_waves in das_kernel.h)."""
    return 1 if (n_frames == 1 and n_tx >= 32) else 4


def _xcd_wedges(scan: torch.Tensor, per_wave: int, groups: int):
    """Wave-sized tiles of a volume grid ordered so that workgroup h (run on XCD h % 8 in the hardware
    block order) beamforms tiles of that XCD's wedge. The lateral tiles are cut, by angle around the
    grid's lateral centre, into 8 sectors of equal tile count (on a symmetric grid each sector has the
    same work at every depth), and XCD x walks sector x of every depth band, band after band (inside a
    band by distance from the centre, so a workgroup's tiles are neighbours). The interleaved runs of
    the other remaps balance the depth-dependent work by sending every band through all eight L2s; here
    each L2 holds only the sample windows of its own sector. Tiles that are not full (grid edges) follow
    in depth-major order, as in ``_tiled``. Returns None for grids without two lateral axes."""
    n = scan.shape[0]
    dev = scan.device
    ranks = [torch.unique(scan[:, a], return_inverse=True)[1] for a in range(3)]
    dims = [int(r.max()) + 1 for r in ranks]
    if dims[0] < 2 or dims[1] < 2 or not _is_grid(dims, n):
        return None
    sz, sx, sy = _VOLUME_TILES[per_wave]
    nz, nx, ny = (dims[2] + sz - 1) // sz, (dims[0] + sx - 1) // sx, (dims[1] + sy - 1) // sy
    lat = (ranks[0] // sx) * ny + ranks[1] // sy
    tid = (ranks[2] // sz) * (nx * ny) + lat
    local = ((ranks[2] % sz) * sx + ranks[0] % sx) * sy + ranks[1] % sy
    count = torch.bincount(tid, minlength=nz * nx * ny).cpu().numpy()
    # sectors of the lateral tiles: angle around the centre, cut into equal counts, each sector ordered
    # by distance from the centre. Tiles on a diagonal (equal angles) go to the two sectors it separates
    # in the pattern A B B A A B B A ... outwards, which gives both the same sum of any work that grows
    # linearly with the distance.
    gx = np.repeat(np.arange(nx), ny)
    gy = np.tile(np.arange(ny), nx)
    cx, cy = gx - (nx - 1) / 2.0, gy - (ny - 1) / 2.0
    ang = np.round(np.arctan2(cy, cx), 9)
    rad = cx * cx + cy * cy
    ring = np.floor(np.maximum(np.abs(cx), np.abs(cy))).astype(np.int64)
    sector = np.empty(nx * ny, dtype=np.int64)
    sector[np.lexsort((rad, (ring + 1) // 2 % 2, ang))] = np.arange(nx * ny) * _N_XCD // (nx * ny)
    within = np.lexsort((ang, rad, sector))  # by sector, then distance, then angle
    rank_in = np.empty(nx * ny, dtype=np.int64)
    rank_in[within] = np.arange(nx * ny) - np.searchsorted(sector[within], sector[within])
    # each sector's full tiles band after band; workgroup h takes the next `groups` tiles of sector
    # h % 8 (of the sector with most tiles left once that one is empty, so workgroups stay whole and
    # sectors of unequal size only move the last workgroups to other XCDs)
    full = np.nonzero(count == per_wave)[0]
    f_lat, f_band = full % (nx * ny), full // (nx * ny)
    f_sec = sector[f_lat]
    o = np.lexsort((rank_in[f_lat], f_band, f_sec))
    lists = np.split(full[o], np.searchsorted(f_sec[o], np.arange(1, _N_XCD)))
    left = np.array([len(t) for t in lists])
    take = np.zeros(_N_XCD, dtype=np.int64)
    seq = []
    h = 0
    if len(set(left.tolist())) == 1 and left[0] % groups == 0:  # equal sectors: one interleave
        seq = np.stack(lists).reshape(_N_XCD, -1, groups).transpose(1, 0, 2).reshape(-1)
        left[:] = 0
    while left.sum():
        for _ in range(groups):
            x = h % _N_XCD if left[h % _N_XCD] else int(np.argmax(left))
            if not left[x]:
                break
            seq.append(lists[x][take[x]])
            take[x] += 1
            left[x] -= 1
        h += 1
    tile_pos = torch.full((nz * nx * ny,), -1, dtype=torch.int64)
    tile_pos[torch.as_tensor(np.asarray(seq, dtype=np.int64))] = torch.arange(len(seq))
    tile_pos = tile_pos.to(dev)[tid]
    dm_rank = torch.empty(n, dtype=torch.int64, device=dev)
    dm_rank[_depth_major(scan)] = torch.arange(n, device=dev)
    key = torch.where(tile_pos >= 0, tile_pos * per_wave + local, len(full) * per_wave + dm_rank)
    return torch.argsort(key, stable=True)


def voxel_order(scan: torch.Tensor, layout: int = 0, groups: int = 4):
    """Processing order of the voxels for a launch with this layout code.

    Depth-major (sorted by z, then x, then y; stable) by default: waves that run together beamform
    voxels at the same depth, whose interpolation taps hit the same few channel-data rows per element,
    so those rows stay in L2. For general-kernel layouts whose waves hold 16 or more voxels
    (``wave_tile``), wave-sized tiles instead (``_tiled``): the elements a wave visits are the union of
    its voxels' apertures, so a compact tile wastes fewer lanes than a row, and its sample windows
    overlap more. Layouts whose bits 16-19 select XCD wedges (``XCD_WEDGE_CODE``) get those tiles in
    XCD wedges on volume grids (``_xcd_wedges``, for launches of ``groups`` voxel groups per workgroup:
    ``groups_per_block``). Any permutation gives the same results (each voxel is computed
    independently, with a fixed summation order); the order only affects speed. Cached per
    scan-coordinate tensor (identity and in-place version), tile and wedge mode.
    ``MACH_AMD_VOXEL_ORDER=0`` disables it (caller order), ``MACH_AMD_VOXEL_ORDER=rows`` keeps
    depth-major rows for every layout.
    """
    n = scan.shape[0]
    mode = os.environ.get("MACH_AMD_VOXEL_ORDER", "1")
    if n < _ORDER_MIN_VOXELS or not scan.is_cuda or mode == "0":
        return None
    per_wave = None if mode == "rows" else wave_tile(layout)
    wedges = groups if per_wave is not None and (layout >> 16) & 0xF == XCD_WEDGE_CODE else 0
    key = (id(scan), scan.data_ptr(), n, per_wave, wedges)
    hit = _ORDER_CACHE.get(key)
    if hit is not None:
        ref, version, order = hit
        if ref() is scan and version == scan._version:
            _ORDER_CACHE.move_to_end(key)
            return order
    idx = _xcd_wedges(scan, per_wave, wedges) if wedges else None
    if idx is None:
        idx = _depth_major(scan) if per_wave is None else _tiled(scan, per_wave)
    order = idx.to(torch.int32).contiguous()
    _ORDER_CACHE[key] = (weakref.ref(scan), scan._version, order)
    while len(_ORDER_CACHE) > _ORDER_CACHE_SIZE:
        _ORDER_CACHE.popitem(last=False)
    return order


def order_on(order, stream):
    """``order`` (from voxel_order, or None) marked as used by work on ``stream``: the order cache is
    bounded and may drop a tensor that a launch on another stream than its allocation stream still reads,
    and the allocator must not hand that memory out again before the launch has finished."""
    if order is not None:
        order.record_stream(stream)
    return order


def _device_for(tensors):
    for t in tensors:
        if t.is_cuda:
            return t.device
    if not torch.cuda.is_available():
        raise RuntimeError(
            "No HIP device is available: mach_amd beamforming runs only on a ROCm GPU (gfx950); "
            "there is no CPU implementation."
        )
    return torch.device("cuda", torch.cuda.current_device())


# mach_output_kind
OUTPUT_SIGNAL, OUTPUT_POWER, OUTPUT_ENVELOPE = 0, 1, 2


def make_args(cd, rx, scan, arrivals, out, *, n_tx, kind, f_number, rx_start_s, sampling_freq_hz,
              sound_speed_m_s, modulation_freq_hz, tukey_alpha, interp_type, accumulate=True, layout=0,
              order=None, output=OUTPUT_SIGNAL, peak=None, tx_waves=None) -> "_lib.BeamformArgs":
    """``mach_beamform_args`` for device tensors (shapes already checked). The frame count is the
    channel data's. ``arrivals`` is None when ``tx_waves`` (packed device wavefronts) is given."""
    core = _core_shape(cd, kind)
    return _lib.BeamformArgs(
        cd.data_ptr(), rx.data_ptr(), scan.data_ptr(), arrivals.data_ptr() if arrivals is not None else None,
        out.data_ptr(),
        n_tx, rx.shape[0], core[-2], core[-1], scan.shape[0],
        kind, int(interp_type),
        f_number, rx_start_s, sampling_freq_hz, sound_speed_m_s, modulation_freq_hz, tukey_alpha,
        1 if accumulate else 0, int(layout),
        order.data_ptr() if order is not None else None,
        int(output), peak.data_ptr() if peak is not None else None,
        tx_waves.data_ptr() if tx_waves is not None else None,
    )


# MACH_AMD_PROFILE=1: the native library times each call's kernels (mach_last_timing) and this layer
# times the host<->device staging of CPU arrays, the reference's CUDA_PROFILE sections
# (kernel.cu:1112, 1146, 1201), printing to stderr.
_PROFILE = os.environ.get("MACH_AMD_PROFILE", "0") not in ("", "0")


def _profile_start(device):
    if not _PROFILE:
        return None
    torch.cuda.synchronize(device)
    return time.perf_counter()


def _profile_end(t0, device, what):
    if t0 is None:
        return
    torch.cuda.synchronize(device)
    print(f"mach_amd profile: {what} {(time.perf_counter() - t0) * 1e3:.4f} ms", file=sys.stderr, flush=True)


def last_timing():
    """(kernel ms, call ms) of the last timed native call on this thread, or None when timing is off
    (set MACH_AMD_PROFILE=1 before the library loads)."""
    k, c = ctypes.c_float(), ctypes.c_float()
    if _LIB.mach_last_timing(ctypes.byref(k), ctypes.byref(c)) != _lib.MACH_OK:
        return None
    return float(k.value), float(c.value)


def _run(cd, rx, scan, arrivals, out, *, n_tx, kind, f_number, rx_start_s, sampling_freq_hz,
         sound_speed_m_s, modulation_freq_hz, tukey_alpha, interp_type, accumulate_device_out=True, layout=0):
    """Stage CPU arrays if needed, call ``mach_beamform`` and write ``out`` back."""
    arrays = [cd, rx, scan, arrivals, out]
    n_cpu = sum(1 for t in arrays if not t.is_cuda)
    for t in arrays:
        if t.device.type not in ("cpu", "cuda"):
            raise RuntimeError(
                f"Found input array on device: {t.device}. Only CPU and ROCm devices are supported."
            )
    device = _device_for(arrays)
    if n_cpu:
        warnings.warn(
            f"Found {n_cpu} input array(s) on CPU. This will add latency due to CPU<->GPU memory "
            "transfers. For optimal performance with HIP beamforming, move arrays to the GPU "
            "(e.g. PyTorch tensors on a ROCm device).",
            UserWarning,
            stacklevel=3,
        )
    t_stage = _profile_start(device) if n_cpu else None
    on_dev = [t if t.is_cuda else t.to(device) for t in arrays[:4]]
    if any(t.is_cuda and t.device != device for t in arrays):
        raise RuntimeError("All GPU arrays must be on the same device")
    host_out = not out.is_cuda
    dev_out = torch.zeros(out.shape, dtype=out.dtype, device=device) if host_out else out
    _profile_end(t_stage, device, "stage CPU inputs to the device")
    if layout == 0:
        from . import autotune

        core = _core_shape(on_dev[0], kind)
        layout = autotune.layout_for(kind, int(interp_type), n_tx, core[-3], dev_out.shape[1], on_dev[2].shape[0],
                                     on_dev[0].data_ptr(), dev_out.data_ptr())
    order = voxel_order(on_dev[2], layout, groups_per_block(n_tx, dev_out.shape[1]))
    args = make_args(*on_dev, dev_out, n_tx=n_tx, kind=kind, f_number=f_number, rx_start_s=rx_start_s,
                     sampling_freq_hz=sampling_freq_hz, sound_speed_m_s=sound_speed_m_s,
                     modulation_freq_hz=modulation_freq_hz, tukey_alpha=tukey_alpha, interp_type=interp_type,
                     accumulate=accumulate_device_out or host_out, layout=layout, order=order)
    with torch.cuda.device(device):
        cur = torch.cuda.current_stream(device)
        order_on(order, cur)
        stream = cur.cuda_stream
        status = _LIB.mach_beamform(ctypes.byref(args), ctypes.c_void_p(stream))
        _lib.check(status)
        if host_out:
            t_copy = _profile_start(device)
            out.copy_(dev_out.cpu())
            _profile_end(t_copy, device, "copy the result to the host")
        elif _SYNC_EACH_CALL:  # the reference synchronises the device after every call (kernel.cu:874)
            torch.cuda.current_stream(device).synchronize()


def beamform(
    channel_data,
    rx_coords_m,
    scan_coords_m,
    tx_wave_arrivals_s,
    out,
    f_number: float,
    rx_start_s: float,
    sampling_freq_hz: float,
    sound_speed_m_s: float,
    modulation_freq_hz: float | None = None,
    tukey_alpha: float = 0.5,
    interp_type: InterpolationType = InterpolationType.Linear,
) -> None:
    """Raw single-transmit beamform, the drop-in for ``_cuda_impl.beamform``.

    Strict like the nanobind overloads (``.noconvert()``, ``kernel.cu:1229-1255``): channel data must
    be complex64 [n_rx, n_samples, n_frames] (``modulation_freq_hz`` required) or float32 (defaults to
    0.0); FP16 IQ storage (float16 [n_rx, n_samples, n_frames, 2] or complex32) is also accepted,
    with a complex64 ``out``. Coordinates are float32 [n, 3], arrivals float32 [n_vox], ``out``
    [n_vox, n_frames]. Arrays must be C-contiguous.
    """
    cd = _as_tensor(channel_data, "channel_data")
    kind = _data_kind(cd)
    if kind != _lib.DATA_F32 and modulation_freq_hz is None:
        raise TypeError("beamform(): modulation_freq_hz is required for complex channel data")
    if modulation_freq_hz is None:
        modulation_freq_hz = 0.0
    rx = _as_tensor(rx_coords_m, "rx_coords_m")
    scan = _as_tensor(scan_coords_m, "scan_coords_m")
    arr = _as_tensor(tx_wave_arrivals_s, "tx_wave_arrivals_s")
    o = _as_tensor(out, "out")
    core_ndim = 4 if (kind == _lib.DATA_C32H and cd.dtype == torch.float16) else 3
    if cd.dim() != core_ndim or not cd.is_contiguous():
        raise TypeError(f"beamform(): incompatible channel_data shape {_shape_str(cd)} (C-contiguous 3-D expected)")
    _require(rx, "rx_coords_m", torch.float32, 2, 3)
    _require(scan, "scan_coords_m", torch.float32, 2, 3)
    _require(arr, "tx_wave_arrivals_s", torch.float32, 1)
    _require(o, "out", torch.float32 if kind == _lib.DATA_F32 else torch.complex64, 2)
    core = _core_shape(cd, kind)
    check_dimensions(core, rx, scan, arr, o, rx.shape[0], scan.shape[0], o.shape[1])
    _check_writable(out)
    _run(cd, rx, scan, arr, o, n_tx=1, kind=kind, f_number=f_number, rx_start_s=rx_start_s,
         sampling_freq_hz=sampling_freq_hz, sound_speed_m_s=sound_speed_m_s,
         modulation_freq_hz=modulation_freq_hz, tukey_alpha=tukey_alpha, interp_type=interp_type)


def beamform_compound(
    channel_data,
    rx_coords_m,
    scan_coords_m,
    tx_wave_arrivals_s,
    out,
    f_number: float,
    rx_start_s: float,
    sampling_freq_hz: float,
    sound_speed_m_s: float,
    modulation_freq_hz: float | None = None,
    tukey_alpha: float = 0.5,
    interp_type: InterpolationType = InterpolationType.Linear,
    accumulate: bool = True,
    layout: int = 0,
) -> None:
    """Coherent compounding in one launch: channel data [n_tx, n_rx, n_samples, n_frames], arrivals
    [n_tx, n_vox]. Replaces the per-transmit loop of ``src/mach/experimental.py:80-97``.
    ``accumulate=False`` overwrites a device ``out`` instead of adding to it; ``layout`` forces a
    kernel layout ((lanes_per_voxel << 4) | frames_per_lane, 0 = automatic)."""
    cd = _as_tensor(channel_data, "channel_data")
    kind = _data_kind(cd)
    if kind != _lib.DATA_F32 and modulation_freq_hz is None:
        raise TypeError("beamform_compound(): modulation_freq_hz is required for complex channel data")
    if modulation_freq_hz is None:
        modulation_freq_hz = 0.0
    rx = _as_tensor(rx_coords_m, "rx_coords_m")
    scan = _as_tensor(scan_coords_m, "scan_coords_m")
    arr = _as_tensor(tx_wave_arrivals_s, "tx_wave_arrivals_s")
    o = _as_tensor(out, "out")
    core_ndim = 5 if (kind == _lib.DATA_C32H and cd.dtype == torch.float16) else 4
    if cd.dim() != core_ndim or not cd.is_contiguous():
        raise TypeError(f"beamform_compound(): incompatible channel_data shape {_shape_str(cd)}")
    _require(rx, "rx_coords_m", torch.float32, 2, 3)
    _require(scan, "scan_coords_m", torch.float32, 2, 3)
    _require(arr, "tx_wave_arrivals_s", torch.float32, 2)
    _require(o, "out", torch.float32 if kind == _lib.DATA_F32 else torch.complex64, 2)
    core = _core_shape(cd, kind)
    if arr.shape[0] != core[0]:
        raise RuntimeError(
            f"Dimension mismatch in transmits: channel_data.shape {list(core)} vs "
            f"tx_wave_arrivals_s.shape {_shape_str(arr)}"
        )
    check_dimensions(core, rx, scan, arr, o, rx.shape[0], scan.shape[0], o.shape[1])
    _check_writable(out)
    _run(cd, rx, scan, arr, o, n_tx=core[0], kind=kind, f_number=f_number, rx_start_s=rx_start_s,
         sampling_freq_hz=sampling_freq_hz, sound_speed_m_s=sound_speed_m_s,
         modulation_freq_hz=modulation_freq_hz, tukey_alpha=tukey_alpha, interp_type=interp_type,
         accumulate_device_out=accumulate, layout=layout)
