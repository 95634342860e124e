"""Seeded synthetic acquisitions for the benchmark and parity workloads.

IQ echoes are synthesized in closed form: sample k (time t_k = t0 + k / fs) of receive element e under
transmit t is

    D[t][e][k][f] = sum_p a_p * g(t_k - tau_p) * exp(-j 2 pi f0 tau_p),  tau_p = T_t(p_f) + |p_f - r_e| / c

with g a Gaussian envelope of the stated fractional bandwidth (-6 dB), T_t the transmit arrival of
the same wavefront model the beamformer is given (``wavefront.plane`` / ``wavefront.spherical``), and
p_f the scatterer positions in frame f. Complex Gaussian noise is added relative to the signal RMS.
Random draws (positions, amplitudes, noise) come from ``numpy.random.default_rng(seed)`` so inputs
are identical on every machine; the echo sums run in PyTorch (float64) on the requested device.

Workloads (grid/array/pulse conventions follow ``src/mach/io/must.py:71-104`` of the reference):
  plane_wave         128-element linear array, 0 deg plane wave, [128 x 2048 x 1], 256x256 grid
  compound           75 plane waves -16..+16 deg, [75 x 128 x 2048 x 1], 387x609 grid, 3 cysts
  doppler            rotating disk + static clutter, [128 x 1024 x 256], 256x256 grid, per-frame output
  volumetric         32x32 matrix, 5 diverging waves, [5 x 1024 x 1536 x 8], 128^3 grid, FP16 IQ
  pymust_disk        the shape of the reference's own published benchmark (BENCHMARKS.md:22-33,
                     tests/compare/conftest.py:25-35, test_pymust.py:129-178): L7-4-like 128-element
                     array, rotating disk, [128 x 1024 x 32], 251x251 grid, linear, f-number 1, no
                     apodization taper; synthetic data (the PyMUST recording is a download)
Each takes a ``scale`` that shrinks grid axes, element and scatterer counts for quick tests (the
sample record, hence the imaged depth, is kept).
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch

from . import geometry, probes, wavefront

F0_HZ = 5.2e6
FS_HZ = 20.8e6
C_M_S = 1540.0
PITCH_M = 0.3e-3
FRACTIONAL_BW = 0.6
NOISE_DB = -40.0


@dataclasses.dataclass
class Acquisition:
    """Beamformer inputs for one workload (arrays are torch tensors on ``device``)."""

    name: str
    channel_data: torch.Tensor  # [T, E, S, F] complex64
    rx_coords_m: torch.Tensor  # [E, 3] float32
    scan_coords_m: torch.Tensor  # [V, 3] float32
    tx_wave_arrivals_s: torch.Tensor  # [T, V] float32
    grid_shape: tuple
    params: dict  # f_number, rx_start_s, sampling_freq_hz, sound_speed_m_s, modulation_freq_hz, tukey_alpha, interp_type
    seed: int

    @property
    def n_tx(self):
        return int(self.channel_data.shape[0])

    @property
    def n_rx(self):
        return int(self.channel_data.shape[1])

    @property
    def n_samples(self):
        return int(self.channel_data.shape[2])

    @property
    def n_frames(self):
        return int(self.channel_data.shape[3])

    @property
    def n_vox(self):
        return int(self.scan_coords_m.shape[0])

    @property
    def points(self) -> int:
        """Nominal points = voxels x receive elements x frames x transmits."""
        return self.n_vox * self.n_rx * self.n_frames * self.n_tx

    def fp16_channel_data(self) -> torch.Tensor:
        """Channel data as FP16 IQ storage: float16 [T, E, S, F, 2]."""
        return torch.view_as_real(self.channel_data).to(torch.float16).contiguous()


def gaussian_sigma_s(f0: float = F0_HZ, bw: float = FRACTIONAL_BW) -> float:
    """Envelope sigma (s) of a Gaussian pulse whose -6 dB spectral width is bw * f0."""
    return math.sqrt(2.0 * math.log(2.0)) / (math.pi * bw * f0)


def plane_wave_arrivals(scan: np.ndarray, angles_rad, c: float = C_M_S) -> np.ndarray:
    """[T, V] float32 arrival times of plane waves steered in azimuth, timed through the origin."""
    out = []
    for a in np.atleast_1d(angles_rad):
        d = np.asarray(geometry.ultrasound_angles_to_cartesian(float(a), 0.0), dtype=np.float64)
        out.append(wavefront.plane(np.zeros(3), scan.astype(np.float64), d) / c)
    return np.stack(out).astype(np.float32)


def diverging_wave_arrivals(scan: np.ndarray, sources: np.ndarray, c: float = C_M_S) -> np.ndarray:
    """[T, V] float32 arrival times of diverging waves from virtual sources behind the origin."""
    out = [wavefront.spherical(np.zeros(3), scan.astype(np.float64), src.astype(np.float64)) / c for src in sources]
    return np.stack(out).astype(np.float32)


def _tx_delay_plane(points: torch.Tensor, angle: float, c: float) -> torch.Tensor:
    d = torch.tensor(geometry.ultrasound_angles_to_cartesian(float(angle), 0.0), dtype=points.dtype,
                     device=points.device)
    return (points * d).sum(-1) / c


def _tx_delay_spherical(points: torch.Tensor, source: np.ndarray, c: float) -> torch.Tensor:
    src = torch.as_tensor(source, dtype=points.dtype, device=points.device)
    return wavefront.spherical(torch.zeros(3, dtype=points.dtype, device=points.device), points, src) / c


def synthesize_echoes(points, amps, rx, tx_delay_fn, n_tx, n_samples, *, fs=FS_HZ, f0=F0_HZ, c=C_M_S,
                      t0=0.0, bw=FRACTIONAL_BW, device="cpu", max_entries=1 << 24):
    """Closed-form IQ echoes.

    Args:
        points: [F, P, 3] scatterer positions per frame (F = 1 for a static medium), float64 numpy.
        amps: [P] complex amplitudes.
        rx: [E, 3] element positions.
        tx_delay_fn: (points_tensor [..., 3], t) -> transmit arrival times [...] (s).
        n_tx, n_samples: output sizes.

    Returns:
        [n_tx, E, n_samples, F] complex64 tensor on ``device``.
    """
    dev = torch.device(device)
    pts = torch.as_tensor(np.asarray(points, dtype=np.float64), device=dev)
    n_frames, P, _ = pts.shape
    a = torch.as_tensor(np.asarray(amps, dtype=np.complex128), device=dev)
    r = torch.as_tensor(np.asarray(rx, dtype=np.float64), device=dev)
    E = r.shape[0]
    sigma = gaussian_sigma_s(f0, bw)
    W = int(math.ceil(4.0 * sigma * fs)) + 1
    offs = torch.arange(-W, W + 1, device=dev)
    out = torch.zeros(n_tx * E * n_samples * n_frames, dtype=torch.complex128, device=dev)
    e_idx = torch.arange(E, device=dev)
    per_point = E * (2 * W + 1)
    p_chunk = max(1, max_entries // max(1, per_point))
    for t in range(n_tx):
        for f in range(n_frames):
            for p0 in range(0, P, p_chunk):
                p = pts[f, p0:p0 + p_chunk]  # [Pc, 3]
                txd = tx_delay_fn(p, t)  # [Pc]
                dist = torch.linalg.vector_norm(p[:, None, :] - r[None, :, :], dim=-1)  # [Pc, E]
                tau = txd[:, None] + dist / c
                kc = torch.round((tau - t0) * fs).long()
                k = kc[..., None] + offs  # [Pc, E, 2W+1]
                tk = t0 + k.to(torch.float64) / fs
                env = torch.exp(-0.5 * ((tk - tau[..., None]) / sigma) ** 2)
                ph = torch.exp(-2j * math.pi * f0 * tau)
                val = a[p0:p0 + p_chunk, None, None] * env * ph[..., None]
                valid = (k >= 0) & (k < n_samples)
                flat = (((t * E + e_idx[None, :, None]) * n_samples + k) * n_frames + f)
                out.index_add_(0, flat[valid], val[valid])
    return out.reshape(n_tx, E, n_samples, n_frames).to(torch.complex64)


def _add_noise(data: torch.Tensor, rng: np.random.Generator, noise_db: float = NOISE_DB) -> torch.Tensor:
    rms = float(torch.sqrt(torch.mean(data.abs() ** 2)).item())
    std = rms * 10.0 ** (noise_db / 20.0) / math.sqrt(2.0)
    n = data.numel()
    noise = (rng.standard_normal(n, dtype=np.float32) + 1j * rng.standard_normal(n, dtype=np.float32)) * std
    return data + torch.as_tensor(noise.astype(np.complex64).reshape(data.shape), device=data.device)


def _grid(nx, nz, x_span, z_span, ny=None, y_span=None):
    x = np.linspace(x_span[0], x_span[1], nx)
    z = np.linspace(z_span[0], z_span[1], nz)
    y = np.array([0.0]) if ny is None else np.linspace(y_span[0], y_span[1], ny)
    return probes.scan_grid(x, y, z).astype(np.float32), ((nx, nz) if ny is None else (nx, ny, nz))


def _compound_angles(n_tx):
    return np.deg2rad(np.linspace(-16.0, 16.0, n_tx))


VOLUMETRIC_TILTS_DEG = [(0.0, 0.0), (8.0, 0.0), (-8.0, 0.0), (0.0, 8.0), (0.0, -8.0)]


def _volumetric_sources():
    """Virtual sources 10 mm behind the aperture centre, along the tilted axes."""
    return np.array([-10e-3 * np.asarray(geometry.ultrasound_angles_to_cartesian(np.deg2rad(a), np.deg2rad(b)))
                     for a, b in VOLUMETRIC_TILTS_DEG])


def workload_grid(name: str, grid_scale: float = 1.0, n_tx: int | None = None):
    """Scan grid [V, 3] float32, grid shape and transmit arrivals [T, V] float32 of a workload with
    every grid axis scaled by ``grid_scale``."""
    if name in ("plane_wave", "doppler"):
        n = max(8, int(round(256 * grid_scale)))
        scan, gshape = _grid(n, n, (-19e-3, 19e-3), (5e-3, 45e-3))
        return scan, gshape, plane_wave_arrivals(scan, [0.0])
    if name == "compound":
        nx, nz = max(8, int(round(387 * grid_scale))), max(8, int(round(609 * grid_scale)))
        scan, gshape = _grid(nx, nz, (-19e-3, 19e-3), (5e-3, 50e-3))
        return scan, gshape, plane_wave_arrivals(scan, _compound_angles(75 if n_tx is None else n_tx))
    if name == "pymust_disk":
        n = max(8, int(round(251 * grid_scale)))
        scan, gshape = _grid(n, n, (-12.5e-3, 12.5e-3), (10e-3, 35e-3))
        return scan, gshape, plane_wave_arrivals(scan, [0.0], c=1480.0)
    if name == "volumetric":
        n = max(4, int(round(128 * grid_scale)))
        x = np.linspace(-10e-3, 10e-3, n)
        z = np.linspace(5e-3, 45e-3, n)
        scan = probes.scan_grid(x, x, z).astype(np.float32)
        return scan, (n, n, n), diverging_wave_arrivals(scan, _volumetric_sources())
    raise KeyError(name)


def regrid(acq: "Acquisition", grid_scale: float) -> "Acquisition":
    """Same acquisition (channel data, array, parameters) beamformed on a grid scaled by ``grid_scale``."""
    scan, gshape, arrivals = workload_grid(acq.name, grid_scale, acq.n_tx)
    dev = acq.channel_data.device
    return dataclasses.replace(acq, scan_coords_m=torch.as_tensor(scan, device=dev).contiguous(),
                               tx_wave_arrivals_s=torch.as_tensor(arrivals, device=dev).contiguous(),
                               grid_shape=gshape)


def _finish(name, data, rx, scan, arrivals, grid_shape, params, seed, device):
    dev = torch.device(device)
    return Acquisition(
        name=name,
        channel_data=data.contiguous(),
        rx_coords_m=torch.as_tensor(rx, device=dev).contiguous(),
        scan_coords_m=torch.as_tensor(scan, device=dev).contiguous(),
        tx_wave_arrivals_s=torch.as_tensor(arrivals, device=dev).contiguous(),
        grid_shape=grid_shape,
        params=params,
        seed=seed,
    )


def _base_params(f_number, interp_type):
    return dict(f_number=f_number, rx_start_s=0.0, sampling_freq_hz=FS_HZ, sound_speed_m_s=C_M_S,
                modulation_freq_hz=F0_HZ, tukey_alpha=0.5, interp_type=interp_type)


def plane_wave(device="cpu", seed=0, scale=1.0) -> Acquisition:
    """Single 0 deg plane wave, 128 elements, 200 point scatterers, [128 x 2048 x 1], 256x256 grid."""
    rng = np.random.default_rng(seed)
    n_el = max(8, int(round(128 * scale)))
    n_s = 2048
    n_grid = max(8, int(round(256 * scale)))
    n_sc = max(4, int(round(200 * scale)))
    rx = probes.linear_probe_positions(n_el, PITCH_M)
    pts = np.stack([rng.uniform(-19e-3, 19e-3, n_sc), np.zeros(n_sc), rng.uniform(5e-3, 45e-3, n_sc)], -1)
    amps = (rng.standard_normal(n_sc) + 1j * rng.standard_normal(n_sc)) / math.sqrt(2)
    data = synthesize_echoes(pts[None], amps, rx, lambda p, t: _tx_delay_plane(p, 0.0, C_M_S), 1, n_s,
                             device=device)
    data = _add_noise(data, rng)
    scan, gshape, arrivals = workload_grid("plane_wave", n_grid / 256)
    return _finish("plane_wave", data, rx, scan, arrivals, gshape, _base_params(1.7, 1), seed, device)


def compound(device="cpu", seed=1, scale=1.0, n_angles=75) -> Acquisition:
    """75 steered plane waves (-16..+16 deg), 500 scatterers with 3 anechoic 4 mm cysts, 387x609 grid."""
    rng = np.random.default_rng(seed)
    n_el = max(8, int(round(128 * scale)))
    n_s = 2048
    n_tx = max(2, int(round(n_angles * scale)))
    nx, nz = max(8, int(round(387 * scale))), max(8, int(round(609 * scale)))
    n_sc = max(8, int(round(500 * scale)))
    rx = probes.linear_probe_positions(n_el, PITCH_M)
    cysts = np.array([[-8e-3, 20e-3], [0.0, 30e-3], [8e-3, 40e-3]])
    pts = []
    while len(pts) < n_sc:
        x, z = rng.uniform(-19e-3, 19e-3), rng.uniform(5e-3, 50e-3)
        if np.all(np.hypot(cysts[:, 0] - x, cysts[:, 1] - z) > 2e-3):
            pts.append((x, 0.0, z))
    pts = np.asarray(pts)
    amps = (rng.standard_normal(n_sc) + 1j * rng.standard_normal(n_sc)) / math.sqrt(2)
    angles = _compound_angles(n_tx)
    data = synthesize_echoes(pts[None], amps, rx, lambda p, t: _tx_delay_plane(p, angles[t], C_M_S), n_tx, n_s,
                             device=device)
    data = _add_noise(data, rng)
    scan, gshape = _grid(nx, nz, (-19e-3, 19e-3), (5e-3, 50e-3))
    arrivals = plane_wave_arrivals(scan, angles)
    return _finish("compound", data, rx, scan, arrivals, gshape, _base_params(1.7, 1), seed, device)


def doppler(device="cpu", seed=2, scale=1.0, n_frames=256) -> Acquisition:
    """Rotating disk (2000 scatterers, r = 10 mm, centre z = 25 mm, 2 rad/s, PRF 5 kHz) in static clutter
    at +20 dB; [128 x 1024 x 256]; 256x256 grid; quadratic interpolation; per-frame output."""
    rng = np.random.default_rng(seed)
    n_el = max(8, int(round(128 * scale)))
    n_s = 1024
    n_grid = max(8, int(round(256 * scale)))
    n_disk = max(8, int(round(2000 * scale)))
    n_clutter = max(8, int(round(2000 * scale)))
    rx = probes.linear_probe_positions(n_el, PITCH_M)
    radius, cz, omega, prf = 10e-3, 25e-3, 2.0, 5000.0
    rr = radius * np.sqrt(rng.uniform(0, 1, n_disk))
    th = rng.uniform(0, 2 * np.pi, n_disk)
    amps_d = (rng.standard_normal(n_disk) + 1j * rng.standard_normal(n_disk)) / math.sqrt(2)
    fr = np.arange(n_frames)
    ang = th[None, :] + omega * fr[:, None] / prf
    disk = np.stack([rr * np.cos(ang), np.zeros_like(ang), cz + rr * np.sin(ang)], -1)  # [F, P, 3]
    clutter = np.stack([rng.uniform(-19e-3, 19e-3, n_clutter), np.zeros(n_clutter), rng.uniform(5e-3, 45e-3, n_clutter)], -1)
    amps_c = 10.0 * (rng.standard_normal(n_clutter) + 1j * rng.standard_normal(n_clutter)) / math.sqrt(2)
    txd = lambda p, t: _tx_delay_plane(p, 0.0, C_M_S)  # noqa: E731
    data = synthesize_echoes(disk, amps_d, rx, txd, 1, n_s, device=device)
    data = data + synthesize_echoes(clutter[None], amps_c, rx, txd, 1, n_s, device=device)  # broadcast over frames
    data = _add_noise(data, rng)
    scan, gshape, arrivals = workload_grid("doppler", n_grid / 256)
    return _finish("doppler", data, rx, scan, arrivals, gshape, _base_params(1.7, 2), seed, device)


def volumetric(device="cpu", seed=3, scale=1.0, n_frames=8) -> Acquisition:
    """32x32 matrix array, 5 diverging waves (virtual sources 10 mm behind, tilted 0 and +-8 deg in x and
    y), 5000 scatterers in a 20x20x40 mm volume, [5 x 1024 x 1536 x 8], 128^3 grid, f# 1.5."""
    rng = np.random.default_rng(seed)
    n_side = max(4, int(round(32 * scale)))
    n_s = 1536
    n_grid = max(4, int(round(128 * scale)))
    n_sc = max(8, int(round(5000 * scale)))
    rx = probes.matrix_probe_positions(n_side, n_side, PITCH_M)
    tilts = VOLUMETRIC_TILTS_DEG
    sources = _volumetric_sources()
    pts = np.stack([rng.uniform(-10e-3, 10e-3, n_sc), rng.uniform(-10e-3, 10e-3, n_sc),
                    rng.uniform(5e-3, 45e-3, n_sc)], -1)
    amps = (rng.standard_normal(n_sc) + 1j * rng.standard_normal(n_sc)) / math.sqrt(2)
    data = synthesize_echoes(pts[None], amps, rx, lambda p, t: _tx_delay_spherical(p, sources[t], C_M_S),
                             len(tilts), n_s, device=device)
    data = data.expand(-1, -1, -1, n_frames).contiguous()
    data = _add_noise(data, rng)
    scan, _, arrivals = workload_grid("volumetric", n_grid / 128)
    return _finish("volumetric", data, rx, scan, arrivals, (n_grid, n_grid, n_grid), _base_params(1.5, 1), seed, device)


def pymust_disk(device="cpu", seed=5, scale=1.0, n_frames=32) -> Acquisition:
    """Synthetic stand-in with the shape of the reference's published benchmark (PyMUST rotating-disk
    recording, BENCHMARKS.md:22-33): 128 elements at 0.298 mm pitch (L7-4), f0 5.208 MHz, fs = 4 f0,
    c 1480 m/s, one 0 deg plane wave per frame, 32 frames, 1024 samples (35 mm and back), grid 251 x 251
    over x +-12.5 mm, z 10-35 mm (tests/compare/conftest.py:25-35), linear interpolation, f-number 1.0,
    tukey_alpha 0 (test_pymust.py:146-175). Rotating disk (r = 8 mm, centre z = 22.5 mm) of 2000
    scatterers in +20 dB static clutter."""
    rng = np.random.default_rng(seed)
    f0, c, pitch = 5.208e6, 1480.0, 0.298e-3
    fs = 4 * f0
    n_el = max(8, int(round(128 * scale)))
    n_s = 1024
    n_grid = max(8, int(round(251 * scale)))
    n_disk = max(8, int(round(2000 * scale)))
    n_clutter = max(8, int(round(2000 * scale)))
    rx = probes.linear_probe_positions(n_el, pitch)
    radius, cz, omega, prf = 8e-3, 22.5e-3, 2.0, 5000.0
    rr = radius * np.sqrt(rng.uniform(0, 1, n_disk))
    th = rng.uniform(0, 2 * np.pi, n_disk)
    amps_d = (rng.standard_normal(n_disk) + 1j * rng.standard_normal(n_disk)) / math.sqrt(2)
    fr = np.arange(n_frames)
    ang = th[None, :] + omega * fr[:, None] / prf
    disk = np.stack([rr * np.cos(ang), np.zeros_like(ang), cz + rr * np.sin(ang)], -1)
    clutter = np.stack([rng.uniform(-12.5e-3, 12.5e-3, n_clutter), np.zeros(n_clutter),
                        rng.uniform(10e-3, 35e-3, n_clutter)], -1)
    amps_c = 10.0 * (rng.standard_normal(n_clutter) + 1j * rng.standard_normal(n_clutter)) / math.sqrt(2)
    txd = lambda p, t: _tx_delay_plane(p, 0.0, c)  # noqa: E731
    kw = dict(fs=fs, f0=f0, c=c, device=device)
    data = synthesize_echoes(disk, amps_d, rx, txd, 1, n_s, **kw)
    data = data + synthesize_echoes(clutter[None], amps_c, rx, txd, 1, n_s, **kw)
    data = _add_noise(data, rng)
    scan, gshape, arrivals = workload_grid("pymust_disk", n_grid / 251)
    params = dict(f_number=1.0, rx_start_s=0.0, sampling_freq_hz=fs, sound_speed_m_s=c, modulation_freq_hz=f0,
                  tukey_alpha=0.0, interp_type=1)
    return _finish("pymust_disk", data, rx, scan, arrivals, gshape, params, seed, device)


WORKLOADS = {"plane_wave": plane_wave, "compound": compound, "doppler": doppler, "volumetric": volumetric,
             "pymust_disk": pymust_disk}


def make(name: str, device="cpu", **kw) -> Acquisition:
    return WORKLOADS[name](device=device, **kw)
