"""FP64 NumPy delay-and-sum oracle.

TEST INFRASTRUCTURE ONLY. This module is the checker for the HIP beamformer. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only to
check or time against. The product path (``mach_amd``) never imports, links or calls anything under
``oracle/``.

What it restates (reference file:line, all under ``src/mach/``):

* scalar handling: every scalar reaches the kernel as float32 (``kernel.cu:1039-1044``, nanobind
  ``float`` parameters); the oracle rounds scalars to float32 and then works in float64.
* input casts: coordinates and arrivals to float32, complex to complex64, real to float32
  (``kernel.py:230-251``). The oracle consumes exactly those float32/complex64 values, upcast.
* aperture: R = z_v / (2 F) (``kernel.cu:583``); skip when h^2 > R^2 with h^2 = dx^2 + dy^2
  (``kernel.cu:456-463``).
* delay: tau = T[v] + |r_e - r_v| * (1 / c) (``kernel.cu:466-471``; the reference multiplies by the
  reciprocal formed once, ``kernel.cu:750``); tau excludes rx_start_s.
* apodization: Tukey half-window on r = h / R when alpha > 0, else 1 (``kernel.cu:333-349, 473-482,
  749``).
* skips: tau < 0 or w == 0 (``kernel.cu:634``).
* sample index s = (tau - t0) * fs (``kernel.cu:637``).
* interpolation (``kernel.cu:106-252``): nearest valid on [-0.5, n-0.5] at round-half-even(s);
  linear valid on [0, n-1], D[floor] + (s - floor)(D[ceil] - D[floor]); quadratic valid on [1, n-2],
  3-point Lagrange about c = round-half-even(s).
* phase rotation: complex data with f_mod != 0 is multiplied by exp(+j 2 pi f_mod tau)
  (``kernel.cu:648-661``); real data is never rotated (``kernel.cu:662-669``).
* accumulation: out[v, f] += sum over transmits and elements (``kernel.cu:674-680``); compounding is
  the sum over transmits (``experimental.py:80-97``).

Post-processing the reference leaves to its callers, restated for the kernel's epilogues:

* power Doppler: sum over frames of |y|^2, ``np.square(np.abs(result)).sum(axis=-1)``
  (``examples/doppler.py:252``; also ``tests/compare/test_pymust.py:317-318``) -> ``power_doppler64``.
* envelope |y| and its dB image: ``db`` replaces exact zeros by ``np.finfo(float).eps`` and takes
  10 log10 (power) or 20 log10 (amplitude) of the magnitude; ``db_zero`` subtracts the image's maximum
  dB value (``src/mach/_vis.py:17-40``) -> ``envelope64``, ``db64``, ``db_zero64``.

Transmit arrivals from wavefront descriptors (``wave_arrivals64``): float32(distance / c + offset),
distance per ``src/mach/wavefront.py:8-50`` (plane: (p - o) . d) and ``53-116`` (spherical:
|o - f| sign(f_z - o_z) - |f - p| sign(f_z - p_z)), evaluated in float64 from the float32 voxel
coordinates with an explicit summation order ((x + y) + z), as the reference's callers do on the host
(``src/mach/io/uff.py:37-67, 156-158``).

RF -> IQ preprocessing (``analytic_signal64``): the analytic signal along the sample axis as
``scipy.signal.hilbert`` computes it (FFT, keep DC and the even-length Nyquist bin, double the
positive frequencies, zero the negative ones, inverse FFT), then the rearrangement from
(n_samples, n_elements, n_waves, n_frames) to (n_waves, n_elements, n_samples, n_frames) of the
reference's UFF preprocessing (``src/mach/io/uff.py:70-103``).

Two edge decisions shared with the HIP kernel (the reference's behaviour there is a bug or undefined):

* nearest neighbour at s in (n-1, n-0.5] rounds to n when n is even (``kernel.cu:117-122`` reads one
  past the end); both this oracle and the kernel clamp the index to n-1.
* a voxel at z = 0 directly under an element gives r = 0/0 in the reference; both use r = 0 there.

Pinning: every function here is checked by ``tests/test_oracle.py`` against closed-form known
answers (Pythagorean geometries, exact interpolation of affine/quadratic signals, Tukey closed forms,
aperture straddles, single-scatterer echoes), invariants (linearity, frame permutation, compounding
equals the sum of single transmits, accumulation) and a scalar brute-force loop; and against the
independent C restatement in ``oracle/das_oracle.c``. The post-processing functions are pinned by
closed forms (Parseval for the power sum of a pure tone, Pythagorean magnitudes, exact decades in dB,
the epsilon substitution at zero).
"""

from __future__ import annotations

import numpy as np

NEAREST, LINEAR, QUADRATIC = 0, 1, 2


def f32(x) -> float:
    """Round a Python scalar to float32 and return it as a float64 value."""
    return float(np.float32(x))


def _as_channel_complex(channel_data: np.ndarray):
    """Upcast stored channel data to float64 / complex128.

    Accepts float32 (RF), complex64 (IQ) and float16 with a trailing IQ axis of length 2 (FP16 IQ
    storage, a layout the reference does not have).
    """
    cd = np.asarray(channel_data)
    if cd.dtype == np.float16:
        if cd.shape[-1] != 2:
            raise ValueError("float16 channel data must carry a trailing IQ axis of length 2")
        c = cd.astype(np.float64)
        return c[..., 0] + 1j * c[..., 1], True
    if np.iscomplexobj(cd):
        return cd.astype(np.complex64).astype(np.complex128), True
    return cd.astype(np.float32).astype(np.float64), False


def tukey(r: np.ndarray, alpha: float) -> np.ndarray:
    """Half Tukey window on r in [0, 1] (``kernel.cu:333-349``)."""
    r = np.asarray(r, dtype=np.float64)
    w = np.where(r <= 1.0 - alpha, 1.0, 0.0)
    if alpha > 0:
        taper = (r > 1.0 - alpha) & (r <= 1.0)
        with np.errstate(invalid="ignore", divide="ignore"):
            w = np.where(taper, 0.5 - 0.5 * np.cos(np.pi * (1.0 - r) / alpha), w)
    w = np.where((r < 0.0) | (r > 1.0), 0.0, w)
    return w


def delays_and_weights(rx, vox, tx_arrival, *, f_number, sound_speed_m_s, tukey_alpha):
    """Per (voxel, element) delay tau [V, E], apodization weight w [V, E] and in-aperture mask.

    ``rx``/``vox`` are float64 views of the float32 API coordinates; ``tx_arrival`` is [V].
    Restates ``calculateTxRxDelayAndApodization`` (``kernel.cu:444-483``) and the aperture radius
    (``kernel.cu:580-584``) in float64.
    """
    F = f32(f_number)
    c = f32(sound_speed_m_s)
    alpha = f32(tukey_alpha)
    xv, yv, zv = vox[:, 0:1], vox[:, 1:2], vox[:, 2:3]
    dx = rx[None, :, 0] - xv
    dy = rx[None, :, 1] - yv
    dz = rx[None, :, 2] - zv
    h2 = dx * dx + dy * dy
    R = zv / (2.0 * F)
    in_ap = ~(h2 > R * R)
    d = np.sqrt(h2 + dz * dz)
    tau = tx_arrival[:, None] + d * (1.0 / c)
    if alpha > 0:
        h = np.sqrt(h2)
        with np.errstate(invalid="ignore", divide="ignore"):
            r = np.where(h2 == 0.0, 0.0, h / R)
        w = tukey(r, alpha)
    else:
        w = np.ones_like(tau)
    return tau, w, in_ap


def interp_taps(s: np.ndarray, n: int, interp: int):
    """Tap indices [..., K], tap weights [..., K] and validity for fractional sample indices ``s``.

    Restates ``interpolate_nearest`` (``kernel.cu:106-131``), ``interpolate_linear`` (145-180) and
    ``interpolate_quadratic`` (194-252). Invalid positions get index 0 and weight 0.
    """
    s = np.asarray(s, dtype=np.float64)
    if interp == NEAREST:
        valid = (s >= -0.5) & (s <= n - 0.5)
        idx = np.minimum(np.rint(np.where(valid, s, 0.0)), n - 1).astype(np.int64)
        idx = np.maximum(idx, 0)
        return idx[..., None], np.where(valid, 1.0, 0.0)[..., None], valid
    if interp == LINEAR:
        valid = (s >= 0.0) & (s <= n - 1)
        sv = np.where(valid, s, 0.0)
        i0 = np.floor(sv)
        i1 = np.ceil(sv)
        t = sv - i0
        idx = np.stack([i0, i1], axis=-1).astype(np.int64)
        # D[i0] + t (D[i1] - D[i0]) = (1 - t) D[i0] + t D[i1]; kept as lerp form in gather_interp
        wts = np.stack([1.0 - t, t], axis=-1) * valid[..., None]
        return idx, wts, valid
    if interp == QUADRATIC:
        valid = (s >= 1.0) & (s <= n - 2.0)
        sv = np.where(valid, s, 1.0)
        c = np.rint(sv)
        x = sv - c
        idx = np.clip(np.stack([c - 1, c, c + 1], axis=-1), 0, max(n - 1, 0)).astype(np.int64)
        wts = np.stack([0.5 * x * (x - 1.0), (1.0 - x) * (1.0 + x), 0.5 * x * (x + 1.0)], axis=-1)
        return idx, wts * valid[..., None], valid
    raise ValueError(f"unknown interpolation type {interp}")


def _gather_interp(D2, idx, wts, interp):
    """Interpolated channel data [..., F] from D2 = D.reshape(E * S, F) at flat tap indices."""
    if interp == LINEAR:
        a = D2[idx[..., 0]]
        b = D2[idx[..., 1]]
        t = wts[..., 1:2]
        valid = (wts[..., 0:1] + wts[..., 1:2]) != 0
        return np.where(valid, a + t * (b - a), 0.0)
    acc = 0.0
    for k in range(idx.shape[-1]):
        acc = acc + wts[..., k : k + 1] * D2[idx[..., k]]
    return acc


def das64(
    channel_data,
    rx_coords_m,
    scan_coords_m,
    tx_wave_arrivals_s,
    *,
    f_number,
    rx_start_s,
    sampling_freq_hz,
    sound_speed_m_s,
    modulation_freq_hz=0.0,
    tukey_alpha=0.5,
    interp_type=LINEAR,
    out=None,
    voxel_chunk=None,
):
    """FP64 delay-and-sum.

    Args:
        channel_data: [E, S, F] or [T, E, S, F]; float32 (RF), complex64 (IQ) or float16 [..., 2] (IQ).
        rx_coords_m: [E, 3]. scan_coords_m: [V, 3]. tx_wave_arrivals_s: [V] or [T, V].
        out: optional [V, F] array added to (reference device-``out`` semantics).

    Returns:
        [V, F] complex128 (IQ) or float64 (RF).
    """
    D, is_complex = _as_channel_complex(channel_data)
    arr = np.asarray(tx_wave_arrivals_s)
    if D.ndim == 3:
        D = D[None]
        arr = arr[None]
    T, E, S, F = D.shape
    rx = np.asarray(rx_coords_m).astype(np.float32).astype(np.float64)
    vox = np.asarray(scan_coords_m).astype(np.float32).astype(np.float64)
    arr = arr.astype(np.float32).astype(np.float64)
    V = vox.shape[0]
    t0 = f32(rx_start_s)
    fs = f32(sampling_freq_hz)
    fmod = f32(modulation_freq_hz)
    interp = int(interp_type)
    rotate = is_complex and fmod != 0.0
    res = np.zeros((V, F), dtype=np.complex128 if is_complex else np.float64)
    if voxel_chunk is None:
        voxel_chunk = max(1, (1 << 22) // max(1, E * F * 3))
    e_ar = np.arange(E)
    for t in range(T):
        D2 = D[t].reshape(E * S, F)
        for v0 in range(0, V, voxel_chunk):
            v1 = min(V, v0 + voxel_chunk)
            tau, w, in_ap = delays_and_weights(
                rx, vox[v0:v1], arr[t, v0:v1],
                f_number=f_number, sound_speed_m_s=sound_speed_m_s, tukey_alpha=tukey_alpha,
            )
            keep = in_ap & ~(tau < 0.0) & (w != 0.0)
            s = (tau - t0) * fs
            idx, wts, valid = interp_taps(s, S, interp)
            keep &= valid
            coef = np.where(keep, w, 0.0)
            if rotate:
                coef = coef * np.exp(1j * 2.0 * np.pi * fmod * tau)
            flat = idx + (e_ar * S)[None, :, None]
            vals = _gather_interp(D2, flat, wts * keep[..., None], interp)
            res[v0:v1] += np.einsum("ve,vef->vf", coef, vals)
    if out is not None:
        res = res + np.asarray(out)
    return res


def das_bruteforce(
    channel_data, rx_coords_m, scan_coords_m, tx_wave_arrivals_s, *, f_number, rx_start_s,
    sampling_freq_hz, sound_speed_m_s, modulation_freq_hz=0.0, tukey_alpha=0.5, interp_type=LINEAR,
):
    """Scalar-loop FP64 delay-and-sum for tiny inputs (an independent check of ``das64``).

    Written directly from the per-point description above with Python math only.
    """
    import math

    D, is_complex = _as_channel_complex(channel_data)
    arr = np.asarray(tx_wave_arrivals_s).astype(np.float32).astype(np.float64)
    if D.ndim == 3:
        D = D[None]
        arr = arr[None]
    T, E, S, F = D.shape
    rx = np.asarray(rx_coords_m).astype(np.float32).astype(np.float64)
    vox = np.asarray(scan_coords_m).astype(np.float32).astype(np.float64)
    Fn, c, alpha = f32(f_number), f32(sound_speed_m_s), f32(tukey_alpha)
    t0, fs, fmod = f32(rx_start_s), f32(sampling_freq_hz), f32(modulation_freq_hz)
    out = np.zeros((vox.shape[0], F), dtype=np.complex128 if is_complex else np.float64)

    def round_half_even(x):
        fl = math.floor(x)
        diff = x - fl
        if diff > 0.5:
            return fl + 1
        if diff < 0.5:
            return fl
        return fl if fl % 2 == 0 else fl + 1

    for v in range(vox.shape[0]):
        xv, yv, zv = vox[v]
        R = zv / (2 * Fn)
        for t in range(T):
            for e in range(E):
                dx, dy, dz = rx[e, 0] - xv, rx[e, 1] - yv, rx[e, 2] - zv
                h2 = dx * dx + dy * dy
                if h2 > R * R:
                    continue
                tau = arr[t, v] + math.sqrt(h2 + dz * dz) * (1.0 / c)
                w = 1.0
                if alpha > 0:
                    r = 0.0 if h2 == 0 else math.sqrt(h2) / R
                    if r < 0 or r > 1:
                        w = 0.0
                    elif r <= 1 - alpha:
                        w = 1.0
                    else:
                        w = 0.5 - 0.5 * math.cos(math.pi * (1 - r) / alpha)
                if tau < 0 or w == 0:
                    continue
                s = (tau - t0) * fs
                for f in range(F):
                    if interp_type == NEAREST:
                        if s < -0.5 or s > S - 0.5:
                            continue
                        val = D[t, e, min(round_half_even(s), S - 1), f]
                    elif interp_type == LINEAR:
                        if s < 0 or s > S - 1:
                            continue
                        lo, hi = math.floor(s), math.ceil(s)
                        a, b = D[t, e, lo, f], D[t, e, hi, f]
                        val = a + (s - lo) * (b - a)
                    else:
                        if s < 1 or s > S - 2:
                            continue
                        cc = round_half_even(s)
                        x = s - cc
                        val = (0.5 * x * (x - 1) * D[t, e, cc - 1, f] + (1 - x) * (1 + x) * D[t, e, cc, f]
                               + 0.5 * x * (x + 1) * D[t, e, cc + 1, f])
                    term = w * val
                    if is_complex and fmod != 0:
                        ph = 2 * math.pi * fmod * tau
                        term = term * complex(math.cos(ph), math.sin(ph))
                    out[v, f] += term
    return out


def rel_l2(y, y_ref) -> float:
    """||y - y_ref||_2 / ||y_ref||_2 over all elements (the parity metric)."""
    y = np.asarray(y).astype(np.complex128)
    y_ref = np.asarray(y_ref).astype(np.complex128)
    den = np.linalg.norm(y_ref.ravel())
    num = np.linalg.norm((y - y_ref).ravel())
    return float(num / den) if den > 0 else float(num)


def bmode_db_diff(y, y_ref, floor_db=-60.0) -> float:
    """max |20 log10(|y|/max|y|) - 20 log10(|y_ref|/max|y_ref|)| over voxels above ``floor_db``.

    The dB convention is the reference's ``_vis.db`` (``src/mach/_vis.py:17-28``) applied per image.
    """
    a = np.abs(np.asarray(y))
    b = np.abs(np.asarray(y_ref))
    with np.errstate(divide="ignore"):
        da = 20 * np.log10(a / a.max())
        db = 20 * np.log10(b / b.max())
    mask = db > floor_db
    return float(np.max(np.abs(da[mask] - db[mask]))) if mask.any() else 0.0


def power_doppler64(y) -> np.ndarray:
    """sum_f |y[v, f]|^2 in float64 (``examples/doppler.py:252``)."""
    y = np.asarray(y).astype(np.complex128)
    return (y.real ** 2 + y.imag ** 2).sum(axis=-1)


def envelope64(y) -> np.ndarray:
    """|y| in float64: complex magnitude for IQ, absolute value for RF (``_vis.py:24-28``)."""
    return np.abs(np.asarray(y).astype(np.complex128))


def db64(x, power_mode=False) -> np.ndarray:
    """``_vis.db`` (``src/mach/_vis.py:17-28``): exact zeros -> float64 eps, then 10 or 20 log10 |x|."""
    x = np.asarray(x, dtype=np.float64)
    x = np.where(x == 0, np.finfo(float).eps, x)
    return (10.0 if power_mode else 20.0) * np.log10(np.abs(x))


def db_zero64(x, power_mode=False) -> np.ndarray:
    """``_vis.db_zero`` (``src/mach/_vis.py:30-40``): ``db64`` minus its maximum."""
    d = db64(x, power_mode)
    return d - d.max()


def wave_arrivals64(waves, scan) -> np.ndarray:
    """[T, V] float32 transmit arrivals of wavefronts ``waves`` = [(kind, origin[3], vector[3], c, offset)]
    (kind 0 plane with unit direction ``vector``, 1 spherical with focus ``vector``) at float32 voxels."""
    p = np.asarray(scan, dtype=np.float32).astype(np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    out = []
    for kind, origin, vec, c, offset in waves:
        o = np.asarray(origin, dtype=np.float64)
        v = np.asarray(vec, dtype=np.float64)
        if kind == 0:
            d = ((x - o[0]) * v[0] + (y - o[1]) * v[1]) + (z - o[2]) * v[2]
        else:
            a = o - v
            origin_focus = np.sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2])
            bx, by, bz = v[0] - x, v[1] - y, v[2] - z
            focus_point = np.sqrt((bx * bx + by * by) + bz * bz)
            d = origin_focus * np.sign(v[2] - o[2]) - focus_point * np.sign(v[2] - z)
        out.append((d / float(c) + float(offset)).astype(np.float32))
    return np.stack(out) if out else np.zeros((0, len(p)), dtype=np.float32)


def analytic_signal64(rf) -> np.ndarray:
    """[S, E, T(, F)] real -> [T, E, S, F] complex128 analytic signal along S (``io/uff.py:70-103``)."""
    x = np.asarray(rf, dtype=np.float64)
    if x.ndim == 3:
        x = x[..., None]
    n = x.shape[0]
    X = np.fft.fft(x, axis=0)
    h = np.zeros(n)
    if n % 2 == 0:
        h[0] = h[n // 2] = 1.0
        h[1:n // 2] = 2.0
    else:
        h[0] = 1.0
        h[1:(n + 1) // 2] = 2.0
    xa = np.fft.ifft(X * h[:, None, None, None], axis=0)
    return np.ascontiguousarray(xa.transpose(2, 1, 0, 3))
