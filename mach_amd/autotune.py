"""Autotuner for the delay-and-sum kernel layout.

The kernel's tunable shape is the frame-lane layout: how many lanes of a wave share one voxel
(``lanes_per_voxel``: the voxel tile per wave is 64 / lanes_per_voxel voxels) and how many
consecutive frames each lane carries in registers (``frames_per_lane``: the vector width of every
channel-data load). All layouts give bit-identical results (same per-voxel summation order); they
differ in speed with the frame count, voxel count and data type.

``tune`` times every valid candidate on a request with HIP events and returns the fastest;
``TuningTable`` stores winners per shape class in JSON. ``mach_amd`` consults the table at call time
(``MACH_AMD_TUNING`` names the file; default ``mach_amd/tuning/gfx950.json``), and falls back to the
built-in heuristic when a shape class has no entry.
"""

from __future__ import annotations

import json
import math
import os
import threading

import torch

LANES = (1, 2, 4, 8, 16, 32, 64)  # 0 in a layout code selects the LDS-staged band kernel
FRAMES_PER_LANE = (1, 2, 4)
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "gfx950.json")


# (lanes_per_voxel, frames_per_lane) layouts also compiled with 2 or 4 frame groups per lane
MULTI_GROUP = {(16, 4), (32, 2), (64, 1), (8, 4), (1, 4)}


# lateral groups: 64 lanes, every lane holding its frames of 2 or 4 voxels that share sample windows
LATERAL_FLAG = 0x8000
LATERAL = ((2, 2), (2, 4), (4, 2), (4, 4))  # (voxels_per_lane, frames_per_lane)


def encode(lanes_per_voxel: int, frames_per_lane: int, chunks_per_lane: int = 1) -> int:
    """C-ABI layout code: (log2(chunks_per_lane) << 12) | (lanes_per_voxel << 4) | frames_per_lane."""
    return ((chunks_per_lane.bit_length() - 1) << 12) | (lanes_per_voxel << 4) | frames_per_lane


def encode_lateral(voxels_per_lane: int, frames_per_lane: int) -> int:
    """C-ABI code of a lateral-group layout: 0x8000 | (log2(voxels_per_lane) << 12) | (64 << 4) | frames."""
    return LATERAL_FLAG | ((voxels_per_lane.bit_length() - 1) << 12) | (64 << 4) | frames_per_lane


# bits 16-19 of a layout code: automatic, contiguous, runs of 8 .. 128, and XCD wedges (the hardware
# block order over a voxel order that gives each XCD one angular sector of every depth band of a volume)
XCD_CODES = (0, 1, 2, 3, 4, 5, 6, 15)


def with_xcd(layout: int, code: int) -> int:
    """``layout`` with its block-to-XCD remap fixed by ``code`` (bits 16-19: 0 automatic, 1 contiguous
    eighths, 2 <= v <= 10 runs of 2^(v+1) blocks, 15 XCD wedges)."""
    return (layout & 0xFFFF) | (code << 16)


def xcd_code(layout: int) -> int:
    return (layout >> 16) & 0xF


def decode(layout: int):
    """(lanes_per_voxel, frames_per_lane) of a layout code."""
    return (layout >> 4) & 0xFF, layout & 0xF


def chunks_per_lane(layout: int) -> int:
    return 1 if layout & LATERAL_FLAG else 1 << ((layout >> 12) & 0xF)


def voxels_per_lane(layout: int) -> int:
    return 1 << ((layout >> 12) & 0x7) if layout & LATERAL_FLAG else 1


def _bucket(n: int) -> int:
    """Next power of two (shape classes group sizes by octave)."""
    return 1 << max(0, math.ceil(math.log2(max(1, n))))


def _parse_key(key: str):
    """Fields of a shape_key string: frames f, voxel bucket v and the rest (storage, interpolation,
    transmit and element buckets, output kind) as a tuple; None for a malformed key."""
    try:
        parts = {p[0]: p[1:] for p in key.split(":")}
        rest = (parts["d"], parts["i"], parts["t"], parts["e"], parts.get("o", "0"))
        return {"f": int(parts["f"]), "v": int(parts["v"]), "rest": rest}
    except (KeyError, ValueError, IndexError):
        return None


def shape_key(kind: int, interp: int, n_tx: int, n_rx: int, n_frames: int, n_vox: int, output: int = 0) -> str:
    """Shape class of a request; ``output`` (mach_output_kind) other than the signal gets its own entry."""
    key = f"d{kind}:i{int(interp)}:t{_bucket(n_tx)}:e{_bucket(n_rx)}:f{n_frames}:v{_bucket(n_vox)}"
    return key + (f":o{int(output)}" if output else "")


def candidates(n_frames: int, band: bool = True, lateral: bool = False):
    """Layouts worth timing for ``n_frames``: frames-per-lane divides the frame count and a voxel's
    lanes cover no more than twice the frames they have to share (or 8 lanes: extra lanes still split
    the per-pair geometry of setup-bound, few-frame problems). With ``band``, the LDS-staged band
    kernel (lanes_per_voxel code 0) is included where its 64-frame rows divide the frame count; with
    ``lateral`` (single-transmit requests), the lateral-group layouts whose 64 lanes are at least half
    busy."""
    out = []
    if lateral:
        out += [encode_lateral(vpl, k) for vpl, k in LATERAL if n_frames % k == 0 and 32 * k <= n_frames]
    if band:
        out += [encode(0, k) for k in FRAMES_PER_LANE if n_frames % (64 * k) == 0]
    for k in FRAMES_PER_LANE:
        if n_frames % k:
            continue
        for lf in LANES:
            # lanes beyond the frames only help setup-bound problems (few frames): allow up to 8
            if lf * k <= max(2 * n_frames, 8) and (lf * k >= min(n_frames, 64) // 4 or lf == 1):
                out.append(encode(lf, k))
            # several frame groups per lane share one setup when the frames fill them
            if (lf, k) in MULTI_GROUP:
                out += [encode(lf, k, ch) for ch in (2, 4) if lf * k * ch <= n_frames]
    return out


def time_layout(run, layout: int, reps: int = 20, warmup: int = 3) -> float:
    """Average milliseconds of ``run(layout)`` (one launch) over ``reps``, HIP events on the stream."""
    for _ in range(warmup):
        run(layout)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        run(layout)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(run, n_frames: int, reps: int = 20, band: bool = True, lateral: bool = False, rounds: int = 3,
         cands=None):
    """Times every candidate layout with ``run(layout)``; returns (best_layout, {layout: ms}).

    Each candidate is timed in ``rounds`` passes over the list, alternately forwards and backwards, and
    keeps its fastest pass: one short pass per candidate let clock ramps and noise decide between
    layouts within a few percent of each other."""
    times, rejected = {}, set()
    cands = candidates(n_frames, band=band, lateral=lateral) if cands is None else list(cands)
    for r in range(rounds):
        for lay in cands if r % 2 == 0 else cands[::-1]:
            if lay in rejected:
                continue
            try:
                t = time_layout(run, lay, reps)
            except RuntimeError:  # layout rejected for this request (alignment)
                rejected.add(lay)
                continue
            times[lay] = min(t, times.get(lay, t))
    best = min(times, key=times.get)
    return best, times


class TuningTable:
    """Shape class -> layout, persisted as JSON."""

    def __init__(self, path: str | None = None):
        self.path = path or os.environ.get("MACH_AMD_TUNING", DEFAULT_TABLE)
        self.entries: dict = {}
        self._lock = threading.Lock()
        if os.path.exists(self.path):
            with open(self.path) as fh:
                data = json.load(fh)
            self.entries = {k: int(v["layout"]) for k, v in data.get("entries", {}).items()}

    def lookup(self, key: str) -> int:
        """The layout of shape class ``key``; without an entry, that of the nearest tuned class
        (``nearest``), or 0 (the library's heuristic)."""
        hit = self.entries.get(key)
        return hit if hit is not None else self.nearest(key)

    def nearest(self, key: str, max_octaves: int = 3) -> int:
        """Layout of the tuned class closest to ``key``: the same storage, interpolation, transmit and
        element octaves, frame count and output kind, at the nearest voxel octave at most
        ``max_octaves`` away (ties to the larger): a grid scaled from a tuned one. 0 when there is none."""
        want = _parse_key(key)
        if want is None:
            return 0
        best, best_rank = 0, None
        for k, lay in self.entries.items():
            have = _parse_key(k)
            if have is None or have["rest"] != want["rest"] or have["f"] != want["f"]:
                continue
            dv = abs(have["v"].bit_length() - want["v"].bit_length())
            if dv > max_octaves:
                continue
            rank = (dv, -have["v"])
            if best_rank is None or rank < best_rank:
                best, best_rank = lay, rank
        return best

    def record(self, key: str, layout: int, times: dict | None = None, meta: dict | None = None) -> None:
        with self._lock:
            self.entries[key] = int(layout)
            self._meta = getattr(self, "_meta", {})
            self._meta[key] = {"layout": int(layout), "layout_hex": hex(int(layout)),
                               "times_ms": {hex(k): round(v, 5) for k, v in (times or {}).items()},
                               **(meta or {})}

    def save(self, path: str | None = None) -> str:
        path = path or self.path
        os.makedirs(os.path.dirname(path), exist_ok=True)
        meta = getattr(self, "_meta", {})
        data = {"arch": "gfx950", "entries": {k: meta.get(k, {"layout": v, "layout_hex": hex(v)})
                                              for k, v in sorted(self.entries.items())}}
        with open(path, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        return path


_TABLE = None


def table() -> TuningTable:
    global _TABLE
    if _TABLE is None:
        _TABLE = TuningTable()
    return _TABLE


def layout_for(kind: int, interp: int, n_tx: int, n_rx: int, n_frames: int, n_vox: int,
               data_ptr: int = 0, out_ptr: int = 0, output: int = 0) -> int:
    """Tuned layout for a request, or 0 (let the library choose) when the shape class has no entry or
    the tuned frames-per-lane does not fit this call's pointer alignment."""
    lay = table().lookup(shape_key(kind, interp, n_tx, n_rx, n_frames, n_vox, output))
    if lay:
        _, k = decode(lay)
        elem = 8 if kind == 1 else 4
        out_elem = (4 if kind == 0 else 8) if output == 0 else 1  # epilogues store float32 scalars
        if n_frames % k or data_ptr % (k * elem) or out_ptr % (k * out_elem):
            return 0
    return lay


def tune_acquisition(acq, storage: str = "c64", reps: int = 20, record: bool = True, output: int = 0):
    """Tunes the layout for a synthetic acquisition (``mach_amd.phantoms.Acquisition``) and one output
    kind (mach_output_kind: 0 signal, 1 power Doppler image, 2 envelope)."""
    import ctypes

    from . import _hip_impl, _lib

    p = acq.params
    cd = acq.fp16_channel_data() if storage == "fp16" else acq.channel_data
    kind = _hip_impl._data_kind(cd)
    if output == 1:
        out = torch.empty((acq.n_vox,), dtype=torch.float32, device=cd.device)
    elif output == 2:
        out = torch.empty((acq.n_vox, acq.n_frames), dtype=torch.float32, device=cd.device)
    else:
        out = torch.empty((acq.n_vox, acq.n_frames), dtype=torch.complex64, device=cd.device)
    lib = _lib.load()

    def run(layout):
        # each candidate with its own order
        order = _hip_impl.voxel_order(acq.scan_coords_m, layout, _hip_impl.groups_per_block(acq.n_tx, acq.n_frames))
        args = _hip_impl.make_args(cd, acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s, out,
                                   n_tx=acq.n_tx, kind=kind, f_number=p["f_number"], rx_start_s=p["rx_start_s"],
                                   sampling_freq_hz=p["sampling_freq_hz"], sound_speed_m_s=p["sound_speed_m_s"],
                                   modulation_freq_hz=p["modulation_freq_hz"], tukey_alpha=p["tukey_alpha"],
                                   interp_type=p["interp_type"], accumulate=False, layout=layout, order=order,
                                   output=output)
        _lib.check(lib.mach_beamform(ctypes.byref(args), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    # the band kernel sums transmits in a different order and has no epilogues: keep it to
    # single-transmit signal requests, so that a tuned layout never changes results
    best, times = tune(run, acq.n_frames, reps, band=acq.n_tx == 1 and output == 0, lateral=acq.n_tx == 1)
    # then the block order over the XCDs of the two fastest layouts (it interacts with the voxel order a
    # layout gets, so it is chosen per layout: e.g. runs of 8 blocks for the 4x4x4-tiled volume, of 64
    # for Doppler rows; the runner-up can overtake the winner once both have their best order)
    if times[best] > 0:
        bases = [lay for lay in sorted(times, key=times.get) if (lay >> 4) & 0xFF != 0][:2]  # no band kernel
        for base in bases:
            # XCD wedges only where the layout has wave tiles (the order of the other layouts ignores them)
            codes = [c for c in XCD_CODES if c != _hip_impl.XCD_WEDGE_CODE or _hip_impl.wave_tile(base) is not None]
            _, xtimes = tune(run, acq.n_frames, reps, cands=[with_xcd(base, c) for c in codes])
            times.update(xtimes)
        best = min(times, key=times.get)
    key = shape_key(kind, p["interp_type"], acq.n_tx, acq.n_rx, acq.n_frames, acq.n_vox, output)
    if record:
        table().record(key, best, times, {"workload": acq.name, "storage": storage,
                                          "output": ("signal", "power", "envelope")[output]})
    return key, best, times


def _main():
    import argparse

    from . import phantoms

    ap = argparse.ArgumentParser(description="Tune kernel layouts for the benchmark workloads and save the table")
    ap.add_argument("--workloads", default="plane_wave,compound,doppler,volumetric,pymust_disk")
    ap.add_argument("--out", default=None)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--power", default="doppler,pymust_disk", help="workloads also tuned for the power Doppler output")
    ap.add_argument("--envelope", default="doppler,pymust_disk", help="workloads also tuned for the envelope output")
    ap.add_argument("--outputs", default="signal,power,envelope",
                    help="output kinds to tune (power and envelope only for the workloads named above)")
    ap.add_argument("--volumetric-frames", default="",
                    help="extra frame counts of the FP16 volume to tune (longer volumetric ensembles)")
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    for name in args.workloads.split(","):
        frame_counts = [None]
        if name == "volumetric" and args.volumetric_frames:
            frame_counts += [int(f) for f in args.volumetric_frames.split(",")]
        for nf in frame_counts:
            acq = phantoms.make(name, device=dev) if nf is None else phantoms.make(name, device=dev, n_frames=nf)
            storages = ("c64", "fp16") if nf is None else ("fp16",)
            kinds = args.outputs.split(",")
            outputs = [k for k, (kind, names) in enumerate((("signal", None), ("power", args.power),
                                                             ("envelope", args.envelope)))
                       if kind in kinds and (names is None or name in names.split(","))]
            reps = args.reps if nf is None else max(3, args.reps * 8 // nf)  # long launches: fewer repetitions
            for st in storages:
                for output in outputs:
                    key, best, times = tune_acquisition(acq, st, reps, output=output)
                    ranked = sorted(times.items(), key=lambda kv: kv[1])
                    print(f"{name:11s} {st:5s} {key}: best {hex(best)} {times[best]:.4f} ms; "
                          + ", ".join(f"{hex(k)}={v:.4f}" for k, v in ranked[:6]), flush=True)
            del acq
            torch.cuda.empty_cache()
    print("saved", table().save(args.out), flush=True)


if __name__ == "__main__":
    _main()
