"""Continuous-batching beamform server over device-resident ring buffers.

Requests (any workload shape, transmit count, frame count, storage type) arrive over time. The
server:

* keeps precomputed geometry on the device, per scan geometry: element and voxel coordinates,
  transmit arrivals and the depth-major voxel order (``GeometryCache``);
* keeps channel data and outputs in device ring buffers (``DeviceRing``). A request's channel data
  lands in its ring slot when it arrives, the way an acquisition front end DMAs into device memory;
* admits arrived requests in arrival order while their channel-data and output bytes fit the HBM
  budget (the rings' capacity). Later arrivals wait until completed work frees space;
* packs every admitted, not-yet-running request into one batch. Single-transmit requests of a batch
  that share a geometry (probe, grid, dtype) are fused into one ensemble along the frame axis, so their
  float64 pair geometry is computed once (``fusable``: multi-transmit requests gain nothing from it). The ensemble's concatenated channel data and output are
  reserved in the rings too (the members run separately when there is no room), so HBM in use never
  exceeds the budget. Small requests (and fused groups) go into one
  persistent-kernel submission (``BatchRunner``, graph replay); a request or group big enough to fill
  the GPU by itself (``plain_min_voxel_frames``) is launched as its own plain kernel on the batch's
  stream, where it gets the launch-shape choices (frame groups, balanced XCD remap) a work queue
  cannot make. Either way nothing synchronises per request. Up to two batches are in flight, on two
  streams, so requests that arrive while a batch runs are admitted and start while the previous
  batch drains;
* while the GPU is busy, holds large requests (``hold_min_points``) back for up to ``max_hold_s`` so
  that same-geometry requests arriving meanwhile fuse with them; requests therefore complete out of
  admission order, and the rings release slots in any order;
* isolates failures: a request that fails validation or launch completes as failed (``error`` set,
  no output) while the others of its batch run, and the stream goes on;
* reports per-request latency (arrival -> result ready on the device), p50/p99, and throughput in
  points/s over the makespan (completed requests only).

``serve_plain`` runs the same arrival stream one request at a time through the plain kernel (one
launch and one synchronisation per request): the baseline the batched server is compared with.
Each request's output is identical in both (same kernel code and per-voxel summation order), which
``tests/test_gpu_serving.py`` checks.
"""

from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np
import torch

from . import autotune
from . import batch as _batch
from . import phantoms
from . import _hip_impl
from ._hip_impl import voxel_order

GIB = 1 << 30


@dataclasses.dataclass
class Geometry:
    rx_coords_m: torch.Tensor
    scan_coords_m: torch.Tensor
    tx_wave_arrivals_s: torch.Tensor
    voxel_order: torch.Tensor | None
    params: dict
    orders: dict = dataclasses.field(default_factory=dict)  # voxel order per wave tile, held for the geometry

    def order_for(self, layout: int, groups: int = 4):
        """The voxel processing order for a launch with this layout code (and ``groups`` voxel groups per
        workgroup), computed once per wave tile and wedge mode and held by the geometry: batches in flight
        on other streams read it, so it must outlive them (the library's own order cache is bounded and may
        drop it)."""
        key = (_hip_impl.wave_tile(layout), (layout >> 16) & 0xF == _hip_impl.XCD_WEDGE_CODE, groups)
        if key not in self.orders:
            self.orders[key] = voxel_order(self.scan_coords_m, layout, groups)
        return self.orders[key]


@dataclasses.dataclass
class ServeRequest:
    rid: int
    arrival_s: float
    geometry: Geometry
    source: torch.Tensor  # resident template channel data [T, E, S, F] the front end "DMAs" from
    kind: str  # workload name
    points: int
    # filled by the server
    admit_s: float = float("nan")
    done_s: float = float("nan")
    out: torch.Tensor | None = None
    data_slot: tuple | None = None
    out_slot: tuple | None = None
    error: str | None = None  # why the request failed (it then has no output)
    service_s: float = float("nan")  # plain path: copy + launch + kernel + synchronisation of this request

    @property
    def data_bytes(self) -> int:
        return self.source.numel() * self.source.element_size()

    @property
    def out_bytes(self) -> int:
        return self.geometry.scan_coords_m.shape[0] * self.source.shape[3] * 8

    @property
    def latency_s(self) -> float:
        return self.done_s - self.arrival_s


class DeviceRing:
    """Byte ring on the device. Allocations are 256-byte aligned and contiguous (an allocation that
    would cross the end starts again at 0). Allocations may be released in any order (``free``);
    space is reclaimed from the oldest end, so a released slot behind a live older one stays occupied
    until that one is released too."""

    ALIGN = 256

    def __init__(self, capacity_bytes: int, device):
        self.capacity = int(capacity_bytes) // self.ALIGN * self.ALIGN
        self.buf = torch.empty(self.capacity, dtype=torch.uint8, device=device)
        self.head = 0
        self.live = collections.deque()  # [start, size, released] in allocation order
        self._undo = None

    @property
    def used(self) -> int:
        """Bytes not yet reclaimable (live allocations and released ones behind a live older one)."""
        return sum(e[1] for e in self.live)

    def alloc(self, nbytes: int):
        """Reserve ``nbytes``; returns (start, size) or None if it does not fit now. A zero-byte
        reservation returns the slot (0, 0), which holds no space and is never live (a live empty
        entry at ``head`` would make the ring look full until it is released)."""
        n = (int(nbytes) + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        if n == 0:
            self._undo = None
            return (0, 0)
        if n > self.capacity:
            raise ValueError(f"request of {nbytes} bytes exceeds the ring capacity {self.capacity}")
        if not self.live:
            start = 0
        else:
            tail = self.live[0][0]
            if self.head > tail:  # occupied bytes are [tail, head)
                if self.head + n <= self.capacity:
                    start = self.head
                elif n <= tail:
                    start = 0
                else:
                    return None
            elif self.head + n <= tail:  # wrapped: the free gap is [head, tail)
                start = self.head
            else:
                return None
        self._undo = self.head
        self.live.append([start, n, False])
        self.head = start + n
        return (start, n)

    def undo_last(self):
        """Release the most recent allocation (admission of a request that did not fully fit)."""
        if self._undo is None:  # the last reservation was of zero bytes
            return
        self.live.pop()
        self.head = self._undo if self.live else 0
        self._undo = None

    def _reclaim(self):
        while self.live and self.live[0][2]:
            self.live.popleft()
        if not self.live:
            self.head = 0

    def free(self, slot):
        """Release the allocation ``slot`` (as returned by ``alloc``), in any order."""
        if slot[1] == 0:  # zero-byte slots hold nothing
            return
        for e in self.live:
            if e[0] == slot[0] and e[1] == slot[1] and not e[2]:
                e[2] = True
                break
        else:
            raise KeyError(f"slot {slot} is not allocated")
        self._reclaim()

    def free_oldest(self):
        self.live[0][2] = True
        self._reclaim()

    def view(self, slot, dtype, shape):
        start, _ = slot
        nbytes = int(np.prod(shape)) * torch.empty((), dtype=dtype).element_size()
        return self.buf[start:start + nbytes].view(dtype).view(shape)


class GeometryCache:
    """Device-resident geometry per workload template and grid scale."""

    def __init__(self, device):
        self.device = device
        self._cache = {}

    def get(self, acq: phantoms.Acquisition, key) -> Geometry:
        g = self._cache.get(key)
        if g is None:
            g = Geometry(acq.rx_coords_m, acq.scan_coords_m, acq.tx_wave_arrivals_s,
                         voxel_order(acq.scan_coords_m), dict(acq.params))
            self._cache[key] = g
        return g


def _frames_tile(data: torch.Tensor, n_frames: int) -> torch.Tensor:
    """[T, E, S, F0] -> [T, E, S, n_frames] by repeating frames (contiguous)."""
    reps = (n_frames + data.shape[3] - 1) // data.shape[3]
    return data.repeat(1, 1, 1, reps)[..., :n_frames].contiguous()


def make_stream(device, n_requests=120, qps=200.0, seed=4, scales=(0.5, 0.625, 0.75, 0.875, 1.0),
                frames=(1, 2, 4, 8, 16, 32, 64, 128, 256), workloads=("plane_wave", "compound", "doppler", "volumetric"),
                template_scale=1.0):
    """Seeded mixed request stream: Poisson arrivals at ``qps``. Each request draws a workload
    (the four benchmark configurations), a grid scale (every grid axis x0.5-1.0) and, for the
    single-plane ensembles (plane_wave, doppler), a frame count 1-256 (powers of two); compound keeps
    its single frame and volumetric its 8 FP16 frames. One channel-data template per workload is
    synthesized (at ``template_scale``) and stays resident; geometry is built once per grid scale."""
    rng = np.random.default_rng(seed)
    cache = GeometryCache(device)
    templates, grids, sources = {}, {}, {}
    reqs = []
    t = 0.0
    for rid in range(n_requests):
        t += rng.exponential(1.0 / qps)
        name = workloads[rng.integers(len(workloads))]
        scale = float(scales[rng.integers(len(scales))])
        if name in ("plane_wave", "doppler"):
            nf = int(frames[rng.integers(len(frames))])
        elif name == "compound":
            nf = 1
        else:
            nf = 8
        if name not in templates:
            kw = {"scale": template_scale}
            if name == "doppler":
                kw["n_frames"] = 1
            templates[name] = phantoms.make(name, device=device, **kw)
        key = (name, scale)
        if key not in grids:
            grids[key] = phantoms.regrid(templates[name], scale * template_scale)
        acq = grids[key]
        geo = cache.get(acq, key)
        skey = (name, nf)
        if skey not in sources:
            if name in ("plane_wave", "doppler"):
                sources[skey] = _frames_tile(acq.channel_data, nf)
            elif name == "volumetric":
                sources[skey] = acq.fp16_channel_data()
            else:
                sources[skey] = acq.channel_data
        src = sources[skey]
        pts = geo.scan_coords_m.shape[0] * acq.n_rx * nf * acq.n_tx
        reqs.append(ServeRequest(rid, t, geo, src, name, pts))
    torch.cuda.synchronize()
    return reqs


def _out_view(req: ServeRequest, ring: DeviceRing, slot):
    return ring.view(slot, torch.complex64, (req.geometry.scan_coords_m.shape[0], req.source.shape[3]))


def _data_view(req: ServeRequest, ring: DeviceRing, slot):
    return ring.view(slot, req.source.dtype, tuple(req.source.shape))


def _to_request(req: ServeRequest, data: torch.Tensor, out: torch.Tensor) -> _batch.Request:
    g = req.geometry
    p = g.params
    r = _batch.Request(data, g.rx_coords_m, g.scan_coords_m, g.tx_wave_arrivals_s, out,
                       f_number=p["f_number"], rx_start_s=p["rx_start_s"], sampling_freq_hz=p["sampling_freq_hz"],
                       sound_speed_m_s=p["sound_speed_m_s"], modulation_freq_hz=p["modulation_freq_hz"],
                       tukey_alpha=p["tukey_alpha"], interp_type=p["interp_type"], accumulate=False,
                       voxel_order=g.voxel_order)
    r.layout = autotune.layout_for(r.kind(), r.interp_type, r.n_tx(), g.rx_coords_m.shape[0], out.shape[1],
                                   g.scan_coords_m.shape[0], data.data_ptr(), out.data_ptr())
    r.voxel_order = g.order_for(r.layout, _hip_impl.groups_per_block(r.n_tx(), out.shape[1]))
    return r


@dataclasses.dataclass
class ServeReport:
    mode: str
    n_requests: int
    makespan_s: float
    points: int
    p50_ms: float
    p99_ms: float
    mean_ms: float
    max_batch: int = 1
    n_batches: int = 0
    n_fused: int = 0  # requests that ran fused with others of the same geometry
    n_direct: int = 0  # requests (or fused groups' members) launched as their own plain kernel in a batch
    n_failed: int = 0  # requests that failed (validation or launch); not counted in points or latency
    by_kind: dict = dataclasses.field(default_factory=dict)  # workload -> requests, points, service time

    @property
    def gpts_per_s(self) -> float:
        return self.points / self.makespan_s / 1e9

    def as_dict(self):
        d = dataclasses.asdict(self)
        d["gpts_per_s"] = self.gpts_per_s
        return d


def _report(mode, reqs, t_first, max_batch=1, n_batches=0) -> ServeReport:
    """Throughput over the points of the requests that completed; latencies of those requests."""
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.latency_s for r in ok]) * 1e3 if ok else np.array([np.nan])
    end = max(r.done_s for r in reqs)
    rep = ServeReport(mode, len(reqs), end - t_first, sum(r.points for r in ok),
                      float(np.percentile(lat, 50)), float(np.percentile(lat, 99)), float(lat.mean()),
                      max_batch, n_batches)
    rep.n_failed = len(reqs) - len(ok)
    for r in ok:
        k = rep.by_kind.setdefault(r.kind, {"requests": 0, "points": 0, "service_s": 0.0})
        k["requests"] += 1
        k["points"] += r.points
        if r.service_s == r.service_s:  # set by the plain path only
            k["service_s"] = round(k["service_s"] + r.service_s, 6)
    return rep


def _submit_isolated(submit_batch, run_direct, batched, direct, after_batched=None):
    """Launches one batch's work so that a request that fails does not take the others with it.

    ``batched`` and ``direct`` are lists of (request, members): ``submit_batch`` takes a list of
    requests as one persistent-kernel submission, ``run_direct`` one request as its own launch. A batch
    submission validates every request before it enqueues anything (mach_batch_submit is all or
    nothing), so when it is refused each request is submitted alone and only those refused again fail.
    ``after_batched`` (if given) is called between the batched and the direct launches. Returns
    [(members, message)] of the failed requests."""
    failed = []
    if batched:
        try:
            submit_batch([r for r, _ in batched])
        except (RuntimeError, ValueError):
            for r, members in batched:
                try:
                    submit_batch([r])
                except (RuntimeError, ValueError) as e:
                    failed.append((members, str(e)))
    if after_batched is not None:
        after_batched()
    for r, members in direct:
        try:
            run_direct(r)
        except (RuntimeError, ValueError) as e:
            failed.append((members, str(e)))
    return failed


def fusable(r) -> bool:
    """Whether fusing the request with same-geometry partners pays: single-transmit ensembles, whose
    wider frame groups share one float64 setup per (voxel, element) across more lanes. Multi-transmit
    requests (compounding, the volume) run alone: their tuned layouts give a lane all of a voxel's
    frames, so a fused ensemble only adds frame blocks that each repeat the setup (measured on the
    8-frame FP16 volume: 23.8 ms alone, 24.8 ms per 8 frames fused to 16, 29.7 ms fused to 32)."""
    return r.source.shape[0] == 1


def _fusion_groups(reqs, fuse=True, max_frames=256):
    """Groups of requests that can run as one ensemble: fusable requests (``fusable``) with the same
    geometry object (probe, grid, transmits, parameters), storage dtype and [n_tx, n_rx, n_samples];
    frames concatenated up to ``max_frames``."""
    if not fuse:
        return [[r] for r in reqs]
    groups, open_ = [], {}
    for r in reqs:
        if not fusable(r):
            groups.append([r])
            continue
        key = (id(r.geometry), r.source.dtype, tuple(r.source.shape[:3]), tuple(r.source.shape[4:]))
        g = open_.get(key)
        nf = r.source.shape[3]
        if g is not None and sum(x.source.shape[3] for x in g) + nf <= max_frames:
            g.append(r)
        else:
            g = [r]
            open_[key] = g
            groups.append(g)
    return groups


def _select(admitted, gpu_busy, now, hold_min_points, max_hold_s, max_fused_frames):
    """Requests to launch now. With the GPU idle, everything admitted. Otherwise large requests
    (``hold_min_points``) that share a geometry are held back so that partners arriving meanwhile fuse
    with them: of those, only the largest group, groups that already fill ``max_fused_frames`` and
    groups whose oldest member has waited ``max_hold_s`` go now."""
    if not gpu_busy or hold_min_points is None:
        return admitted, []
    groups = collections.OrderedDict()
    for item in admitted:
        r = item[0]
        if r.points >= hold_min_points and fusable(r):
            groups.setdefault((id(r.geometry), r.source.dtype, tuple(r.source.shape[:3])), []).append(item)
    if not groups:
        return admitted, []
    frames = {k: sum(it[0].source.shape[3] for it in g) for k, g in groups.items()}
    largest = max(groups, key=lambda k: frames[k])
    go_keys = {k for k, g in groups.items()
               if k == largest or frames[k] >= max_fused_frames or now - min(it[0].admit_s for it in g) >= max_hold_s}
    held = {id(it) for k, g in groups.items() if k not in go_keys for it in g}
    return [it for it in admitted if id(it) not in held], [it for it in admitted if id(it) in held]


def _reserve_fusion(group, data_ring, out_ring, extra):
    """Splits a fusion group into the groups that run: the whole group as one ensemble over its
    concatenated frames when the rings have room for that ensemble's channel data and output (their
    slots are appended to ``extra`` and released with the batch, so the HBM budget covers them), else
    each member on its own. Returns [(members, (data_slot, out_slot) or None)]."""
    if len(group) == 1:
        return [(group, None)]
    n_frames = sum(r.source.shape[3] for r in group)
    n_vox = group[0].geometry.scan_coords_m.shape[0]
    data_bytes, out_bytes = sum(r.data_bytes for r in group), n_vox * n_frames * 8
    ds = os_ = None
    if data_bytes <= data_ring.capacity and out_bytes <= out_ring.capacity:  # else it can never fit
        ds = data_ring.alloc(data_bytes)
        os_ = out_ring.alloc(out_bytes) if ds is not None else None
        if ds is not None and os_ is None:
            data_ring.undo_last()
    if ds is None or os_ is None:
        return [([r], None) for r in group]
    extra += [(data_ring, ds), (out_ring, os_)]
    return [(group, (ds, os_))]


def _prepare(members, slots, data_ring, out_ring, stream):
    """(request, fused channel data or None) of a group from _reserve_fusion; sets each member's out."""
    if slots is None:
        r = members[0]
        r.out = _out_view(r, out_ring, r.out_slot)
        return _to_request(r, _data_view(r, data_ring, r.data_slot), r.out), None
    shape = list(members[0].source.shape)
    shape[3] = sum(r.source.shape[3] for r in members)
    fused = data_ring.view(slots[0], members[0].source.dtype, tuple(shape))
    out_f = out_ring.view(slots[1], torch.complex64, (members[0].geometry.scan_coords_m.shape[0], shape[3]))
    with torch.cuda.stream(stream):
        torch.cat([_data_view(r, data_ring, r.data_slot) for r in members], dim=3, out=fused)
    a = 0
    for r in members:
        nf = r.source.shape[3]
        r.out = out_f[:, a:a + nf]
        a += nf
    return _to_request(members[0], fused, out_f), fused


def serve_batched(reqs, device, budget_bytes=200 * GIB, max_batch=64, use_graph=True, keep_outputs=False,
                  fuse=True, max_fused_frames=256, plain_min_voxel_frames=1 << 22, plain_min_points=5e8,
                  hold_min_points=4e9, max_hold_s=0.5, n_streams=4):
    """Continuous batching over the persistent kernel. Requests of one batch that share a geometry are
    fused into one ensemble (``fuse``); a request or fused group of at least ``plain_min_voxel_frames``
    voxel-frames or ``plain_min_points`` points (about 0.1 ms of GPU work: it fills the GPU alone, and
    its plain launch gets its tuned layout, voxel tiles and XCD order, which the persistent kernel's
    shared register budget and compiled layout set cannot give it) runs as its own plain launch within
    the batch (``plain_min_voxel_frames=None``: never). With
    ``hold_min_points``, large requests wait (at most ``max_hold_s`` while the GPU is busy) for
    same-geometry partners to fuse with. A batch's persistent-kernel submission runs on one of
    ``n_streams`` compute streams and its direct launches, largest first, go each to the stream with
    the least direct work outstanding, so large requests overlap one another's tails. Returns a ServeReport; with ``keep_outputs`` each request's
    ``out`` is a private copy of its result."""
    free, _total = torch.cuda.mem_get_info(device)
    budget = min(int(budget_bytes), int(free * 0.8))
    data_ring = DeviceRing(budget * 3 // 4, device)
    out_ring = DeviceRing(budget - budget * 3 // 4, device)
    copy_stream = torch.cuda.Stream(device)
    compute = [torch.cuda.Stream(device) for _ in range(max(1, n_streams))]
    # direct launches not yet known to be complete, per compute stream: (completion event, points)
    stream_work = [collections.deque() for _ in compute]
    runner = _batch.BatchRunner(max_requests=max_batch, use_graph=use_graph)
    pending = collections.deque(sorted(reqs, key=lambda r: r.arrival_s))
    waiting = collections.deque()
    admitted = []  # (req, copy_event)
    # completion marks of launched work: [batch, members, event, kept objects, extra ring slots]. A batch
    # leaves one mark after its persistent-kernel submission, one per direct launch and a final one
    # (failed members, fused buffers, references); requests complete at their own mark.
    marks = []
    live_batches = collections.Counter()  # batch -> marks not yet complete
    n_done, n_batches, biggest, n_fused, n_direct = 0, 0, 0, 0, 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while n_done < len(reqs):
        now = time.perf_counter() - t0
        while pending and pending[0].arrival_s <= now:
            waiting.append(pending.popleft())
        # admission by HBM budget, in arrival order; a request larger than its ring can never be
        # admitted, so it fails alone instead of blocking the queue
        while waiting:
            r = waiting[0]
            if r.data_bytes > data_ring.capacity or r.out_bytes > out_ring.capacity:
                waiting.popleft()
                r.error = (f"request needs {r.data_bytes} B of channel data and {r.out_bytes} B of output; the HBM "
                           f"budget's rings hold {data_ring.capacity} B and {out_ring.capacity} B")
                r.out, r.admit_s, r.done_s = None, now, now
                n_done += 1
                continue
            ds = data_ring.alloc(r.data_bytes)
            if ds is None:
                break
            os_ = out_ring.alloc(r.out_bytes)
            if os_ is None:
                data_ring.undo_last()
                break
            waiting.popleft()
            r.data_slot, r.out_slot, r.admit_s = ds, os_, now
            with torch.cuda.stream(copy_stream):  # the front end's DMA into the ring slot
                _data_view(r, data_ring, ds).copy_(r.source, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(copy_stream)
            admitted.append((r, ev))
        # completions: every mark whose work has finished
        still = []
        for mk in marks:
            batch_id, done_reqs, ev, _, extra = mk
            if not ev.query():
                still.append(mk)
                continue
            t_done = time.perf_counter() - t0
            for r in done_reqs:
                r.done_s = t_done
                if keep_outputs and r.out is not None:  # private copy, complete before its slot is reused
                    r.out = r.out.clone()
                    torch.cuda.current_stream(device).synchronize()
                data_ring.free(r.data_slot)
                out_ring.free(r.out_slot)
            for ring, slot in extra:  # fused groups' data and output
                ring.free(slot)
            n_done += len(done_reqs)
            live_batches[batch_id] -= 1
            if live_batches[batch_id] == 0:
                del live_batches[batch_id]
        marks = still
        # launch: what is admitted (less the large requests held for fusion partners) forms a batch
        take = []
        if admitted and len(live_batches) < max(2, len(compute)):
            go, admitted = _select(admitted, len(live_batches) > 0, time.perf_counter() - t0, hold_min_points,
                                   max_hold_s, max_fused_frames)
            take = go[:max_batch]
            admitted = go[max_batch:] + admitted
        if take:
            stream = compute[n_batches % len(compute)]
            for _r, ev in take:
                stream.wait_event(ev)
            batched, direct, keep, extra = [], [], [], []
            failed = []
            groups = [sub for g in _fusion_groups([r for r, _ in take], fuse, max_fused_frames)
                      for sub in _reserve_fusion(g, data_ring, out_ring, extra)]
            for members, slots in groups:
                try:
                    req, fused = _prepare(members, slots, data_ring, out_ring, stream)
                except (RuntimeError, ValueError) as e:
                    failed.append((members, str(e)))
                    continue
                # every tensor a launch reads (channel data, output, coordinates, arrivals, voxel order)
                # stays referenced until its batch completes: the kernels run on another stream than
                # the one these tensors were allocated on, so a tensor freed early could be handed out
                # again by the allocator while a kernel still reads it
                keep.append(req)
                if fused is not None:
                    n_fused += len(members)
                points = sum(r.points for r in members)
                if plain_min_voxel_frames is not None and (req.out.numel() >= plain_min_voxel_frames or
                                                           (plain_min_points is not None and points >= plain_min_points)):
                    direct.append((req, members))
                    n_direct += len(members)
                else:
                    batched.append((req, members))
            # small requests first: they finish sooner; GPU-filling ones get their own launch (tuned
            # layout, balanced XCD remap). An event after the submission and after each direct launch
            # lets those requests complete without waiting for the rest of the batch.
            new_marks = []

            def mark(members, on=None):
                ev = torch.cuda.Event()
                ev.record(stream if on is None else on)
                new_marks.append([n_batches, members, ev, None, []])
                return ev

            # direct launches, largest first, each on the compute stream with the fewest points of direct
            # work still outstanding (so the streams' tails end together), after the batch stream's work
            # so far (the copies that fuse groups) and their members' data copies
            prepared = torch.cuda.Event()
            prepared.record(stream)
            direct_done = []

            def run_direct(req):
                for q in stream_work:
                    while q and q[0][0].query():
                        q.popleft()
                i = min(range(len(compute)), key=lambda k: sum(pts for _, pts in stream_work[k]))
                st = compute[i]
                if st is not stream:
                    st.wait_event(prepared)
                _batch.run_plain(req, stream=st)
                ev = mark(direct_members[id(req)], on=st)
                direct_done.append(ev)
                stream_work[i].append((ev, req.points()))

            direct.sort(key=lambda rm: -rm[0].points())
            direct_members = {id(req): members for req, members in direct}
            fails = _submit_isolated(lambda rs: runner.submit(rs, stream=stream), run_direct, batched, direct,
                                     after_batched=lambda: mark([]))
            failed += fails
            failed_ids = {id(members) for members, _ in fails}
            if batched:  # the submission's mark covers the batched requests that were not refused
                sub = next(mk for mk in new_marks if mk[1] == [])
                sub[1] = [r for req, members in batched if id(members) not in failed_ids for r in members]
            else:
                new_marks = [mk for mk in new_marks if mk[1]]
            for members, msg in failed:
                for r in members:
                    r.error, r.out = msg, None
            # the final mark, after every launch of the batch on every stream: failed members, fused
            # buffers, and every launched request's tensors
            for ev_d in direct_done:
                stream.wait_event(ev_d)
            mark([r for members, _ in failed for r in members])
            new_marks[-1][3], new_marks[-1][4] = keep, extra
            marks += new_marks
            live_batches[n_batches] = len(new_marks)
            n_batches += 1
            biggest = max(biggest, len(take))
        elif not admitted and not marks and pending and not waiting:
            # idle until the next arrival
            dt = pending[0].arrival_s - (time.perf_counter() - t0)
            if dt > 2e-4:
                time.sleep(dt - 1e-4)
    torch.cuda.synchronize()
    runner.close()
    rep = _report("batched", reqs, min(r.arrival_s for r in reqs) if reqs else 0.0, biggest, n_batches)
    rep.n_fused = n_fused
    rep.n_direct = n_direct
    del data_ring, out_ring
    return rep


def serve_plain(reqs, device, keep_outputs=False):
    """Baseline: each request, in arrival order, is copied into its own device buffer and beamformed
    by one plain-kernel launch followed by a synchronisation."""
    ordered = sorted(reqs, key=lambda r: r.arrival_s)
    stream = torch.cuda.current_stream(device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in ordered:
        dt = r.arrival_s - (time.perf_counter() - t0)
        if dt > 2e-4:
            time.sleep(dt - 1e-4)
        while time.perf_counter() - t0 < r.arrival_s:
            pass
        r.admit_s = t_start = time.perf_counter() - t0
        data = torch.empty_like(r.source)
        data.copy_(r.source, non_blocking=True)
        out = torch.empty((r.geometry.scan_coords_m.shape[0], r.source.shape[3]), dtype=torch.complex64,
                          device=device)
        try:
            _batch.run_plain(_to_request(r, data, out), stream=stream)
        except (RuntimeError, ValueError) as e:  # this request fails; the stream goes on
            r.error, out = str(e), None
        stream.synchronize()
        r.done_s = time.perf_counter() - t0
        r.service_s = r.done_s - t_start
        r.out = out if keep_outputs else None
    return _report("plain", ordered, ordered[0].arrival_s if ordered else 0.0)
