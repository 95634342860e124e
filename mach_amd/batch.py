"""Batched execution: many beamform requests in one persistent-kernel launch (``mach_batch_*``).

A ``Request`` holds device tensors and parameters for one beamform (any shape, transmit count,
storage type or interpolation). ``BatchRunner.submit`` packs a list of requests into one launch of
the persistent kernel, replayed from a captured hipGraph. Results equal those of the plain
``mach_beamform`` on each request alone: every voxel is computed by the same code with the same
summation order, only the scheduling of work differs.
"""

from __future__ import annotations

import ctypes
import dataclasses

import torch

from . import _hip_impl, _lib
from ._hip_impl import InterpolationType


@dataclasses.dataclass
class Request:
    """One beamform request. Arrays are device tensors; ``channel_data`` is [n_tx, n_rx, n_samples,
    n_frames] (complex64, float32, or float16 [..., 2] FP16 IQ); ``tx_wave_arrivals_s`` [n_tx, n_vox];
    ``out`` [n_vox, n_frames] (complex64 for IQ, float32 for RF)."""

    channel_data: torch.Tensor
    rx_coords_m: torch.Tensor
    scan_coords_m: torch.Tensor
    tx_wave_arrivals_s: torch.Tensor
    out: torch.Tensor
    f_number: float
    rx_start_s: float
    sampling_freq_hz: float
    sound_speed_m_s: float
    modulation_freq_hz: float = 0.0
    tukey_alpha: float = 0.5
    interp_type: int = int(InterpolationType.Linear)
    accumulate: bool = False
    layout: int = 0
    voxel_order: torch.Tensor | None = None

    def kind(self) -> int:
        return _hip_impl._data_kind(self.channel_data)

    def n_tx(self) -> int:
        return int(self.tx_wave_arrivals_s.shape[0])

    def points(self) -> int:
        core = _hip_impl._core_shape(self.channel_data, self.kind())
        return int(core[0] * core[1] * core[3] * self.scan_coords_m.shape[0])

    def check(self) -> None:
        kind = self.kind()
        core = _hip_impl._core_shape(self.channel_data, kind)
        if len(core) != 4:
            raise ValueError(f"channel_data must be [n_tx, n_rx, n_samples, n_frames], got {tuple(core)}")
        for t, name in ((self.channel_data, "channel_data"), (self.rx_coords_m, "rx_coords_m"),
                        (self.scan_coords_m, "scan_coords_m"), (self.tx_wave_arrivals_s, "tx_wave_arrivals_s"),
                        (self.out, "out")):
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous device tensor")
        if self.tx_wave_arrivals_s.shape[0] != core[0]:
            raise RuntimeError("Dimension mismatch in transmits")
        _hip_impl.check_dimensions(core, self.rx_coords_m, self.scan_coords_m, self.tx_wave_arrivals_s, self.out,
                                   self.rx_coords_m.shape[0], self.scan_coords_m.shape[0], self.out.shape[1])
        want = torch.float32 if kind == _lib.DATA_F32 else torch.complex64
        if self.out.dtype != want:
            raise ValueError(f"out must be {want}")

    def args(self) -> _lib.BeamformArgs:
        return _hip_impl.make_args(
            self.channel_data, self.rx_coords_m, self.scan_coords_m, self.tx_wave_arrivals_s, self.out,
            n_tx=self.n_tx(), kind=self.kind(), f_number=self.f_number, rx_start_s=self.rx_start_s,
            sampling_freq_hz=self.sampling_freq_hz, sound_speed_m_s=self.sound_speed_m_s,
            modulation_freq_hz=self.modulation_freq_hz, tukey_alpha=self.tukey_alpha,
            interp_type=self.interp_type, accumulate=self.accumulate, layout=self.layout,
            order=self.voxel_order,
        )


def run_plain(req: Request, stream=None) -> None:
    """One ``mach_beamform`` launch for ``req`` (the per-request baseline of the batched path)."""
    a = req.args()
    st = torch.cuda.current_stream() if stream is None else stream
    _hip_impl.order_on(req.voxel_order, st)
    _lib.check(_hip_impl._LIB.mach_beamform(ctypes.byref(a), ctypes.c_void_p(st.cuda_stream)))


def submit_noalloc(requests, stream=None) -> None:
    """All requests as one persistent launch on ``stream`` through ``mach_batch_submit_noalloc``: no
    batch object, no allocation (at most 64 requests; calls overlapping on one GPU race). Tensors must
    stay alive until that work completes."""
    n = len(requests)
    arr = (_lib.BeamformArgs * max(n, 1))()
    for i, r in enumerate(requests):
        arr[i] = r.args()
    st = torch.cuda.current_stream() if stream is None else stream
    for r in requests:
        _hip_impl.order_on(r.voxel_order, st)
    _lib.check(_hip_impl._LIB.mach_batch_submit_noalloc(arr, n, ctypes.c_void_p(st.cuda_stream)),
               "mach_batch_submit_noalloc")


class BatchRunner:
    """Owns a ``mach_batch`` (two descriptor slots, graph replay)."""

    def __init__(self, max_requests: int = 256, use_graph: bool = True):
        self.max_requests = max_requests
        self._h = ctypes.c_void_p()
        _lib.check(_hip_impl._LIB.mach_batch_create(max_requests, 1 if use_graph else 0, ctypes.byref(self._h)),
                   "mach_batch_create")

    def submit(self, requests, stream=None) -> None:
        """Enqueue all requests as one persistent launch on ``stream`` (default: current stream).
        Tensors must stay alive until that work completes."""
        n = len(requests)
        if n > self.max_requests:
            raise ValueError(f"{n} requests exceed the batch capacity {self.max_requests}")
        arr = (_lib.BeamformArgs * max(n, 1))()
        for i, r in enumerate(requests):
            arr[i] = r.args()
        st = torch.cuda.current_stream() if stream is None else stream
        for r in requests:
            _hip_impl.order_on(r.voxel_order, st)
        _lib.check(_hip_impl._LIB.mach_batch_submit(self._h, arr, n, ctypes.c_void_p(st.cuda_stream)),
                   "mach_batch_submit")

    def close(self) -> None:
        if self._h:
            _hip_impl._LIB.mach_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
