"""RF -> IQ on the device: the analytic signal of real channel data, already in the beamformer's layout.

The reference's UFF loader prepares channel data on the host (``src/mach/io/uff.py:70-103``): for a
modulation frequency of 0 it takes ``scipy.signal.hilbert`` along the sample axis and rearranges
(n_samples, n_elements, n_waves, n_frames) to (n_waves, n_elements, n_samples, n_frames). Here one
kernel (``mach_analytic_signal``: an LDS mixed-radix FFT per column) does both on the GPU and can
store the result as FP16 IQ pairs directly. Beamform the result with ``modulation_freq_hz=0``: the
analytic signal keeps its carrier.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._hip_impl import _as_tensor, _shape_str

__all__ = ["analytic_signal", "preprocess_signal"]


def analytic_signal(rf, storage: str = "c64") -> torch.Tensor:
    """Analytic signal along the sample axis.

    Args:
        rf: float32 device array [n_samples, n_elements, n_waves] or
            [n_samples, n_elements, n_waves, n_frames] (the reference's UFF order).
        storage: "c64" (complex64) or "fp16" (float16 [..., 2] re/im pairs; values must fit binary16).

    Returns:
        [n_waves, n_elements, n_samples, n_frames] on the device of ``rf``.

    Raises:
        NotImplementedError: if n_samples is not 2^a 3^b 5^c <= 4096 (the FFT's supported lengths).
    """
    x = _as_tensor(rf, "rf")
    if x.dtype != torch.float32 or x.dim() not in (3, 4) or not x.is_contiguous() or not x.is_cuda:
        raise TypeError(f"analytic_signal(): rf must be a contiguous float32 device array of 3 or 4 dimensions, "
                        f"got {x.dtype} {_shape_str(x)}")
    if storage not in ("c64", "fp16"):
        raise ValueError("storage must be 'c64' or 'fp16'")
    n_s, n_e, n_t = x.shape[:3]
    n_f = x.shape[3] if x.dim() == 4 else 1
    if storage == "c64":
        out = torch.empty((n_t, n_e, n_s, n_f), dtype=torch.complex64, device=x.device)
        code = _lib.DATA_C64
    else:
        out = torch.empty((n_t, n_e, n_s, n_f, 2), dtype=torch.float16, device=x.device)
        code = _lib.DATA_C32H
    lib = _lib.load()
    with torch.cuda.device(x.device):
        stream = torch.cuda.current_stream(x.device).cuda_stream
        st = lib.mach_analytic_signal(ctypes.c_void_p(x.data_ptr()), n_s, n_e, n_t, n_f,
                                      ctypes.c_void_p(out.data_ptr()), code, ctypes.c_void_p(stream))
    if st == _lib.ERR_UNSUPPORTED:
        raise NotImplementedError(_lib.last_error())
    _lib.check(st)
    return out


def preprocess_signal(signal_data, modulation_frequency: float, storage: str = "c64") -> torch.Tensor:
    """Device counterpart of the reference's ``preprocess_signal`` (``src/mach/io/uff.py:70-103``).

    The reference computes the analytic signal only for ``modulation_frequency == 0`` (for any other
    value its result variable is never assigned); the same holds here, with a clear error.
    """
    if modulation_frequency != 0:
        raise NotImplementedError("preprocess_signal(): only RF data (modulation_frequency == 0) is converted; "
                                  "demodulated IQ data is already in the form the beamformer takes")
    return analytic_signal(signal_data, storage)
