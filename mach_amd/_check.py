"""Array checks used by the Python entry points (reference: ``src/mach/_check.py:10-41``)."""

from __future__ import annotations

import warnings

import numpy as np

from . import _xp


def is_contiguous(array) -> bool:
    """C-contiguity for NumPy arrays and PyTorch tensors; other arrays are assumed contiguous."""
    if isinstance(array, np.ndarray):
        return bool(array.flags["C_CONTIGUOUS"])
    if _xp.is_torch(array):
        return bool(array.is_contiguous())
    return True


def ensure_contiguous(array, *, warn: bool = True):
    """Return ``array`` or a C-contiguous copy, warning as the reference does (``_check.py:35-38``)."""
    if is_contiguous(array):
        return array
    if warn:
        warnings.warn("array is not contiguous, rearranging will add latency", stacklevel=2)
    if isinstance(array, np.ndarray):
        return np.ascontiguousarray(array)
    return array.contiguous()


def is_writeable(array) -> bool:
    if isinstance(array, np.ndarray):
        return bool(array.flags.writeable)
    return True
