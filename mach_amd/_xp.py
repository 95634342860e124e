"""Minimal array-namespace helpers for the host-side geometry code.

The host helpers (`wavefront`, `geometry`, `probes`) accept NumPy arrays and PyTorch tensors. The
reference resolves namespaces through ``array_api_compat`` (``src/mach/_array_api.py:120-145``); that
package is not part of this build, so the two namespaces this build supports are handled here directly.
"""

from __future__ import annotations

import numpy as np

try:  # torch is optional for the host helpers; the GPU path requires it
    import torch
except ImportError:  # pragma: no cover - torch is present in every supported environment
    torch = None


def is_torch(x) -> bool:
    return torch is not None and isinstance(x, torch.Tensor)


def namespace(*arrays):
    """Return ``torch`` if any argument is a tensor, else ``numpy``."""
    for a in arrays:
        if is_torch(a):
            return torch
    return np


def asarray_like(value, like):
    """Convert ``value`` to the namespace/device/dtype family of ``like``."""
    if is_torch(like):
        if is_torch(value):
            return value.to(like.device)
        dtype = like.dtype if like.is_floating_point() else torch.float64
        return torch.as_tensor(value, dtype=dtype, device=like.device)
    return np.asarray(value)


def vector_norm(x, axis=-1):
    if is_torch(x):
        return torch.linalg.vector_norm(x, dim=axis)
    return np.linalg.norm(x, axis=axis)


def vecdot(a, b, axis=-1):
    if is_torch(a) or is_torch(b):
        return (a * b).sum(dim=axis)
    return np.sum(a * b, axis=axis)


def stack(arrays, axis=-1):
    if any(is_torch(a) for a in arrays):
        return torch.stack(list(arrays), dim=axis)
    return np.stack(arrays, axis=axis)
