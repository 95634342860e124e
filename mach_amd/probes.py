"""Probe-element and scan-grid helpers.

``linear_probe_positions`` and ``scan_grid`` follow ``src/mach/io/must.py:71-104`` of the reference
(element x = (i - mean) * pitch, y = z = 0, float32; ``meshgrid(indexing="ij")`` flattened so the
voxel index is (ix * n_y + iy) * n_z + iz). ``matrix_probe_positions`` is the 2-D analogue used for
matrix arrays (no reference counterpart; same centring convention on both axes).
"""

from __future__ import annotations

import numpy as np


def linear_probe_positions(n_elements: int, pitch: float) -> np.ndarray:
    """[n_elements, 3] float32 positions of a linear array along x, centred at the origin."""
    x = np.arange(n_elements) * pitch
    x = x - np.mean(x)
    zeros = np.zeros(n_elements)
    return np.column_stack([x, zeros, zeros]).astype(np.float32)


def matrix_probe_positions(n_x: int, n_y: int, pitch: float) -> np.ndarray:
    """[n_x * n_y, 3] float32 positions of a centred matrix array in the z = 0 plane (x-major)."""
    x = np.arange(n_x) * pitch
    x = x - np.mean(x)
    y = np.arange(n_y) * pitch
    y = y - np.mean(y)
    gx, gy = np.meshgrid(x, y, indexing="ij")
    return np.column_stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)]).astype(np.float32)


def scan_grid(*axes: np.ndarray) -> np.ndarray:
    """Flattened ``meshgrid(*axes, indexing="ij")`` with shape [prod(len(axes)), len(axes)]."""
    n = len(axes)
    return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, n)
