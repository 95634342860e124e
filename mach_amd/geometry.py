"""Angle -> Cartesian conversions used to build steered transmits.

Behaviour follows ``src/mach/geometry.py`` of the reference:
``ultrasound_angles_to_cartesian`` (lines 31-86), ``spherical_to_cartesian`` (lines 89-132) and the
scalar/array dispatch of ``_prepare_inputs_and_namespace`` (lines 135-154). Scalars return a tuple of
Python floats; array inputs return an array of shape [..., 3].
"""

from __future__ import annotations

import math
import numbers

import numpy as np

from . import _xp


def _prepare(*inputs):
    if all(isinstance(x, numbers.Real) for x in inputs):
        return math, inputs
    arrays = [x for x in inputs if hasattr(x, "shape")]
    xp = _xp.namespace(*arrays)
    like = arrays[0]
    out = tuple(_xp.asarray_like(x, like) if isinstance(x, numbers.Real) else x for x in inputs)
    return xp, out


def ultrasound_angles_to_cartesian(azimuth_rad, elevation_rad, radius_m=1):
    """Ultrasound convention: azimuth rotates about y, then elevation about x.

    x = r sin(az); y = r sin(el) cos(az); z = r cos(el) cos(az)
    (reference ``geometry.py:31-86``).
    """
    xp, (az, el, r) = _prepare(azimuth_rad, elevation_rad, radius_m)
    x = r * xp.sin(az)
    y = r * xp.sin(el) * xp.cos(az)
    z = r * xp.cos(el) * xp.cos(az)
    if xp is math:
        return (x, y, z)
    return _xp.stack(np.broadcast_arrays(x, y, z) if xp is np else _broadcast_torch(x, y, z), axis=-1)


def spherical_to_cartesian(theta_rad, phi_rad, radius_m=1):
    """ISO 80000-2 physics convention: theta = polar angle, phi = azimuth about z.

    x = r sin(theta) cos(phi); y = r sin(theta) sin(phi); z = r cos(theta)
    (reference ``geometry.py:89-132``).
    """
    xp, (theta, phi, r) = _prepare(theta_rad, phi_rad, radius_m)
    x = r * xp.sin(theta) * xp.cos(phi)
    y = r * xp.sin(theta) * xp.sin(phi)
    z = r * xp.cos(theta)
    if xp is math:
        return (x, y, z)
    return _xp.stack(np.broadcast_arrays(x, y, z) if xp is np else _broadcast_torch(x, y, z), axis=-1)


def _broadcast_torch(*arrays):
    import torch

    return torch.broadcast_tensors(*arrays)
