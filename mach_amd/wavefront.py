"""Transmit-wavefront arrival *distances* (divide by the sound speed to get ``tx_wave_arrivals_s``).

Behaviour follows ``src/mach/wavefront.py`` of the reference: ``plane`` at lines 8-50 and
``spherical`` at lines 53-116. Works on NumPy arrays and PyTorch tensors (CPU or GPU).
"""

from __future__ import annotations

from . import _xp


def plane(origin_m, points_m, direction):
    """Signed distance travelled by a plane wave from ``origin_m`` to each point.

    distance = (point - origin) . direction, with ``direction`` a unit vector
    (reference ``wavefront.py:8-50``; the unit-norm check with tolerance 1e-6 is at lines 36-43).

    Args:
        origin_m: [3] origin of the wavefront (m).
        points_m: [..., 3] points (m).
        direction: [3] unit propagation direction.

    Returns:
        [...] distances in metres; negative for points behind the origin.

    Raises:
        ValueError: if ``direction`` is not a unit vector.
    """
    norm = _xp.vector_norm(direction, axis=-1)
    if not abs(float(norm) - 1.0) < 1e-6:
        raise ValueError("direction must be a unit vector")
    diff = points_m - origin_m
    return _xp.vecdot(diff, direction, axis=-1)


def spherical(origin_m, points_m, focus_m):
    """Signed arrival distance of a focused (focus in front) or diverging (focus behind) spherical wave.

    distance = |origin - focus| * sign(focus_z - origin_z) - |focus - point| * sign(focus_z - point_z)
    (reference ``wavefront.py:53-116``; the z-sign convention is at lines 109-116). The wave is timed
    so that it passes ``origin_m`` at distance 0.

    Args:
        origin_m: [3] transmitting element / reference position (m).
        points_m: [..., 3] points (m).
        focus_m: [3] convergence point (virtual source for a diverging wave) (m).

    Returns:
        [...] distances in metres.
    """
    xp = _xp.namespace(origin_m, points_m, focus_m)
    origin_focus = _xp.vector_norm(origin_m - focus_m, axis=-1)
    focus_point = _xp.vector_norm(focus_m - points_m, axis=-1)
    origin_sign = xp.sign(focus_m[2] - origin_m[2])
    point_sign = xp.sign(focus_m[2] - points_m[..., 2])
    return origin_focus * origin_sign - focus_point * point_sign
