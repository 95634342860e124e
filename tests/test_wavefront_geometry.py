"""Known-answer tests for the transmit-arrival and angle helpers.

The input/expected pairs are the closed-form cases of the reference's own tests
(``tests/test_wavefront.py:22-146`` plane waves, ``:153-288`` spherical waves,
``tests/test_geometry.py:22-341`` angle conversions, ``tests/io/test_must.py:43-63`` probe/grid).
"""

import math

import numpy as np
import pytest
import torch

from mach_amd import geometry, probes, wavefront


def A(x):
    return np.asarray(x, dtype=np.float64)


@pytest.mark.parametrize(
    "origin, direction, points, expected",
    [
        ([0, 0, 0], [1, 0, 0], [5, 0, 0], 5.0),
        ([0, 0, 0], [1, 0, 0], [1, 1, 0], 1.0),
        ([0, 0, 0], [1, 0, 0], [0, 5, 0], 0.0),
        ([2, 3, 1], [0, 1, 0], [2, 8, 1], 5.0),
        ([0, 0, 0], [1 / math.sqrt(3)] * 3, [math.sqrt(3)] * 3, 3.0),
        ([5, 5, 5], [1, 0, 0], [5, 5, 5], 0.0),
    ],
)
def test_plane_single_points(origin, direction, points, expected):
    assert float(wavefront.plane(A(origin), A(points), A(direction))) == pytest.approx(expected)


def test_plane_batches():
    res = wavefront.plane(A([0, 0, 0]), A([[1, 0, 0], [2, 1, 0], [0, 3, 0], [-1, 0, 0]]), A([1, 0, 0]))
    assert res.shape == (4,)
    np.testing.assert_allclose(res, [1.0, 2.0, 0.0, -1.0])
    d = A([1 / math.sqrt(2), 1 / math.sqrt(2), 0])
    res = wavefront.plane(A([1, 1, 0]), A([[1, 1, 0], [2, 2, 0], [0, 0, 0]]), d)
    np.testing.assert_allclose(res, [0.0, math.sqrt(2), -math.sqrt(2)], atol=1e-12)
    assert wavefront.plane(A([0, 0, 0]), A([[2, 0, 0]]), A([1, 0, 0])).shape == (1,)


def test_plane_rejects_non_unit_direction():
    with pytest.raises(ValueError, match="direction must be a unit vector"):
        wavefront.plane(A([0, 0, 0]), A([1, 0, 0]), A([2, 0, 0]))


@pytest.mark.parametrize(
    "origin, focus, points, expected",
    [
        ([0, 0, -5], [0, 0, 0], [0, 0, 3], 8.0),  # focused, point beyond focus
        ([0, 0, 0], [0, 0, -5], [0, 0, 3], 3.0),  # diverging, virtual source behind
        ([0, 0, -2], [0, 0, 0], [0, 0, 0], 2.0),  # point at the focus
        ([0, 0, 0], [0, 0, 5], [0, 0, 0], 0.0),  # point at the origin
        ([2, 0, -1], [0, 0, 2], [1, 0, 4], math.sqrt(13) + math.sqrt(5)),
        ([1, 1, 1], [2, 3, 4], [3, 2, 5], math.sqrt(14) + math.sqrt(3)),
        ([0, 0, -5], [0, 0, 0], [0, 0, 5], 10.0),
    ],
)
def test_spherical_single_points(origin, focus, points, expected):
    assert float(wavefront.spherical(A(origin), A(points), A(focus))) == pytest.approx(expected)


def test_spherical_batches():
    res = wavefront.spherical(A([0, 0, -3]), A([[0, 0, 0], [0, 0, 2], [0, 0, -1], [0, 2, 0]]), A([0, 0, 0]))
    np.testing.assert_allclose(res, [3.0, 5.0, 2.0, 3.0])
    res = wavefront.spherical(A([0, 0, 0]), A([[0, 0, 6]]), A([0, 0, 3]))
    assert res.shape == (1,) and float(res[0]) == pytest.approx(6.0)


def test_wavefront_torch_matches_numpy():
    pts = np.random.default_rng(0).normal(size=(16, 3))
    d = A([0.0, 0.6, 0.8])
    np.testing.assert_allclose(
        wavefront.plane(torch.zeros(3, dtype=torch.float64), torch.as_tensor(pts), torch.as_tensor(d)).numpy(),
        wavefront.plane(np.zeros(3), pts, d),
    )
    f = A([0.0, 0.0, -0.01])
    np.testing.assert_allclose(
        wavefront.spherical(torch.zeros(3, dtype=torch.float64), torch.as_tensor(pts), torch.as_tensor(f)).numpy(),
        wavefront.spherical(np.zeros(3), pts, f),
    )


def test_ultrasound_angles_scalars():
    assert geometry.ultrasound_angles_to_cartesian(0.0, 0.0) == pytest.approx((0.0, 0.0, 1.0))
    assert geometry.ultrasound_angles_to_cartesian(math.pi / 2, 0.0, 2.0) == pytest.approx((2.0, 0.0, 0.0), abs=1e-12)
    assert geometry.ultrasound_angles_to_cartesian(0.0, math.pi / 2, 1.5) == pytest.approx((0.0, 1.5, 0.0), abs=1e-12)
    az, el, r = math.radians(30), math.radians(20), 2.5
    x, y, z = geometry.ultrasound_angles_to_cartesian(az, el, r)
    assert (x, y, z) == pytest.approx(
        (r * math.sin(az), r * math.sin(el) * math.cos(az), r * math.cos(el) * math.cos(az))
    )
    assert math.sqrt(x * x + y * y + z * z) == pytest.approx(r)


def test_ultrasound_angles_arrays():
    res = geometry.ultrasound_angles_to_cartesian(A([0.0, math.pi / 2, -math.pi / 4]), A([0.0, 0.0, math.pi / 6]),
                                                  A([1.0, 2.0, 1.5]))
    assert res.shape == (3, 3)
    np.testing.assert_allclose(res[0], [0, 0, 1], atol=1e-12)
    np.testing.assert_allclose(res[1], [2, 0, 0], atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(res, axis=-1), [1.0, 2.0, 1.5])
    mixed = geometry.ultrasound_angles_to_cartesian(A([0.0, math.pi / 2]), 0.0, 1.0)
    assert mixed[0, 2] == pytest.approx(1.0) and mixed[1, 0] == pytest.approx(1.0)


def test_spherical_to_cartesian():
    assert geometry.spherical_to_cartesian(0.0, 0.0) == pytest.approx((0.0, 0.0, 1.0))
    x, y, z = geometry.spherical_to_cartesian(math.pi / 4, math.pi / 4)
    assert (x, y, z) == pytest.approx((0.5, 0.5, math.cos(math.pi / 4)))
    res = geometry.spherical_to_cartesian(A([0.0, math.pi / 4, math.pi / 3]), A([0.0, math.pi / 4, math.pi / 6]),
                                          A([1.0, 2.0, 1.5]))
    np.testing.assert_allclose(res[0], [0, 0, 1], atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(res, axis=-1), [1.0, 2.0, 1.5])
    assert geometry.spherical_to_cartesian(0, 0, 2)[2] == pytest.approx(2.0)


def test_linear_probe_positions_and_grid():
    pos = probes.linear_probe_positions(10, 0.1)
    assert pos.shape == (10, 3) and pos.dtype == np.float32
    np.testing.assert_allclose(pos.mean(axis=0), np.zeros(3), atol=1e-7)
    np.testing.assert_allclose(np.diff(pos[:, 0]), 0.1, rtol=1e-6)
    grid = probes.scan_grid(np.linspace(0, 1, 3), np.linspace(0, 1, 4), np.linspace(0, 1, 5))
    assert grid.shape == (60, 3)
    # ij ordering: voxel index = (ix * n_y + iy) * n_z + iz
    assert tuple(grid[(2 * 4 + 1) * 5 + 3]) == pytest.approx((1.0, 1 / 3, 0.75))
    m = probes.matrix_probe_positions(4, 4, 0.3e-3)
    assert m.shape == (16, 3)
    np.testing.assert_allclose(m.mean(axis=0), 0.0, atol=1e-9)
