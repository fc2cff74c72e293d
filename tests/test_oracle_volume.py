"""Pins of the oracle's volume painters against float64 geometry on square
volumes, where the kernels' index arithmetic is the plain x-fastest order."""

import numpy as np
import pytest

import volume as V


def _brute(dims, es, shape, pos, size):
    dx, dy, dz = dims
    c = [((np.arange(n) + 0.5) - n / 2.0) * e - p for n, e, p in zip(dims, es, pos)]
    z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    if shape == "box":
        m = (np.abs(y) <= size[0] / 2) & (np.abs(x) <= size[1] / 2) & (np.abs(z) <= size[2] / 2)
    elif shape == "tube":
        m = (np.abs(z) <= size[0] / 2) & ((x / size[1]) ** 2 + (y / size[2]) ** 2 <= 1.0)
    else:
        m = x * x + y * y + z * z <= size[0] ** 2
    return m.reshape(-1)


@pytest.mark.parametrize("shape,size", [("box", (23.3, 41.7, 12.9)), ("tube", (30.3, 17.7, 9.3)),
                                        ("sphere", (13.3,))])
def test_painters_match_float64_geometry_on_square_volumes(shape, size):
    dims, es, pos = (48, 48, 40), (1.0, 1.0, 0.5), (3.1, -2.3, 1.7)
    vol = np.zeros(48 * 48 * 40, np.int32)
    V.draw(vol, dims, es, shape, pos, 7, size)
    want = _brute(dims, es, shape, pos, size)
    assert want.sum() > 100
    assert np.array_equal(vol == 7, want)
    # the index decomposition is x fastest (the MHD order)
    ix, iy, iz = V.voxel_indices(dims)
    g = np.arange(48 * 48 * 40)
    assert np.array_equal(ix + 48 * iy + 48 * 48 * iz, g)


def test_later_shapes_overwrite_and_labels_convert():
    dims, es = (32, 32, 32), (1.0, 1.0, 1.0)
    vol = np.zeros(32 ** 3, np.uint8)
    V.draw(vol, dims, es, "box", (0, 0, 0), 1, (20.0, 20.0, 20.0))
    V.draw(vol, dims, es, "sphere", (0, 0, 0), 2, (5.2,))
    n_sphere = int((_brute(dims, es, "sphere", (0, 0, 0), (5.2,))).sum())
    assert (vol == 2).sum() == n_sphere
    assert (vol == 1).sum() == 20 ** 3 - n_sphere
    f = np.zeros(8, np.float32)
    V.draw(f, (2, 2, 2), es, "sphere", (0, 0, 0), 2.5, (10.0,))
    assert (f == np.float32(2.5)).all()
