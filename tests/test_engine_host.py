"""The engine's own arithmetic, built for the host from the same source
(libggems_engine.so), against the oracle and IEEE arithmetic.  These pin the
division-free and branch-free forms the HIP kernels use."""

import ctypes as C

import numpy as np

import ggems_oracle as O
from ggems_amd import engine as E

F = np.float32


def test_uniform_mapping_against_literal_formula():
    rng = np.random.default_rng(1)
    s = np.concatenate([rng.integers(0, 2 ** 32, 2_000_000, dtype=np.uint64).astype(np.uint32),
                        np.arange(0, 70000, dtype=np.uint32), (2 ** 32 - 1 - np.arange(70000)).astype(np.uint32)])
    lit = ((s.astype(np.float32).astype(np.float64) / 4294967295.0) * np.float64(F(1.0) - F(2.0 ** -23))).astype(F)
    assert np.array_equal(E.host_uniform_from_sum(s), lit)


def test_division_free_quotient_is_ieee():
    rng = np.random.default_rng(2)
    # voxel indices: offsets in a 410 mm phantom over voxel sizes; exponential
    # distances: -log(u) over cross sections in mm^-1
    a = np.concatenate([rng.uniform(0, 420, 1_000_000), -np.log(rng.uniform(2 ** -32, 1, 1_000_000))]).astype(F)
    b = np.concatenate([rng.choice([0.8, 1.0, 3.2, 0.5, 2.0, 0.3125, 1.7], 1_000_000),
                        np.exp(rng.uniform(np.log(1e-6), np.log(50.0), 1_000_000))]).astype(F)
    got = E.host_div_rn(a, b)
    assert np.array_equal(got, (a / b).astype(F))
    # integer parts of the voxel index on a dense grid of offsets
    a = (np.arange(0, 409.6, 0.0001, dtype=np.float64)).astype(F)
    for v in (0.8, 3.2, 1.0):
        b = np.full(a.shape, v, F)
        assert np.array_equal(E.host_div_rn(a, b).astype(np.int32), (a / b).astype(np.int32))


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-10, 10, (n, 3)).astype(F)
    hi = (lo + rng.uniform(0.01, 5, (n, 3))).astype(F)
    p = rng.uniform(-15, 15, (n, 3)).astype(F)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(F)
    k = n // 8
    # axis-parallel and nearly parallel directions
    d[:k, rng.integers(0, 3, k)] = 0.0
    d[k:2 * k, 0] = F(5e-7)
    # points inside, on faces, within the tolerance of faces
    p[2 * k:3 * k] = (lo[2 * k:3 * k] + hi[2 * k:3 * k]) * F(0.5)
    p[3 * k:4 * k, 0] = lo[3 * k:4 * k, 0]
    p[4 * k:5 * k, 1] = hi[4 * k:5 * k, 1] - F(2e-4)
    return p, d, lo, hi


def test_slab_test_matches_oracle():
    p, d, lo, hi = _cases(200_000, 3)
    box = np.concatenate([lo, hi], axis=1)
    got = E.host_distance_to_aabb(p, d, box)
    fn = O.lib().ggo_distance_to_aabb
    exp = np.empty_like(got)
    for i in range(len(got)):
        exp[i] = fn(O.ptr(p[i]), O.ptr(d[i]), O.ptr(lo[i]), O.ptr(hi[i]))
    assert np.array_equal(got, exp)
    assert (got == F(3.402823466e38)).sum() > 1000 and (got < 1e30).sum() > 1000


def _reference_binary_search(key, a):
    """BinarySearch (include/GGEMS/maths/GGEMSMathAlgorithms.hh:46-60), as written"""
    low, high = 0, len(a) - 1
    while high - low > 1:
        mid = (high + low) // 2
        v = a[mid]
        if v == key:
            return mid
        if v < key:
            low = mid
        else:
            high = mid
    return low


def test_energy_bin_is_the_reference_binary_search():
    n = 220
    i = np.arange(n, dtype=F)
    slope = np.log(F(1.0) / F(0.001)).astype(F)
    log_grid = (F(0.001) * np.exp(slope * (i / F(n - 1)))).astype(F)  # the cross-section grid's form
    rng = np.random.default_rng(5)
    odd_grid = np.sort(rng.uniform(0.001, 1.0, n)).astype(F)  # not log-spaced: the guess is poor
    for grid in (log_grid, odd_grid):
        keys = np.concatenate([
            np.exp(rng.uniform(np.log(1e-4), np.log(3.0), 20000)).astype(F),
            grid, np.nextafter(grid, F(0)), np.nextafter(grid, F(10)),
            np.array([0.0, -1.0, 1e-30, 1e30, np.inf, -np.inf, np.nan], F)])
        got = E.host_energy_bin(keys, grid)
        want = np.array([_reference_binary_search(k, grid) for k in keys], np.int32)
        assert np.array_equal(got, want)
