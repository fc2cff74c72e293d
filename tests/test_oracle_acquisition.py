"""Pins of the oracle's CT acquisition (projections as consecutive runs):
exact equalities with plain runs where the slot/stream bookkeeping makes
them equal, and the rotation symmetry of a cylinder for the placements."""

import copy
import math

import numpy as np
import pytest

from scene import Scene, Transform, cfg1_scene, mat_mult

F = np.float32


def _small_scene(aperture_deg=0.4):
    sc = Scene()
    sc.add_phantom(np.zeros((10, 10, 10), np.uint8), ["Water"], (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (16, 16, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), aperture_deg, [0.06], mono=True)
    return sc.build()


def test_single_projection_equals_run():
    sc = _small_scene()
    n = 150_000
    acq = sc.run_oracle_acquisition([(0.0, 0.0, 0.0)], [n], 11, threads=8)
    ref = sc.run_oracle(n, 11, threads=8)
    assert np.array_equal(acq["hist"][0][0], ref["hist"][0])
    assert np.array_equal(acq["scatter"][0][0], ref["scatter"][0])
    assert acq["stats"] == ref["stats"]


def test_rotated_projection_equals_run_of_rotated_scene():
    """budgets [0, n]: projection 1 starts from the seeded streams, so it is
    the plain run of the scene with source and module placed at R*M"""
    sc = _small_scene()
    n = 100_000
    r = (0.0, 0.0, math.radians(37.0))
    acq = sc.run_oracle_acquisition([(0.0, 0.0, 0.0), r], [0, n], 5, threads=8)
    assert not acq["hist"][0][0].any()
    rot = copy.copy(sc)
    R = Transform().rotate(r).m
    rot.source = dict(sc.source, m=mat_mult(R, sc.source["m"]))
    navs = []
    for kind, mats, pay in sc.navigators:
        if kind == "ct":
            pay = dict(pay, modules=[dict(m, m=mat_mult(R, m["m"])) for m in pay["modules"]])
        navs.append((kind, mats, pay))
    rot.navigators = navs
    ref = rot.run_oracle(n, 5, threads=8)
    assert np.array_equal(acq["hist"][0][1], ref["hist"][0])
    assert ref["hist"][0].sum() > 0.3 * n


def test_consecutive_projections_continue_the_streams():
    """two projections of 2^20 primaries at one placement: slot s runs one
    photon of each, continuing its stream - the layout of one run of 2^21
    primaries (two batches, slot s takes photon s of each)"""
    sc = _small_scene()
    n = 1 << 20
    acq = sc.run_oracle_acquisition([(0.0, 0.0, 0.0)] * 2, [n, n], 3, threads=8)
    ref = sc.run_oracle(2 * n, 3, threads=8)
    assert np.array_equal(acq["hist"][0].sum(axis=0), ref["hist"][0])
    assert acq["stats"] == ref["stats"]


def test_cylinder_projections_agree_under_rotation():
    """the cfg1 cylinder is symmetric under quarter turns about z: the
    projections at 0 and 90 degrees agree within Poisson noise"""
    sc = cfg1_scene(n_side=60, detector=32, dosimetry=False, aperture_deg=1.5)
    n = 200_000
    acq = sc.run_oracle_acquisition([(0.0, 0.0, 0.0), (0.0, 0.0, math.pi / 2)], [n, n], 9, threads=8)
    a, b = acq["hist"][0][0].astype(np.float64), acq["hist"][0][1].astype(np.float64)
    assert a.sum() > 0.01 * n
    assert abs(a.sum() - b.sum()) < 5 * math.sqrt(a.sum() + b.sum())
    # column profiles (along the cylinder axis the phantom is uniform)
    pa, pb = a.sum(axis=0), b.sum(axis=0)
    chi2 = float((((pa - pb) ** 2) / np.maximum(pa + pb, 1)).sum())
    assert chi2 < len(pa) + 6 * math.sqrt(2 * len(pa)), chi2


def _two_phantom_scene(keep=(0, 1)):
    """two different phantoms at the isocentre (a water cube and a bone-cored
    cylinder), of which `keep` are registered, and one flat panel"""
    from scene import cylinder_phantom
    sc = Scene()
    if 0 in keep:
        sc.add_phantom(np.zeros((12, 12, 12), np.uint8), ["Water"], (1.5, 1.5, 1.5))
    if 1 in keep:
        sc.add_phantom(cylinder_phantom(24, radius=6.0), ["Water", "RibBone"], (1.0, 1.0, 1.0),
                       rotation=(0.0, 0.3, 0.0), dosimetry={})
    sc.add_ct_flat("GOS", (1, 1), (20, 20, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 0.6, [0.06], mono=True)
    return sc.build()


def test_projection_sees_only_its_phantom():
    """A projection over phantom solid k of a world holding both phantoms is
    the plain acquisition of the world holding phantom k alone: the same
    histories, bit for bit, including the dose map of the phantom kept."""
    both = _two_phantom_scene()
    rot = [(0.0, 0.0, 0.0), (0.0, 0.0, 0.9)]
    n = [60_000, 90_000]
    for k in (0, 1):
        acq = both.run_oracle_acquisition(rot, n, 21, threads=8, phantoms=[k, k])
        alone = _two_phantom_scene(keep=(k,)).run_oracle_acquisition(rot, n, 21, threads=8)
        assert np.array_equal(acq["hist"][0], alone["hist"][0]), k
        assert np.array_equal(acq["scatter"][0], alone["scatter"][0]), k
        assert acq["stats"] == alone["stats"], k
        if k == 1:
            assert alone["edep"].sum() > 0 and np.array_equal(acq["edep"], alone["edep"])
        else:
            assert not acq["edep"].any()  # the dose phantom is absent from every projection
    # -1 keeps every solid: the plain acquisition of the world with both
    acq = both.run_oracle_acquisition(rot, n, 21, threads=8, phantoms=[-1, -1])
    plain = both.run_oracle_acquisition(rot, n, 21, threads=8)
    assert np.array_equal(acq["hist"][0], plain["hist"][0])


def test_projection_phantom_must_name_a_phantom():
    both = _two_phantom_scene()
    for bad in ([2, 0], [5, 0], [-2, 0]):  # solid 2 is the panel
        with pytest.raises(RuntimeError):
            both.run_oracle_acquisition([(0.0, 0.0, 0.0)] * 2, [10, 10], 1, threads=2, phantoms=bad)
