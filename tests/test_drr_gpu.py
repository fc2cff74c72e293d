"""The ray-traced primary transmission on the GPU: against the oracle's
Siddon restatement, against the Monte Carlo primary counts (scatter flag
clear) normalised by an open field, and per acquisition projection."""

import math

import numpy as np
import pytest

import drr as R
from ggems_amd import authoring, configs
from scene import Scene, Transform, cfg1_scene, mat_mult

pytestmark = pytest.mark.gpu


def test_drr_matches_oracle_mono_and_spectrum(tmp_path):
    sc = cfg1_scene(n_side=60, detector=24, dosimetry=False)
    eng = sc.make_engine()
    k = sc.engine_solids["box"][0]
    got = eng.drr(k)
    ref = R.drr(sc)
    assert np.abs(got - ref).max() < 2e-5 * max(1.0, ref.max()), np.abs(got - ref).max()
    assert ref.min() < 0.5
    eng.close()
    # polyenergetic: the reduced voxel-CT body phantom with the 120 kVp spectrum
    mhd, rng, spec = configs.voxel_ct_inputs(str(tmp_path), 7, (32, 64, 64), 6.4)
    lab, spacing = authoring.read_mhd(mhd)
    e, w = np.loadtxt(spec, unpack=True)
    sc2 = Scene()
    sc2.add_phantom(lab, authoring.BODY_MATERIALS, tuple(spacing))
    sc2.add_ct_flat("GOS", (1, 1), (20, 16, 1), (12.0, 12.0, 1.0), 600.0, 1000.0, 10e-3)
    sc2.set_source((-600.0, 0.0, 0.0), 10.0, e, intensities=w)
    sc2.build()
    eng2 = sc2.make_engine()
    got2 = eng2.drr(sc2.engine_solids["box"][0])
    ref2 = R.drr(sc2)
    assert got2.shape == (16, 20)
    assert np.abs(got2 - ref2).max() < 2e-5, np.abs(got2 - ref2).max()
    assert ref2.min() < 0.2
    eng2.close()


def _counts(sc, n, seed):
    out = sc.run_engine(n, seed)
    h, s = out["hist"][0].astype(np.float64), out["scatter"][0].astype(np.float64)
    return h - s


def test_drr_predicts_monte_carlo_primaries():
    """60 keV through the cfg1 water/bone cylinder: primary counts over
    open-field counts = transmission, pixel by pixel (chi^2), the DRR taken
    as the mean of 4x4 rays per pixel"""
    n = 40_000_000
    det = 32
    sc = cfg1_scene(n_side=60, detector=det, dosimetry=False)
    primary = _counts(sc, n, 3)
    flat_sc = Scene()
    flat_sc.add_ct_flat("GOS", (1, 1), (det, det, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    flat_sc.set_source(*_source_args(sc))
    flat_sc.build()
    flat = _counts(flat_sc, n, 4)
    fine = cfg1_scene(n_side=60, detector=det, dosimetry=False)
    # the same 32 mm module cut into 0.25 mm pixels
    fine.navigators[1][2]["modules"][0].update(ne=np.array([4 * det, 4 * det, 1], np.int32))
    eng = fine.make_engine()
    t = eng.drr(fine.engine_solids["box"][0]).astype(np.float64).reshape(det, 4, det, 4).mean(axis=(1, 3))
    eng.close()
    assert flat.min() > 2000
    var = primary + t * t * flat
    chi2 = float((((primary - t * flat) ** 2) / var).sum())
    npx = det * det
    assert chi2 < 1.3 * npx + 6 * math.sqrt(2 * npx), chi2 / npx
    assert t.min() < 0.5


def _source_args(sc):
    S = sc.source
    pos = tuple(float(v) for v in sc.source["m"][:3, 3])
    aperture = math.degrees(float(S["aperture"]))
    return pos, aperture, [float(S["energy"][0])], None, (0, 0, 0), (0, 0, 0), True


def test_drr_of_an_acquisition_projection():
    sc = cfg1_scene(n_side=60, detector=24, dosimetry=False)
    rot = [(0.0, 0.0, 0.0), (0.0, 0.0, 0.7)]
    eng = sc.engine_acquisition(rot, [1000, 1000], 5)
    k = sc.engine_solids["box"][0]
    p1 = eng.drr(k, 1)
    src, box = sc.acquisition_matrices(rot)
    ref = R.drr(sc, src_m=src[1], box_m=box[1][0])
    assert np.abs(p1 - ref).max() < 2e-5
    assert np.abs(eng.drr(k, 0) - eng.drr(k)).max() == 0.0
    eng.close()
