"""Dose finalisation on the GPU (needs an MI355X).

ggh_compute_dose runs compute_dose_ggems_voxelized_solid
(src/kernels/ComputeDoseGGEMSVoxelizedSolid.cl:53-121; for a meshed solid
ComputeDoseGGEMSMeshSolid.cl:53-114).  Here the engine's own edep, edep^2 and
hit tallies are fed to the oracle's ggo_compute_dose, which
tests/test_oracle_pins.py pins to the closed forms, and the two dose and
uncertainty maps must agree bit for bit: both evaluate the same FP32 dosel
centres, the same IEEE FP64 quotients and square roots, and round once to
FP32.  The unit constants, the water reference, the minimum density and the
uncertainty formula are also checked against a closed form computed here in
FP64 on chosen dosels.
"""

import ctypes as C
import math
import os

import numpy as np
import pytest

import ggems_oracle as O
from ggems_amd import authoring
from ggems_amd.engine import library
from scene import REPO, Scene, cfg1_scene, voxel_ct_scene

pytestmark = pytest.mark.gpu
F = np.float32

G_CM3 = 6.241509704e21 / 1e3        # internal density unit per g/cm3
MEV_J, GY = 1.602176634e-13, 1.0e-12  # MeV in J; Gy in internal units (MeV / (internal mass))


def _oracle_dose(sc, edep, edep2, hit, scale, water, min_density, solid=0):
    """ggo_compute_dose on the given tallies of the scene's dosimetry solid."""
    t = {}
    _, solids, _ = sc._oracle_structs(t)
    s = solids[solid]
    edep = np.ascontiguousarray(edep, np.float64)
    edep2 = np.ascontiguousarray(edep2, np.float64)
    hit = np.ascontiguousarray(hit, np.int32)
    s.edep, s.edep2, s.hit = O.ptr(edep), O.ptr(edep2), O.ptr(hit)
    dens = np.ascontiguousarray(sc.arrays["density"], F)
    dose = np.zeros(edep.shape, F)
    unc = np.zeros(edep.shape, F)
    O.lib().ggo_compute_dose(C.byref(s), O.ptr(dens), F(scale), int(water), F(min_density), O.ptr(dose),
                             O.ptr(unc))
    return dose, unc


def _check_against_oracle(sc, eng, cases):
    edep, edep2, hit, _ = eng.dose_tallies()
    assert hit.sum() > 1000 and (hit > 1).any()
    for scale, water, min_density in cases:
        d_gpu, u_gpu = eng.compute_dose(scale, water, min_density)
        d_ref, u_ref = _oracle_dose(sc, edep, edep2, hit, scale, water, min_density)
        assert np.array_equal(d_gpu, d_ref), (scale, water, min_density,
                                              int((d_gpu != d_ref).sum()))
        assert np.array_equal(u_gpu, u_ref), (scale, water, min_density, int((u_gpu != u_ref).sum()))
    return edep, edep2, hit


def _closed_form(edep, edep2, hit, rho_g_cm3, dosel_mm3, scale):
    """FP64 closed form of one dosel: dose = scale edep / (rho V) in Gy, and
    the Ma et al. relative uncertainty sqrt((N sum e^2 - (sum e)^2) / ((N - 1) (sum e)^2))."""
    mass_kg = rho_g_cm3 * 1e-3 * dosel_mm3 * 1e-3  # g/cm3 * mm3 -> kg
    dose = scale * edep * MEV_J / mass_kg
    unc = math.sqrt((hit * edep2 - edep * edep) / ((hit - 1) * edep * edep)) if hit > 1 and edep else 1.0
    return dose, unc


def test_voxel_dose_matches_oracle_and_closed_form():
    """cfg1's 100^3 water phantom with its bone cylinder: scale 1 and 2.5,
    the water reference, and a minimum density between water (1.0) and bone
    (1.92 g/cm3) that zeroes every water dosel."""
    sc = cfg1_scene(n_side=100, detector=64)
    eng = sc.make_engine()
    sc.run_engine(400_000, 515, eng)
    cases = [(1.0, False, 0.0), (2.5, False, 0.0), (1.0, True, 0.0), (1.0, False, F(1.5 * G_CM3))]
    edep, edep2, hit = _check_against_oracle(sc, eng, cases)
    dose, unc = eng.compute_dose(1.0, False, 0.0)
    rho = {0: 1.0, 1: 1.920}  # Water, RibBone (data/materials.txt)
    labels = sc.navigators[0][2]["labels"]
    flat = np.flatnonzero(hit.reshape(-1) > 1)
    checked = {0: 0, 1: 0}
    for i in flat[:: max(1, flat.size // 400)]:
        lab = int(labels.reshape(-1)[i])
        want_d, want_u = _closed_form(edep.reshape(-1)[i], edep2.reshape(-1)[i], int(hit.reshape(-1)[i]), rho[lab],
                                      1.0, 1.0)
        assert abs(dose.reshape(-1)[i] / want_d - 1.0) < 1e-5, (i, lab)
        assert abs(unc.reshape(-1)[i] - want_u) < 1e-5 * max(want_u, 1e-3), (i, lab)
        checked[lab] += 1
    assert checked[0] > 50 and checked[1] > 50
    # dosels with fewer than two hits carry uncertainty 1
    few = hit <= 1
    assert np.all(unc[few] == 1.0)
    # the minimum density keeps bone and zeroes water
    d_min, _ = eng.compute_dose(1.0, False, F(1.5 * G_CM3))
    assert not d_min[labels == 0].any() and d_min[(labels == 1) & (edep > 0)].all()
    # the water reference: every dosel as water
    d_w, _ = eng.compute_dose(1.0, True, 0.0)
    bone = (labels == 1) & (edep > 0)
    np.testing.assert_allclose(d_w[bone] / dose[bone], 1.920, rtol=1e-5)
    eng.close()


def test_mesh_dose_matches_oracle():
    """A soft-tissue icosphere mesh with a 1 mm dose map over its octree box:
    one material, so dose = scale edep / (rho V) everywhere."""
    sc = Scene()
    sc.add_mesh(authoring.icosphere(40.0, 3), "SoftTissus_15", rotation=(0.2, 0.1, 0.3), depth=4,
                dosimetry={"dosel_size": (1.0, 1.0, 1.0)})
    sc.add_ct_flat("GOS", (1, 1), (64, 64, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 3.0, [0.07], mono=True)
    sc.build()
    eng = sc.make_engine()
    sc.run_engine(300_000, 616, eng)
    edep, _, _ = _check_against_oracle(sc, eng, [(1.0, False, 0.0), (3.0, True, 0.0)])
    dose, _ = eng.compute_dose(1.0, False, 0.0)
    rho = float(sc.arrays["density"][0]) / G_CM3
    nz = edep > 0
    assert nz.sum() > 1000
    np.testing.assert_allclose(dose[nz], edep[nz] * MEV_J / (rho * 1e-6), rtol=1e-5)
    eng.close()


def test_full_grid_dose_matches_oracle():
    """The voxel-CT body phantom's whole 512x512x256 dose grid (67,108,864
    dosels, five materials) after 10^7 primaries."""
    body = voxel_ct_scene(os.path.join(REPO, "ggems_amd", "data", "_gen", "voxel_ct"))
    eng = body.make_engine()
    body.run_engine(10_000_000, 4711, eng)
    _check_against_oracle(body, eng, [(1.0, False, 0.0), (1.0, True, F(0.5 * G_CM3))])
    eng.close()


def test_edep_only_dosimetry_tallies_edep_and_refuses_the_rest():
    """Dosimetry with edep squared and hits off (GGEMSDosimetryCalculator's
    defaults for a script that activates neither) tallies the same edep as
    a run with both on, and the readback refuses the inactive tallies and
    the uncertainty that needs them."""
    sc = cfg1_scene(n_side=60, detector=64)
    full = sc.run_engine(200_000, 91)
    eng = sc.make_engine()
    k = sc.engine_solids["voxel"][0]
    d = sc.dose_grid(sc.navigators[0][2])
    pay = sc.navigators[0][2]
    eng.set_dosimetry(k, d["size"], d["n"], pay["bmin"], pay["bmax"], edep_squared=False, hit=False)
    part = sc.run_engine(200_000, 91, eng)
    assert part["edep2"] is None and part["hit"] is None
    assert part["stats"] == full["stats"]
    # the same deposits, summed by FP64 atomics in another order
    np.testing.assert_allclose(part["edep"], full["edep"], rtol=1e-12, atol=1e-18)
    assert part["edep"].sum() > 0
    lib = library()
    nx, ny, nz = eng.n_dosels
    buf = np.zeros(nx * ny * nz, np.float64)
    ibuf = np.zeros(nx * ny * nz, np.int32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert lib.ggh_get_dose_tallies(eng.h, None, p(buf), None, None) == -2
    assert lib.ggh_get_dose_tallies(eng.h, None, None, p(ibuf), None) == -2
    fbuf = np.zeros(nx * ny * nz, np.float32)
    assert lib.ggh_compute_dose(eng.h, C.c_float(1.0), 0, C.c_float(0.0), p(fbuf), p(fbuf.copy())) == -2
    assert lib.ggh_compute_dose(eng.h, C.c_float(1.0), 0, C.c_float(0.0), p(fbuf), None) == 0
    assert fbuf.sum() > 0
    eng.close()
