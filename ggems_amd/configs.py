"""The standard scenes, written as GGEMS scripts against ggems_amd's GGEMS API.

cfg1       100^3 voxels of 1 mm water with a centred 20 mm-radius cortical
           bone (RibBone) cylinder along z; 60 keV point cone beam at
           SID 600 mm; flat 256x256 detector of 1 mm GOS pixels at SDD 1000 mm.
voxel_ct   the 512x512x256 body of 0.8 mm voxels (air, lung, adipose,
           soft tissue, cortical bone; 40 seeded ellipsoids) with the cfg1
           source/detector geometry, a 120 kVp Kramers spectrum with 2.5 mm Al
           filtration and FP64 dose + uncertainty tallies.

Physics in both: Compton, photoelectric, Rayleigh; cross-section tables of
220 bins over 1 keV - 1 MeV; photon range cut 0.1 mm; detector threshold
10 keV (as examples/2_CT_Scanner/ct_scanner.py of GGEMS).
"""

import math
import os

from . import authoring
from .ggems import (GGEMS, GGEMSCTSystem, GGEMSDosimetryCalculator, GGEMSMaterialsDatabaseManager,
                    GGEMSOpenCLManager, GGEMSProcessesManager, GGEMSRangeCutsManager, GGEMSVerbosity,
                    GGEMSVoxelizedPhantom, GGEMSXRaySource, clean_safely)

HERE = os.path.dirname(os.path.abspath(__file__))
MATERIALS_TXT = os.path.join(HERE, "data", "_gen", "materials.txt")

SID, SDD, DETECTOR = 600.0, 1000.0, 256


def _common_physics():
    pm = GGEMSProcessesManager()
    pm.add_process("Compton", "gamma", "all")
    pm.add_process("Photoelectric", "gamma", "all")
    pm.add_process("Rayleigh", "gamma", "all")
    pm.set_cross_section_table_number_of_bins(220)
    pm.set_cross_section_table_energy_min(1.0, "keV")
    pm.set_cross_section_table_energy_max(1.0, "MeV")
    GGEMSRangeCutsManager().set_cut("gamma", 0.1, "mm", "all")


def _detector(workdir, n=DETECTOR, thickness_mm=1.0):
    ct = GGEMSCTSystem("detector")
    ct.set_ct_type("flat")
    ct.set_number_of_modules(1, 1)
    ct.set_number_of_detection_elements(n, n, 1)
    ct.set_size_of_detection_elements(1.0, 1.0, thickness_mm, "mm")
    ct.set_material("GOS")
    ct.set_source_detector_distance(SDD, "mm")
    ct.set_source_isocenter_distance(SID, "mm")
    ct.set_rotation(0.0, 0.0, 0.0, "deg")
    ct.set_threshold(10.0, "keV")
    ct.save(os.path.join(workdir, "projection"))
    ct.store_scatter(True)
    return ct


def _source(n_particles, detector=DETECTOR):
    src = GGEMSXRaySource("point_source")
    src.set_source_particle_type("gamma")
    src.set_number_of_particles(n_particles)
    src.set_position(-SID, 0.0, 0.0, "mm")
    src.set_rotation(0.0, 0.0, 0.0, "deg")
    # cone half-angle reaching the detector corners
    src.set_beam_aperture(math.degrees(math.atan(detector * 0.5 * math.sqrt(2.0) / SDD)), "deg")
    src.set_focal_spot_size(0.0, 0.0, 0.0, "mm")
    return src


def setup(workdir):
    os.makedirs(workdir, exist_ok=True)
    clean_safely()
    GGEMSVerbosity(0)
    GGEMSOpenCLManager().set_device_to_activate("gpu")
    GGEMSMaterialsDatabaseManager().set_materials(MATERIALS_TXT)


def cfg1(workdir, n_particles=1_000_000, n_side=100, dosimetry=False):
    """Builds the cfg1 scene; returns the objects that must stay alive."""
    setup(workdir)
    lab = authoring.cylinder_phantom(n_side, 20.0, 1.0)
    mhd = authoring.write_mhd(os.path.join(workdir, "phantom.mhd"), lab, (1.0, 1.0, 1.0))
    rng = authoring.write_range(os.path.join(workdir, "range_phantom.txt"), ["Water", "RibBone"])
    phantom = GGEMSVoxelizedPhantom("phantom")
    phantom.set_phantom(mhd, rng)
    phantom.set_rotation(0.0, 0.0, 0.0, "deg")
    phantom.set_position(0.0, 0.0, 0.0, "mm")
    keep = [phantom]
    if dosimetry:
        dose = GGEMSDosimetryCalculator()
        dose.attach_to_navigator("phantom")
        dose.set_output_basename(os.path.join(workdir, "dosimetry"))
        dose.uncertainty(True)
        dose.edep(True)
        dose.hit(True)
        dose.edep_squared(True)
        keep.append(dose)
    keep.append(_detector(workdir))
    _common_physics()
    src = _source(n_particles)
    src.set_monoenergy(60.0, "keV")
    keep.append(src)
    return keep


def voxel_ct_inputs(workdir, seed=20240, shape=(256, 512, 512), voxel=0.8):
    """Writes the body phantom, its range file and the 120 kVp spectrum (cached)."""
    os.makedirs(workdir, exist_ok=True)
    mhd = os.path.join(workdir, "body_%d_%dx%dx%d.mhd" % (seed, shape[2], shape[1], shape[0]))
    rng = os.path.join(workdir, "range_body.txt")
    spec = os.path.join(workdir, "spectrum_120kVp_2.5mmAl.dat")
    if not os.path.exists(mhd):
        lab = authoring.ellipsoid_body_phantom(seed, shape=shape, voxel=voxel)
        authoring.write_mhd(mhd, lab, (voxel, voxel, voxel))
    authoring.write_range(rng, authoring.BODY_MATERIALS)
    if not os.path.exists(spec):
        GGEMSMaterialsDatabaseManager().set_materials(MATERIALS_TXT)
        e, w = authoring.kramers_spectrum(120.0, 20.0, 2.5, authoring.aluminium_mu())
        authoring.write_spectrum(spec, e, w)
    return mhd, rng, spec


def voxel_ct(workdir, n_particles=100_000_000, seed=20240, dosimetry=True, shape=(256, 512, 512), voxel=0.8):
    mhd, rng, spec = voxel_ct_inputs(workdir, seed, shape, voxel)
    setup(workdir)
    phantom = GGEMSVoxelizedPhantom("phantom")
    phantom.set_phantom(mhd, rng)
    phantom.set_rotation(0.0, 0.0, 0.0, "deg")
    phantom.set_position(0.0, 0.0, 0.0, "mm")
    keep = [phantom]
    if dosimetry:
        dose = GGEMSDosimetryCalculator()
        dose.attach_to_navigator("phantom")
        dose.set_output_basename(os.path.join(workdir, "dosimetry"))
        dose.water_reference(False)
        dose.minimum_density(0.1, "g/cm3")
        dose.uncertainty(True)
        dose.edep(True)
        dose.hit(True)
        dose.edep_squared(True)
        keep.append(dose)
    keep.append(_detector(workdir))
    _common_physics()
    src = _source(n_particles)
    src.set_polyenergy(spec)
    keep.append(src)
    return keep


def run(seed=777):
    g = GGEMS()
    g.initialize(seed)
    g.run()
    return g
