"""The standard scenes, written as GGEMS scripts against ggems_amd's GGEMS API.

cfg1       100^3 voxels of 1 mm water with a centred 20 mm-radius cortical
           bone (RibBone) cylinder along z; 60 keV point cone beam at
           SID 600 mm; flat 256x256 detector of 1 mm GOS pixels at SDD 1000 mm.
voxel_ct   the 512x512x256 body of 0.8 mm voxels (air, lung, adipose,
           soft tissue, cortical bone; 40 seeded ellipsoids) with the cfg1
           source/detector geometry, a 120 kVp Kramers spectrum with 2.5 mm Al
           filtration and FP64 dose + uncertainty tallies.
acquisition  a CT acquisition in one launch: n_angles gantry positions (1
           degree apart by default) around a seeded ellipsoid phantom of
           side^3 1 mm voxels, per-projection photon budgets drawn uniformly
           in a range, a flat 512x384 detector of 0.78 mm pixels at SDD
           1000 mm (SID 600 mm) and the 120 kVp spectrum.
mesh_ct    two meshed phantoms (an icosphere of radius 50 mm and 81,920
           triangles in soft tissue, one of radius 15 mm in cortical bone) in
           a 300 mm water box, an 80 keV point cone beam at SID 600 mm and a
           curved CT detector of 888 columns x 64 rows on a 1040 mm arc.

Physics in all: Compton, photoelectric, Rayleigh; cross-section tables of
220 bins over 1 keV - 1 MeV; photon range cut 0.1 mm; detector threshold
10 keV (as examples/2_CT_Scanner/ct_scanner.py of GGEMS).
"""

import math
import os

from . import authoring
from .ggems import (GGEMS, GGEMSCTSystem, GGEMSDosimetryCalculator, GGEMSMaterialsDatabaseManager,
                    GGEMSMeshedPhantom, GGEMSOpenCLManager, GGEMSProcessesManager, GGEMSRangeCutsManager,
                    GGEMSVerbosity, GGEMSVoxelizedPhantom, GGEMSXRaySource, clean_safely)

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


MESH_SID, MESH_SDD = 600.0, 1040.0
MESH_MODULES, MESH_ELEMENTS = (1, 37), (64, 24, 1)  # 37 x 24 = 888 columns along the arc, 64 rows


def mesh_ct_inputs(workdir, outer_subdivisions=6, inner_subdivisions=4, box_mm=300.0, box_voxel=2.0):
    """Writes the two icosphere STL files and the water-box phantom (cached)."""
    os.makedirs(workdir, exist_ok=True)
    outer = os.path.join(workdir, "icosphere_r50_s%d.stl" % outer_subdivisions)
    inner = os.path.join(workdir, "icosphere_r15_s%d.stl" % inner_subdivisions)
    n = int(round(box_mm / box_voxel))
    box = os.path.join(workdir, "water_box_%d_%g.mhd" % (n, box_voxel))
    rng = os.path.join(workdir, "range_water.txt")
    if not os.path.exists(outer):
        authoring.write_stl(outer, authoring.icosphere(50.0, outer_subdivisions))
    if not os.path.exists(inner):
        authoring.write_stl(inner, authoring.icosphere(15.0, inner_subdivisions))
    if not os.path.exists(box):
        import numpy as np
        authoring.write_mhd(box, np.zeros((n, n, n), np.uint8), (box_voxel, box_voxel, box_voxel))
    authoring.write_range(rng, ["Water"])
    return outer, inner, box, rng


def mesh_ct(workdir, n_particles=100_000_000, water_box=True, dosimetry=False, outer_subdivisions=6,
            inner_subdivisions=4, box_voxel=2.0, aperture_deg=5.0):
    """The meshed-phantom scene.  The meshes are registered before the water
    box.  GGEMS gives a particle to the first solid whose OBB holds it and
    tracks it there until it leaves that OBB, so with water_box=True every
    photon is claimed by the enclosing box and crosses the meshes as water;
    water_box=False leaves the meshes in vacuum, where they are navigated."""
    outer_stl, inner_stl, box_mhd, box_rng = mesh_ct_inputs(workdir, outer_subdivisions, inner_subdivisions,
                                                            box_voxel=box_voxel)
    setup(workdir)
    keep = []
    for name, stl, material in (("outer_sphere", outer_stl, "SoftTissus_15"), ("inner_sphere", inner_stl, "RibBone")):
        m = GGEMSMeshedPhantom(name)
        m.set_phantom(stl)
        m.set_rotation(0.0, 0.0, 0.0, "deg")
        m.set_position(0.0, 0.0, 0.0, "mm")
        m.set_mesh_octree_depth(4)
        m.set_material(material)
        keep.append(m)
    if water_box:
        box = GGEMSVoxelizedPhantom("water_box")
        box.set_phantom(box_mhd, box_rng)
        box.set_rotation(0.0, 0.0, 0.0, "deg")
        box.set_position(0.0, 0.0, 0.0, "mm")
        keep.append(box)
    if dosimetry:
        dose = GGEMSDosimetryCalculator()
        dose.attach_to_navigator("outer_sphere")
        dose.set_output_basename(os.path.join(workdir, "dosimetry"))
        dose.set_dosel_size(1.0, 1.0, 1.0, "mm")
        dose.water_reference(False)
        dose.minimum_density(0.1, "g/cm3")
        dose.uncertainty(True)
        dose.edep(True)
        dose.hit(True)
        dose.edep_squared(True)
        keep.append(dose)
    ct = GGEMSCTSystem("detector")
    ct.set_ct_type("curved")
    ct.set_number_of_modules(*MESH_MODULES)
    ct.set_number_of_detection_elements(*MESH_ELEMENTS)
    ct.set_size_of_detection_elements(1.0, 1.0, 1.0, "mm")
    ct.set_material("GOS")
    ct.set_source_detector_distance(MESH_SDD, "mm")
    ct.set_source_isocenter_distance(MESH_SID, "mm")
    ct.set_rotation(0.0, 0.0, 0.0, "deg")
    ct.set_threshold(10.0, "keV")
    ct.save(os.path.join(workdir, "projection"))
    ct.store_scatter(True)
    keep.append(ct)
    _common_physics()
    src = GGEMSXRaySource("point_source")
    src.set_source_particle_type("gamma")
    src.set_number_of_particles(n_particles)
    src.set_position(-MESH_SID, 0.0, 0.0, "mm")
    src.set_rotation(0.0, 0.0, 0.0, "deg")
    src.set_beam_aperture(aperture_deg, "deg")
    src.set_focal_spot_size(0.0, 0.0, 0.0, "mm")
    src.set_monoenergy(80.0, "keV")
    keep.append(src)
    return keep


ACQ_DETECTOR = (384, 512)  # rows (axial, local x) x columns (transaxial, local y)
ACQ_PIXEL = 0.78


def acquisition_inputs(workdir, side=128, seed=1):
    """The ellipsoid phantom of side^3 1 mm voxels (body and ellipsoids scaled
    to the side), its range file and the 120 kVp spectrum (cached)."""
    os.makedirs(workdir, exist_ok=True)
    mhd = os.path.join(workdir, "ellipsoids_%d_%d.mhd" % (seed, side))
    rng = os.path.join(workdir, "range_body.txt")
    spec = os.path.join(workdir, "spectrum_120kVp_2.5mmAl.dat")
    if not os.path.exists(mhd):
        lab = authoring.ellipsoid_body_phantom(seed, shape=(side, side, side), voxel=1.0, n_ellipsoids=40,
                                               body_axes=(0.45 * side, 0.35 * side),
                                               semi_axis_range=(side / 50.0, side / 4.0))
        authoring.write_mhd(mhd, lab, (1.0, 1.0, 1.0))
    authoring.write_range(rng, authoring.BODY_MATERIALS)
    if not os.path.exists(spec):
        GGEMSMaterialsDatabaseManager().set_materials(MATERIALS_TXT)
        e, w = authoring.kramers_spectrum(120.0, 20.0, 2.5, authoring.aluminium_mu())
        authoring.write_spectrum(spec, e, w)
    return mhd, rng, spec


def acquisition_plan(n_angles=360, budget_range=(1_000_000, 30_000_000), seed=5, step_deg=1.0):
    """Gantry rotations about z (degrees) and seeded per-projection budgets."""
    import numpy as np
    r = np.zeros((n_angles, 3), np.float32)
    r[:, 2] = np.arange(n_angles, dtype=np.float32) * np.float32(step_deg)
    budgets = np.random.default_rng(seed).integers(budget_range[0], budget_range[1] + 1, n_angles).astype(np.uint64)
    return r, budgets


def acquisition(workdir, side=128, seed=1, sides=None, seeds=None):
    """The acquisition scene (phantoms, flat panel, source); transport with
    GGEMS().run_acquisition(rotations, budgets) after initialize.  With
    `sides` (and `seeds`, one per side), one ellipsoid phantom per side, all
    at the isocentre, for an acquisition whose projections each see one of
    them (run_acquisition's `phantoms`, in this order)."""
    if sides is None:
        sides, seeds = (side,), (seed,)
    keep = []
    for k, (sd, sdseed) in enumerate(zip(sides, seeds)):
        mhd, rng, spec = acquisition_inputs(workdir, sd, sdseed)
        if k == 0:
            setup(workdir)
        phantom = GGEMSVoxelizedPhantom("phantom" if len(sides) == 1 else "phantom_%d" % k)
        phantom.set_phantom(mhd, rng)
        phantom.set_rotation(0.0, 0.0, 0.0, "deg")
        phantom.set_position(0.0, 0.0, 0.0, "mm")
        keep.append(phantom)
    ct = GGEMSCTSystem("detector")
    ct.set_ct_type("flat")
    ct.set_number_of_modules(1, 1)
    ct.set_number_of_detection_elements(ACQ_DETECTOR[0], ACQ_DETECTOR[1], 1)
    ct.set_size_of_detection_elements(ACQ_PIXEL, ACQ_PIXEL, 1.0, "mm")
    ct.set_material("GOS")
    ct.set_source_detector_distance(SDD, "mm")
    ct.set_source_isocenter_distance(SID, "mm")
    ct.set_rotation(0.0, 0.0, 0.0, "deg")
    ct.set_threshold(10.0, "keV")
    ct.save(os.path.join(workdir, "projections"))
    ct.store_scatter(True)
    _common_physics()
    src = GGEMSXRaySource("point_source")
    src.set_source_particle_type("gamma")
    src.set_number_of_particles(1)
    src.set_position(-SID, 0.0, 0.0, "mm")
    src.set_rotation(0.0, 0.0, 0.0, "deg")
    half_diag = 0.5 * ACQ_PIXEL * math.hypot(*ACQ_DETECTOR)
    src.set_beam_aperture(math.degrees(math.atan(half_diag / SDD)), "deg")
    src.set_focal_spot_size(0.0, 0.0, 0.0, "mm")
    src.set_polyenergy(spec)
    return keep + [ct, src]


def scatter_scene(workdir, phantom_seed=None, angle_deg=0.0, shape=(64, 96, 96), voxel=2.0, detector=DETECTOR):
    """One view of the scatter-surrogate data: a seeded body phantom of
    shape voxels (2 mm by default, body and ellipsoids scaled to the volume)
    turned by angle_deg about z, the 120 kVp spectrum from a point source at
    SID 600 mm and a detector of detector^2 1 mm GOS pixels at SDD 1000 mm
    that stores the scatter image.  phantom_seed None: the open field."""
    os.makedirs(workdir, exist_ok=True)
    spec = os.path.join(workdir, "spectrum_120kVp_2.5mmAl.dat")
    if not os.path.exists(spec):
        GGEMSMaterialsDatabaseManager().set_materials(MATERIALS_TXT)
        e, w = authoring.kramers_spectrum(120.0, 20.0, 2.5, authoring.aluminium_mu())
        authoring.write_spectrum(spec, e, w)
    keep = []
    if phantom_seed is not None:
        nz, ny, nx = shape
        mhd = os.path.join(workdir, "body_%d_%dx%dx%d.mhd" % (phantom_seed, nx, ny, nz))
        if not os.path.exists(mhd):
            lab = authoring.ellipsoid_body_phantom(phantom_seed, shape=shape, voxel=voxel,
                                                   body_axes=(0.45 * nx * voxel, 0.35 * ny * voxel),
                                                   semi_axis_range=(nx * voxel / 50.0, nx * voxel / 5.0))
            authoring.write_mhd(mhd, lab, (voxel, voxel, voxel))
        rng = authoring.write_range(os.path.join(workdir, "range_body.txt"), authoring.BODY_MATERIALS)
    setup(workdir)
    if phantom_seed is not None:
        phantom = GGEMSVoxelizedPhantom("phantom")
        phantom.set_phantom(mhd, rng)
        phantom.set_rotation(0.0, 0.0, angle_deg, "deg")
        phantom.set_position(0.0, 0.0, 0.0, "mm")
        keep.append(phantom)
    keep.append(_detector(workdir, n=detector))
    _common_physics()
    src = _source(1, detector=detector)
    src.set_polyenergy(spec)
    keep.append(src)
    return keep


def scatter_scenes(workdir, phantom_seeds, shape=(64, 96, 96), voxel=2.0, detector=DETECTOR):
    """The scatter-surrogate scene with several seeded body phantoms, all at
    the isocentre, for one acquisition launch in which each projection sees
    one of them (GGEMS.run_acquisition with `phantoms`): the phantoms of
    scatter_scene, registered in the order of phantom_seeds, then the
    detector and the source.  Returns the objects that must stay alive."""
    os.makedirs(workdir, exist_ok=True)
    spec = os.path.join(workdir, "spectrum_120kVp_2.5mmAl.dat")
    if not os.path.exists(spec):
        GGEMSMaterialsDatabaseManager().set_materials(MATERIALS_TXT)
        e, w = authoring.kramers_spectrum(120.0, 20.0, 2.5, authoring.aluminium_mu())
        authoring.write_spectrum(spec, e, w)
    nz, ny, nx = shape
    mhds = []
    for seed in phantom_seeds:
        mhd = os.path.join(workdir, "body_%d_%dx%dx%d.mhd" % (seed, nx, ny, nz))
        if not os.path.exists(mhd):
            lab = authoring.ellipsoid_body_phantom(seed, shape=shape, voxel=voxel,
                                                   body_axes=(0.45 * nx * voxel, 0.35 * ny * voxel),
                                                   semi_axis_range=(nx * voxel / 50.0, nx * voxel / 5.0))
            authoring.write_mhd(mhd, lab, (voxel, voxel, voxel))
        mhds.append(mhd)
    rng = authoring.write_range(os.path.join(workdir, "range_body.txt"), authoring.BODY_MATERIALS)
    setup(workdir)
    keep = []
    for k, mhd in enumerate(mhds):
        phantom = GGEMSVoxelizedPhantom("phantom_%d" % k)
        phantom.set_phantom(mhd, rng)
        phantom.set_rotation(0.0, 0.0, 0.0, "deg")
        phantom.set_position(0.0, 0.0, 0.0, "mm")
        keep.append(phantom)
    keep.append(_detector(workdir, n=detector))
    _common_physics()
    src = _source(1, detector=detector)
    src.set_polyenergy(spec)
    keep.append(src)
    return keep


def run(seed=777):
    g = GGEMS()
    g.initialize(seed)
    g.run()
    return g
