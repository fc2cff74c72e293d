"""Python binding of libggems.so with GGEMS's own class and method names.

A GGEMS script written against the reference's ``ggems`` package runs
against this module with its import line changed (``from ggems_amd import *``):
the classes, methods, argument order and unit strings are the same
(python_module/*.py of GGEMS).  Unlike GGEMS, a failed call raises
GGEMSError in Python instead of ending the process.

Methods that only drive GGEMS's OpenGL display are accepted and ignored.
"""

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(os.environ.get("GGEMS_AMD_LIB_DIR") or os.path.join(HERE, "lib"), "libggems.so")


class GGEMSError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise GGEMSError("libggems.so not built: %s (run __graft_entry__.build())" % HOST_LIB)
        _lib = C.CDLL(HOST_LIB)
        _lib.get_last_error_ggems.restype = C.c_char_p
    return _lib


def _b(s):
    return s.encode() if isinstance(s, str) else s


def _call(name, restype, argtypes, *args):
    f = getattr(lib(), name)
    f.restype = restype
    f.argtypes = argtypes
    out = f(*args)
    err = lib().get_last_error_ggems()
    if err:
        raise GGEMSError(err.decode(errors="replace"))
    return out


VP, F, S, B, I, U32, SZ = C.c_void_p, C.c_float, C.c_char_p, C.c_bool, C.c_int, C.c_uint32, C.c_size_t


def clean_safely():
    """Deletes every navigator and source so that a new scene can be set up."""
    _call("clean_safely_ggems", None, [])


def GGEMSVerbosity(level):
    _call("set_ggems_verbose", None, [I], int(level))


class GGEMSOpenCLManager:
    """HIP device selection with GGEMS's OpenCL-manager interface."""

    def __init__(self):
        self.obj = _call("get_instance_ggems_opencl_manager", VP, [])

    def print_infos(self):
        _call("print_infos_opencl_manager", None, [VP], self.obj)

    def set_device_index(self, device_id):
        _call("set_device_index_ggems_opencl_manager", None, [VP, SZ], self.obj, device_id)

    def set_device_to_activate(self, device_type, device_vendor=""):
        _call("set_device_to_activate_opencl_manager", None, [VP, S, S], self.obj, _b(device_type), _b(device_vendor))

    def set_device_balancing(self, device_balancing):
        _call("set_device_balancing_opencl_manager", None, [VP, S], self.obj, _b(device_balancing))

    def clean(self):
        _call("clean_opencl_manager", None, [VP], self.obj)


class GGEMSMaterialsDatabaseManager:
    def __init__(self):
        self.obj = _call("get_instance_materials_manager", VP, [])

    def set_materials(self, filename):
        _call("set_materials_ggems_materials_manager", None, [VP, S], self.obj, _b(filename))

    def print_available_chemical_elements(self):
        _call("print_available_chemical_elements_ggems_materials_manager", None, [VP], self.obj)

    def print_available_materials(self):
        _call("print_available_materials_ggems_materials_manager", None, [VP], self.obj)


class GGEMSMaterials:
    def __init__(self):
        self.obj = _call("create_ggems_materials", VP, [])

    def add_material(self, material):
        _call("add_material_ggems_materials", None, [VP, S], self.obj, _b(material))

    def initialize(self):
        _call("initialize_ggems_materials", None, [VP], self.obj)

    def print_material_properties(self):
        _call("print_material_properties_ggems_materials", None, [VP], self.obj)

    def get_density(self, material_name):
        return _call("get_density_ggems_materials", F, [VP, S], self.obj, _b(material_name))

    def get_energy_cut(self, material_name, particle_name, value, unit):
        return _call("get_energy_cut_ggems_materials", F, [VP, S, S, F, S], self.obj, _b(material_name),
                     _b(particle_name), value, _b(unit))

    def get_atomic_number_density(self, material_name):
        return _call("get_atomic_number_density_ggems_materials", F, [VP, S], self.obj, _b(material_name))

    def get_photon_cut(self, material_name):
        """photon energy cut in MeV once initialised (no GGEMS counterpart)"""
        return _call("get_photon_cut_ggems_materials", F, [VP, S], self.obj, _b(material_name))

    def clean(self):
        if self.obj:
            _call("clean_ggems_materials", None, [VP], self.obj)
            self.obj = None


class GGEMSCrossSections:
    def __init__(self, materials):
        self.materials = materials
        self.obj = _call("create_ggems_cross_sections", VP, [VP], materials.obj)

    def add_process(self, process_name, particle_name, is_secondary=False):
        _call("add_process_ggems_cross_sections", None, [VP, S, S, B], self.obj, _b(process_name),
              _b(particle_name), is_secondary)

    def initialize(self):
        _call("initialize_ggems_cross_sections", None, [VP], self.obj)

    def get_cs(self, process_name, material_name, energy, unit):
        return _call("get_cs_ggems_cross_sections", F, [VP, S, S, F, S], self.obj, _b(process_name),
                     _b(material_name), energy, _b(unit))

    def get_table(self, process_name, material_name, nbins):
        """raw mm^-1 table row of a material (no GGEMS counterpart)"""
        import numpy as np
        out = np.zeros(nbins, np.float32)
        _call("get_cross_section_table_ggems_cross_sections", None, [VP, S, S, VP, SZ], self.obj, _b(process_name),
              _b(material_name), out.ctypes.data_as(VP), nbins)
        return out

    def get_energy_bins(self, nbins):
        import numpy as np
        out = np.zeros(nbins, np.float32)
        _call("get_energy_bins_ggems_cross_sections", None, [VP, VP, SZ], self.obj, out.ctypes.data_as(VP), nbins)
        return out

    def clean(self):
        if self.obj:
            _call("clean_ggems_cross_sections", None, [VP], self.obj)
            self.obj = None


class GGEMSAttenuations:
    def __init__(self, materials, cross_sections):
        self.obj = _call("create_ggems_attenuations", VP, [VP, VP], materials.obj, cross_sections.obj)

    def initialize(self):
        _call("initialize_ggems_attenuations", None, [VP], self.obj)

    def get_mu(self, material_name, energy, unit):
        return _call("get_mu_ggems_attenuations", F, [VP, S, F, S], self.obj, _b(material_name), energy, _b(unit))

    def get_mu_en(self, material_name, energy, unit):
        return _call("get_mu_en_ggems_attenuations", F, [VP, S, F, S], self.obj, _b(material_name), energy, _b(unit))

    def clean(self):
        if self.obj:
            _call("clean_ggems_attenuations", None, [VP], self.obj)
            self.obj = None


class GGEMSProcessesManager:
    def __init__(self):
        self.obj = _call("get_instance_processes_manager", VP, [])

    def set_cross_section_table_number_of_bins(self, number_of_bins):
        _call("set_cross_section_table_number_of_bins_processes_manager", None, [VP, SZ], self.obj, number_of_bins)

    def set_cross_section_table_energy_min(self, energy, unit):
        _call("set_cross_section_table_minimum_energy_processes_manager", None, [VP, F, S], self.obj, energy, _b(unit))

    def set_cross_section_table_energy_max(self, energy, unit):
        _call("set_cross_section_table_maximum_energy_processes_manager", None, [VP, F, S], self.obj, energy, _b(unit))

    def print_available_processes(self):
        _call("print_available_processes_manager", None, [VP], self.obj)

    def print_infos(self):
        _call("print_infos_processes_manager", None, [VP], self.obj)

    def add_process(self, process_name, particle_name, phantom_name="all", is_secondary=False):
        _call("add_process_processes_manager", None, [VP, S, S, S], self.obj, _b(process_name), _b(particle_name),
              _b(phantom_name))

    def print_tables(self, flag):
        _call("print_tables_processes_manager", None, [VP, B], self.obj, flag)


class GGEMSRangeCutsManager:
    def __init__(self):
        self.obj = _call("get_instance_range_cuts_manager", VP, [])

    def set_cut(self, particle, value, unit, phantom="all"):
        _call("set_cut_range_cuts_manager", None, [VP, S, S, F, S], self.obj, _b(phantom), _b(particle), value,
              _b(unit))


class GGEMSSourceManager:
    def __init__(self):
        self.obj = _call("get_instance_ggems_source_manager", VP, [])

    def initialize(self, seed):
        _call("initialize_source_manager", None, [VP, U32], self.obj, seed)

    def print_infos(self):
        _call("print_infos_source_manager", None, [VP], self.obj)

    def clean(self):
        _call("clean_source_manager", None, [VP], self.obj)


class GGEMSXRaySource:
    def __init__(self, source_name):
        self.obj = _call("create_ggems_xray_source", VP, [S], _b(source_name))

    def set_position(self, x, y, z, unit):
        _call("set_position_ggems_xray_source", None, [VP, F, F, F, S], self.obj, x, y, z, _b(unit))

    def set_rotation(self, rx, ry, rz, unit):
        _call("set_rotation_ggems_xray_source", None, [VP, F, F, F, S], self.obj, rx, ry, rz, _b(unit))

    def set_beam_aperture(self, beam_aperture, unit):
        _call("set_beam_aperture_ggems_xray_source", None, [VP, F, S], self.obj, beam_aperture, _b(unit))

    def set_focal_spot_size(self, width, height, depth, unit):
        _call("set_focal_spot_size_ggems_xray_source", None, [VP, F, F, F, S], self.obj, width, height, depth, _b(unit))

    def set_number_of_particles(self, number_of_particles):
        _call("set_number_of_particles_xray_source", None, [VP, SZ], self.obj, int(number_of_particles))

    def set_source_particle_type(self, particle_type):
        _call("set_source_particle_type_ggems_xray_source", None, [VP, S], self.obj, _b(particle_type))

    def set_monoenergy(self, e, unit):
        _call("set_monoenergy_ggems_xray_source", None, [VP, F, S], self.obj, e, _b(unit))

    def set_polyenergy(self, file):
        _call("set_polyenergy_ggems_xray_source", None, [VP, S], self.obj, _b(file))

    def set_energy_peak(self, energy, unit, intensity):
        _call("set_energy_peak_ggems_xray_source", None, [VP, F, S, F], self.obj, energy, _b(unit), intensity)


class GGEMSVoxelizedPhantom:
    def __init__(self, voxelized_phantom_name):
        self.obj = _call("create_ggems_voxelized_phantom", VP, [S], _b(voxelized_phantom_name))

    def set_phantom(self, phantom_filename, range_data_filename):
        _call("set_phantom_file_ggems_voxelized_phantom", None, [VP, S, S], self.obj, _b(phantom_filename),
              _b(range_data_filename))

    def set_position(self, pos_x, pos_y, pos_z, unit):
        _call("set_position_ggems_voxelized_phantom", None, [VP, F, F, F, S], self.obj, pos_x, pos_y, pos_z, _b(unit))

    def set_rotation(self, rx, ry, rz, unit):
        _call("set_rotation_ggems_voxelized_phantom", None, [VP, F, F, F, S], self.obj, rx, ry, rz, _b(unit))

    def set_visible(self, flag):
        pass

    def set_material_visible(self, material_name, flag):
        pass

    def set_material_color(self, material_name, red=0, green=0, blue=0, color_name=""):
        pass


class GGEMSVolumeCreatorManager:
    """Label volume painted on the device, written as MHD/raw plus a range file"""

    def __init__(self):
        self.obj = _call("get_instance_volume_creator_manager", VP, [])

    def set_dimensions(self, width, height, depth):
        _call("set_volume_dimension_volume_creator_manager", None, [VP, SZ, SZ, SZ], self.obj, width, height, depth)

    def set_element_sizes(self, width, height, depth, unit):
        _call("set_element_sizes_volume_creator_manager", None, [VP, F, F, F, S], self.obj, width, height, depth,
              _b(unit))

    def set_output(self, output):
        _call("set_output_image_filename_volume_creator_manager", None, [VP, S], self.obj, _b(output))

    def set_range_output(self, output):
        _call("set_output_range_to_material_filename_volume_creator_manager", None, [VP, S], self.obj, _b(output))

    def initialize(self):
        _call("initialize_volume_creator_manager", None, [VP], self.obj)

    def write(self):
        _call("write_volume_creator_manager", None, [VP], self.obj)

    def set_material(self, material):
        _call("set_material_volume_creator_manager", None, [VP, S], self.obj, _b(material))

    def set_data_type(self, data_type):
        _call("set_data_type_volume_creator_manager", None, [VP, S], self.obj, _b(data_type))

    def clean(self):
        _call("clean_volume_creator_manager", None, [VP], self.obj)


class GGEMSTube:
    """GGEMSVolume tube painted into the volume creator's label volume"""

    def __init__(self, radius_x, radius_y, height, unit):
        self.obj = _call("create_tube", VP, [F, F, F, S], radius_x, radius_y, height, _b(unit))

    def delete(self):
        _call("delete_tube", None, [VP], self.obj)

    def set_label_value(self, label_value):
        _call("set_label_value_tube", None, [VP, F], self.obj, label_value)

    def set_position(self, pos_x, pos_y, pos_z, unit):
        _call("set_position_tube", None, [VP, F, F, F, S], self.obj, pos_x, pos_y, pos_z, _b(unit))

    def set_material(self, material):
        _call("set_material_tube", None, [VP, S], self.obj, _b(material))

    def initialize(self):
        _call("initialize_tube", None, [VP], self.obj)

    def draw(self):
        _call("draw_tube", None, [VP], self.obj)

class GGEMSBox:
    """GGEMSVolume box painted into the volume creator's label volume"""

    def __init__(self, width, height, depth, unit):
        self.obj = _call("create_box", VP, [F, F, F, S], width, height, depth, _b(unit))

    def delete(self):
        _call("delete_box", None, [VP], self.obj)

    def set_label_value(self, label_value):
        _call("set_label_value_box", None, [VP, F], self.obj, label_value)

    def set_position(self, pos_x, pos_y, pos_z, unit):
        _call("set_position_box", None, [VP, F, F, F, S], self.obj, pos_x, pos_y, pos_z, _b(unit))

    def set_material(self, material):
        _call("set_material_box", None, [VP, S], self.obj, _b(material))

    def initialize(self):
        _call("initialize_box", None, [VP], self.obj)

    def draw(self):
        _call("draw_box", None, [VP], self.obj)

class GGEMSSphere:
    """GGEMSVolume sphere painted into the volume creator's label volume"""

    def __init__(self, radius, unit):
        self.obj = _call("create_sphere", VP, [F, S], radius, _b(unit))

    def delete(self):
        _call("delete_sphere", None, [VP], self.obj)

    def set_label_value(self, label_value):
        _call("set_label_value_sphere", None, [VP, F], self.obj, label_value)

    def set_position(self, pos_x, pos_y, pos_z, unit):
        _call("set_position_sphere", None, [VP, F, F, F, S], self.obj, pos_x, pos_y, pos_z, _b(unit))

    def set_material(self, material):
        _call("set_material_sphere", None, [VP, S], self.obj, _b(material))

    def initialize(self):
        _call("initialize_sphere", None, [VP], self.obj)

    def draw(self):
        _call("draw_sphere", None, [VP], self.obj)


class GGEMSMeshedPhantom:
    def __init__(self, meshed_phantom_name):
        self.obj = _call("create_ggems_meshed_phantom", VP, [S], _b(meshed_phantom_name))

    def set_phantom(self, phantom_filename):
        _call("set_phantom_file_ggems_meshed_phantom", None, [VP, S], self.obj, _b(phantom_filename))

    def set_position(self, pos_x, pos_y, pos_z, unit):
        _call("set_position_ggems_meshed_phantom", None, [VP, F, F, F, S], self.obj, pos_x, pos_y, pos_z, _b(unit))

    def set_rotation(self, rx, ry, rz, unit):
        _call("set_rotation_ggems_meshed_phantom", None, [VP, F, F, F, S], self.obj, rx, ry, rz, _b(unit))

    def set_visible(self, flag):
        pass

    def set_material(self, material_name):
        _call("set_material_name_ggems_meshed_phantom", None, [VP, S], self.obj, _b(material_name))

    def set_mesh_octree_depth(self, depth):
        _call("set_mesh_octree_depth_ggems_meshed_phantom", None, [VP, I], self.obj, int(depth))

    def set_material_color(self, material_name, red=0, green=0, blue=0, color_name=""):
        pass


class GGEMSCTSystem:
    def __init__(self, ct_system_name):
        self.obj = _call("create_ggems_ct_system", VP, [S], _b(ct_system_name))

    def set_number_of_modules(self, module_x, module_y):
        _call("set_number_of_modules_ggems_ct_system", None, [VP, SZ, SZ], self.obj, module_x, module_y)

    def set_ct_type(self, ct_system_type):
        _call("set_ct_system_type_ggems_ct_system", None, [VP, S], self.obj, _b(ct_system_type))

    def set_number_of_detection_elements(self, element_x, element_y, element_z):
        _call("set_number_of_detection_elements_ggems_ct_system", None, [VP, SZ, SZ, SZ], self.obj, element_x,
              element_y, element_z)

    def set_size_of_detection_elements(self, size_x, size_y, size_z, unit):
        _call("set_size_of_detection_elements_ggems_ct_system", None, [VP, F, F, F, S], self.obj, size_x, size_y,
              size_z, _b(unit))

    def set_material(self, material_name):
        _call("set_material_name_ggems_ct_system", None, [VP, S], self.obj, _b(material_name))

    def set_source_detector_distance(self, sdd, unit):
        _call("set_source_detector_distance_ggems_ct_system", None, [VP, F, S], self.obj, sdd, _b(unit))

    def set_source_isocenter_distance(self, sid, unit):
        _call("set_source_isocenter_distance_ggems_ct_system", None, [VP, F, S], self.obj, sid, _b(unit))

    def set_rotation(self, rx, ry, rz, unit):
        _call("set_rotation_ggems_ct_system", None, [VP, F, F, F, S], self.obj, rx, ry, rz, _b(unit))

    def set_global_system_position(self, x, y, z, unit):
        _call("set_global_system_position_ggems_ct_system", None, [VP, F, F, F, S], self.obj, x, y, z, _b(unit))

    def set_threshold(self, threshold, unit):
        _call("set_threshold_ggems_ct_system", None, [VP, F, S], self.obj, threshold, _b(unit))

    def save(self, basename):
        _call("set_save_ggems_ct_system", None, [VP, S], self.obj, _b(basename))

    def store_scatter(self, flag):
        _call("store_scatter_ggems_ct_system", None, [VP, B], self.obj, flag)

    def set_visible(self, flag):
        pass

    def set_material_color(self, material_name, red=0, green=0, blue=0, color_name=""):
        pass


class GGEMSDosimetryCalculator:
    def __init__(self):
        self.obj = _call("create_ggems_dosimetry_calculator", VP, [])

    def set_dosel_size(self, dose_x, dose_y, dose_z, unit):
        _call("set_dosel_size_dosimetry_calculator", None, [VP, F, F, F, S], self.obj, dose_x, dose_y, dose_z, _b(unit))

    def delete(self):
        if self.obj:
            _call("delete_dosimetry_calculator", None, [VP], self.obj)
            self.obj = None

    def set_output_basename(self, output):
        _call("set_dose_output_dosimetry_calculator", None, [VP, S], self.obj, _b(output))

    def photon_tracking(self, activate):
        _call("dose_photon_tracking_dosimetry_calculator", None, [VP, B], self.obj, activate)

    def edep(self, activate):
        _call("dose_edep_dosimetry_calculator", None, [VP, B], self.obj, activate)

    def hit(self, activate):
        _call("dose_hit_dosimetry_calculator", None, [VP, B], self.obj, activate)

    def edep_squared(self, activate):
        _call("dose_edep_squared_dosimetry_calculator", None, [VP, B], self.obj, activate)

    def uncertainty(self, activate):
        _call("dose_uncertainty_dosimetry_calculator", None, [VP, B], self.obj, activate)

    def set_tle(self, activate):
        _call("dose_tle_navigator", None, [VP, B], self.obj, activate)

    def scale_factor(self, scale):
        _call("scale_factor_dosimetry_calculator", None, [VP, F], self.obj, scale)

    def water_reference(self, activate):
        _call("water_reference_dosimetry_calculator", None, [VP, B], self.obj, activate)

    def minimum_density(self, density, unit):
        _call("minimum_density_dosimetry_calculator", None, [VP, F, S], self.obj, density, _b(unit))

    def attach_to_navigator(self, name):
        _call("attach_to_navigator_dosimetry_calculator", None, [VP, S], self.obj, _b(name))


class GGEMSProfilerManager:
    def __init__(self):
        self.obj = _call("get_instance_profiler_manager", VP, [])

    def print_summary_profile(self):
        _call("print_summary_profiler_manager", None, [VP], self.obj)


class GGEMSRAMManager:
    def __init__(self):
        self.obj = _call("get_instance_ggems_ram_manager", VP, [])

    def print_infos(self):
        _call("print_infos_ram_manager", None, [VP], self.obj)


class GGEMS:
    def __init__(self):
        self.obj = _call("create_ggems", VP, [])

    def delete(self):
        if self.obj:
            _call("delete_ggems", None, [VP], self.obj)
            self.obj = None

    def initialize(self, seed=0):
        _call("initialize_ggems", None, [VP, U32], self.obj, seed)

    def run(self):
        _call("run_ggems", None, [VP], self.obj)

    def run_acquisition(self, rotations, n_particles, unit="deg", phantoms=None):
        """CT acquisition in one launch: projection p turns the source and every
        CT system by rotations[p] = (rx, ry, rz) and transports n_particles[p]
        primaries; each CT system saves a stack of projections.  phantoms[p]
        (optional): projection p sees only that phantom navigator (its rank
        among the voxelized and meshed phantoms in creation order, -1 all)."""
        import numpy as np
        r = np.ascontiguousarray(rotations, np.float32).reshape(-1, 3)
        n = np.ascontiguousarray(n_particles, np.uint64).reshape(-1)
        if n.size != r.shape[0]:
            raise GGEMSError("one photon budget per projection")
        rx, ry, rz = (np.ascontiguousarray(r[:, i]) for i in range(3))
        fp = C.POINTER(C.c_float)
        args = [self.obj, r.shape[0], rx.ctypes.data_as(fp), ry.ctypes.data_as(fp), rz.ctypes.data_as(fp),
                unit.encode(), n.ctypes.data_as(C.POINTER(C.c_uint64))]
        types = [VP, C.c_size_t, fp, fp, fp, C.c_char_p, C.POINTER(C.c_uint64)]
        if phantoms is None:
            _call("run_acquisition_ggems", None, types, *args)
            return
        ph = np.ascontiguousarray(phantoms, np.int32).reshape(-1)
        if ph.size != r.shape[0]:
            raise GGEMSError("one phantom per projection")
        _call("run_acquisition_phantoms_ggems", None, types + [C.POINTER(C.c_int32)], *args,
              ph.ctypes.data_as(C.POINTER(C.c_int32)))

    def repeat_acquisition(self):
        _call("repeat_acquisition_ggems", None, [VP], self.obj)

    def set_sampling(self, mode):
        """How voxel steps sample interactions (an addition with no GGEMS
        counterpart; call after initialize): "reference", GGEMS's three
        competing exponentials per step (the default), or "tau", one
        exponential number of mean free paths per flight"""
        self.engine().set_sampling(mode)

    def autotune(self, n_trial):
        _call("autotune_ggems", None, [VP, C.c_uint64], self.obj, int(n_trial))

    def opencl_verbose(self, flag):
        _call("set_opencl_verbose_ggems", None, [VP, B], self.obj, flag)

    def material_database_verbose(self, flag):
        _call("set_material_database_verbose_ggems", None, [VP, B], self.obj, flag)

    def navigator_verbose(self, flag):
        _call("set_navigator_ggems", None, [VP, B], self.obj, flag)

    def source_verbose(self, flag):
        _call("set_source_ggems", None, [VP, B], self.obj, flag)

    def memory_verbose(self, flag):
        _call("set_memory_ram_ggems", None, [VP, B], self.obj, flag)

    def process_verbose(self, flag):
        _call("set_process_ggems", None, [VP, B], self.obj, flag)

    def profiling_verbose(self, flag):
        _call("set_profiling_ggems", None, [VP, B], self.obj, flag)

    def range_cuts_verbose(self, flag):
        _call("set_range_cuts_ggems", None, [VP, B], self.obj, flag)

    def random_verbose(self, flag):
        _call("set_random_ggems", None, [VP, B], self.obj, flag)

    def tracking_verbose(self, flag, particle_id):
        _call("set_tracking_ggems", None, [VP, B, I], self.obj, flag, particle_id)

    # -- additions with no GGEMS counterpart ------------------------------------
    def run_time_ms(self):
        """wall time of the last run's transport (ms)"""
        return _call("get_run_time_ms_ggems", C.c_double, [VP], self.obj)

    def engine(self):
        """non-owning ggems_amd.engine.Engine view of this simulation's engine context"""
        from .engine import Engine
        return Engine.attach(_call("get_engine_context_ggems", VP, [VP], self.obj))

    def engine_stats(self):
        out = (C.c_uint64 * 8)()
        _call("get_engine_stats_ggems", None, [VP, VP], self.obj, C.cast(out, VP))
        keys = ("primaries", "voxel_steps", "interactions", "detector_steps", "detector_interactions", "hops",
                "mesh_steps", "mesh_interactions")
        return {k: int(v) for k, v in zip(keys, out)}


__all__ = [n for n in dir() if n.startswith("GGEMS")] + ["GGEMSError", "clean_safely"]
