"""ctypes binding of the CPU oracle (oracle/_build/libggems_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never uses it.
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libggems_oracle.so")
REPO = os.path.dirname(HERE)
PHYSICS_BLOB = os.path.join(REPO, "ggems_amd", "data", "_gen", "ggems_physics.bin")
MATERIALS_TXT = os.path.join(REPO, "ggems_amd", "data", "_gen", "materials.txt")

MAX_MAT = 255
MAX_ELEM = 32
MAX_SLOTS = 1048576


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class MaterialTables(C.Structure):
    _fields_ = [
        ("n_materials", C.c_int),
        ("n_elements", C.c_int * MAX_MAT),
        ("elem_index", C.c_int * MAX_MAT),
        ("density", C.c_float * MAX_MAT),
        ("photon_cut", C.c_float * MAX_MAT),
        ("z", C.c_uint8 * (MAX_MAT * MAX_ELEM)),
        ("atom_density", C.c_float * (MAX_MAT * MAX_ELEM)),
        ("mass_fraction", C.c_float * (MAX_MAT * MAX_ELEM)),
    ]


class OBB(C.Structure):
    _fields_ = [("m", (C.c_float * 4) * 4), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3)]


class Mesh(C.Structure):
    _fields_ = [
        ("n_triangles", C.c_int),
        ("tri", C.c_void_p),
        ("n_nodes", C.c_int),
        ("node_box", C.c_void_p),
        ("node_first", C.c_void_p),
        ("tri_next", C.c_void_p),
    ]


class Solid(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("obb", OBB),
        ("mat_offset", C.c_int),
        ("voxel_size", C.c_float * 3),
        ("nvox", C.c_int * 3),
        ("labels", C.c_void_p),
        ("dosimetry", C.c_int),
        ("dosel_size", C.c_float * 3),
        ("inv_dosel", C.c_float * 3),
        ("dose_bmin", C.c_float * 3),
        ("dose_bmax", C.c_float * 3),
        ("ndosel", C.c_int * 3),
        ("edep", C.c_void_p),
        ("edep2", C.c_void_p),
        ("hit", C.c_void_p),
        ("photon_tracking", C.c_void_p),
        ("nelem", C.c_int * 3),
        ("box_size", C.c_float * 3),
        ("threshold", C.c_float),
        ("histogram", C.c_void_p),
        ("scatter", C.c_void_p),
        ("mesh", Mesh),
    ]


class Physics(C.Structure):
    _fields_ = [
        ("nbins", C.c_int),
        ("energy_bins", C.c_void_p),
        ("n_processes", C.c_int),
        ("process_id", C.c_int * 3),
        ("cs", C.c_void_p * 3),
        ("cs_atom_rayleigh", C.c_void_p),
        ("n_elements", C.c_void_p),
        ("elem_index", C.c_void_p),
        ("z", C.c_void_p),
        ("atom_density", C.c_void_p),
        ("photon_cut", C.c_void_p),
        ("tau_mode", C.c_int),
    ]


class Source(C.Structure):
    _fields_ = [
        ("m", (C.c_float * 4) * 4),
        ("aperture", C.c_float),
        ("focal", C.c_float * 3),
        ("n_energy", C.c_int),
        ("energy", C.c_void_p),
        ("cdf", C.c_void_p),
        ("is_interp", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("primaries", "voxel_steps", "interactions", "detector_steps",
                                          "detector_interactions", "hops", "mesh_steps", "mesh_interactions")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ggo_load_physics_data.argtypes = [C.c_char_p]
        L.ggo_load_materials.argtypes = [C.c_char_p]
        L.ggo_build_material_tables.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.c_float, C.c_float,
                                                C.POINTER(MaterialTables)]
        L.ggo_energy_bins.argtypes = [C.c_int, C.c_float, C.c_float, C.c_void_p]
        for f in ("ggo_cs_compton_atom", "ggo_cs_photoelectric_atom", "ggo_cs_rayleigh_atom"):
            getattr(L, f).argtypes = [C.c_float, C.c_int]
            getattr(L, f).restype = C.c_float
        L.ggo_cs_per_material.argtypes = [C.c_int, C.POINTER(MaterialTables), C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        L.ggo_photon_cut.argtypes = [C.POINTER(MaterialTables), C.c_int, C.c_float, C.c_float]
        L.ggo_photon_cut.restype = C.c_float
        L.ggo_attenuation.argtypes = [C.POINTER(MaterialTables), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ggo_attenuation_element_search.argtypes = [C.POINTER(MaterialTables), C.c_int, C.c_void_p, C.c_void_p,
                                                     C.c_void_p]
        L.ggo_mt19937_stream.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p]
        L.ggo_seed_slots.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p]
        L.ggo_kiss_uniform.argtypes = [C.c_void_p]
        L.ggo_kiss_uniform.restype = C.c_float
        L.ggo_distance_to_aabb.argtypes = [C.c_void_p] * 4
        L.ggo_distance_to_aabb.restype = C.c_float
        L.ggo_run.argtypes = [C.POINTER(Physics), C.POINTER(Solid), C.c_int, C.POINTER(Source), C.c_uint64,
                              C.c_uint32, C.c_int, C.POINTER(Stats)]
        L.ggo_run_slots.argtypes = [C.POINTER(Physics), C.POINTER(Solid), C.c_int, C.POINTER(Source), C.c_uint64,
                                    C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(Stats)]
        L.ggo_run_slots.restype = C.c_uint64
        L.ggo_run_acquisition.argtypes = [C.POINTER(Physics), C.POINTER(Solid), C.c_int, C.POINTER(Source), C.c_int,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                          C.POINTER(Stats)]
        L.ggo_batch_layout.argtypes = [C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.ggo_voxel_walk.argtypes = [C.POINTER(Solid), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.ggo_sample_compton.argtypes = [C.c_float, C.c_void_p, C.c_void_p]
        L.ggo_sample_compton.restype = C.c_float
        L.ggo_sample_rayleigh.argtypes = [C.POINTER(Physics), C.c_int, C.c_float, C.c_void_p, C.c_void_p,
                                          C.POINTER(C.c_int)]
        L.ggo_sample_rayleigh.restype = C.c_float
        L.ggo_sample_primary.argtypes = [C.POINTER(Source), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ggo_compute_dose.argtypes = [C.POINTER(Solid), C.c_void_p, C.c_float, C.c_int, C.c_float, C.c_void_p,
                                       C.c_void_p]
        L.ggo_mesh_build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
        L.ggo_mesh_build.restype = C.c_int
        L.ggo_mesh_hits.argtypes = [C.POINTER(Mesh), C.c_void_p, C.c_void_p, C.c_void_p]
        L.ggo_mesh_hits.restype = C.c_int
        L.ggo_data.argtypes = [C.c_char_p, C.POINTER(C.c_uint64)]
        L.ggo_data.restype = C.c_void_p
        if L.ggo_load_physics_data(PHYSICS_BLOB.encode()) != 0:
            raise RuntimeError("cannot load physics data blob %s" % PHYSICS_BLOB)
        if L.ggo_load_materials(MATERIALS_TXT.encode()) <= 0:
            raise RuntimeError("cannot load materials %s" % MATERIALS_TXT)
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def data_array(name, dtype):
    n = C.c_uint64()
    p = lib().ggo_data(name.encode(), C.byref(n))
    if not p:
        raise KeyError(name)
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), shape=(n.value,)).copy()


# --------------------------------------------------------------------------
# tables
# --------------------------------------------------------------------------
def material_tables(names, cut_distance=0.0, cut_emin=1e-3):
    mt = MaterialTables()
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    rc = lib().ggo_build_material_tables(arr, len(names), cut_distance, cut_emin, C.byref(mt))
    if rc != 0:
        raise ValueError("material table build failed (%d)" % rc)
    return mt


def energy_bins(nbins, emin, emax):
    out = np.zeros(nbins, np.float32)
    lib().ggo_energy_bins(nbins, emin, emax, ptr(out))
    return out


def cs_tables(mt, nbins, ebins, processes=(0, 1, 2)):
    """Returns cs[3][n_mat][nbins] (zeros for inactive processes) and the
    per-atom Rayleigh table [101][nbins]."""
    nm = mt.n_materials
    cs = np.zeros((3, nm, nbins), np.float32)
    atom_r = np.zeros((101, nbins), np.float32)
    for p in processes:
        tmp = np.zeros((nm, nbins), np.float32)
        atom = atom_r if p == 2 else np.zeros((101, nbins), np.float32)
        lib().ggo_cs_per_material(p, C.byref(mt), nbins, ptr(ebins), ptr(tmp), ptr(atom))
        cs[p] = tmp
    return cs, atom_r


def mt_arrays(mt):
    nm = mt.n_materials
    n_el = np.array(mt.n_elements[:nm], np.int32)
    idx = np.array(mt.elem_index[:nm], np.int32)
    tot = int(idx[-1] + n_el[-1])
    return dict(n_elements=n_el, elem_index=idx, z=np.array(mt.z[:tot], np.uint8),
                atom_density=np.array(mt.atom_density[:tot], np.float32),
                density=np.array(mt.density[:nm], np.float32), photon_cut=np.array(mt.photon_cut[:nm], np.float32))


def seed_slots(seed, n=MAX_SLOTS):
    st = np.zeros((5, n), np.uint32)
    lib().ggo_seed_slots(seed, n, ptr(st))
    return st


def kiss_stream(state5, n):
    s = np.ascontiguousarray(state5, dtype=np.uint32).copy()
    out = np.empty(n, np.float32)
    for i in range(n):
        out[i] = lib().ggo_kiss_uniform(ptr(s))
    return out, s


def mesh_build(triangles, depth):
    """GGEMS octree over triangles [n][3][3] (global coordinates): returns a
    dict with the arrays ggo_mesh points into and the octree box."""
    tri = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
    n_nodes = (8 ** depth - 1) // 7
    out = dict(tri=tri, node_box=np.zeros((n_nodes, 6), np.float32), node_first=np.zeros(n_nodes, np.int32),
               tri_next=np.full(len(tri), -1, np.int32), obb_min=np.zeros(3, np.float32),
               obb_max=np.zeros(3, np.float32))
    rc = lib().ggo_mesh_build(ptr(tri), len(tri), depth, ptr(out["node_box"]), ptr(out["node_first"]),
                              ptr(out["tri_next"]), ptr(out["obb_min"]), ptr(out["obb_max"]))
    if rc != n_nodes:
        raise ValueError("mesh build failed (%d)" % rc)
    return out


def mesh_struct(m):
    s = Mesh()
    s.n_triangles = len(m["tri"])
    s.tri = ptr(m["tri"])
    s.n_nodes = len(m["node_first"])
    s.node_box = ptr(m["node_box"])
    s.node_first = ptr(m["node_first"])
    s.tri_next = ptr(m["tri_next"])
    return s


def mesh_hits(mesh_struct_, p, d):
    pts = np.zeros((16, 3), np.float32)
    n = lib().ggo_mesh_hits(C.byref(mesh_struct_), ptr(np.asarray(p, np.float32)), ptr(np.asarray(d, np.float32)),
                            ptr(pts))
    return n, pts[:n]


def batch_layout(n):
    nb, ns = C.c_uint64(), C.c_uint64()
    lib().ggo_batch_layout(n, C.byref(nb), C.byref(ns))
    return nb.value, ns.value
