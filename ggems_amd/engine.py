"""ctypes binding of the HIP transport engine (libggems_engine.so).

Mirrors include/ggems_engine.h one to one.  The engine has no CPU
fallback: loading fails loudly when the library has not been built, and
every call raises EngineError with the engine's message on failure.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.environ.get("GGEMS_AMD_LIB_DIR") or os.path.join(HERE, "lib")
ENGINE_LIB = os.path.join(LIB_DIR, "libggems_engine.so")

MAX_SLOTS = 1048576
MAX_HOPS = 100


class EngineError(RuntimeError):
    pass


class OBB(C.Structure):
    _fields_ = [("matrix", C.c_float * 16), ("border_min", C.c_float * 3), ("border_max", C.c_float * 3)]


class PhysicsDesc(C.Structure):
    _fields_ = [
        ("nbins", C.c_int32), ("n_materials", C.c_int32), ("n_processes", C.c_int32),
        ("process_id", C.c_int32 * 3),
        ("energy_bins", C.c_void_p), ("cross_sections", C.c_void_p), ("cs_atom_rayleigh", C.c_void_p),
        ("total_elements", C.c_int32),
        ("n_elements", C.c_void_p), ("element_index", C.c_void_p), ("atomic_number", C.c_void_p),
        ("atom_density", C.c_void_p), ("photon_cut", C.c_void_p), ("density", C.c_void_p),
        ("rayleigh_pp", C.c_void_p), ("rayleigh_factor", C.c_float),
    ]


class VoxelSolidDesc(C.Structure):
    _fields_ = [("obb", OBB), ("material_offset", C.c_int32), ("n_voxels", C.c_int32 * 3),
                ("voxel_size", C.c_float * 3), ("labels", C.c_void_p)]


class BoxSolidDesc(C.Structure):
    _fields_ = [("obb", OBB), ("material_offset", C.c_int32), ("n_elements", C.c_int32 * 3),
                ("box_size", C.c_float * 3), ("threshold", C.c_float), ("store_scatter", C.c_int32)]


class MeshSolidDesc(C.Structure):
    _fields_ = [("obb", OBB), ("material_offset", C.c_int32), ("n_triangles", C.c_int32),
                ("triangles", C.c_void_p)]


class DosimetryDesc(C.Structure):
    _fields_ = [("solid", C.c_int32), ("dosel_size", C.c_float * 3), ("n_dosels", C.c_int32 * 3),
                ("border_min", C.c_float * 3), ("border_max", C.c_float * 3),
                ("edep_squared", C.c_int32), ("hit", C.c_int32), ("photon_tracking", C.c_int32)]


class XRaySourceDesc(C.Structure):
    _fields_ = [("matrix", C.c_float * 16), ("aperture", C.c_float), ("focal_spot", C.c_float * 3),
                ("n_energy", C.c_int32), ("energy", C.c_void_p), ("cdf", C.c_void_p), ("is_interp", C.c_int32)]


class LaunchDesc(C.Structure):
    _fields_ = [("round_steps", C.c_int32), ("cold_passes", C.c_int32), ("interact_batch", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("general_kernel", C.c_int32)]


class AcquisitionDesc(C.Structure):
    _fields_ = [("n_projections", C.c_int32), ("source_matrices", C.c_void_p), ("box_matrices", C.c_void_p),
                ("n_particles", C.c_void_p), ("phantoms", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("primaries", "voxel_steps", "interactions", "detector_steps",
                                          "detector_interactions", "hops", "mesh_steps", "mesh_interactions")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(ENGINE_LIB):
            raise EngineError("HIP engine library not built: %s (run __graft_entry__.build())" % ENGINE_LIB)
        L = C.CDLL(ENGINE_LIB)
        vp, i32 = C.c_void_p, C.c_int32
        sig = {
            "ggh_last_error": ([], C.c_char_p),
            "ggh_device_count": ([C.POINTER(C.c_int)], C.c_int),
            "ggh_context_create": ([C.c_int, C.POINTER(vp)], C.c_int),
            "ggh_context_destroy": ([vp], C.c_int),
            "ggh_set_physics": ([vp, C.POINTER(PhysicsDesc)], C.c_int),
            "ggh_add_voxel_solid": ([vp, C.POINTER(VoxelSolidDesc)], C.c_int),
            "ggh_add_box_solid": ([vp, C.POINTER(BoxSolidDesc)], C.c_int),
            "ggh_set_dosimetry": ([vp, C.POINTER(DosimetryDesc)], C.c_int),
            "ggh_set_xray_source": ([vp, C.POINTER(XRaySourceDesc)], C.c_int),
            "ggh_seed": ([vp, C.c_uint32], C.c_int),
            "ggh_run": ([vp, C.c_uint64], C.c_int),
            "ggh_synchronize": ([vp], C.c_int),
            "ggh_last_kernel_ms": ([vp, C.POINTER(C.c_float)], C.c_int),
            "ggh_reset_tallies": ([vp], C.c_int),
            "ggh_get_stats": ([vp, C.POINTER(Stats)], C.c_int),
            "ggh_get_histogram": ([vp, C.c_int, vp, vp], C.c_int),
            "ggh_get_dose_tallies": ([vp, vp, vp, vp, vp], C.c_int),
            "ggh_compute_dose": ([vp, C.c_float, i32, C.c_float, vp, vp], C.c_int),
            "ggh_debug_voxel_walk": ([vp, C.c_int, C.c_int, vp, vp, C.c_int, vp, vp], C.c_int),
            "ggh_debug_primaries": ([vp, C.c_int, vp, vp, vp], C.c_int),
            "ggh_debug_uniforms": ([vp, C.c_int, C.c_int, vp], C.c_int),
            "ggh_host_uniform_from_sum": ([vp, vp, C.c_size_t], None),
            "ggh_host_div_rn": ([vp, vp, vp, C.c_size_t], None),
            "ggh_host_distance_to_aabb": ([vp, vp, vp, vp, C.c_size_t], None),
            "ggh_host_energy_bin": ([vp, C.c_size_t, vp, C.c_int32, vp], None),
            "ggh_set_launch": ([vp, C.POINTER(LaunchDesc)], C.c_int),
            "ggh_last_grid": ([vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)], C.c_int),
            "ggh_debug_neg_log": ([vp, C.c_int, vp, vp, vp], C.c_int),
            "ggh_add_mesh_solid": ([vp, C.POINTER(MeshSolidDesc)], C.c_int),
            "ggh_debug_mesh_hits": ([vp, C.c_int, C.c_int, vp, vp, vp, vp], C.c_int),
            "ggh_set_acquisition": ([vp, C.POINTER(AcquisitionDesc)], C.c_int),
            "ggh_run_acquisition": ([vp], C.c_int),
            "ggh_get_projection": ([vp, C.c_int, C.c_int, vp, vp], C.c_int),
            "ggh_capture_acquisition": ([vp], C.c_int),
            "ggh_replay_acquisition": ([vp], C.c_int),
            "ggh_autotune": ([vp, C.c_uint64, C.POINTER(LaunchDesc)], C.c_int),
            "ggh_drr": ([vp, C.c_int, C.c_int, vp], C.c_int),
            "ggh_drr_noalloc": ([vp, C.c_int, C.c_int, vp, vp], C.c_int),
            "ggh_set_sampling": ([vp, C.c_int32], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc < 0:
        raise EngineError(library().ggh_last_error().decode(errors="replace") + " (code %d)" % rc)
    return rc


def host_uniform_from_sum(sums):
    sums = np.ascontiguousarray(sums, dtype=np.uint32)
    out = np.empty(sums.shape, np.float32)
    library().ggh_host_uniform_from_sum(_ptr(sums), _ptr(out), sums.size)
    return out


def host_div_rn(a, b):
    """a / b by the engine's division-free sequence, host build."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.empty(a.shape, np.float32)
    library().ggh_host_div_rn(_ptr(a), _ptr(b), _ptr(out), a.size)
    return out


def host_energy_bin(e, ebins):
    """the engine's energy-bin lookup (host build of the device function)"""
    e = np.ascontiguousarray(e, np.float32)
    ebins = np.ascontiguousarray(ebins, np.float32)
    out = np.zeros(e.shape, np.int32)
    library().ggh_host_energy_bin(_ptr(e), e.size, _ptr(ebins), ebins.size, _ptr(out))
    return out


def host_distance_to_aabb(p, d, box):
    """The engine's slab test, host build: p, d [n, 3]; box [n, 6] = min xyz, max xyz."""
    p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(d, dtype=np.float32).reshape(-1, 3)
    box = np.ascontiguousarray(box, dtype=np.float32).reshape(-1, 6)
    out = np.empty(p.shape[0], np.float32)
    library().ggh_host_distance_to_aabb(_ptr(p), _ptr(d), _ptr(box), _ptr(out), p.shape[0])
    return out


def device_count():
    n = C.c_int(0)
    _check(library().ggh_device_count(C.byref(n)))
    return n.value


def _obb(matrix, bmin, bmax):
    o = OBB()
    m = np.asarray(matrix, np.float32).reshape(16)
    o.matrix[:] = [float(v) for v in m]
    o.border_min[:] = [float(v) for v in bmin]
    o.border_max[:] = [float(v) for v in bmax]
    return o


class Engine:
    """One device context of the transport engine."""

    def __init__(self, device=0):
        self._keep = []
        h = C.c_void_p()
        _check(library().ggh_context_create(device, C.byref(h)))
        self.h = h
        self.solid_kinds = []
        self.box_elements = {}
        self.n_dosels = None

    @classmethod
    def attach(cls, handle):
        """Non-owning view of a context created elsewhere (e.g. by libggems.so)."""
        if not handle:
            raise EngineError("no engine context (GGEMS not initialised)")
        e = cls.__new__(cls)
        e._keep, e.solid_kinds, e.box_elements, e.n_dosels = [], [], {}, None
        e.h = C.c_void_p(handle)
        e._owned = False
        return e

    def close(self):
        if self.h and getattr(self, "_owned", True):
            library().ggh_context_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_physics(self, energy_bins, cross_sections, cs_atom_rayleigh, n_elements, element_index, atomic_number,
                    atom_density, photon_cut, density, rayleigh_pp, rayleigh_factor, process_ids):
        arrs = dict(
            ebins=np.ascontiguousarray(energy_bins, np.float32),
            cs=np.ascontiguousarray(cross_sections, np.float32),
            atom=np.ascontiguousarray(cs_atom_rayleigh, np.float32),
            nel=np.ascontiguousarray(n_elements, np.int32),
            idx=np.ascontiguousarray(element_index, np.int32),
            z=np.ascontiguousarray(atomic_number, np.uint8),
            ad=np.ascontiguousarray(atom_density, np.float32),
            cut=np.ascontiguousarray(photon_cut, np.float32),
            rho=np.ascontiguousarray(density, np.float32),
            pp=np.ascontiguousarray(rayleigh_pp, np.float32),
        )
        nbins = arrs["ebins"].size
        nm = arrs["nel"].size
        if arrs["cs"].shape != (3, nm, nbins):
            raise ValueError("cross_sections must be [3][n_materials][nbins]")
        d = PhysicsDesc()
        d.nbins, d.n_materials, d.n_processes = nbins, nm, len(process_ids)
        for i, p in enumerate(process_ids):
            d.process_id[i] = int(p)
        d.energy_bins, d.cross_sections, d.cs_atom_rayleigh = _ptr(arrs["ebins"]), _ptr(arrs["cs"]), _ptr(arrs["atom"])
        d.total_elements = arrs["z"].size
        d.n_elements, d.element_index = _ptr(arrs["nel"]), _ptr(arrs["idx"])
        d.atomic_number, d.atom_density = _ptr(arrs["z"]), _ptr(arrs["ad"])
        d.photon_cut, d.density = _ptr(arrs["cut"]), _ptr(arrs["rho"])
        d.rayleigh_pp, d.rayleigh_factor = _ptr(arrs["pp"]), float(rayleigh_factor)
        _check(library().ggh_set_physics(self.h, C.byref(d)))

    def add_voxel_solid(self, matrix, border_min, border_max, labels, voxel_size, material_offset):
        labels = np.ascontiguousarray(labels, np.uint8)  # [nz][ny][nx]
        nz, ny, nx = labels.shape
        d = VoxelSolidDesc()
        d.obb = _obb(matrix, border_min, border_max)
        d.material_offset = int(material_offset)
        d.n_voxels[:] = [nx, ny, nz]
        d.voxel_size[:] = [float(v) for v in voxel_size]
        d.labels = _ptr(labels)
        k = _check(library().ggh_add_voxel_solid(self.h, C.byref(d)))
        self.solid_kinds.append("voxel")
        return k

    def add_box_solid(self, matrix, border_min, border_max, n_elements, box_size, threshold, material_offset,
                      store_scatter=False):
        d = BoxSolidDesc()
        d.obb = _obb(matrix, border_min, border_max)
        d.material_offset = int(material_offset)
        d.n_elements[:] = [int(v) for v in n_elements]
        d.box_size[:] = [float(v) for v in box_size]
        d.threshold = float(threshold)
        d.store_scatter = int(bool(store_scatter))
        k = _check(library().ggh_add_box_solid(self.h, C.byref(d)))
        self.solid_kinds.append("box")
        self.box_elements[k] = (int(n_elements[1]), int(n_elements[0]), bool(store_scatter))
        return k

    def add_mesh_solid(self, triangles, border_min, border_max, material_offset, matrix=None):
        """triangles [n][3][3] in global coordinates; the OBB is the octree box"""
        tri = np.ascontiguousarray(triangles, dtype=np.float32).reshape(-1, 9)
        d = MeshSolidDesc()
        d.obb = _obb(np.eye(4, dtype=np.float32) if matrix is None else matrix, border_min, border_max)
        d.material_offset = int(material_offset)
        d.n_triangles = len(tri)
        d.triangles = tri.ctypes.data_as(C.c_void_p)
        k = _check(library().ggh_add_mesh_solid(self.h, C.byref(d)))
        self.solid_kinds.append("mesh")
        return k

    def debug_mesh_hits(self, solid, pos, dir):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        count = np.zeros(n, np.int32)
        nearest = np.zeros((n, 2, 3), np.float32)
        _check(library().ggh_debug_mesh_hits(self.h, solid, n, _ptr(pos), _ptr(dir), _ptr(count), _ptr(nearest)))
        return count, nearest

    def set_dosimetry(self, solid, dosel_size, n_dosels, border_min, border_max, edep_squared=True, hit=True,
                      photon_tracking=False):
        d = DosimetryDesc()
        d.solid = int(solid)
        d.dosel_size[:] = [float(v) for v in dosel_size]
        d.n_dosels[:] = [int(v) for v in n_dosels]
        d.border_min[:] = [float(v) for v in border_min]
        d.border_max[:] = [float(v) for v in border_max]
        d.edep_squared, d.hit, d.photon_tracking = int(bool(edep_squared)), int(bool(hit)), int(bool(photon_tracking))
        _check(library().ggh_set_dosimetry(self.h, C.byref(d)))
        self.n_dosels = tuple(int(v) for v in n_dosels)
        self.dose_flags = (bool(edep_squared), bool(hit), bool(photon_tracking))

    def set_xray_source(self, matrix, aperture, focal_spot, energy, cdf, is_interp):
        energy = np.ascontiguousarray(energy, np.float32)
        cdf = np.ascontiguousarray(cdf, np.float32)
        if energy.size != cdf.size or energy.size < 2:
            raise ValueError("energy and cdf need n_energy+1 entries")
        d = XRaySourceDesc()
        d.matrix[:] = [float(v) for v in np.asarray(matrix, np.float32).reshape(16)]
        d.aperture = float(aperture)
        d.focal_spot[:] = [float(v) for v in focal_spot]
        d.n_energy = energy.size - 1
        d.energy, d.cdf = _ptr(energy), _ptr(cdf)
        d.is_interp = int(bool(is_interp))
        _check(library().ggh_set_xray_source(self.h, C.byref(d)))

    def seed(self, seed):
        _check(library().ggh_seed(self.h, int(seed) & 0xFFFFFFFF))

    def run(self, n_particles):
        _check(library().ggh_run(self.h, int(n_particles)))

    def last_kernel_ms(self):
        v = C.c_float()
        _check(library().ggh_last_kernel_ms(self.h, C.byref(v)))
        return v.value

    def reset_tallies(self):
        _check(library().ggh_reset_tallies(self.h))

    def stats(self):
        s = Stats()
        _check(library().ggh_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def histogram(self, solid):
        ny, nx, has_scatter = self.box_elements[solid]
        counts = np.zeros((ny, nx), np.int32)
        scatter = np.zeros((ny, nx), np.int32) if has_scatter else None
        _check(library().ggh_get_histogram(self.h, solid, _ptr(counts), _ptr(scatter)))
        return counts, scatter

    def dose_tallies(self):
        nx, ny, nz = self.n_dosels
        e2, hit, pt = self.dose_flags
        edep = np.zeros((nz, ny, nx), np.float64)
        edep2 = np.zeros_like(edep) if e2 else None
        hits = np.zeros((nz, ny, nx), np.int32) if hit else None
        ptrack = np.zeros((nz, ny, nx), np.int32) if pt else None
        _check(library().ggh_get_dose_tallies(self.h, _ptr(edep), _ptr(edep2), _ptr(hits), _ptr(ptrack)))
        return edep, edep2, hits, ptrack

    def compute_dose(self, scale=1.0, water_reference=False, minimum_density=0.0, uncertainty=True):
        nx, ny, nz = self.n_dosels
        dose = np.zeros((nz, ny, nx), np.float32)
        unc = np.zeros_like(dose) if uncertainty else None
        _check(library().ggh_compute_dose(self.h, float(scale), int(bool(water_reference)), float(minimum_density),
                                          _ptr(dose), _ptr(unc)))
        return dose, unc

    def debug_voxel_walk(self, solid, pos, dir, max_steps=4096):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        out = np.full((n, max_steps), -1, np.int32)
        cnt = np.zeros(n, np.int32)
        _check(library().ggh_debug_voxel_walk(self.h, solid, n, _ptr(pos), _ptr(dir), max_steps, _ptr(out), _ptr(cnt)))
        return [out[i, :min(cnt[i], max_steps)].copy() for i in range(n)], cnt

    def debug_primaries(self, n):
        pos = np.zeros((n, 3), np.float32)
        dir = np.zeros((n, 3), np.float32)
        e = np.zeros(n, np.float32)
        _check(library().ggh_debug_primaries(self.h, n, _ptr(pos), _ptr(dir), _ptr(e)))
        return pos, dir, e

    def set_launch(self, round_steps=0, interact_batch=0, blocks_per_cu=0, general_kernel=False, cold_passes=0):
        """launch shape of run(); zero keeps a default (interact_batch 16; round_steps 20 and
        cold_passes 2 with the reference sampling, 16 and 1 with tau sampling)"""
        d = LaunchDesc(int(round_steps), int(cold_passes), int(interact_batch), int(blocks_per_cu),
                       int(bool(general_kernel)))
        _check(library().ggh_set_launch(self.h, C.byref(d)))

    def last_grid(self):
        """(blocks, block size, kernel variant: 0 specialised, 1 general, 2 three
        processes over several solids) of the last run"""
        b, s, v = C.c_int32(), C.c_int32(), C.c_int32()
        _check(library().ggh_last_grid(self.h, C.byref(b), C.byref(s), C.byref(v)))
        return b.value, s.value, v.value

    def debug_neg_log(self, sums):
        sums = np.ascontiguousarray(sums, dtype=np.uint32)
        fast = np.empty(sums.shape, np.float32)
        ref = np.empty(sums.shape, np.float32)
        _check(library().ggh_debug_neg_log(self.h, sums.size, _ptr(sums), _ptr(fast), _ptr(ref)))
        return fast, ref

    # ---- CT acquisition ---------------------------------------------------
    def set_acquisition(self, source_matrices, box_matrices, n_particles, phantoms=None):
        """source_matrices [P][4][4], box_matrices [P][n_box][4][4] (box solids
        in registration order), n_particles [P]: one persistent launch per
        run_acquisition, equal to P consecutive runs with those placements.
        phantoms [P] (optional): projection p sees the box solids and only
        solid phantoms[p] among the others (-1: every solid)."""
        src = np.ascontiguousarray(source_matrices, np.float32).reshape(-1, 16)
        n_proj = src.shape[0]
        box = np.ascontiguousarray(box_matrices, np.float32).reshape(n_proj, -1)
        n = np.ascontiguousarray(n_particles, np.uint64).reshape(-1)
        if n.size != n_proj:
            raise EngineError("one photon budget per projection")
        ph = None
        if phantoms is not None:
            ph = np.ascontiguousarray(phantoms, np.int32).reshape(-1)
            if ph.size != n_proj:
                raise EngineError("one phantom index per projection")
        self._acq = (src, box, n, ph)
        d = AcquisitionDesc(n_proj, _ptr(src).value, _ptr(box).value if box.size else None, _ptr(n).value,
                            _ptr(ph).value if ph is not None else None)
        _check(library().ggh_set_acquisition(self.h, C.byref(d)))
        self.n_projections = n_proj

    def run_acquisition(self):
        _check(library().ggh_run_acquisition(self.h))

    def capture_acquisition(self):
        _check(library().ggh_capture_acquisition(self.h))

    def replay_acquisition(self):
        _check(library().ggh_replay_acquisition(self.h))

    def projection(self, solid, p):
        ny, nx, has_scatter = self.box_elements[solid]
        counts = np.zeros((ny, nx), np.int32)
        scatter = np.zeros((ny, nx), np.int32) if has_scatter else None
        _check(library().ggh_get_projection(self.h, solid, p, _ptr(counts), _ptr(scatter)))
        return counts, scatter

    def drr(self, solid, projection=-1):
        """ray-traced primary transmission [rows][cols] onto box solid `solid`
        (projection < 0: the set-up placement; else that acquisition projection)"""
        ny, nx, _ = self.box_elements[solid]
        img = np.zeros((ny, nx), np.float32)
        _check(library().ggh_drr(self.h, solid, int(projection), _ptr(img)))
        return img

    def set_sampling(self, mode):
        """"reference" (GGEMS's three exponentials per voxel step, the
        default) or "tau" (one exponential number of mean free paths per
        flight; same distributions, different random streams)"""
        _check(library().ggh_set_sampling(self.h, {"reference": 0, "tau": 1}[mode]))

    def synchronize(self):
        """waits for the engine's stream"""
        _check(library().ggh_synchronize(self.h))

    def drr_noalloc(self, solid, image_ptr, projection=-1, stream=None):
        """ggh_drr into the device buffer at image_ptr (rows x cols float32),
        queued on `stream` (a hipStream_t handle; None: the engine's own)
        without allocating; returns before the image is ready"""
        _check(library().ggh_drr_noalloc(self.h, solid, int(projection), C.c_void_p(image_ptr),
                                         None if stream is None else C.c_void_p(stream)))

    def autotune(self, n_trial):
        """fastest launch shape over a trial run; call after seed() and before any run"""
        d = LaunchDesc()
        _check(library().ggh_autotune(self.h, int(n_trial), C.byref(d)))
        return {k: int(getattr(d, k)) for k, _ in LaunchDesc._fields_}

    def debug_uniforms(self, n_slots, n):
        out = np.zeros((n_slots, n), np.float32)
        _check(library().ggh_debug_uniforms(self.h, n_slots, n, _ptr(out)))
        return out
