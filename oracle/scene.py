"""Scene builders for the oracle (TEST INFRASTRUCTURE ONLY).

A Scene holds one voxelized phantom (optional dosimetry), CT modules and an
X-ray cone-beam source, with physics tables built by the CPU oracle.  The
same Scene drives the oracle (oracle.ggems_oracle) and the HIP engine
(ggems_amd.engine), so both see identical inputs.

Matrices follow GGEMSGeometryTransformation (src/maths/GGEMSGeometryTransformation.cc:
143-269): the local axis matrix A is the start, every SetTranslation /
SetRotation left-multiplies the current matrix, R = Rz*Ry*Rx, all in FP32
with libm cosf/sinf.
"""

import ctypes as C
import math
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import ggems_oracle as O  # noqa: E402

_libm = C.CDLL("libm.so.6")
_libm.cosf.restype = C.c_float
_libm.cosf.argtypes = [C.c_float]
_libm.sinf.restype = C.c_float
_libm.sinf.argtypes = [C.c_float]
for _fn in ("sqrtf", "asinf"):
    getattr(_libm, _fn).restype = C.c_float
    getattr(_libm, _fn).argtypes = [C.c_float]

F = np.float32
DEG = F(3.141592653589793238463) / F(180.0)
G_CM3 = F(6.241509704e21) / F(1e3)

AXIS_IDENTITY = ((1, 0, 0), (0, 1, 0), (0, 0, 1))
AXIS_SOURCE = ((0, 0, -1), (0, 1, 0), (1, 0, 0))       # GGEMSXRaySource.cc:45-51
AXIS_SOLID_BOX = ((0, 0, 1), (0, 1, 0), (-1, 0, 0))    # GGEMSSolidBox.cc:71-77


def mat_mult(a, b):
    """GGfloat44MultGGfloat44 (GGEMSMatrixOperations.hh:150-184), FP32 left to right."""
    out = np.zeros((4, 4), F)
    for i in range(4):
        for j in range(4):
            acc = F(a[i][0] * b[0][j])
            acc = F(acc + F(a[i][1] * b[1][j]))
            acc = F(acc + F(a[i][2] * b[2][j]))
            acc = F(acc + F(a[i][3] * b[3][j]))
            out[i, j] = acc
    return out


class Transform:
    def __init__(self, axis=AXIS_IDENTITY):
        self.m = np.eye(4, dtype=F)
        self.m[:3, :3] = np.array(axis, F)

    def translate(self, t):
        T = np.eye(4, dtype=F)
        T[:3, 3] = np.array(t, F)
        self.m = mat_mult(T, self.m)
        return self

    def rotate(self, r):
        rx, ry, rz = (F(v) for v in r)
        c, s = F(_libm.cosf(rx)), F(_libm.sinf(rx))
        Rx = np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]], F)
        c, s = F(_libm.cosf(ry)), F(_libm.sinf(ry))
        Ry = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], F)
        c, s = F(_libm.cosf(rz)), F(_libm.sinf(rz))
        Rz = np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], F)
        R = mat_mult(Rz, mat_mult(Ry, Rx))
        self.m = mat_mult(R, self.m)
        return self


class Scene:
    def __init__(self, nbins=220, emin=1e-3, emax=1.0, processes=(0, 1, 2), cut_distance=0.1):
        self.nbins, self.emin, self.emax = nbins, emin, emax
        self.processes = tuple(processes)
        self.cut_distance = cut_distance
        self.navigators = []   # (kind, materials, payload)
        self.source = None
        self.tau = False       # tau sampling mode (one exponential per flight) in oracle and engine

    # ---- geometry -----------------------------------------------------------
    def add_phantom(self, labels, materials, voxel_size, position=(0, 0, 0), rotation=(0, 0, 0), dosimetry=None):
        """labels [nz][ny][nx] uint8; dosimetry: dict(edep_squared, hit, photon_tracking) or None."""
        labels = np.ascontiguousarray(labels, np.uint8)
        nz, ny, nx = labels.shape
        vs = np.array(voxel_size, F)
        n = np.array([nx, ny, nz], F)
        bmin = -n * vs * F(0.5)
        bmax = n * vs * F(0.5)
        tr = Transform(AXIS_IDENTITY).rotate(rotation).translate(position)  # rotation before position
        self.navigators.append(("voxel", list(materials), dict(labels=labels, vs=vs, bmin=bmin, bmax=bmax, m=tr.m,
                                                              dosimetry=dosimetry)))

    def add_ct_flat(self, material, n_modules, n_elements, element_size, sid, sdd, threshold,
                    store_scatter=True, rotation=None):
        """GGEMSCTSystem::InitializeFlatGeometry (src/navigators/GGEMSCTSystem.cc:186-222)."""
        mods = []
        es = np.array(element_size, F)
        ne = np.array(n_elements, np.int32)
        size = ne.astype(F) * es
        gx = F(sdd) - F(sid)
        for j in range(n_modules[1]):
            gy = F(ne[1]) * es[1] * (F(j) + F(0.5) * (F(1.0) - F(n_modules[1])))
            for i in range(n_modules[0]):
                gz = F(ne[0]) * es[0] * (F(i) + F(0.5) * (F(1.0) - F(n_modules[0])))
                tr = Transform(AXIS_SOLID_BOX).rotate((0, 0, 0)).translate((gx, gy, gz))
                if rotation is not None:
                    tr.rotate(rotation)
                mods.append(dict(m=tr.m, bmin=-size * F(0.5), bmax=size * F(0.5), ne=ne, size=size))
        self.navigators.append(("ct", [material], dict(modules=mods, threshold=F(threshold),
                                                        store_scatter=store_scatter, n_modules=n_modules)))

    def add_ct_curved(self, material, n_modules, n_elements, element_size, sid, sdd, threshold, store_scatter=True):
        """GGEMSCTSystem::InitializeCurvedGeometry (src/navigators/GGEMSCTSystem.cc:131-183): modules
        turned about the source by alpha = 2 asin(c / rho) per module."""
        es = np.array(element_size, F)
        ne = np.array(n_elements, np.int32)
        size = ne.astype(F) * es
        c = F(F(ne[1]) * es[1]) * F(0.5)
        rho = F(_libm.sqrtf(F(F(sdd) * F(sdd) + c * c)))
        alpha = F(2.0) * F(_libm.asinf(F(c / rho)))
        ox, oy, px, py = -F(sid), F(0.0), F(sdd) - F(sid), F(0.0)
        mods = []
        for j in range(n_modules[1]):
            step = F(alpha * F(F(j) + F(0.5) * (F(1.0) - F(n_modules[1]))))
            cs, sn = F(_libm.cosf(step)), F(_libm.sinf(step))
            gx = F(F(F((px - ox) * cs) - F((py - oy) * sn)) + ox)
            gy = F(F(F((px - ox) * sn) + F((py - oy) * cs)) + oy)
            for i in range(n_modules[0]):
                gz = F(ne[0]) * es[0] * (F(i) + F(0.5) * (F(1.0) - F(n_modules[0])))
                tr = Transform(AXIS_SOLID_BOX).rotate((0, 0, step)).translate((gx, gy, gz))
                mods.append(dict(m=tr.m, bmin=-size * F(0.5), bmax=size * F(0.5), ne=ne, size=size))
        self.navigators.append(("ct", [material], dict(modules=mods, threshold=F(threshold),
                                                        store_scatter=store_scatter, n_modules=n_modules)))

    def add_mesh(self, triangles, material, position=(0, 0, 0), rotation=(0, 0, 0), depth=4, dosimetry=None):
        """GGEMSMeshedPhantom: triangles [n][3][3] (STL frame) moved by the
        navigator's rotation then position, as GGEMSMeshedSolid::UpdateTriangles
        does (src/geometries/GGEMSMeshedSolid.cc:205-262), and the GGEMS octree.
        dosimetry: dict(dosel_size=(x, y, z) mm, photon_tracking) or None; the
        dose map covers the octree box (GGEMSDosimetryCalculator.cc:445-500)."""
        tr = Transform(AXIS_IDENTITY).rotate(rotation).translate(position)
        m = tr.m
        tri = np.asarray(triangles, F).reshape(-1, 3, 3)
        out = np.empty_like(tri)
        for a in range(3):
            acc = F(m[a, 0]) * tri[:, :, 0]
            acc = (acc + F(m[a, 1]) * tri[:, :, 1]).astype(F)
            acc = (acc + F(m[a, 2]) * tri[:, :, 2]).astype(F)
            out[:, :, a] = (acc + F(m[a, 3])).astype(F)
        mesh = O.mesh_build(out, depth)
        self.navigators.append(("mesh", [material], dict(tri=out, mesh=mesh, dosimetry=dosimetry,
                                                          bmin=mesh["obb_min"], bmax=mesh["obb_max"])))

    def set_source(self, position, aperture_deg, energies, intensities=None, focal=(0, 0, 0), rotation=(0, 0, 0),
                   mono=False):
        """GGEMSXRaySource + GGEMSSource::FillEnergy (src/sources/GGEMSSource.cc:265-316)."""
        tr = Transform(AXIS_SOURCE).translate(position).rotate(rotation)
        e = np.asarray(energies, np.float64)
        w = np.ones_like(e) if intensities is None else np.asarray(intensities, np.float64)
        nb = e.size
        spectrum = np.zeros(nb + 1, F)
        spectrum[:nb] = e.astype(F)
        spectrum[nb] = spectrum[nb - 1]
        total = float(np.sum(w))
        tmp = np.zeros(nb + 1, np.float64)
        tmp[1:] = w
        cdf = np.zeros(nb + 1, F)
        for i in range(1, nb + 1):
            tmp[i] = tmp[i] + tmp[i - 1]
            cdf[i] = F(tmp[i] / total)
        cdf[nb] = F(1.0)
        self.source = dict(m=tr.m, aperture=F(aperture_deg) * DEG, focal=np.array(focal, F), energy=spectrum,
                           cdf=cdf, is_interp=0 if mono else 1, n_energy=nb)

    # ---- tables ---------------------------------------------------------------
    def build(self):
        self.ebins = O.energy_bins(self.nbins, self.emin, self.emax)
        # a navigator with the material list of an earlier one shares its
        # materials (every navigator here has the same cut distance, so the
        # tables are equal value for value)
        mats = []
        self.offsets = []
        first = {}
        for kind, materials, _ in self.navigators:
            key = tuple(materials)
            if key not in first:
                first[key] = len(mats)
                mats.extend(materials)
            self.offsets.append(first[key])
        self.materials = mats
        self.mt = O.material_tables(mats, cut_distance=self.cut_distance, cut_emin=self.emin)
        self.cs, self.cs_atom = O.cs_tables(self.mt, self.nbins, self.ebins, self.processes)
        self.arrays = O.mt_arrays(self.mt)
        pp = np.stack([O.data_array("rayleigh_pp%d" % k, np.float32) for k in range(9)])
        self.pp = np.ascontiguousarray(pp, np.float32)
        self.factor = float(O.data_array("rayleigh_factor", np.float32)[0])
        return self

    # ---- oracle ---------------------------------------------------------------
    def _oracle_structs(self, tallies, images=None):
        ph = O.Physics()
        ph.nbins = self.nbins
        ph.energy_bins = O.ptr(self.ebins)
        ph.n_processes = len(self.processes)
        for i, p in enumerate(self.processes):
            ph.process_id[i] = p
        self._cs_rows = [np.ascontiguousarray(self.cs[p]) for p in range(3)]
        for p in range(3):
            ph.cs[p] = O.ptr(self._cs_rows[p])
        ph.cs_atom_rayleigh = O.ptr(self.cs_atom)
        a = self.arrays
        ph.n_elements, ph.elem_index = O.ptr(a["n_elements"]), O.ptr(a["elem_index"])
        ph.z, ph.atom_density, ph.photon_cut = O.ptr(a["z"]), O.ptr(a["atom_density"]), O.ptr(a["photon_cut"])
        ph.tau_mode = 1 if self.tau else 0
        solids = []
        for (kind, materials, pay), off in zip(self.navigators, self.offsets):
            if kind == "voxel":
                s = O.Solid()
                s.type = 0
                s.obb.m = ((C.c_float * 4) * 4)(*[(C.c_float * 4)(*row) for row in pay["m"].tolist()])
                s.obb.bmin[:] = pay["bmin"].tolist()
                s.obb.bmax[:] = pay["bmax"].tolist()
                s.mat_offset = off
                s.voxel_size[:] = pay["vs"].tolist()
                nz, ny, nx = pay["labels"].shape
                s.nvox[:] = [nx, ny, nz]
                s.labels = O.ptr(pay["labels"])
                self._attach_dose(s, pay, tallies)
                solids.append(s)
            elif kind == "mesh":
                s = O.Solid()
                s.type = 2
                s.obb.m = ((C.c_float * 4) * 4)(*[(C.c_float * 4)(*row) for row in np.eye(4, dtype=F).tolist()])
                s.obb.bmin[:] = pay["mesh"]["obb_min"].tolist()
                s.obb.bmax[:] = pay["mesh"]["obb_max"].tolist()
                s.mat_offset = off
                s.mesh = O.mesh_struct(pay["mesh"])
                self._attach_dose(s, pay, tallies)
                solids.append(s)
            else:
                for k, mod in enumerate(pay["modules"]):
                    s = O.Solid()
                    s.type = 1
                    s.obb.m = ((C.c_float * 4) * 4)(*[(C.c_float * 4)(*row) for row in mod["m"].tolist()])
                    s.obb.bmin[:] = mod["bmin"].tolist()
                    s.obb.bmax[:] = mod["bmax"].tolist()
                    s.mat_offset = off
                    s.nelem[:] = mod["ne"].tolist()
                    s.box_size[:] = mod["size"].tolist()
                    s.threshold = pay["threshold"]
                    shape = (mod["ne"][1], mod["ne"][0]) if images is None else (images, mod["ne"][1], mod["ne"][0])
                    h = np.zeros(shape, np.int32)
                    tallies.setdefault("hist", []).append(h)
                    s.histogram = O.ptr(h)
                    if pay["store_scatter"]:
                        sc = np.zeros_like(h)
                        tallies.setdefault("scatter", []).append(sc)
                        s.scatter = O.ptr(sc)
                    solids.append(s)
        src = O.Source()
        S = self.source
        src.m = ((C.c_float * 4) * 4)(*[(C.c_float * 4)(*row) for row in S["m"].tolist()])
        src.aperture = S["aperture"]
        src.focal[:] = S["focal"].tolist()
        src.n_energy = S["n_energy"]
        src.energy, src.cdf = O.ptr(S["energy"]), O.ptr(S["cdf"])
        src.is_interp = S["is_interp"]
        return ph, (O.Solid * len(solids))(*solids), src

    def _attach_dose(self, s, pay, tallies):
        if pay["dosimetry"] is None:
            return
        d = self.dose_grid(pay)
        s.dosimetry = 1
        s.dosel_size[:] = d["size"].tolist()
        s.inv_dosel[:] = (F(1.0) / d["size"]).tolist()
        s.dose_bmin[:] = pay["bmin"].tolist()
        s.dose_bmax[:] = pay["bmax"].tolist()
        s.ndosel[:] = d["n"]
        shape = (d["n"][2], d["n"][1], d["n"][0])
        tallies["edep"] = np.zeros(shape, np.float64)
        tallies["edep2"] = np.zeros(shape, np.float64)
        tallies["hit"] = np.zeros(shape, np.int32)
        s.edep, s.edep2, s.hit = O.ptr(tallies["edep"]), O.ptr(tallies["edep2"]), O.ptr(tallies["hit"])
        if pay["dosimetry"].get("photon_tracking"):
            tallies["photon_tracking"] = np.zeros(shape, np.int32)
            s.photon_tracking = O.ptr(tallies["photon_tracking"])

    def dose_grid(self, pay):
        # GGEMSDosimetryCalculator::Initialize (:445-500): dosel size = voxel
        # size of a voxelized phantom, or the one set (required for a mesh)
        size = pay.get("vs")
        if pay["dosimetry"].get("dosel_size") is not None:
            size = np.array(pay["dosimetry"]["dosel_size"], F)
        n = [int(F(F(pay["bmax"][i] - pay["bmin"][i]) / size[i])) for i in range(3)]
        return dict(size=size, n=n)

    def prepare_oracle(self):
        """Builds the oracle structs once, with every tally page faulted in
        (a steady-state run keeps its tallies allocated)."""
        tallies = {}
        structs = self._oracle_structs(tallies)
        for k, v in tallies.items():
            for a in (v if isinstance(v, list) else [v]):
                a.fill(0)
        self._prepared = (structs, tallies)
        return self._prepared

    def run_oracle(self, n, seed, threads=0, slots=None, prepared=None):
        if prepared is None:
            prepared = self.prepare_oracle()
        (ph, solids, src), tallies = prepared
        st = O.Stats()
        if slots is None:
            O.lib().ggo_run(C.byref(ph), solids, len(solids), C.byref(src), n, seed, threads, C.byref(st))
        else:
            O.lib().ggo_run_slots(C.byref(ph), solids, len(solids), C.byref(src), n, seed, slots[0], slots[1],
                                  threads, C.byref(st))
        tallies["stats"] = {k: int(getattr(st, k)) for k, _ in O.Stats._fields_}
        return tallies

    # ---- CT acquisition ---------------------------------------------------------
    def acquisition_matrices(self, rotations):
        """Per projection, the source and every CT module turned by
        R = Rz*Ry*Rx of rotations[p] (radians) on top of the scene's
        placement: R left-multiplies the matrix, as SetRotation does
        (src/maths/GGEMSGeometryTransformation.cc:196-269)."""
        src, box = [], []
        mods = [mod for kind, _, pay in self.navigators if kind == "ct" for mod in pay["modules"]]
        for r in rotations:
            R = Transform().rotate(r).m
            src.append(mat_mult(R, self.source["m"]))
            box.append([mat_mult(R, mod["m"]) for mod in mods])
        return np.array(src, F).reshape(len(rotations), 4, 4), np.array(box, F).reshape(len(rotations), len(mods), 4, 4)

    def run_oracle_acquisition(self, rotations, n_particles, seed, threads=0, phantoms=None):
        """The projections as consecutive runs (slot-major, continuing JKISS
        states); hist/scatter are per-module stacks [P][rows][cols].
        phantoms[p] (optional): the one phantom solid projection p sees."""
        src, box = self.acquisition_matrices(rotations)
        tallies = {}
        ph, solids, source = self._oracle_structs(tallies, images=len(rotations))
        n = np.ascontiguousarray(n_particles, np.uint64)
        only = None if phantoms is None else np.ascontiguousarray(phantoms, np.int32)
        st = O.Stats()
        rc = O.lib().ggo_run_acquisition(C.byref(ph), solids, len(solids), C.byref(source), len(rotations),
                                         O.ptr(np.ascontiguousarray(src)), O.ptr(np.ascontiguousarray(box)),
                                         O.ptr(n), None if only is None else O.ptr(only), seed, threads,
                                         C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle acquisition failed")
        tallies["stats"] = {k: int(getattr(st, k)) for k, _ in O.Stats._fields_}
        return tallies

    def engine_acquisition(self, rotations, n_particles, seed, eng=None, phantoms=None):
        """An engine with the acquisition set (seeded, tallies zero)."""
        eng = eng or self.make_engine()
        src, box = self.acquisition_matrices(rotations)
        eng.seed(seed)
        eng.reset_tallies()
        eng.set_acquisition(src, box, n_particles, phantoms)
        return eng

    def read_acquisition(self, eng):
        out = {"hist": [], "scatter": []}
        for k in self.engine_solids.get("box", []):
            hs = [eng.projection(k, p) for p in range(eng.n_projections)]
            out["hist"].append(np.stack([h for h, _ in hs]))
            if hs[0][1] is not None:
                out["scatter"].append(np.stack([s for _, s in hs]))
        out["stats"] = eng.stats()
        out["kernel_ms"] = eng.last_kernel_ms()
        return out

    # ---- engine ---------------------------------------------------------------
    def make_engine(self, device=0):
        from ggems_amd.engine import Engine
        eng = Engine(device)
        a = self.arrays
        eng.set_physics(self.ebins, self.cs, self.cs_atom, a["n_elements"], a["elem_index"], a["z"],
                        a["atom_density"], a["photon_cut"], a["density"], self.pp, self.factor, self.processes)
        self.engine_solids = {}
        for (kind, materials, pay), off in zip(self.navigators, self.offsets):
            if kind == "voxel":
                k = eng.add_voxel_solid(pay["m"], pay["bmin"], pay["bmax"], pay["labels"], pay["vs"], off)
                self.engine_solids.setdefault("voxel", []).append(k)
                if pay["dosimetry"] is not None:
                    d = self.dose_grid(pay)
                    eng.set_dosimetry(k, d["size"], d["n"], pay["bmin"], pay["bmax"], True, True,
                                      bool(pay["dosimetry"].get("photon_tracking")))
            elif kind == "mesh":
                k = eng.add_mesh_solid(pay["tri"], pay["mesh"]["obb_min"], pay["mesh"]["obb_max"], off)
                self.engine_solids.setdefault("mesh", []).append(k)
                if pay["dosimetry"] is not None:
                    d = self.dose_grid(pay)
                    eng.set_dosimetry(k, d["size"], d["n"], pay["bmin"], pay["bmax"], True, True,
                                      bool(pay["dosimetry"].get("photon_tracking")))
            else:
                for mod in pay["modules"]:
                    k = eng.add_box_solid(mod["m"], mod["bmin"], mod["bmax"], mod["ne"], mod["size"],
                                          pay["threshold"], off, pay["store_scatter"])
                    self.engine_solids.setdefault("box", []).append(k)
        S = self.source
        eng.set_xray_source(S["m"], S["aperture"], S["focal"], S["energy"], S["cdf"], S["is_interp"])
        if self.tau:
            eng.set_sampling("tau")
        return eng

    def run_engine(self, n, seed, eng=None):
        eng = eng or self.make_engine()
        eng.seed(seed)
        eng.reset_tallies()
        eng.run(n)
        out = {"hist": [], "scatter": []}
        for k in self.engine_solids.get("box", []):
            h, s = eng.histogram(k)
            out["hist"].append(h)
            if s is not None:
                out["scatter"].append(s)
        if eng.n_dosels is not None:
            out["edep"], out["edep2"], out["hit"], pt = eng.dose_tallies()
            if pt is not None:
                out["photon_tracking"] = pt
        out["stats"] = eng.stats()
        out["kernel_ms"] = eng.last_kernel_ms()
        return out


# ---- configurations ------------------------------------------------------------
def cylinder_phantom(n=100, radius=20.0, voxel=1.0):
    """cfg1: n^3 voxels of water with a centred cortical-bone cylinder along z."""
    c = (np.arange(n) + 0.5) * voxel - n * voxel / 2
    yy, xx = np.meshgrid(c, c, indexing="ij")
    cyl = (xx ** 2 + yy ** 2) <= radius ** 2
    lab = np.zeros((n, n, n), np.uint8)
    lab[:, cyl] = 1
    return lab


def cfg1_scene(n_side=100, detector=256, dosimetry=True, energy_kev=60.0, materials=("Water", "RibBone"),
               aperture_deg=None, sid=600.0, sdd=1000.0):
    sc = Scene()
    sc.add_phantom(cylinder_phantom(n_side), list(materials), (1.0, 1.0, 1.0),
                   dosimetry={"photon_tracking": False} if dosimetry else None)
    sc.add_ct_flat("GOS", (1, 1), (detector, detector, 1), (1.0, 1.0, 1.0), sid, sdd, 10e-3)
    if aperture_deg is None:
        aperture_deg = math.degrees(math.atan(detector * 0.5 * math.sqrt(2.0) / sdd))
    sc.set_source((-sid, 0.0, 0.0), aperture_deg, [energy_kev * 1e-3], mono=True)
    return sc.build()


def voxel_ct_scene(workdir):
    """The voxel-CT workload: the 512x512x256 body phantom of 0.8 mm voxels
    (five materials, FP64 dosimetry on the voxel grid), the 120 kVp spectrum
    behind 2.5 mm Al, a flat 256x256 GOS detector of 1 mm pixels at SDD
    1000 mm and a point cone beam at SID 600 mm that just covers it."""
    from ggems_amd import authoring, configs
    mhd, _, spec = configs.voxel_ct_inputs(workdir)
    lab, spacing = authoring.read_mhd(mhd)
    e, w = np.loadtxt(spec, unpack=True)
    sc = Scene()
    sc.add_phantom(lab, authoring.BODY_MATERIALS, tuple(spacing), dosimetry={"photon_tracking": False})
    sc.add_ct_flat("GOS", (1, 1), (256, 256, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), math.degrees(math.atan(128 * math.sqrt(2.0) / 1000.0)), e, intensities=w)
    return sc.build()
