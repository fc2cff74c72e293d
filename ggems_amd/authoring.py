"""Input authoring: the on-disk files a GGEMS script consumes.

* MHD/raw images and range files in the formats GGEMSMHDImage::Read and
  GGEMSVoxelizedSolid::ConvertImageToLabel read (src/io/GGEMSMHDImage.cc:129-216,
  include/GGEMS/geometries/GGEMSVoxelizedSolid.hh:161-257);
* seeded phantoms: the water cylinder with a cortical-bone insert and the
  body of random overlapping ellipsoids (GGEMS has no ellipsoid primitive);
* a Kramers X-ray spectrum with aluminium filtration, written in the
  "E[MeV] intensity" lines GGEMSSource::SetPolyenergy reads
  (src/sources/GGEMSSource.cc:220-248).
"""

import os

import numpy as np

MET_TYPES = {
    np.dtype(np.int8): "MET_CHAR", np.dtype(np.uint8): "MET_UCHAR", np.dtype(np.int16): "MET_SHORT",
    np.dtype(np.uint16): "MET_USHORT", np.dtype(np.int32): "MET_INT", np.dtype(np.uint32): "MET_UINT",
    np.dtype(np.float32): "MET_FLOAT", np.dtype(np.float64): "MET_DOUBLE",
}


def write_mhd(path, volume, spacing):
    """volume[z][y][x] -> path (.mhd) + raw next to it."""
    volume = np.ascontiguousarray(volume)
    base = path[:-4] if path.endswith(".mhd") else path
    raw = base + ".raw"
    nz, ny, nx = volume.shape
    with open(base + ".mhd", "w") as f:
        f.write("ObjectType = Image\nBinaryDataByteOrderMSB = False\nNDims = 3\n")
        f.write("ElementSpacing = %g %g %g\n" % tuple(spacing))
        f.write("DimSize = %d %d %d\n" % (nx, ny, nz))
        f.write("ElementType = %s\n" % MET_TYPES[volume.dtype])
        f.write("ElementDataFile = %s\n" % os.path.basename(raw))
    volume.tofile(raw)
    return base + ".mhd"


def read_mhd(path):
    """(volume[z][y][x], spacing) of an MHD image written by GGEMS or write_mhd."""
    keys = {}
    with open(path) as f:
        for line in f:
            if "=" in line:
                k, v = line.split("=", 1)
                keys[k.strip()] = v.strip()
    dims = [int(v) for v in keys["DimSize"].split()]
    spacing = [float(v) for v in keys["ElementSpacing"].split()]
    dtype = {v: k for k, v in MET_TYPES.items()}[keys["ElementType"]]
    raw = os.path.join(os.path.dirname(path), keys["ElementDataFile"])
    vol = np.fromfile(raw, dtype=dtype).reshape(dims[2], dims[1], dims[0])
    return vol, spacing


def write_range(path, materials):
    """One line per label value: 'value value material' (label = line order)."""
    with open(path, "w") as f:
        for i, m in enumerate(materials):
            f.write("%d %d %s\n" % (i, i, m))
    return path


def cylinder_phantom(n=100, radius=20.0, voxel=1.0):
    """n^3 voxels: label 0 everywhere, label 1 inside a centred cylinder along z."""
    c = (np.arange(n) + 0.5) * voxel - n * voxel / 2
    yy, xx = np.meshgrid(c, c, indexing="ij")
    lab = np.zeros((n, n, n), np.uint8)
    lab[:, (xx ** 2 + yy ** 2) <= radius ** 2] = 1
    return lab


BODY_MATERIALS = ["Air", "Lung", "Adipose", "SoftTissus_15", "RibBone"]


def ellipsoid_body_phantom(seed, shape=(256, 512, 512), voxel=0.8, n_ellipsoids=40, body_axes=(180.0, 140.0),
                           semi_axis_range=(5.0, 120.0)):
    """Seeded body phantom [z][y][x] with labels of BODY_MATERIALS.

    An elliptic cylinder along z (semi-axes body_axes, mm) of soft tissue in
    air, then n_ellipsoids axis-aligned ellipsoids with semi-axes uniform in
    semi_axis_range, centres uniform inside the body and labels drawn from
    lung / adipose / soft tissue / cortical bone, painted in order (later
    ones overwrite earlier ones) and clipped to the body.
    """
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    xs = (np.arange(nx) + 0.5) * voxel - nx * voxel / 2
    ys = (np.arange(ny) + 0.5) * voxel - ny * voxel / 2
    zs = (np.arange(nz) + 0.5) * voxel - nz * voxel / 2
    a, b = body_axes
    body2d = (xs[None, :] / a) ** 2 + (ys[:, None] / b) ** 2 <= 1.0
    lab = np.zeros(shape, np.uint8)
    lab[:, body2d] = 3
    for _ in range(n_ellipsoids):
        while True:
            cx, cy = rng.uniform(-a, a), rng.uniform(-b, b)
            if (cx / a) ** 2 + (cy / b) ** 2 <= 1.0:
                break
        cz = rng.uniform(zs[0], zs[-1])
        r = rng.uniform(*semi_axis_range, size=3)
        material = int(rng.integers(1, 5))
        ix = np.nonzero(np.abs(xs - cx) <= r[0])[0]
        iy = np.nonzero(np.abs(ys - cy) <= r[1])[0]
        iz = np.nonzero(np.abs(zs - cz) <= r[2])[0]
        if ix.size == 0 or iy.size == 0 or iz.size == 0:
            continue
        sl = (slice(iz[0], iz[-1] + 1), slice(iy[0], iy[-1] + 1), slice(ix[0], ix[-1] + 1))
        q = (((xs[sl[2]] - cx) / r[0]) ** 2)[None, None, :] + (((ys[sl[1]] - cy) / r[1]) ** 2)[None, :, None] + \
            (((zs[sl[0]] - cz) / r[2]) ** 2)[:, None, None]
        inside = (q <= 1.0) & body2d[sl[1], sl[2]][None, :, :]
        lab[sl][inside] = material
    return lab


def kramers_spectrum(kvp=120.0, e_min_kev=20.0, al_mm=2.5, mu_al_per_cm=None):
    """Kramers bremsstrahlung I(E) ~ (kVp - E)/E on 1 keV bins, times
    exp(-mu_Al(E) * t).  mu_al_per_cm(E_MeV) gives the aluminium linear
    attenuation coefficient in cm^-1.  Returns (energies MeV, intensities)."""
    e = np.arange(e_min_kev, kvp + 0.5, 1.0)
    w = np.clip(kvp - e, 0.0, None) / e
    if al_mm > 0 and mu_al_per_cm is not None:
        mu = np.array([mu_al_per_cm(v * 1e-3) for v in e])
        w = w * np.exp(-mu * al_mm * 0.1)
    return e * 1e-3, w


def write_spectrum(path, energies_mev, intensities):
    with open(path, "w") as f:
        for e, w in zip(energies_mev, intensities):
            f.write("%.6f %.9g\n" % (e, w))
    return path


def icosphere(radius, subdivisions):
    """Closed triangle mesh of a sphere: an icosahedron whose faces are split
    in four `subdivisions` times, vertices pushed onto the sphere.  Returns
    float64 triangles [20 * 4^subdivisions][3][3], outward counter-clockwise."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, np.int64)
    for _ in range(subdivisions):
        cache = {}
        vlist = list(verts)

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (vlist[a] + vlist[b]) * 0.5
                vlist.append(m / np.linalg.norm(m))
                cache[key] = len(vlist) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        verts = np.array(vlist)
        faces = np.array(nf, np.int64)
    return verts[faces] * radius


def write_stl(path, triangles, header=b"ggems_amd mesh"):
    """Binary STL (80-byte header, uint32 count, per triangle: normal, three
    vertices, uint16 attribute), the format GGEMSSTLReader::Read reads
    (src/io/GGEMSSTLReader.cc:71-99)."""
    tri = np.asarray(triangles, np.float64).reshape(-1, 3, 3)
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    rec = np.zeros(len(tri), dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    rec["n"] = n
    rec["v"] = tri
    with open(path, "wb") as fh:
        fh.write(header[:80].ljust(80, b"\0"))
        fh.write(np.uint32(len(tri)).tobytes())
        fh.write(rec.tobytes())
    return path


def read_stl(path):
    """float32 triangles [n][3][3] of a binary STL file"""
    with open(path, "rb") as fh:
        fh.read(80)
        n = int(np.frombuffer(fh.read(4), "<u4")[0])
        rec = np.frombuffer(fh.read(50 * n), dtype=[("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")])
    return np.array(rec["v"], np.float32)


MATERIALS_TXT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "_gen", "materials.txt")


def aluminium_mu(table_bins=220, emin_kev=1.0, emax_mev=1.0, materials_txt=MATERIALS_TXT):
    """mu(E) of aluminium (cm^-1) from the transport's own cross-section tables
    (Compton + photoelectric + Rayleigh, sigma/rho interpolated, times density).

    GGEMSAttenuations is not used: its lookup searches the concatenated
    energy lists of all lighter elements (src/physics/GGEMSAttenuations.cc:201-212),
    which that API reproduces as is but which is no filtration model."""
    from .ggems import GGEMSCrossSections, GGEMSMaterials, GGEMSMaterialsDatabaseManager, GGEMSProcessesManager
    GGEMSMaterialsDatabaseManager().set_materials(materials_txt)  # no-op when already loaded
    pm = GGEMSProcessesManager()
    pm.set_cross_section_table_number_of_bins(table_bins)
    pm.set_cross_section_table_energy_min(emin_kev, "keV")
    pm.set_cross_section_table_energy_max(emax_mev, "MeV")
    mats = GGEMSMaterials()
    mats.add_material("Aluminium")
    mats.initialize()
    cs = GGEMSCrossSections(mats)
    for p in ("Compton", "Photoelectric", "Rayleigh"):
        cs.add_process(p, "gamma")
    cs.initialize()
    rho = mats.get_density("Aluminium")

    def mu(e_mev):
        return rho * sum(cs.get_cs(p, "Aluminium", float(e_mev), "MeV") for p in ("Compton", "Photoelectric",
                                                                                  "Rayleigh"))
    mu._keep = (mats, cs)
    return mu
