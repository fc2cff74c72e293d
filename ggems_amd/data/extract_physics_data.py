"""Build step: extract GGEMS's physics data tables into a binary blob.

GGEMS ships its photon physics data (Sandia photoabsorption coefficients,
Livermore Rayleigh cross sections, Rayleigh form-factor fit parameters,
NIST-style mu/mu_en data and the chemical-element table) as C array
literals inside its source tree.  Those numbers are physical data that the
transport must share bit-for-bit with GGEMS, so this script reads them from
a GGEMS source tree at build time and writes them to
``ggems_amd/data/_gen/ggems_physics.bin``.  The blob is a build artefact:
it is never committed, and nothing reads the GGEMS tree at run time.

Blob layout (little endian):
  magic  b"GGPD", u32 version (=1), u32 n_entries
  n_entries x { char name[48]; u32 dtype (0=f32, 1=i32, 2=f64); u32 pad;
                u64 count; u64 byte_offset }
  payload (each array 16-byte aligned)

Sources read (paths relative to the GGEMS root):
  include/GGEMS/physics/GGEMSSandiaTable.hh           (kIonizationPotentials,
        kNumberOfIntervals, kCumulativeIntervals, kZtoARatio, kSandiaTable)
  include/GGEMS/physics/GGEMSRayleighScattering.hh    (kCrossSection* tables)
  include/GGEMS/physics/GGEMSRayleighScatteringModels.hh:36-153 (FACTOR, PP0..PP8)
  include/GGEMS/physics/GGEMSMuDataConstants.hh       (kMuNbEnergyBins,
        kMuIndexEnergy, kMuData)
  src/materials/GGEMSMaterialsDatabaseManager.cc:337-437 (AddChemicalElements)
  data/materials.txt                                  (copied verbatim next to the blob)
"""

import ctypes
import os
import re
import shutil
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_gen")
BLOB = os.path.join(OUT_DIR, "ggems_physics.bin")

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
_libc.strtod.restype = ctypes.c_double
_libc.strtod.argtypes = [ctypes.c_char_p, ctypes.c_void_p]

_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?[fF]?")


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _f32(tok):
    """Decimal literal -> float32 with a single rounding (as a C compiler does)."""
    return float(_libc.strtof(tok.rstrip("fF").encode(), None))


def _array(text, name):
    """Return the numeric tokens of the brace initialiser of array `name`.

    Elements the initialiser leaves out are zero, as in C, so the result is
    padded with "0" up to the declared size.
    """
    m = re.search(r"\b" + re.escape(name) + r"\s*((?:\[\s*\d*\s*\])*)\s*=\s*\{", text)
    if not m:
        raise KeyError(name)
    depth, i = 1, m.end()
    while depth:
        c = text[i]
        depth += (c == "{") - (c == "}")
        i += 1
    toks = _NUM.findall(text[m.end():i - 1])
    dims = [int(d) for d in re.findall(r"\d+", m.group(1) or "")]
    declared = 1
    for d in dims:
        declared *= d
    if dims and len(toks) > declared:
        raise ValueError("%s: %d initialisers for %d elements" % (name, len(toks), declared))
    return toks + ["0"] * (declared - len(toks) if dims else 0)


def _scalar(text, name):
    m = re.search(r"\b" + re.escape(name) + r"\s*=\s*(" + _NUM.pattern + r")", text)
    if not m:
        raise KeyError(name)
    return m.group(1)


def extract(ggems_root):
    def src(rel):
        with open(os.path.join(ggems_root, rel)) as f:
            return _strip_comments(f.read())

    arrays = {}
    sandia = src("include/GGEMS/physics/GGEMSSandiaTable.hh")
    arrays["sandia_ionization_potential"] = ("f32", [_f32(t) for t in _array(sandia, "kIonizationPotentials")])
    arrays["sandia_number_of_intervals"] = ("i32", [int(t) for t in _array(sandia, "kNumberOfIntervals")])
    arrays["sandia_cumulative_intervals"] = ("i32", [int(t) for t in _array(sandia, "kCumulativeIntervals")])
    arrays["sandia_z_to_a_ratio"] = ("f32", [_f32(t) for t in _array(sandia, "kZtoARatio")])
    arrays["sandia_table"] = ("f32", [_f32(t) for t in _array(sandia, "kSandiaTable")])

    ray = src("include/GGEMS/physics/GGEMSRayleighScattering.hh")
    arrays["rayleigh_cs_cumulative_intervals"] = ("i32", [int(t) for t in _array(ray, "kCrossSectionCumulativeIntervals")])
    arrays["rayleigh_cs_number_of_intervals"] = ("i32", [int(t) for t in _array(ray, "kCrossSectionNumberOfIntervals")])
    arrays["rayleigh_cross_section"] = ("f32", [_f32(t) for t in _array(ray, "kCrossSection")])

    model = src("include/GGEMS/physics/GGEMSRayleighScatteringModels.hh")
    arrays["rayleigh_factor"] = ("f32", [_f32(_scalar(model, "FACTOR"))])
    for k in range(9):
        arrays["rayleigh_pp%d" % k] = ("f32", [_f32(t) for t in _array(model, "PP%d" % k)])

    mu = src("include/GGEMS/physics/GGEMSMuDataConstants.hh")
    arrays["mu_number_of_energy_bins"] = ("i32", [int(t) for t in _array(mu, "kMuNbEnergyBins")])
    arrays["mu_index_energy"] = ("i32", [int(t) for t in _array(mu, "kMuIndexEnergy")])
    arrays["mu_data"] = ("f32", [_f32(t) for t in _array(mu, "kMuData")])

    db = src("src/materials/GGEMSMaterialsDatabaseManager.cc")
    rows = re.findall(r'AddChemicalElements\(\s*"([A-Za-z]+)"\s*,\s*(\d+)\s*,\s*(' + _NUM.pattern + r')\s*,\s*('
                      + _NUM.pattern + r')\s*,\s*(GAS|SOLID)', db)
    if len(rows) < 100:
        raise RuntimeError("chemical element table not found (%d rows)" % len(rows))
    names = "\n".join(r[0] for r in rows)
    arrays["element_z"] = ("i32", [int(r[1]) for r in rows])
    arrays["element_molar_mass"] = ("f32", [_f32(r[2]) for r in rows])
    arrays["element_mean_excitation"] = ("f32", [_f32(r[3]) for r in rows])
    arrays["element_is_gas"] = ("i32", [1 if r[4] == "GAS" else 0 for r in rows])
    arrays["element_names"] = ("i8", list(names.encode()) + [0])

    # consistency checks on shapes the transport relies on
    assert len(arrays["sandia_table"][1]) == 981 * 5
    assert len(arrays["rayleigh_cross_section"][1]) == 213816
    assert len(arrays["mu_data"][1]) == 10401
    for k in range(9):
        assert len(arrays["rayleigh_pp%d" % k][1]) == 101
    return arrays


def write_blob(arrays, path):
    codes = {"f32": (0, "f", 4), "i32": (1, "i", 4), "f64": (2, "d", 8), "i8": (3, "b", 1)}
    header = struct.pack("<4sII", b"GGPD", 1, len(arrays))
    table_size = len(arrays) * (48 + 4 + 4 + 8 + 8)
    offset = (len(header) + table_size + 15) & ~15
    table, payload = b"", b""
    for name, (dt, vals) in arrays.items():
        code, fmt, _ = codes[dt]
        data = struct.pack("<%d%s" % (len(vals), fmt), *vals)
        table += struct.pack("<48sIIQQ", name.encode(), code, 0, len(vals), offset + len(payload))
        payload += data + b"\0" * ((-len(data)) % 16)
    blob = header + table
    blob += b"\0" * (offset - len(blob)) + payload
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(blob)
    os.replace(tmp, path)


def main(ggems_root=None):
    ggems_root = ggems_root or os.environ.get("GGEMS_SOURCE_DIR", "/root/reference")
    if not os.path.isdir(os.path.join(ggems_root, "include", "GGEMS")):
        if os.path.exists(BLOB):
            print("GGEMS source tree not found; keeping existing", BLOB)
            return BLOB
        raise RuntimeError("GGEMS source tree not found at %r (set GGEMS_SOURCE_DIR)" % ggems_root)
    os.makedirs(OUT_DIR, exist_ok=True)
    write_blob(extract(ggems_root), BLOB)
    shutil.copyfile(os.path.join(ggems_root, "data", "materials.txt"), os.path.join(OUT_DIR, "materials.txt"))
    return BLOB


if __name__ == "__main__":
    print(main(sys.argv[1] if len(sys.argv) > 1 else None))
