"""The C-ABI libraries load on a machine without a GPU and export every
function their headers declare (no compute call is made)."""

import ctypes as C
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(REPO, "ggems_amd", "lib")
HEADERS = {"ggems.h": "libggems.so", "ggems_engine.h": "libggems_engine.so"}


def declared_functions(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"#[^\n]*", " ", text)
    text = re.sub(r"typedef\s+struct\s*\w*\s*\{.*?\}\s*\w+\s*;", " ", text, flags=re.S)
    text = re.sub(r"\bstruct\s+\w+\s*\{.*?\}\s*;", " ", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{}()]*(?:\([^()]*\)[^;{}()]*)*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return", "sizeof")))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_header_symbols_exported(header):
    names = declared_functions(header)
    assert len(names) > 10
    lib = C.CDLL(os.path.join(LIBDIR, HEADERS[header]))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


# GGEMS headers whose exports are not part of this library, and why
NOT_EXPORTED = {
    "GGEMS/graphics/GGEMSOpenGLManager.hh": "OpenGL display, out of scope",
    "GGEMS/navigators/GGEMSWorld.hh": "world-grid photon tracking, outside the transport hot path",
    "GGEMS/sources/GGEMSVoxelizedSource.hh": "voxelized (emission) source, outside the transport hot path",
}


def _reference_rows():
    rows = []
    for line in open(os.path.join(REPO, "tests", "golden", "reference_exports.txt")):
        if line.strip():
            h, n, ret, args = line.split()
            rows.append((h, n, ret, () if args == "void" else tuple(args.split(","))))
    return rows


def declared_signatures(header):
    """name -> (canonical return type, canonical argument types) of every
    function declared in include/<header>"""
    from signatures import parse_declaration
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"#[^\n]*", " ", text)
    text = re.sub(r"typedef[^;{]*(\{.*?\})?[^;]*;", " ", text, flags=re.S)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?\b[A-Za-z_]\w*\s*\([^;{}]*\))\s*;", text):
        name, ret, args = parse_declaration(m.group(1))
        out[name] = (ret, args)
    return out


def test_ggems_header_covers_the_reference_exports():
    """Every extern "C" function of the GGEMS headers (listed in
    tests/golden/reference_exports.txt) is declared in include/ggems.h and
    exported, except the headers listed in NOT_EXPORTED."""
    ours = set(declared_functions("ggems.h"))
    rows = _reference_rows()
    assert len(rows) == 177
    missing = [(h, n) for h, n, _, _ in rows if h not in NOT_EXPORTED and n not in ours]
    assert not missing, missing
    assert sum(1 for h, _, _, _ in rows if h not in NOT_EXPORTED) >= 118


def test_ggems_header_signatures_match_the_reference():
    """Each of those functions has the reference's C signature: return type,
    argument count, order and types, with GGEMS's typedefs resolved
    (GGfloat = float, GGsize = size_t, ...) and class pointers as opaque
    handles, so that GGEMS's own ctypes bindings (python_module/*.py) call
    this library unchanged."""
    import sys
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    ours = declared_signatures("ggems.h")
    checked = 0
    for h, n, ret, args in _reference_rows():
        if h in NOT_EXPORTED:
            continue
        assert ours[n] == (ret, args), (h, n, ours[n], (ret, args))
        checked += 1
    assert checked >= 118


def test_engine_reports_errors_without_a_device():
    lib = C.CDLL(os.path.join(LIBDIR, "libggems_engine.so"))
    lib.ggh_last_error.restype = C.c_char_p
    assert isinstance(lib.ggh_last_error(), bytes)


def test_unet_entry_points_validate_before_touching_the_device():
    """the U-Net's C entry points refuse null pointers (-1) and shapes their
    kernels do not take (-2), with a message, before any device call: the
    pointers below are never dereferenced"""
    lib = C.CDLL(os.path.join(LIBDIR, "libggems_engine.so"))
    lib.ggh_unet_last_error.restype = C.c_char_p
    p = C.c_void_p(256)
    i = C.c_int32
    sigs = {
        "ggh_conv3x3_forward": [C.c_void_p] * 4 + [i] * 6 + [C.c_void_p],
        "ggh_conv3x3_weight_grad": [C.c_void_p] * 3 + [i] * 6 + [C.c_void_p],
        "ggh_relu_bias_backward": [C.c_void_p, i, C.c_void_p, i, C.c_void_p, C.c_void_p, C.c_int64, i, C.c_void_p],
        "ggh_pool2x2_forward": [C.c_void_p] * 3 + [i] * 4 + [C.c_void_p],
        "ggh_pool2x2_backward": [C.c_void_p] * 3 + [i] * 4 + [C.c_void_p],
        "ggh_upcat_forward": [C.c_void_p] * 3 + [i] * 5 + [C.c_void_p],
        "ggh_upcat_backward": [C.c_void_p] * 3 + [i] * 5 + [C.c_void_p],
    }
    for name, args in sigs.items():
        getattr(lib, name).argtypes = args
        getattr(lib, name).restype = C.c_int
    cases = [
        (lib.ggh_conv3x3_forward(None, p, None, p, 1, 32, 32, 8, 16, 0, None), -1),
        (lib.ggh_conv3x3_forward(p, p, None, p, 1, 32, 24, 8, 16, 0, None), -2),     # W % 16
        (lib.ggh_conv3x3_forward(p, p, None, p, 1, 32, 32, 12, 16, 0, None), -2),    # Cin % 8
        (lib.ggh_conv3x3_weight_grad(p, p, p, 1, 32, 32, 8, 256, 0, None), -2),      # Cout > 128
        (lib.ggh_relu_bias_backward(p, 0, p, 1, p, p, 100, 24, None), -2),           # C not a power of two
        (lib.ggh_relu_bias_backward(p, 0, None, 1, p, p, 100, 16, None), -1),        # ReLU without y
        (lib.ggh_pool2x2_forward(p, p, None, 1, 4, 4, 8, None), -1),
        (lib.ggh_pool2x2_forward(p, p, p, 1, 3, 4, 8, None), -2),                    # odd height
        (lib.ggh_pool2x2_backward(p, p, p, 1, 4, 4, 12, None), -2),                  # C % 8
        (lib.ggh_upcat_forward(p, p, p, 1, 4, 6, 8, 12, None), -2),                  # Cs % 8
        (lib.ggh_upcat_backward(p, None, p, 1, 4, 4, 8, 8, None), -1),
        (lib.ggh_upcat_backward(p, p, p, 1, 4, 5, 8, 8, None), -2),                  # odd width
    ]
    for got, want in cases:
        assert got == want, (got, want)
        assert lib.ggh_unet_last_error()
