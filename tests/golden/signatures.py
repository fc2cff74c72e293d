"""Canonical C signatures, so that a GGEMS header declaration and the
matching declaration of include/ggems.h can be compared for ABI identity
(argument count, order and C type; return type), whatever the typedef
names, const qualifiers, argument names or default arguments."""

import re

# GGEMS's scalar typedefs (include/GGEMS/tools/GGEMSTypes.hh) and the C
# types they stand for
SCALARS = {
    "GGfloat": "float", "float": "float",
    "GGdouble": "double", "double": "double",
    "GGint": "int32", "int": "int32", "int32_t": "int32", "GGlong": "int64", "int64_t": "int64",
    "GGuint": "uint32", "unsigned int": "uint32", "uint32_t": "uint32", "unsigned": "uint32",
    "GGsize": "size", "size_t": "size",
    "GGshort": "int16", "short": "int16", "int16_t": "int16",
    "GGushort": "uint16", "unsigned short": "uint16", "uint16_t": "uint16",
    "GGchar": "int8", "signed char": "int8", "int8_t": "int8",
    "GGuchar": "uint8", "unsigned char": "uint8", "uint8_t": "uint8",
    "char": "char", "bool": "bool", "void": "void",
    "GGbool": "uint32", "cl_bool": "uint32",  # cl_bool is cl_uint
}


def canonical_type(t):
    """'GGfloat const' -> 'float'; 'char const*' -> 'char*'; 'GGEMSBox*' -> 'handle*'."""
    t = re.sub(r"\bconst\b", " ", t)
    t = re.sub(r"\bstruct\b", " ", t)
    stars = t.count("*")
    refs = t.count("&")
    base = " ".join(t.replace("*", " ").replace("&", " ").split())
    if base in SCALARS:
        base = SCALARS[base]
    elif stars:
        base = "handle"  # an opaque object the C ABI only passes through
    return base + "*" * stars + "&" * refs


def _split_args(args):
    out, depth, cur = [], 0, ""
    for ch in args:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def _drop_name(arg):
    arg = arg.split("=")[0].strip()  # default argument
    m = re.match(r"^(.*?[\s*&])([A-Za-z_]\w*)\s*(\[\s*\d*\s*\])?$", arg)
    if m and m.group(1).strip() and m.group(2) not in ("const",):
        base = m.group(1)
        if m.group(3):
            base += "*"
        if base.strip() in ("unsigned", "signed", "short", "long") or base.strip().endswith(("unsigned", "signed")):
            return arg  # 'unsigned int' without a name: the last word is part of the type
        return base
    return arg


def parse_declaration(decl):
    """'ret name(args)' -> (name, canonical return type, tuple of canonical arg types)."""
    decl = " ".join(decl.split())
    m = re.match(r"^(.*?)\b([A-Za-z_]\w*)\s*\((.*)\)$", decl)
    if not m:
        raise ValueError(decl)
    ret, name, args = m.groups()
    arg_types = []
    for a in _split_args(args):
        a = a.strip()
        if a in ("", "void"):
            continue
        arg_types.append(canonical_type(_drop_name(a)))
    return name, canonical_type(ret), tuple(arg_types)
