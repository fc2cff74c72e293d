"""Writes reference_exports.txt: every extern "C" function that GGEMS
declares under include/, one per line as

    header name return_type arg_type,arg_type,...

with the C types written in a canonical form (signatures.canonical_type):
GGEMS's typedefs resolved to the C types they stand for, const and argument
names dropped, and pointers to GGEMS's classes written as `handle*`, since
the C ABI only passes them through.

Run in a container that holds the reference tree:
    python tests/golden/make_reference_exports.py /root/reference
"""

import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from signatures import parse_declaration  # noqa: E402


def exports(root):
    out = []
    inc = os.path.join(root, "include")
    for dirpath, _, files in os.walk(inc):
        for f in sorted(files):
            if not f.endswith(".hh"):
                continue
            path = os.path.join(dirpath, f)
            text = open(path, errors="replace").read()
            text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
            text = re.sub(r"//[^\n]*", " ", text)
            for m in re.finditer(r'extern\s+"C"\s+(?:GGEMS_EXPORT\s+)?([^;{]*?\))\s*;', text, flags=re.S):
                name, ret, args = parse_declaration(m.group(1))
                out.append((os.path.relpath(path, inc), name, ret, args))
    return sorted(set(out))


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    rows = exports(root)
    with open(os.path.join(HERE, "reference_exports.txt"), "w") as f:
        for h, n, r, a in rows:
            f.write("%s %s %s %s\n" % (h, n, r, ",".join(a) if a else "void"))
    print(len(rows), "exports")
