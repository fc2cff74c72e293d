"""Writes reference_exports.txt: the names of the extern "C" functions that
GGEMS declares under include/ (one "header name" pair per line).

Run in a container that holds the reference tree:
    python tests/golden/make_reference_exports.py /root/reference
"""

import os
import re
import sys


def exports(root):
    out = []
    inc = os.path.join(root, "include")
    for dirpath, _, files in os.walk(inc):
        for f in sorted(files):
            if not f.endswith(".hh"):
                continue
            path = os.path.join(dirpath, f)
            text = open(path, errors="replace").read()
            for m in re.finditer(r'extern\s+"C"\s+(?:GGEMS_EXPORT\s+)?[^;(]*?\b(\w+)\s*\(', text):
                out.append((os.path.relpath(path, inc), m.group(1)))
    return sorted(set(out))


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    rows = exports(root)
    with open(os.path.join(here, "reference_exports.txt"), "w") as f:
        for h, n in rows:
            f.write("%s %s\n" % (h, n))
    print(len(rows), "exports")
