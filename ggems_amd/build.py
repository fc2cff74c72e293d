"""In-tree build of the native libraries.

  libggems_engine.so  HIP transport engine for gfx950 (include/ggems_engine.h)
  libggems.so         GGEMS-compatible C++ host API (include/ggems.h), linked to the engine

Both land in ggems_amd/lib/.  The physics data blob is extracted first
(ggems_amd/data/extract_physics_data.py).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GGEMS_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps a*b+c as two IEEE roundings, as the reference's
# arithmetic is restated; packed FP32 math is disabled on this pool.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

ENGINE_SOURCES = [os.path.join(CSRC, "engine", f) for f in ("engine.hip", "volume.hip", "unet.hip")]
HOST_SOURCES = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host"))
                      if f.endswith(".cc")) if os.path.isdir(os.path.join(CSRC, "host")) else []


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libggems_engine.so")
    _run([HIPCC, "--offload-arch=" + ARCH] + HIP_FLAGS + ["-shared", "-o", out + ".tmp"] + ENGINE_SOURCES)
    os.replace(out + ".tmp", out)
    return out


def build_host():
    if not HOST_SOURCES:
        return None
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libggems.so")
    cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-shared",
           "-I", os.path.join(REPO, "include"), "-o", out + ".tmp"] + HOST_SOURCES + \
          ["-L", LIB, "-lggems_engine", "-Wl,-rpath,$ORIGIN", "-ldl", "-lpthread"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def extract_data():
    sys.path.insert(0, os.path.join(HERE, "data"))
    import extract_physics_data
    return extract_physics_data.main()


def build_oracle():
    subprocess.check_call(["make", "-s", "-B", "-C", os.path.join(REPO, "oracle")])


def build_all(oracle=True):
    extract_data()
    build_engine()
    build_host()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build_all()
