import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "tests"))


ARTIFACTS = [os.path.join(REPO, "ggems_amd", "lib", "libggems_engine.so"),
             os.path.join(REPO, "ggems_amd", "lib", "libggems.so"),
             os.path.join(REPO, "ggems_amd", "data", "_gen", "ggems_physics.bin"),
             os.path.join(REPO, "oracle", "_build", "libggems_oracle.so")]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    # a fresh checkout: build the libraries, the data blob and the oracle once
    if not all(os.path.exists(p) for p in ARTIFACTS):
        from ggems_amd import build
        build.build_all(oracle=True)


def pytest_runtest_setup(item):
    # flush the case name before any GPU step so a lost box still names it
    if "gpu" in item.keywords:
        print("\n[gpu-test] %s" % item.nodeid, flush=True)
