import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def pytest_runtest_setup(item):
    # flush the case name before any GPU step so a lost box still names it
    if "gpu" in item.keywords:
        print("\n[gpu-test] %s" % item.nodeid, flush=True)
