"""The voxel-CT workload at its full size (needs an MI355X): the 512x512x256
body phantom of 0.8 mm voxels with FP64 dosimetry, the 120 kVp spectrum and
the flat 256x256 detector.

A run of a few hundred thousand primaries is compared with the oracle on the
same seeded inputs, history by history as in the small-phantom tests.  The
full step of 10^9 primaries is checked through properties that do not need a
whole oracle run: the launch shape leaves every history unchanged, and the
per-primary figures agree with the oracle's run of the step's first 2048
particle slots (about 1.9 million of its histories) within their statistical
errors.
"""

import os

import numpy as np
import pytest

import parity
from scene import REPO, voxel_ct_scene

pytestmark = pytest.mark.gpu

WORKDIR = os.path.join(REPO, "ggems_amd", "data", "_gen", "voxel_ct")


@pytest.fixture(scope="module")
def body():
    return voxel_ct_scene(WORKDIR)


def test_full_phantom_run_matches_oracle(body):
    n = 300_000
    ref = body.run_oracle(n, 777, threads=0)
    got = body.run_engine(n, 777)
    rs, gs = ref["stats"], got["stats"]
    assert gs["primaries"] == rs["primaries"] == n
    assert rs["voxel_steps"] > 100 * n  # the photons cross the body
    parity.counters("full phantom", rs, gs, ("voxel_steps", "interactions", "detector_interactions", "hops"), 2e-3)
    h_ref, h_got = ref["hist"][0], got["hist"][0]
    assert h_ref.sum() > 100
    # histories differ only where an ulp of a device transcendental flipped one
    parity.l1("full phantom histogram", h_ref, h_got, 5e-3, 10)
    parity.total("full phantom edep", ref["edep"], got["edep"], 1e-3)
    parity.total("full phantom edep2", ref["edep2"], got["edep2"], 1e-3)
    parity.l1("full phantom hit", ref["hit"], got["hit"], 1e-2)


def test_full_step_launch_invariance_and_oracle_sample(body):
    n = 1_000_000_000
    eng = body.make_engine()
    got = body.run_engine(n, 777, eng)
    gs = got["stats"]
    assert gs["primaries"] == n
    # another launch shape runs the same histories
    eng.set_launch(round_steps=32, interact_batch=8, cold_passes=1)
    again = body.run_engine(n, 777, eng)
    assert again["stats"] == gs
    assert np.array_equal(again["hist"][0], got["hist"][0]) and np.array_equal(again["hit"], got["hit"])
    np.testing.assert_allclose(again["edep"], got["edep"], rtol=1e-12, atol=1e-18)
    eng.close()

    # the oracle's run of the first 2048 slots of the same step: a sample of
    # the step's own histories, so the per-primary figures agree to within
    # sampling error
    ref = body.run_oracle(n, 777, threads=0, slots=(0, 2048))
    rs = ref["stats"]
    m = rs["primaries"]
    assert m == 2048 * (n // 1048576) + min(2048, n % 1048576)
    # step and interaction counts per primary: a relative standard error of
    # about 1e-3 over m histories; 1 % is ten of them
    for k in ("voxel_steps", "interactions"):
        assert abs(gs[k] / n - rs[k] / m) <= 0.01 * rs[k] / m, (k, gs[k] / n, rs[k] / m)
    # detector counts: Poisson error of the sample's count
    c_ref, c_got = int(ref["hist"][0].sum()), int(got["hist"][0].sum())
    assert c_ref > 1000
    assert abs(c_got / n - c_ref / m) <= 5.0 * np.sqrt(c_ref) / m, (c_got / n, c_ref / m)
    # deposited energy per primary
    e_ref, e_got = ref["edep"].sum() / m, got["edep"].sum() / n
    assert abs(e_got - e_ref) <= 0.02 * e_ref, (e_got, e_ref)


def test_dose_and_projection_statistics_against_oracle(body):
    """Per-voxel agreement of the dose map and per-pixel agreement of the
    projection with independent seeds (z-score and chi-square checks of
    GGEMSDosimetryCalculator's outputs,
    src/navigators/GGEMSDosimetryCalculator.cc:536-809, and of the detector
    histogram, src/kernels/TrackThroughGGEMSSolidBox.cl:215-225).  10^7
    primaries each: engine seed 1001 against the oracle's seed 1002, and, as
    the null reference for the same statistic, engine seed 1001 against
    engine seed 1003."""
    n = 10_000_000
    eng = body.make_engine()
    a = body.run_engine(n, 1001, eng)
    c = body.run_engine(n, 1003, eng)
    eng.close()
    ref = body.run_oracle(n, 1002, threads=0)
    parity.same_distribution("dose blocks 8^3, engine vs oracle (null: engine vs engine)",
                             parity.dose_z(a, ref), parity.dose_z(a, c), 2000)
    parity.same_distribution("projection 8x8 pixel blocks, engine vs oracle (null: engine vs engine)",
                             parity.pixel_z(a, ref), parity.pixel_z(a, c), 300)
