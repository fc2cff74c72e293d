"""Tau sampling on the GPU (needs an MI355X): the engine against the oracle's
tau mode history by history (both consume one exponential per flight from
the same streams), on every kernel variant that has the mode; the launch
shape leaves histories unchanged; an acquisition equals its per-projection
launches; and at the voxel-CT workload's full size the dose map and the
projection agree with the reference sampling as independent samples."""

import os

import numpy as np
import pytest

import parity
from scene import REPO, Scene, cfg1_scene, cylinder_phantom, voxel_ct_scene

pytestmark = pytest.mark.gpu


def _tau(sc):
    sc.tau = True
    return sc


@pytest.mark.parametrize("general", [False, True])
def test_tau_cfg1_matches_oracle(general):
    sc = _tau(cfg1_scene(n_side=100, detector=256))
    n = 200_000
    ref = sc.run_oracle(n, 777, threads=0)
    eng = sc.make_engine()
    eng.set_launch(general_kernel=general)
    got = sc.run_engine(n, 777, eng)
    assert eng.last_grid()[2] == (1 if general else 0)
    rs, gs = ref["stats"], got["stats"]
    assert gs["primaries"] == rs["primaries"] == n
    tag = "tau cfg1 general=%s" % general
    parity.counters(tag, rs, gs, ("voxel_steps", "interactions", "detector_interactions", "hops"), 1e-3)
    parity.l1(tag + " histogram", ref["hist"][0], got["hist"][0], 1e-3, 10)
    parity.l1(tag + " scatter", ref["scatter"][0], got["scatter"][0], 5e-3, 10)
    parity.total(tag + " edep", ref["edep"], got["edep"], 1e-3)
    parity.l1(tag + " hit", ref["hit"], got["hit"], 2e-3, 10)
    eng.close()


def test_tau_two_phantoms_and_photon_tracking_match_oracle():
    """the several-solid three-process variant (two phantoms, one rotated)
    and the general variant with fluence tracking"""
    sc = Scene()
    sc.add_phantom(cylinder_phantom(40), ["Water", "RibBone"], (1.0, 1.0, 1.0), position=(0.0, -45.0, 0.0))
    sc.add_phantom(cylinder_phantom(32), ["Water", "RibBone"], (1.25, 1.25, 1.25), position=(0.0, 45.0, 0.0),
                   rotation=(0.0, 0.0, 30.0), dosimetry={"photon_tracking": True})
    sc.add_ct_flat("GOS", (1, 1), (128, 128, 1), (1.5, 1.5, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 8.0, [0.06], mono=True)
    _tau(sc.build())
    n = 200_000
    ref = sc.run_oracle(n, 31, threads=0)
    eng = sc.make_engine()
    got = sc.run_engine(n, 31, eng)
    assert eng.last_grid()[2] == 1  # fluence tracking runs the general variant
    parity.counters("tau two phantoms", ref["stats"], got["stats"], ("voxel_steps", "interactions", "hops"), 1e-3)
    parity.l1("tau two phantoms histogram", ref["hist"][0], got["hist"][0], 1e-3, 10)
    parity.l1("tau two phantoms fluence", ref["photon_tracking"], got["photon_tracking"], 1e-3, 10)
    parity.total("tau two phantoms edep", ref["edep"], got["edep"], 1e-3)
    eng.close()
    sc.navigators[1][2]["dosimetry"] = None  # no tally: the three-process several-solid variant
    ref = sc.run_oracle(n, 32, threads=0)
    eng = sc.make_engine()
    got = sc.run_engine(n, 32, eng)
    assert eng.last_grid()[2] == 2
    parity.counters("tau two phantoms variant 2", ref["stats"], got["stats"], ("voxel_steps", "interactions", "hops"),
                    1e-3)
    parity.l1("tau two phantoms variant 2 histogram", ref["hist"][0], got["hist"][0], 1e-3, 10)
    eng.close()


def test_tau_ragged_phantom_matches_oracle():
    """37 x 50 x 23 voxels of 1.1 x 0.7 x 1.6 mm, rotated: the tau step reads
    the labels from 8 x 4 x 4-voxel bricks, and every axis ends in a partly
    filled brick here (5, 13 and 6 bricks), so a wrong brick stride or
    padding changes histories against the oracle, which reads the labels
    row-major.  Three materials in seeded blobs make every label matter."""
    g = np.random.default_rng(11)
    lab = np.zeros((23, 50, 37), np.uint8)
    zz, yy, xx = np.meshgrid(np.arange(23), np.arange(50), np.arange(37), indexing="ij")
    for _ in range(12):
        c = g.uniform([0, 0, 0], [23, 50, 37])
        r = g.uniform(3.0, 9.0)
        lab[(zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2 < r * r] = g.integers(1, 3)
    sc = Scene()
    sc.add_phantom(lab, ["Water", "RibBone", "Lung"], (1.1, 0.7, 1.6), rotation=(10.0, 0.0, 25.0),
                   dosimetry={"photon_tracking": False})
    sc.add_ct_flat("GOS", (1, 1), (96, 96, 1), (1.0, 1.0, 1.0), 300.0, 500.0, 10e-3)
    sc.set_source((-300.0, 0.0, 0.0), 4.0, [0.05], mono=True)
    _tau(sc.build())
    n = 200_000
    ref = sc.run_oracle(n, 4242, threads=0)
    eng = sc.make_engine()
    got = sc.run_engine(n, 4242, eng)
    assert eng.last_grid()[2] == 0  # the specialised kernel: the bricks of the only voxel solid
    assert ref["stats"]["voxel_steps"] > 10 * n
    parity.counters("tau ragged phantom", ref["stats"], got["stats"], ("voxel_steps", "interactions", "hops"), 1e-3)
    parity.l1("tau ragged phantom histogram", ref["hist"][0], got["hist"][0], 1e-3, 10)
    parity.l1("tau ragged phantom hit", ref["hit"], got["hit"], 2e-3, 10)
    parity.total("tau ragged phantom edep", ref["edep"], got["edep"], 1e-3)
    eng.close()


def test_tau_launch_shape_does_not_change_histories():
    sc = _tau(cfg1_scene(n_side=60, detector=64))
    eng = sc.make_engine()
    base = sc.run_engine(300_000, 5, eng)
    for shape in ({"round_steps": 1, "interact_batch": 1, "cold_passes": 1},
                  {"round_steps": 64, "interact_batch": 64, "cold_passes": 16}):
        eng.set_launch(**shape)
        again = sc.run_engine(300_000, 5, eng)
        assert again["stats"] == base["stats"], shape
        assert np.array_equal(again["hist"][0], base["hist"][0]) and np.array_equal(again["hit"], base["hit"])
        np.testing.assert_allclose(again["edep"], base["edep"], rtol=1e-12, atol=0.0)
    eng.close()


def test_tau_acquisition_matches_oracle_and_consecutive_launches():
    rot = [(0.0, 0.0, 0.0), (0.0, 0.0, 0.5), (0.0, 0.0, 1.4)]
    budget = [120_000, 1_200_000, 50_000]
    sc = _tau(cfg1_scene(n_side=60, detector=48, dosimetry=True, aperture_deg=2.0))
    eng = sc.engine_acquisition(rot, budget, 17)
    eng.run_acquisition()
    got = sc.read_acquisition(eng)
    ref = sc.run_oracle_acquisition(rot, budget, 17)
    for p in range(3):
        parity.l1("tau acquisition projection %d" % p, ref["hist"][0][p], got["hist"][0][p], 2e-3, 10)
    step = sc.make_engine()
    step.seed(17)
    step.reset_tallies()
    src, box = sc.acquisition_matrices(rot)
    for p in range(3):
        step.set_acquisition(src[p:p + 1], box[p:p + 1], budget[p:p + 1])
        step.run_acquisition()
        h, _ = step.projection(sc.engine_solids["box"][0], 0)
        assert np.array_equal(h, got["hist"][0][p]), p
    assert step.stats() == got["stats"]
    eng.close()
    step.close()


WORKDIR = os.path.join(REPO, "ggems_amd", "data", "_gen", "voxel_ct")


@pytest.fixture(scope="module")
def body():
    return voxel_ct_scene(WORKDIR)


def test_tau_full_phantom_matches_oracle(body):
    """the 512x512x256 body, 120 kVp: tau engine against tau oracle"""
    n = 300_000
    body.tau = True
    try:
        ref = body.run_oracle(n, 777, threads=0)
        got = body.run_engine(n, 777)
    finally:
        body.tau = False
    parity.counters("tau full phantom", ref["stats"], got["stats"],
                    ("voxel_steps", "interactions", "detector_interactions", "hops"), 2e-3)
    parity.l1("tau full phantom histogram", ref["hist"][0], got["hist"][0], 5e-3, 10)
    parity.total("tau full phantom edep", ref["edep"], got["edep"], 1e-3)
    parity.l1("tau full phantom hit", ref["hit"], got["hit"], 1e-2)


def test_tau_against_reference_sampling_at_full_size(body):
    """10^7 primaries each: tau sampling (seed 2001) against the reference
    sampling (seed 2002), with the reference against itself (seed 2003) as
    the null: dose in 8^3-dosel blocks and the projection in 8x8-pixel
    blocks, plus the per-primary counters."""
    n = 10_000_000
    eng = body.make_engine()
    ref = body.run_engine(n, 2002, eng)
    null = body.run_engine(n, 2003, eng)
    eng.set_sampling("tau")
    tau = body.run_engine(n, 2001, eng)
    eng.close()
    parity.same_distribution("dose blocks 8^3, tau vs reference sampling (null: reference vs reference)",
                             parity.dose_z(tau, ref), parity.dose_z(null, ref), 2000)
    parity.same_distribution("projection 8x8 pixel blocks, tau vs reference sampling (null: reference vs reference)",
                             parity.pixel_z(tau, ref), parity.pixel_z(null, ref), 300)
    rs, ts, ns = ref["stats"], tau["stats"], null["stats"]
    for k in ("voxel_steps", "interactions", "detector_interactions", "hops"):
        print("[stat] %s per primary: reference %.6g, tau %.6g, null %.6g" % (k, rs[k] / n, ts[k] / n, ns[k] / n))
        assert abs(ts[k] - rs[k]) <= max(6 * abs(ns[k] - rs[k]), 2e-3 * rs[k]), (k, rs[k], ts[k], ns[k])


def test_tau_against_reference_sampling_at_1e8(body):
    """10^8 primaries each (ten times the test above): tau sampling (seed
    3001) against the reference sampling (seed 3002), the reference against
    itself (seed 3003) as the null; dose in 8^3-dosel blocks, the projection
    in 4x4-pixel blocks, and the per-primary counters within their
    statistical spread."""
    n = 100_000_000
    eng = body.make_engine()
    ref = body.run_engine(n, 3002, eng)
    null = body.run_engine(n, 3003, eng)
    eng.set_sampling("tau")
    tau = body.run_engine(n, 3001, eng)
    eng.close()
    parity.same_distribution("dose blocks 8^3 at 1e8, tau vs reference sampling (null: reference vs reference)",
                             parity.dose_z(tau, ref), parity.dose_z(null, ref), 20000)
    parity.same_distribution("projection 4x4 pixel blocks at 1e8, tau vs reference sampling (null: reference vs "
                             "reference)", parity.pixel_z(tau, ref, b=4), parity.pixel_z(null, ref, b=4), 2000)
    rs, ts, ns = ref["stats"], tau["stats"], null["stats"]
    for k in ("voxel_steps", "interactions", "detector_interactions", "hops"):
        print("[stat] %s per primary at 1e8: reference %.7g, tau %.7g, null %.7g"
              % (k, rs[k] / n, ts[k] / n, ns[k] / n))
        assert abs(ts[k] - rs[k]) <= max(6 * abs(ns[k] - rs[k]), 5e-4 * rs[k]), (k, rs[k], ts[k], ns[k])
