"""CT acquisition in one persistent launch on the GPU: against the oracle's
consecutive runs, bit-exact against consecutive single-projection launches,
hipGraph replays against eager launches, the autotuner's state restore, and
the GGEMS-level script."""

import math
import os

import numpy as np
import pytest

import parity
from ggems_amd import authoring, configs
from scene import Scene, cfg1_scene

pytestmark = pytest.mark.gpu

ROT = [(0.0, 0.0, 0.0), (0.0, 0.0, 0.4), (0.0, 0.0, 1.3), (0.2, 0.0, 2.0)]
BUDGET = [150_000, 0, 1_300_000, 40_000]  # an empty projection and a two-batch one


def _scene():
    return cfg1_scene(n_side=60, detector=48, dosimetry=True, aperture_deg=2.0)


def _images(sc, eng):
    out = sc.read_acquisition(eng)
    return out["hist"][0], out["scatter"][0], out["stats"]


def test_acquisition_matches_oracle():
    sc = _scene()
    eng = sc.engine_acquisition(ROT, BUDGET, 17)
    eng.run_acquisition()
    hist, scat, stats = _images(sc, eng)
    ref = sc.run_oracle_acquisition(ROT, BUDGET, 17)
    assert stats["primaries"] == ref["stats"]["primaries"] == sum(BUDGET)
    assert not hist[1].any()
    for p in (0, 2, 3):
        parity.l1("acquisition projection %d" % p, ref["hist"][0][p], hist[p], 2e-3, 10)
    parity.counters("acquisition", ref["stats"], stats, ("voxel_steps",), 1e-4, 0)
    edep, _, _, _ = eng.dose_tallies()
    parity.total("acquisition edep", ref["edep"], edep, 1e-4)
    eng.close()


def test_acquisition_equals_consecutive_launches():
    """one launch over all projections == one launch per projection, with the
    random streams continuing: identical integer images and counters"""
    sc = _scene()
    eng = sc.engine_acquisition(ROT, BUDGET, 23)
    eng.run_acquisition()
    hist, scat, stats = _images(sc, eng)
    eng2 = sc.make_engine()
    eng2.seed(23)
    eng2.reset_tallies()
    src, box = sc.acquisition_matrices(ROT)
    for p in range(len(ROT)):
        eng2.set_acquisition(src[p:p + 1], box[p:p + 1], BUDGET[p:p + 1])
        eng2.run_acquisition()
        h, s = eng2.projection(sc.engine_solids["box"][0], 0)
        assert np.array_equal(h, hist[p]), p
        assert np.array_equal(s, scat[p]), p
    assert eng2.stats() == stats
    e1, _, hit1, _ = eng.dose_tallies()
    e2, _, hit2, _ = eng2.dose_tallies()
    assert np.array_equal(hit1, hit2)
    assert np.allclose(e1, e2, rtol=1e-12, atol=0.0)
    eng.close()
    eng2.close()


def test_graph_replays_match_eager_launches():
    """each replay of the captured acquisition equals one more eager launch,
    with eager work and synchronisations between the replays"""
    sc = _scene()
    eager = sc.engine_acquisition(ROT, BUDGET, 31)
    graph = sc.engine_acquisition(ROT, BUDGET, 31)
    graph.capture_acquisition()
    for rep in range(3):
        eager.run_acquisition()
        graph.replay_acquisition()
        graph.stats()  # eager work between replays
        h_e, s_e, st_e = _images(sc, eager)
        h_g, s_g, st_g = _images(sc, graph)
        assert np.array_equal(h_e, h_g), rep
        assert np.array_equal(s_e, s_g), rep
        assert st_e == st_g, rep
        assert st_g["primaries"] == (rep + 1) * sum(BUDGET)
    eager.close()
    graph.close()


def test_autotune_restores_streams_and_tallies():
    sc = _scene()
    tuned = sc.make_engine()
    tuned.seed(41)
    best = tuned.autotune(2_000_000)
    assert best["blocks_per_cu"] in (1, 2) and best["round_steps"] in (12, 20, 32)
    assert best["interact_batch"] in (8, 16, 32)
    assert tuned.stats()["primaries"] == 0
    tuned.reset_tallies()
    tuned.run(300_000)
    h_t, _ = tuned.histogram(sc.engine_solids["box"][0])
    fresh = sc.make_engine()
    fresh.seed(41)
    fresh.reset_tallies()
    fresh.run(300_000)
    h_f, _ = fresh.histogram(sc.engine_solids["box"][0])
    assert np.array_equal(h_t, h_f)
    tuned.close()
    fresh.close()


def test_autotune_tau_sampling_tunes_round_length():
    """with tau sampling the grid is blocks x round steps around its best,
    the interaction batch and the navigation passes keep their values, and
    the streams are restored"""
    sc = _scene()
    sc.tau = True
    tuned = sc.make_engine()
    tuned.seed(43)
    best = tuned.autotune(2_000_000)
    assert best["blocks_per_cu"] in (1, 2) and best["round_steps"] in (12, 16, 20, 24)
    assert best["cold_passes"] == 0 and best["interact_batch"] == 16
    assert tuned.stats()["primaries"] == 0
    tuned.reset_tallies()
    tuned.run(300_000)
    h_t, _ = tuned.histogram(sc.engine_solids["box"][0])
    fresh = sc.make_engine()
    fresh.seed(43)
    fresh.reset_tallies()
    fresh.run(300_000)
    h_f, _ = fresh.histogram(sc.engine_solids["box"][0])
    assert np.array_equal(h_t, h_f)
    tuned.close()
    fresh.close()


def test_acquisition_script_matches_oracle(tmp_path):
    """configs.acquisition through the GGEMS API: the saved projection stack
    against the oracle's consecutive runs on the same inputs"""
    wd = str(tmp_path)
    keep = configs.acquisition(wd, side=64, seed=3)
    rot, budgets = configs.acquisition_plan(n_angles=3, budget_range=(100_000, 300_000), seed=2, step_deg=40.0)
    from ggems_amd.ggems import GGEMS
    g = GGEMS()
    g.initialize(123)
    g.run_acquisition(rot, budgets, "deg")
    stack, spacing = authoring.read_mhd(os.path.join(wd, "projections.mhd"))
    assert stack.shape == (3, 512, 384)
    mhd, rng, spec = configs.acquisition_inputs(wd, 64, 3)
    lab, _ = authoring.read_mhd(mhd)
    e, w = np.loadtxt(spec, unpack=True)
    sc = Scene()
    sc.add_phantom(lab, authoring.BODY_MATERIALS, (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (384, 512, 1), (0.78, 0.78, 1.0), 600.0, 1000.0, 10e-3)
    half_diag = 0.5 * 0.78 * math.hypot(384, 512)
    sc.set_source((-600.0, 0.0, 0.0), math.degrees(math.atan(half_diag / 1000.0)), e, intensities=w)
    sc.build()
    ref = sc.run_oracle_acquisition(np.radians(rot.astype(np.float64)).astype(np.float32), budgets, 123)
    for p in range(3):
        parity.l1("acquisition script projection %d" % p, ref["hist"][0][p], stack[p], 1e-3, 10)
    g.delete()
    del keep


def test_setters_drop_the_captured_acquisition():
    """A captured acquisition holds the device addresses of the tables,
    tallies and source.  Every setter that reallocates one of them drops the
    graph, so a later replay reports that nothing is captured instead of
    replaying stale addresses; capturing again works as before."""
    from ggems_amd.engine import EngineError
    sc = _scene()
    eng = sc.engine_acquisition(ROT, BUDGET, 5)
    S = sc.source
    a = sc.arrays
    setters = [
        lambda: eng.set_xray_source(S["m"], S["aperture"], S["focal"], S["energy"], S["cdf"], S["is_interp"]),
        lambda: eng.set_physics(sc.ebins, sc.cs, sc.cs_atom, a["n_elements"], a["elem_index"], a["z"],
                                a["atom_density"], a["photon_cut"], a["density"], sc.pp, sc.factor, sc.processes),
    ]
    for setter in setters:
        eng.capture_acquisition()
        eng.replay_acquisition()
        setter()
        with pytest.raises(EngineError, match="no captured acquisition"):
            eng.replay_acquisition()
    eng.close()
    # the dose tallies: a fresh dosimetry reallocates them
    eng = sc.engine_acquisition(ROT, BUDGET, 5)
    eng.capture_acquisition()
    k = sc.engine_solids["voxel"][0]
    pay = sc.navigators[0][2]
    d = sc.dose_grid(pay)
    eng.set_dosimetry(k, d["size"], d["n"], pay["bmin"], pay["bmax"], True, True, False)
    with pytest.raises(EngineError, match="no captured acquisition"):
        eng.replay_acquisition()
    eng.capture_acquisition()
    eng.replay_acquisition()
    assert eng.stats()["primaries"] == sum(BUDGET)
    eng.close()


# ---- several phantoms in one launch ----------------------------------------
PH_ROT = [(0.0, 0.0, 0.1 * p) for p in range(6)]
PH_BUDGET = [120_000, 80_000, 0, 150_000, 60_000, 100_000]
PH_ONLY = [0, 1, 1, 2, 0, -1]  # solid indices; -1: every phantom at once


def _phantoms_scene():
    """three different phantoms at the isocentre, two of them with the same
    material list (they share their materials), and one flat panel"""
    from scene import cylinder_phantom
    sc = Scene()
    sc.add_phantom(cylinder_phantom(40, radius=12.0), ["Water", "RibBone"], (1.0, 1.0, 1.0))
    sc.add_phantom(cylinder_phantom(30, radius=5.0), ["Water", "RibBone"], (1.5, 1.5, 1.5),
                   rotation=(0.0, 0.4, 0.0), dosimetry={})
    sc.add_phantom(np.zeros((16, 16, 16), np.uint8), ["Lung"], (2.0, 2.0, 2.0))
    sc.add_ct_flat("GOS", (1, 1), (48, 48, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 2.0, [0.06], mono=True)
    return sc.build()


def _l1(a, b):
    return int(np.abs(a.astype(np.int64) - b).sum())


def test_phantom_acquisition_matches_oracle():
    """Projections over different phantoms in one launch against the oracle's
    consecutive runs, history by history: each projection sees the panel and
    its one phantom."""
    sc = _phantoms_scene()
    assert sc.offsets[0] == sc.offsets[1]  # the shared material list
    eng = sc.engine_acquisition(PH_ROT, PH_BUDGET, 71, phantoms=PH_ONLY)
    eng.run_acquisition()
    assert eng.last_grid()[2] == 2  # three processes over several voxel solids
    got = sc.read_acquisition(eng)
    ref = sc.run_oracle_acquisition(PH_ROT, PH_BUDGET, 71, phantoms=PH_ONLY)
    assert got["stats"]["primaries"] == ref["stats"]["primaries"] == sum(PH_BUDGET)
    for k in ("voxel_steps", "interactions", "hops"):
        assert abs(got["stats"][k] - ref["stats"][k]) <= 1e-3 * ref["stats"][k] + 10, k
    for p in range(len(PH_ROT)):
        h, r = got["hist"][0][p], ref["hist"][0][p]
        l1 = _l1(h, r)
        print("[acq] projection %d phantom %d: %d counts, L1 %d" % (p, PH_ONLY[p], int(r.sum()), l1), flush=True)
        assert l1 <= 2e-3 * r.sum() + 4, (p, l1, int(r.sum()))
    # the dose phantom (solid 1) is crossed only in projections 1 and 5
    edep, _, hit, _ = eng.dose_tallies()
    assert ref["edep"].sum() > 0
    assert abs(edep.sum() - ref["edep"].sum()) <= 2e-3 * ref["edep"].sum()
    assert _l1(hit, ref["hit"]) <= 2e-3 * ref["hit"].sum() + 4
    eng.close()


def test_phantom_acquisition_equals_consecutive_launches():
    """One launch over projections of three phantoms == one launch per
    projection with the streams continuing, bit for bit; the captured graph
    replays it exactly."""
    sc = _phantoms_scene()
    eng = sc.engine_acquisition(PH_ROT, PH_BUDGET, 73, phantoms=PH_ONLY)
    eng.run_acquisition()
    one = sc.read_acquisition(eng)
    step = sc.make_engine()
    step.seed(73)
    step.reset_tallies()
    src, box = sc.acquisition_matrices(PH_ROT)
    for p in range(len(PH_ROT)):
        step.set_acquisition(src[p:p + 1], box[p:p + 1], PH_BUDGET[p:p + 1], PH_ONLY[p:p + 1])
        step.run_acquisition()
        h, s = step.projection(sc.engine_solids["box"][0], 0)
        assert np.array_equal(h, one["hist"][0][p]), p
        assert np.array_equal(s, one["scatter"][0][p]), p
    assert step.stats() == one["stats"]
    e1, _, hit1, _ = eng.dose_tallies()
    e2, _, hit2, _ = step.dose_tallies()
    assert np.array_equal(hit1, hit2)
    np.testing.assert_allclose(e1, e2, rtol=1e-12, atol=0.0)
    graph = sc.engine_acquisition(PH_ROT, PH_BUDGET, 73, phantoms=PH_ONLY)
    graph.capture_acquisition()
    graph.replay_acquisition()
    g = sc.read_acquisition(graph)
    assert np.array_equal(np.stack(g["hist"]), np.stack(one["hist"])) and g["stats"] == one["stats"]
    for e in (eng, step, graph):
        e.close()


def test_phantom_acquisition_rejects_a_panel_as_phantom():
    from ggems_amd.engine import EngineError
    sc = _phantoms_scene()
    eng = sc.make_engine()
    panel = sc.engine_solids["box"][0]
    src, box = sc.acquisition_matrices(PH_ROT[:2])
    for bad in ([panel, 0], [7, 0], [-2, 0]):
        with pytest.raises(EngineError):
            eng.set_acquisition(src, box, [10, 10], bad)
    eng.close()


def _stated_acquisition(budget_range, seed=7):
    """The batched CT acquisition at its stated shape: seeded ellipsoid
    phantoms of 128^3, 256^3 and 384^3 1 mm voxels, 360 one-degree
    projections each, the 120 kVp spectrum and the flat 512x384 GOS panel
    of 0.78 mm pixels at SDD 1000 mm; per-projection budgets drawn uniformly
    in budget_range."""
    sc = Scene()
    for k, side in enumerate((128, 256, 384)):
        lab = authoring.ellipsoid_body_phantom(1 + k, shape=(side, side, side), voxel=1.0, n_ellipsoids=40,
                                               body_axes=(0.45 * side, 0.35 * side),
                                               semi_axis_range=(side / 50.0, side / 4.0))
        sc.add_phantom(lab, authoring.BODY_MATERIALS, (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (384, 512, 1), (0.78, 0.78, 1.0), 600.0, 1000.0, 10e-3)
    e, w = authoring.kramers_spectrum(120.0, 20.0, 2.5, authoring.aluminium_mu())
    half_diag = 0.5 * 0.78 * math.hypot(384, 512)
    sc.set_source((-600.0, 0.0, 0.0), math.degrees(math.atan(half_diag / 1000.0)), e, intensities=w)
    sc.build()
    rot, only, budgets = [], [], []
    for k in range(3):
        r, b = configs.acquisition_plan(n_angles=360, budget_range=budget_range, seed=seed + k)
        rot.append(np.radians(r.astype(np.float64)).astype(np.float32))
        budgets.append(b)
        only += [k] * 360
    return sc, np.concatenate(rot), np.concatenate(budgets), np.array(only, np.int32)


def _binned_chi2(a, b, block):
    """Two-sample chi-square of two count images summed over block x block
    pixels: sum (a - b)^2 / (a + b) over bins with counts, and its degrees
    of freedom."""
    rows, cols = a.shape[-2] // block * block, a.shape[-1] // block * block
    def binned(x):
        x = x[..., :rows, :cols].astype(np.float64)
        return x.reshape(-1, rows // block, block, cols // block, block).sum(axis=(0, 2, 4))
    A, B = binned(a), binned(b)
    m = (A + B) > 0
    return float(((A - B)[m] ** 2 / (A + B)[m]).sum()), int(m.sum())


def test_stated_acquisition_in_one_launch_matches_oracle():
    """The three stated phantoms x 360 projections (1,080 projections of
    uneven budgets, here 10^4 to 3*10^4 primaries each) in ONE captured
    launch.

    A particle slot runs its photons of projection 0, then 1, ... 1,079 in
    sequence, so once an ulp of a device transcendental changes one history
    of a slot, that slot's later histories differ from the oracle's as well:
    the images of the first projections agree history by history, and later
    ones only statistically.  So the launch is checked three ways:
    - the first ten projections against the oracle's, image by image;
    - the whole stack against the oracle by a two-sample chi-square per
      phantom (its 360 images summed, 16 x 16-pixel bins) and per projection
      (counts), plus the run counters;
    - bit for bit against three consecutive launches, one per phantom, with
      the random streams continuing."""
    sc, rot, budgets, only = _stated_acquisition((10_000, 30_000))
    eng = sc.engine_acquisition(rot, budgets, 19, phantoms=only)
    eng.capture_acquisition()
    eng.replay_acquisition()
    assert eng.last_grid()[2] == 2
    got = sc.read_acquisition(eng)
    ref = sc.run_oracle_acquisition(rot, budgets, 19, threads=0, phantoms=only)
    n = int(budgets.sum())
    assert got["stats"]["primaries"] == ref["stats"]["primaries"] == n
    for k in ("voxel_steps", "interactions", "hops"):
        assert abs(got["stats"][k] - ref["stats"][k]) <= 1e-3 * ref["stats"][k], k
    h, r = got["hist"][0], ref["hist"][0]
    per = np.abs(h.astype(np.int64) - r).reshape(len(budgets), -1).sum(axis=1)
    counts = r.reshape(len(budgets), -1).sum(axis=1)
    mine = h.reshape(len(budgets), -1).sum(axis=1)
    tot = int(counts.sum())
    assert (counts > 0).all()
    head = int(per[:10].sum()), int(counts[:10].sum())
    print("[acq] stated acquisition: %d primaries, %d counts, kernel %.1f ms; first 10 projections L1 %d of %d "
          "counts; whole stack L1 %d (%.2e of the counts)" % (n, tot, got["kernel_ms"], head[0], head[1],
                                                              int(per.sum()), per.sum() / tot), flush=True)
    assert head[0] <= 5e-3 * head[1] + 4
    for k in range(3):
        sel = only == k
        chi2, dof = _binned_chi2(h[sel], r[sel], 16)
        pc = mine[sel].astype(np.float64), counts[sel].astype(np.float64)
        chi2_p = float(((pc[0] - pc[1]) ** 2 / (pc[0] + pc[1])).sum())
        print("[acq] phantom %d: image chi2/dof %.3f (dof %d), per-projection counts chi2/dof %.3f (dof %d)"
              % (k, chi2 / dof, dof, chi2_p / sel.sum(), int(sel.sum())), flush=True)
        assert chi2 / dof < 1.3 and chi2_p / sel.sum() < 1.3
    # the same acquisition as three launches, one phantom each
    step = sc.engine_acquisition(rot[:360], budgets[:360], 19, phantoms=only[:360])
    src, box = sc.acquisition_matrices(rot)
    panel = sc.engine_solids["box"][0]
    for k in range(3):
        sel = slice(360 * k, 360 * (k + 1))
        if k:
            step.set_acquisition(src[sel], box[sel], budgets[sel], only[sel])
        step.run_acquisition()
        for p in range(360):
            hp, sp = step.projection(panel, p)
            assert np.array_equal(hp, h[360 * k + p]), (k, p)
            if got["scatter"]:
                assert np.array_equal(sp, got["scatter"][0][360 * k + p]), (k, p)
    assert step.stats() == got["stats"]
    eng.close()
    step.close()
