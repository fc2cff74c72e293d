"""HIP engine vs CPU oracle on the same seeded inputs (needs an MI355X).

Per-history bit parity is expected wherever the path uses only IEEE
+,-,*,/ and sqrt (RNG streams, voxel walks, geometry); device log/exp/cos/sin
differ from the host libm by an ulp now and then, so whole-run tallies are
compared to a stated tolerance and, independently, statistically.
"""

import ctypes as C

import numpy as np
import pytest

import ggems_oracle as O
import parity
from scene import Scene, cfg1_scene, cylinder_phantom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1():
    return cfg1_scene(n_side=100, detector=256)


def test_uniform_streams_bit_exact(cfg1):
    eng = cfg1.make_engine()
    eng.seed(777)
    got = eng.debug_uniforms(2048, 48)
    st = O.seed_slots(777)
    for s in list(range(0, 2048, 97)) + [2047]:
        exp, _ = O.kiss_stream(st[:, s], 48)
        np.testing.assert_array_equal(got[s], exp)
    eng.close()


@pytest.mark.parametrize("mono", [True, False])
def test_primaries_match_oracle(mono):
    sc = Scene()
    sc.add_phantom(cylinder_phantom(20), ["Water", "RibBone"], (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (64, 64, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    if mono:
        sc.set_source((-600.0, 0.0, 0.0), 10.0, [0.06], mono=True, focal=(0.6, 0.6, 0.0))
    else:
        e = np.arange(20, 121) * 1e-3
        sc.set_source((-600.0, 0.0, 0.0), 10.0, e, intensities=(120 - e * 1e3 + 1) / e, focal=(0.6, 0.6, 0.0))
    sc.build()
    eng = sc.make_engine()
    eng.seed(1234)
    n = 4096
    pos, dirs, en = eng.debug_primaries(n)
    st = O.seed_slots(1234)
    _, _, src = sc._oracle_structs({})
    opos = np.zeros((n, 3), np.float32)
    odir = np.zeros((n, 3), np.float32)
    oen = np.zeros(n, np.float32)
    p3 = (C.c_float * 3)()
    d3 = (C.c_float * 3)()
    e1 = C.c_float()
    for s in range(n):
        rng = np.ascontiguousarray(st[:, s])
        O.lib().ggo_sample_primary(C.byref(src), O.ptr(rng), p3, d3, C.byref(e1))
        opos[s], odir[s], oen[s] = p3[:], d3[:], e1.value
    # double-precision trig of the cone angles may differ by an ulp
    np.testing.assert_allclose(dirs, odir, atol=2e-7, rtol=0)
    np.testing.assert_allclose(pos, opos, atol=1e-4, rtol=0)
    np.testing.assert_array_equal(en, oen)
    assert np.mean(np.all(dirs == odir, axis=1)) > 0.99
    eng.close()


def _walk_rays(rng, n, half):
    pos = rng.uniform(-half * 0.98, half * 0.98, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # axis-aligned and diagonal rays through voxel centres and corners
    special_p = np.array([[0.5, 0.5, 0.5], [0, 0, 0], [-7.5, 3.5, 1.0], [2.0, 2.0, 2.0]], np.float32)
    special_d = np.array([[1, 0, 0], [0, 1, 0], [0, 0, -1], [1, 1, 1]], np.float64)
    special_d /= np.linalg.norm(special_d, axis=1, keepdims=True)
    return np.concatenate([special_p, pos]), np.concatenate([special_d, d]).astype(np.float32)


def test_voxel_walk_bit_exact():
    sc = Scene()
    sc.add_phantom(cylinder_phantom(40), ["Water", "RibBone"], (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (16, 16, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 5.0, [0.06], mono=True)
    sc.build()
    eng = sc.make_engine()
    rng = np.random.default_rng(5)
    pos, dirs = _walk_rays(rng, 2000, 20.0)
    walks, counts = eng.debug_voxel_walk(0, pos, dirs, max_steps=512)
    tallies = {}
    _, solids, _ = sc._oracle_structs(tallies)
    out = np.zeros(512, np.int32)
    mism = 0
    for i in range(len(pos)):
        n = O.lib().ggo_voxel_walk(C.byref(solids[0]), O.ptr(pos[i]), O.ptr(dirs[i]), O.ptr(out), 512)
        if n != counts[i] or not np.array_equal(out[:min(n, 512)], walks[i]):
            mism += 1
    assert mism == 0
    assert counts.min() >= 1
    eng.close()


@pytest.mark.parametrize("general,interact_batch", [(False, 0), (True, 0), (False, 8), (True, 4)])
def test_cfg1_run_matches_oracle(cfg1, general, interact_batch):
    """The same seeded run through the specialised and the general kernel,
    with interactions resolved in the navigation phase or inside the voxel loop."""
    n = 200_000
    ref = cfg1.run_oracle(n, 777, threads=0)
    eng = cfg1.make_engine()
    eng.set_launch(interact_batch=interact_batch, general_kernel=general)
    got = cfg1.run_engine(n, 777, eng)
    assert eng.last_grid()[2] == (1 if general else 0)
    rs, gs = ref["stats"], got["stats"]
    assert gs["primaries"] == rs["primaries"] == n
    tag = "cfg1 general=%s batch=%d" % (general, interact_batch)
    parity.counters(tag, rs, gs, ("voxel_steps", "interactions", "detector_interactions", "hops"), 1e-3)
    # per-history agreement: histograms differ only where an ulp of a device
    # transcendental flipped a history
    parity.l1(tag + " histogram", ref["hist"][0], got["hist"][0], 1e-3, 10)
    parity.l1(tag + " scatter", ref["scatter"][0], got["scatter"][0], 5e-3, 10)
    parity.total(tag + " edep", ref["edep"], got["edep"], 1e-3)
    parity.l1(tag + " hit", ref["hit"], got["hit"], 2e-3, 10)
    eng.close()


@pytest.mark.parametrize("n", [1, 3, 1_048_576 + 4_097, (1 << 24) + 1])
def test_batch_layout_edge_cases_match_oracle(cfg1, n):
    """Runs of 1 and 3 primaries, and one just over a batch: GGEMS splits it
    into two batches of 526,337 and 526,336 photons
    (src/sources/GGEMSSource.cc:403-425), so slot 526,336 carries one photon
    and every lower slot two.  At 2^24 + 1 primaries the float ceiling of the
    batch count gives 16 batches whose first would overrun the slots; the
    exact ceiling gives 17, two of 986,896 photons and 15 of 986,895."""
    ref = cfg1.run_oracle(n, 4242, threads=0)
    eng = cfg1.make_engine()
    got = cfg1.run_engine(n, 4242, eng)
    rs, gs = ref["stats"], got["stats"]
    assert gs["primaries"] == rs["primaries"] == n
    h_ref, h_got = ref["hist"][0], got["hist"][0]
    if n < 10:
        # a few histories: identical unless an ulp of a transcendental flips one
        assert gs["voxel_steps"] == rs["voxel_steps"] and np.array_equal(h_ref, h_got)
    else:
        parity.counters("batch layout n=%d" % n, rs, gs, ("voxel_steps", "interactions", "hops"), 1e-3)
        parity.l1("batch layout n=%d histogram" % n, h_ref, h_got, 3e-3, 10)
    eng.close()


def test_photon_tracking_dose_matches_oracle():
    """Dosimetry with photon tracking (the fluence tally at each transport
    step's point of departure, dose_photon_tracking,
    include/GGEMS/navigators/GGEMSDoseRecording.hh:99-118), which runs the
    general kernel: tallies against the oracle's, history by history."""
    sc = Scene()
    sc.add_phantom(cylinder_phantom(60, radius=12.0), ["Water", "RibBone"], (1.0, 1.0, 1.0),
                   dosimetry={"photon_tracking": True})
    sc.add_ct_flat("GOS", (1, 1), (96, 96, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 3.0, [0.05], mono=True)
    sc.build()
    n = 100_000
    ref = sc.run_oracle(n, 606, threads=0)
    eng = sc.make_engine()
    got = sc.run_engine(n, 606, eng)
    assert eng.last_grid()[2] == 1
    pt_ref, pt_got = ref["photon_tracking"], got["photon_tracking"]
    assert pt_ref.sum() > 10 * n  # every transport step is counted
    assert abs(int(pt_got.sum()) - int(pt_ref.sum())) <= 0.01 * pt_ref.sum()
    parity.l1("photon tracking fluence", pt_ref, pt_got, 1e-3, 10)
    parity.l1("photon tracking hit", ref["hit"], got["hit"], 1e-3, 10)
    parity.total("photon tracking edep", ref["edep"], got["edep"], 1e-3)
    parity.total("photon tracking edep2", ref["edep2"], got["edep2"], 1e-3)
    eng.close()


def test_runs_beyond_2_32_primaries_count_every_one(cfg1):
    """2^32 + 12,345 primaries in one run: 4,097 batches, counted in 64-bit
    totals, with the detector counts per primary of an independent shorter
    run to within 5 Poisson errors."""
    n, m = (1 << 32) + 12_345, 1 << 26
    eng = cfg1.make_engine()
    big = cfg1.run_engine(n, 99, eng)
    small = cfg1.run_engine(m, 100, eng)
    eng.close()
    assert big["stats"]["primaries"] == n and small["stats"]["primaries"] == m
    cb, cs = int(big["hist"][0].astype(np.int64).sum()), int(small["hist"][0].astype(np.int64).sum())
    assert cs > 10_000 and big["hist"][0].min() >= 0
    assert abs(cb / n - cs / m) <= 5.0 * np.sqrt(cs) / m, (cb / n, cs / m)


@pytest.mark.parametrize("general", [False, True])
def test_launch_shape_does_not_change_histories(cfg1, general):
    """Rounds of 1 or 1,024 steps, 1 to 16 navigation passes, interaction
    batches of 1 to 64 lanes: the regrouping only moves photons between
    lanes, so every history, and with it every count, is the same."""
    n = 150_000
    eng = cfg1.make_engine()
    eng.set_launch(general_kernel=general)
    base = cfg1.run_engine(n, 2024, eng)
    for rs, cp, ib in [(1, 1, 1), (1024, 16, 64), (7, 3, 5)]:
        eng.set_launch(round_steps=rs, cold_passes=cp, interact_batch=ib, general_kernel=general)
        got = cfg1.run_engine(n, 2024, eng)
        assert got["stats"] == base["stats"], (rs, cp, ib)
        assert np.array_equal(got["hist"][0], base["hist"][0]) and np.array_equal(got["hit"], base["hit"])
        # FP64 sums whose order follows the atomics
        np.testing.assert_allclose(got["edep"], base["edep"], rtol=1e-12, atol=1e-18)
    eng.close()


def test_launch_rejects_invalid_parameters(cfg1):
    from ggems_amd.engine import EngineError
    eng = cfg1.make_engine()
    for kw in (dict(round_steps=-1), dict(round_steps=1025), dict(cold_passes=17), dict(interact_batch=65),
               dict(blocks_per_cu=-2), dict(blocks_per_cu=65)):
        with pytest.raises(EngineError):
            eng.set_launch(**kw)
    eng.close()


@pytest.mark.parametrize("general", [False, True])
def test_two_phantoms_match_oracle(general):
    """Two voxel phantoms side by side (one of them rotated), so that one wave
    holds photons in different voxel solids: the regrouping and the per-lane
    phantom geometry of the three-process several-solid kernel (variant 2)
    and of the general kernel (variant 1) against the oracle, history by
    history."""
    sc = Scene()
    sc.add_phantom(cylinder_phantom(40), ["Water", "RibBone"], (1.0, 1.0, 1.0), position=(0.0, -45.0, 0.0))
    sc.add_phantom(cylinder_phantom(32), ["Water", "RibBone"], (1.25, 1.25, 1.25), position=(0.0, 45.0, 0.0),
                   rotation=(0.0, 0.0, 30.0), dosimetry={})
    sc.add_ct_flat("GOS", (1, 1), (128, 128, 1), (1.5, 1.5, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 8.0, [0.06], mono=True)
    sc.build()
    n = 200_000
    ref = sc.run_oracle(n, 31, threads=0)
    eng = sc.make_engine()
    eng.set_launch(general_kernel=general)
    got = sc.run_engine(n, 31, eng)
    assert eng.last_grid()[2] == (1 if general else 2)
    rs, gs = ref["stats"], got["stats"]
    assert gs["primaries"] == rs["primaries"] == n
    assert rs["voxel_steps"] > 2 * n and ref["edep"].sum() > 0  # both phantoms are crossed
    tag = "two phantoms variant %d" % eng.last_grid()[2]
    parity.counters(tag, rs, gs, ("voxel_steps", "interactions", "hops"), 1e-3)
    parity.l1(tag + " histogram", ref["hist"][0], got["hist"][0], 1e-3, 10)
    parity.total(tag + " edep", ref["edep"], got["edep"], 1e-3)
    eng.close()


def _chi2_two_sample(a, b):
    """chi-square of two independent Poisson samples of one distribution, p-value."""
    from scipy.stats import chi2
    a = a.astype(np.float64).ravel()
    b = b.astype(np.float64).ravel()
    m = (a + b) > 0
    x2 = float(np.sum((a[m] - b[m]) ** 2 / (a[m] + b[m])))
    return x2, int(m.sum()), chi2.sf(x2, int(m.sum()))


def test_cfg1_statistical_agreement_independent_seeds(cfg1):
    """Engine and oracle with different seeds: per-pixel (8x8 binned) detector
    counts, scatter counts and per-slab dose agree as samples of one
    distribution (two-sample chi-square), and the mean z-score is ~0."""
    n = 400_000
    ref = cfg1.run_oracle(n, 1001, threads=0)
    got = cfg1.run_engine(n, 2002)

    def binned(h, k=8):
        return h.reshape(h.shape[0] // k, k, h.shape[1] // k, k).sum(axis=(1, 3))

    for key in ("hist", "scatter"):
        a, b = binned(ref[key][0]), binned(got[key][0])
        x2, dof, p = _chi2_two_sample(a, b)
        assert p > 1e-4, (key, x2, dof)
        z = (b.astype(np.float64) - a) / np.sqrt(np.maximum(a + b, 1))
        assert abs(z.mean()) < 4.0 / np.sqrt(z.size)
    # hits per 4-voxel slab along x: counts of energy-deposit events
    ha = ref["hit"].reshape(ref["hit"].shape[0], ref["hit"].shape[1], -1, 4).sum(axis=(0, 1, 3))
    hb = got["hit"].reshape(got["hit"].shape[0], got["hit"].shape[1], -1, 4).sum(axis=(0, 1, 3))
    x2, dof, p = _chi2_two_sample(ha, hb)
    assert p > 1e-4, (x2, dof)


def test_step_log_matches_device_logf(cfg1):
    """The voxel step's -log(u) is bit-identical to -logf(u) on the device."""
    eng = cfg1.make_engine()
    rng = np.random.default_rng(9)
    sums = np.concatenate([rng.integers(1, 2 ** 32, 4_000_000, dtype=np.uint64).astype(np.uint32),
                           np.arange(1, 200_000, dtype=np.uint32),
                           (2 ** 32 - 1 - np.arange(200_000)).astype(np.uint32)])
    fast, ref = eng.debug_neg_log(sums)
    np.testing.assert_array_equal(fast, ref)
    f0, r0 = eng.debug_neg_log(np.zeros(1, np.uint32))
    assert np.isnan(f0[0]) and r0[0] == np.inf  # both lose every comparison with `<`
    eng.close()


def test_engine_reproduces_golden_vectors():
    import os
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz"),
                   allow_pickle=False)
    sc = cfg1_scene(n_side=100, detector=256, dosimetry=False)
    eng = sc.make_engine()
    eng.seed(777)
    np.testing.assert_array_equal(eng.debug_uniforms(4, 64), gold["kiss_777"])
    eng.seed(777)
    pos, dirs, en = eng.debug_primaries(64)
    np.testing.assert_array_equal(en, gold["primary_e"])
    np.testing.assert_allclose(dirs, gold["primary_dir"], atol=2e-7, rtol=0)
    np.testing.assert_allclose(pos, gold["primary_pos"], atol=1e-4, rtol=0)
    walks, counts = eng.debug_voxel_walk(0, gold["walk_p"], gold["walk_d"], max_steps=400)
    for k in range(len(walks)):
        exp = gold["walks"][k]
        np.testing.assert_array_equal(walks[k], exp[exp >= 0])
    eng.close()
