"""Pins of the oracle's ray tracer: Siddon path lengths against analytic
chords, the energy weights against the source's own sampling, and the
transmission of a uniform slab against Beer-Lambert."""

import math

import numpy as np

import drr as R
from scene import Scene, cfg1_scene

F = np.float32


def _chord(p0, p1, lo, hi):
    d = np.asarray(p1, float) - np.asarray(p0, float)
    t0, t1 = 0.0, 1.0
    for a in range(3):
        ta, tb = (lo[a] - p0[a]) / d[a], (hi[a] - p0[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return max(0.0, t1 - t0) * float(np.linalg.norm(d))


def test_siddon_lengths_match_analytic_chords():
    """a 20x16x12 grid of 1.5 mm voxels with an insert of whole voxels: the
    lengths in the insert and in the whole grid are the chords of the two boxes"""
    dims, vs = (20, 16, 12), (1.5, 1.5, 1.5)
    lo = np.array([-15.0, -12.0, -9.0])
    hi = lo + np.array(dims) * 1.5
    ins_lo, ins_hi = lo + np.array([4, 3, 2]) * 1.5, lo + np.array([13, 9, 7]) * 1.5
    inside = np.zeros(dims[::-1], bool)
    inside[2:7, 3:9, 4:13] = True
    rng = np.random.default_rng(3)
    for _ in range(200):
        p0 = rng.normal(size=3) * 40
        p1 = rng.uniform(ins_lo, ins_hi) * 2 - p0  # through the insert region, mostly
        segs = R.siddon_lengths(p0, p1, lo, hi, vs, dims)
        total = sum(l for _, l in segs)
        ins = sum(l for v, l in segs if inside.reshape(-1)[v])
        assert abs(total - _chord(p0, p1, lo, hi)) < 1e-9
        assert abs(ins - _chord(p0, p1, ins_lo, ins_hi)) < 1e-9


def test_energy_weights_follow_the_source_sampling():
    sc = cfg1_scene(n_side=8, detector=4, dosimetry=False)
    w = R.energy_weights(sc.ebins, sc.source["energy"], sc.source["cdf"], sc.source["is_interp"])
    assert np.count_nonzero(w) == 1 and abs(w.sum() - 1.0) < 1e-12
    assert R.left_bin(sc.ebins, F(0.06)) == int(np.nonzero(w)[0][0])
    # interpolated spectrum: the weights are the bin frequencies of energies
    # drawn the way the source draws them
    e = np.arange(20.0, 121.0, 1.0) * 1e-3
    inten = (120.0 - e * 1e3).clip(0) / (e * 1e3)
    sc2 = Scene()
    sc2.add_phantom(np.zeros((2, 2, 2), np.uint8), ["Water"], (1.0, 1.0, 1.0))
    sc2.set_source((-600.0, 0, 0), 1.0, e, intensities=inten)
    sc2.build()
    S = sc2.source
    w2 = R.energy_weights(sc2.ebins, S["energy"], S["cdf"], S["is_interp"])
    assert abs(w2.sum() - 1.0) < 1e-9
    rng = np.random.default_rng(9)
    u = rng.random(400_000).astype(F)
    cdf = S["cdf"].astype(np.float64)
    idx = np.searchsorted(cdf, u, side="right") - 1
    idx = np.clip(idx, 0, len(cdf) - 2)
    frac = (u - cdf[idx]) / (cdf[idx + 1] - cdf[idx])
    en = S["energy"][idx] + frac * (S["energy"][idx + 1] - S["energy"][idx])
    b = np.searchsorted(sc2.ebins, en, side="right") - 1
    counts = np.bincount(b, minlength=len(sc2.ebins))
    exp = w2 * len(u)
    m = exp > 50
    chi2 = float((((counts[m] - exp[m]) ** 2) / exp[m]).sum())
    assert chi2 < m.sum() + 6 * math.sqrt(2 * m.sum()), chi2


def test_uniform_slab_transmission_is_beer_lambert():
    """a water block at 60 keV: every pixel's transmission is
    exp(-sigma_tot * chord) along its own ray"""
    sc = Scene()
    sc.add_phantom(np.zeros((10, 20, 30), np.uint8), ["Water"], (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (6, 6, 1), (4.0, 4.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 1.0, [0.06], mono=True)
    sc.build()
    img = R.drr(sc)
    sig = sum(float(sc.cs[p, 0, R.left_bin(sc.ebins, F(0.06))]) for p in sc.processes)
    mod = sc.navigators[1][2]["modules"][0]
    src = R._apply(sc.source["m"], (0, 0, 0))
    pay = sc.navigators[0][2]
    for iy in range(6):
        for ix in range(6):
            pc = (mod["bmin"][0] + (ix + 0.5) * 4.0, mod["bmin"][1] + (iy + 0.5) * 4.0, 0.0)
            pe = R._apply(mod["m"], pc)
            L = _chord(R._inverse_apply(pay["m"], src), R._inverse_apply(pay["m"], pe), pay["bmin"], pay["bmax"])
            assert abs(img[iy, ix] - math.exp(-sig * L)) < 1e-12
    assert 0.2 < img.min() < 0.9 and img.max() == 1.0  # edge rays miss the 10 mm tall block
