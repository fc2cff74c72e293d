"""The oracle's opt-in tau sampling mode (one exponential number of mean free
paths per flight instead of three competing exponentials per voxel step),
pinned by what the physics fixes: Beer-Lambert transmission with the
left-bin cross section, and the reference sampling's distributions, as two
independent samples (two-sample chi-square of detector counts, scatter and
dose, and the per-primary interaction and step counts)."""

import math

import numpy as np

import parity
from scene import Scene, cfg1_scene

F = np.float32


def _slab(processes, tau):
    sc = Scene(processes=processes)
    sc.tau = tau
    sc.add_phantom(np.zeros((20, 20, 50), np.uint8), ["Water"], (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (8, 8, 1), (2.0, 2.0, 30.0), 600.0, 1000.0, 1e-3, store_scatter=False)
    sc.set_source((-600.0, 0.0, 0.0), 0.01, [0.030], mono=True)
    return sc.build()


def test_tau_mode_beer_lambert_left_bin():
    """Photoelectric only: detector counts ~ Binomial(N, exp(-sigma_left L))."""
    n, L, e0 = 100_000, 50, 0.030
    sc = _slab((1,), True)
    r = sc.run_oracle(n, 4242, threads=4)
    counts = int(r["hist"][0].sum())
    eid = int(np.searchsorted(sc.ebins, F(e0), side="right") - 1)
    p = math.exp(-float(sc.cs[1, 0, eid]) * L)
    sd = math.sqrt(n * p * (1 - p))
    assert abs(counts - n * p) < 4 * sd, (counts, n * p, sd)


def test_tau_mode_matches_reference_sampling_statistically():
    """cfg1 at 60 keV (water with a bone cylinder, Compton + photoelectric +
    Rayleigh), 600k primaries: tau sampling (seed 12) against the reference
    sampling (seed 11), with the reference against itself (seed 13) as the
    null: dose in 4^3-dosel blocks, detector counts and scatter counts in
    4 x 4-pixel blocks, and the per-primary counters."""
    n = 600_000
    ref = cfg1_scene(n_side=60, detector=64).run_oracle(n, 11, threads=0)
    null = cfg1_scene(n_side=60, detector=64).run_oracle(n, 13, threads=0)
    sc = cfg1_scene(n_side=60, detector=64)
    sc.tau = True
    tau = sc.run_oracle(n, 12, threads=0)
    parity.same_distribution("oracle tau vs reference sampling: dose 4^3 blocks",
                             parity.dose_z(tau, ref, 4), parity.dose_z(null, ref, 4), 1000)
    parity.same_distribution("oracle tau vs reference sampling: detector 4x4 blocks",
                             parity.pixel_z(tau, ref, 4), parity.pixel_z(null, ref, 4), 100)
    parity.same_distribution("oracle tau vs reference sampling: scatter 4x4 blocks",
                             parity.pixel_z(tau, ref, 4, 10, "scatter"), parity.pixel_z(null, ref, 4, 10, "scatter"),
                             50)
    rs, ts, ns = ref["stats"], tau["stats"], null["stats"]
    for k in ("interactions", "voxel_steps", "detector_interactions", "hops"):
        # tau against reference within 6 times the reference-to-reference spread, and within 0.5 %
        assert abs(ts[k] - rs[k]) <= max(6 * abs(ns[k] - rs[k]), 5e-3 * rs[k]), (k, rs[k], ts[k], ns[k])
