"""What the trained scatter surrogate gains over the method without it (needs
an MI355X).

Seeded body phantoms (96x96x64 voxels of 2 mm) are projected onto a 128x128
detector by the batched acquisition: 64 training views (8 phantoms x 8
gantry angles) with 2*10^7-primary scatter targets, and 16 held-out views of
4 other phantoms with 2*10^8-primary targets, each target run in two
independent halves so that its noise can be subtracted.  The U-Net trains
for 1,500 Adam steps on the MFMA convolutions.  On the held-out views its
error against the noise-corrected truth must beat both
- the low-count (10^6-primary) Monte Carlo's own scatter image (one of the
  network's three inputs, with that run's counts and the ray-traced
  transmission), and
- the method without the network: scatter = a constant times the ray-traced
  transmission, the constant fitted by least squares on the training views.
"""

import os

import numpy as np
import pytest

from ggems_amd import surrogate as S
from scene import REPO

pytestmark = pytest.mark.gpu

WORKDIR = os.path.join(REPO, "ggems_amd", "data", "_gen", "surrogate_gain")
DET = 128


def test_surrogate_beats_low_count_mc_and_the_drr_baseline():
    rate = S.open_field_rate(WORKDIR, 20_000_000, seed=3, detector=DET)
    tr = S.dataset_batched(WORKDIR, range(1, 9), 8, 100, 1_000_000, 20_000_000, rate, detector=DET)
    te = S.dataset_batched(WORKDIR, range(101, 105), 4, 200, 1_000_000, 200_000_000, rate, detector=DET)
    assert tr["x"].shape == (64, DET, DET, 3) and te["y"].shape == (16, DET, DET)
    net, seconds, losses = S.train(tr["x"], tr["y"], steps=1500, batch=8, seed=1)
    assert losses[-1] < 0.5 * losses[0]
    pred = S.predict(net, te["x"])
    err = S.scatter_errors(pred, te["y"], te["var"], te["low"])
    base = S.scatter_errors(S.drr_baseline(tr["x"], tr["y"])(te["x"]), te["y"], te["var"], te["low"])
    print("[surrogate] held-out error against the noise-corrected truth: network %.4f, low-count MC %.4f, "
          "k * DRR %.4f (target noise %.4f); training %.1f s"
          % (err["pred_vs_truth"], err["low_vs_truth"], base["pred_vs_truth"], err["target_noise"], seconds),
          flush=True)
    assert err["pred_vs_truth"] < 0.5 * err["low_vs_truth"]
    assert err["pred_vs_truth"] < 0.8 * base["pred_vs_truth"]
