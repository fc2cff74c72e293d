"""Host-side scoring of the scatter surrogate: the noise-corrected errors and
the equal-error primary factor, on synthetic images whose truth is known."""

import numpy as np

from ggems_amd import surrogate as S


def _poisson_view(rng, truth, n):
    """scatter image of n primaries at `truth` counts per primary, as the
    engine's two halves and their mean, on the counts-per-primary scale"""
    a = rng.poisson(truth * (n // 2)) / (n // 2)
    b = rng.poisson(truth * (n - n // 2)) / (n - n // 2)
    return (a + b) / 2.0, (a - b) ** 2 / 4.0


def test_noise_corrected_errors_recover_the_truth():
    rng = np.random.default_rng(7)
    yy, xx = np.mgrid[0:256, 0:256] / 256.0
    truth = 1e-4 * (1.0 + 0.5 * np.sin(6.0 * xx) * np.cos(4.0 * yy))  # counts per primary
    truth = np.broadcast_to(truth, (12, 256, 256)).copy()
    n_high, n_low = 10 ** 8, 10 ** 6
    y, var = _poisson_view(rng, truth, n_high)
    low = rng.poisson(truth * n_low) / n_low
    bias = 0.03 * truth.mean() * np.cos(9.0 * xx)  # a prediction off by a known pattern
    pred = truth + bias

    e = S.scatter_errors(pred, y, var, low)
    mean = y.mean()
    # the target's noise: Poisson variance truth / n
    assert abs(e["target_noise"] - np.sqrt((truth / n_high).mean()) / mean) < 0.02 * e["target_noise"]
    # against the truth the prediction is off by exactly the planted bias
    want = np.sqrt((bias ** 2).mean()) / mean
    assert abs(e["pred_vs_truth"] - want) < 0.03 * want
    # the raw error also holds the target's noise
    assert e["pred"] > e["pred_vs_truth"]
    # the low-count image's error against the truth is its own Poisson noise
    low_noise = np.sqrt((truth / n_low).mean()) / mean
    assert abs(e["low_vs_truth"] - low_noise) < 0.02 * low_noise
    # plain MC matches the prediction at (low noise / bias)^2 times the primaries
    assert abs(e["mc_primaries_factor"] - (low_noise / want) ** 2) < 0.08 * (low_noise / want) ** 2


def test_perfect_prediction_has_no_finite_factor():
    rng = np.random.default_rng(3)
    truth = np.full((2, 64, 64), 2e-4)
    y, var = _poisson_view(rng, truth, 10 ** 7)
    low = rng.poisson(truth * 10 ** 5) / 10 ** 5
    e = S.scatter_errors(y.copy(), y, var, low)
    assert e["pred"] == 0.0 and e["pred_vs_truth"] == 0.0
    assert e["mc_primaries_factor"] == float("inf")
