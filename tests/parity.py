"""Per-history comparisons of an engine run with the oracle's run of the same
seeded inputs, each printed with its measured value and bound.

Engine and oracle consume the same random streams in the same order, so a
tally differs only where an ulp of a device transcendental (log, exp, cos,
sin) changed one history.  The bounds sit a few times above the measured
distances, so a real bias of a fraction of a percent fails.
"""

import numpy as np


def l1(tag, ref, got, rel, absolute=0):
    """L1 distance of two integer tallies, bounded by rel * (ref total) + absolute."""
    ref = np.asarray(ref).astype(np.int64)
    got = np.asarray(got).astype(np.int64)
    d = int(np.abs(ref - got).sum())
    tot = int(ref.sum())
    bound = rel * tot + absolute
    print("[parity] %s: L1 %d of %d counts (%.2e), bound %.0f" % (tag, d, tot, d / max(tot, 1), bound), flush=True)
    assert d <= bound, (tag, d, tot)
    return d


def total(tag, ref, got, rel):
    """Relative difference of two sums (FP64 energy tallies)."""
    r = float(np.sum(ref, dtype=np.float64))
    g = float(np.sum(got, dtype=np.float64))
    d = abs(g - r) / max(abs(r), 1e-300)
    print("[parity] %s: sum %.6g vs %.6g (rel %.2e), bound %.1e" % (tag, g, r, d, rel), flush=True)
    assert d <= rel, (tag, g, r)
    return d


def counters(tag, rs, gs, keys, rel, absolute=10):
    """Run counters (voxel steps, interactions, ...) within rel * ref + absolute."""
    worst = 0.0
    for k in keys:
        d = abs(gs[k] - rs[k])
        worst = max(worst, d / max(rs[k], 1))
        assert d <= rel * rs[k] + absolute, (tag, k, gs[k], rs[k])
    print("[parity] %s: counters %s worst rel %.2e, bound %.1e" % (tag, ",".join(keys), worst, rel), flush=True)
    return worst


# ---- independent samples (different seeds or sampling modes) ---------------
def _block(a, b):
    nz, ny, nx = a.shape
    return a[: nz // b * b, : ny // b * b, : nx // b * b].reshape(nz // b, b, ny // b, b, nx // b, b).sum(
        axis=(1, 3, 5), dtype=np.float64)


def dose_z(x, y, b=8, min_hits=50):
    """Two-sample z-scores of the energy deposited in blocks of b^3 dosels:
    (E_x - E_y) / sqrt(V_x + V_y), V = the block's sum of squared deposits
    (the variance of a compound-Poisson sum).  Deposits of one history that
    fall in one block are correlated, so V is a slight underestimate; a
    comparison of two samples of one kind measures that inflation."""
    ex, vx, hx = _block(x["edep"], b), _block(x["edep2"], b), _block(x["hit"], b)
    ey, vy, hy = _block(y["edep"], b), _block(y["edep2"], b), _block(y["hit"], b)
    keep = (hx >= min_hits) & (hy >= min_hits)
    return (ex[keep] - ey[keep]) / np.sqrt(vx[keep] + vy[keep])


def pixel_z(x, y, b=8, min_counts=20, key="hist"):
    """z-scores of detector counts in blocks of b x b pixels."""
    def blocks(h):
        ny, nx = h.shape
        return h[: ny // b * b, : nx // b * b].astype(np.float64).reshape(ny // b, b, nx // b, b).sum(axis=(1, 3))
    hx, hy = blocks(x[key][0]), blocks(y[key][0])
    keep = hx + hy >= min_counts
    return (hx[keep] - hy[keep]) / np.sqrt(hx[keep] + hy[keep])


def same_distribution(tag, z_ab, z_null, min_cells):
    """z_ab: the comparison under test; z_null: two samples of one kind (the
    null reference of the same statistic).  Mean z ~ 0 within 6 standard
    errors, the null's chi^2/dof below 1.5, and the tested chi^2/dof within
    6 standard deviations of the null's."""
    m = z_ab.size
    chi_ab, chi_null = float(np.mean(z_ab ** 2)), float(np.mean(z_null ** 2))
    print("[stat] %s: %d cells, mean z %.4f chi2/dof %.4f; null mean z %.4f chi2/dof %.4f"
          % (tag, m, z_ab.mean(), chi_ab, z_null.mean(), chi_null), flush=True)
    assert m > min_cells, (tag, m)
    se = np.sqrt(max(chi_ab, 1.0) / m)
    assert abs(z_ab.mean()) < 6 * se, (tag, z_ab.mean(), se)
    assert chi_null < 1.5, (tag, chi_null)
    assert chi_ab < chi_null + 6 * np.sqrt(2.0 / m) * max(chi_null, 1.0), (tag, chi_ab, chi_null)
