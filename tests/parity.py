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
