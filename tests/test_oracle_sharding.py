"""A run split into disjoint particle-slot ranges, as several GPUs would
split it (each GPU seeds and runs its own range of slots, then the tallies
are summed), equals the whole run: the same counters and integer tallies,
and the FP64 energy tallies to the rounding of the summation order."""

import numpy as np

import ggems_oracle as O
from scene import cfg1_scene


def test_disjoint_slot_ranges_sum_to_the_whole_run():
    n, seed = 2_500_000, 31
    nb, slots = O.batch_layout(n)
    assert nb == 3 and 800_000 < slots < 1 << 20
    sc = cfg1_scene(n_side=60, detector=64)
    whole = sc.run_oracle(n, seed, threads=0)
    prepared = sc.prepare_oracle()
    parts, total = [], {}
    for lo, hi in ((0, 100_000), (100_000, 500_000), (500_000, slots)):
        r = sc.run_oracle(n, seed, threads=0, slots=(lo, hi), prepared=prepared)
        parts.append(dict(r["stats"]))
    for k in whole["stats"]:
        total[k] = sum(p[k] for p in parts)
    assert total == whole["stats"]
    assert total["primaries"] == n
    assert np.array_equal(r["hist"][0], whole["hist"][0]) and np.array_equal(r["scatter"][0], whole["scatter"][0])
    assert np.array_equal(r["hit"], whole["hit"])
    np.testing.assert_allclose(r["edep"], whole["edep"], rtol=1e-12, atol=1e-18)
