"""ggh_drr_noalloc, the ray tracer that allocates nothing, through the C ABI
(needs an MI355X): its device image equals ggh_drr's bit for bit on two calls
in a row on one GPU, the second with a smaller input (fewer energy bins with
weight, a smaller panel), so the second reads none of the first call's
staged weights; and it refuses, allocating nothing, physics tables beyond its
staging limit and a solid table not yet on the device."""

import numpy as np
import pytest
import torch

from ggems_amd import authoring
from ggems_amd.engine import EngineError
from scene import Scene, cylinder_phantom

pytestmark = pytest.mark.gpu


def _scene(spectrum, detector, nbins=220):
    sc = Scene(nbins=nbins)
    sc.add_phantom(cylinder_phantom(48, radius=14.0), ["Water", "RibBone"], (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (detector[0], detector[1], 1), (1.2, 1.2, 1.0), 600.0, 1000.0, 10e-3)
    if spectrum:
        e, w = authoring.kramers_spectrum(120.0, 20.0, 2.5, authoring.aluminium_mu())
        sc.set_source((-600.0, 0.0, 0.0), 6.0, e, intensities=w)
    else:
        sc.set_source((-600.0, 0.0, 0.0), 6.0, [0.06], mono=True)
    return sc.build()


def _both(sc):
    eng = sc.make_engine()
    k = sc.engine_solids["box"][0]
    want = eng.drr(k)  # the allocating entry point, which also uploads the solid table
    img = torch.full(want.shape, float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    eng.drr_noalloc(k, img.data_ptr())
    eng.synchronize()
    got = img.cpu().numpy()
    eng.close()
    return want, got


def test_noalloc_drr_equals_drr_twice_second_smaller():
    big, big_got = _both(_scene(True, (96, 80)))
    assert big.shape == (80, 96) and big.min() < 0.5 < big.max()
    assert np.array_equal(big_got, big)
    small, small_got = _both(_scene(False, (40, 24)))
    assert small.shape == (24, 40)
    assert np.array_equal(small_got, small)
    assert not np.array_equal(small[:8, :8], big[:8, :8])


def test_noalloc_drr_refuses_without_allocating():
    sc = _scene(False, (16, 16), nbins=1100)
    eng = sc.make_engine()
    k = sc.engine_solids["box"][0]
    img = torch.zeros((16, 16), dtype=torch.float32, device="cuda")
    # the solid table is not on the device before a run or an allocating ggh_drr
    with pytest.raises(EngineError, match="not on the device"):
        eng.drr_noalloc(k, img.data_ptr())
    eng.drr(k)
    with pytest.raises(EngineError, match="at most 1024 energy bins"):
        eng.drr_noalloc(k, img.data_ptr())
    eng.close()
