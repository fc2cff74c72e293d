"""GGEMS scripts through libggems.so (the drop-in boundary) vs the CPU oracle.

The scene is set up exactly as a GGEMS user script does (ggems_amd.configs),
the outputs are the MHD files GGEMS writes, and the oracle transports the
same inputs with the same seed.
"""

import os

import numpy as np
import pytest

import parity
from ggems_amd import authoring, configs
from scene import Scene, cfg1_scene

pytestmark = pytest.mark.gpu


def test_cfg1_script_outputs_match_oracle(tmp_path):
    n = 300_000
    keep = configs.cfg1(str(tmp_path), n_particles=n, dosimetry=True)
    g = configs.run(seed=777)
    proj, spacing = authoring.read_mhd(os.path.join(tmp_path, "projection.mhd"))
    scat, _ = authoring.read_mhd(os.path.join(tmp_path, "projection-scatter.mhd"))
    assert proj.shape == (1, 256, 256) and spacing == [1.0, 1.0, 1.0]
    stats = g.engine_stats()
    assert stats["primaries"] == n
    ref = cfg1_scene(n_side=100, detector=256, dosimetry=True).run_oracle(n, 777)
    parity.l1("cfg1 script projection", ref["hist"][0], proj[0], 1e-3, 10)
    parity.l1("cfg1 script scatter", ref["scatter"][0], scat[0], 5e-3, 10)
    edep, _ = authoring.read_mhd(os.path.join(tmp_path, "dosimetry_edep.mhd"))
    hit, _ = authoring.read_mhd(os.path.join(tmp_path, "dosimetry_hit.mhd"))
    parity.total("cfg1 script edep", ref["edep"], edep, 1e-3)
    parity.l1("cfg1 script hit", ref["hit"], hit, 5e-3, 10)
    dose, _ = authoring.read_mhd(os.path.join(tmp_path, "dosimetry_dose.mhd"))
    unc, _ = authoring.read_mhd(os.path.join(tmp_path, "dosimetry_uncertainty.mhd"))
    assert dose.dtype == np.float32 and unc.dtype == np.float32
    assert np.all(unc[hit <= 1] == 1.0)
    g.delete()
    del keep


def test_voxel_ct_small_script_runs_and_matches_oracle(tmp_path):
    """The voxel-CT scene on a reduced body phantom (64x128x128 of 3.2 mm)."""
    n = 200_000
    shape, voxel = (64, 128, 128), 3.2
    keep = configs.voxel_ct(str(tmp_path), n_particles=n, seed=7, shape=shape, voxel=voxel)
    g = configs.run(seed=99)
    proj, _ = authoring.read_mhd(os.path.join(tmp_path, "projection.mhd"))
    edep, _ = authoring.read_mhd(os.path.join(tmp_path, "dosimetry_edep.mhd"))
    mhd, rng, spec = configs.voxel_ct_inputs(str(tmp_path), 7, shape, voxel)
    lab, _ = authoring.read_mhd(mhd)
    e, w = np.loadtxt(spec, unpack=True)
    sc = Scene()
    sc.add_phantom(lab, authoring.BODY_MATERIALS, (voxel, voxel, voxel), dosimetry={"photon_tracking": False})
    sc.add_ct_flat("GOS", (1, 1), (256, 256, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    import math
    sc.set_source((-600.0, 0.0, 0.0), math.degrees(math.atan(128 * math.sqrt(2.0) / 1000.0)), e, intensities=w)
    sc.build()
    ref = sc.run_oracle(n, 99)
    parity.l1("voxel-CT script projection", ref["hist"][0], proj[0], 5e-3, 10)
    parity.total("voxel-CT script edep", ref["edep"], edep, 1e-3)
    g.delete()
    del keep
