"""The volume creator through the GGEMS API on the GPU: painted volumes
bit-exact against the oracle's painters in every element type, the range
file, and a phantom made this way (the tube of GGEMS's dosimetry example)
used by a simulation, against the oracle on the same files."""

import math
import os

import numpy as np
import pytest

import parity
import volume as V
from ggems_amd import authoring, configs
from ggems_amd.ggems import (GGEMS, GGEMSBox, GGEMSSphere, GGEMSTube, GGEMSVolumeCreatorManager,
                             GGEMSVoxelizedPhantom)
from scene import Scene

pytestmark = pytest.mark.gpu

SHAPES = [("tube", GGEMSTube, (12.0, 9.5, 30.0), (30.0, 12.0, 9.5), (1.5, -2.0, 0.0), 1, "Water"),
          ("box", GGEMSBox, (20.0, 14.0, 11.0), (14.0, 20.0, 11.0), (-3.0, 2.0, 4.0), 2, "RibBone"),
          ("sphere", GGEMSSphere, (6.3,), (6.3,), (2.0, 3.0, -4.0), 3, "Lung")]


def _paint(wd, dims, es, dtype):
    vc = GGEMSVolumeCreatorManager()
    vc.clean()
    vc.set_dimensions(*dims)
    vc.set_element_sizes(*es, "mm")
    vc.set_output(os.path.join(wd, "phantom.mhd"))
    vc.set_range_output(os.path.join(wd, "range.txt"))
    vc.set_material("Air")
    vc.set_data_type(dtype)
    vc.initialize()
    for name, cls, args, size, pos, label, mat in SHAPES:
        s = cls(*args, "mm")
        s.set_position(*pos, "mm")
        s.set_label_value(label)
        s.set_material(mat)
        s.initialize()
        s.draw()
        s.delete()
    vc.write()
    expect = np.zeros(int(np.prod(dims)), V.MET[dtype])
    for name, cls, args, size, pos, label, mat in SHAPES:
        V.draw(expect, dims, es, name, pos, label, size)
    return expect


@pytest.mark.parametrize("dtype", sorted(V.MET))
@pytest.mark.parametrize("dims", [(40, 40, 36), (48, 32, 20)])  # the second is not square in x/y
def test_painted_volume_matches_oracle(tmp_path, dtype, dims):
    es = (1.0, 1.0, 1.5)
    expect = _paint(str(tmp_path), dims, es, dtype)
    got, spacing = authoring.read_mhd(os.path.join(tmp_path, "phantom.mhd"))
    assert got.dtype == expect.dtype and got.shape == (dims[2], dims[1], dims[0])
    assert np.array_equal(got.reshape(-1), expect)
    assert spacing == [1.0, 1.0, 1.5]
    with open(os.path.join(tmp_path, "range.txt")) as f:
        assert f.read() == "0 0 Air\n1 1 Water\n2 2 RibBone\n3 3 Lung\n"
    if dims[0] == dims[1]:
        assert (got == 1).sum() > 1000 and (got == 3).sum() > 500


def test_tube_phantom_simulation_matches_oracle(tmp_path):
    """GGEMS's dosimetry example builds its phantom as an air volume holding
    a water tube (examples/4_Dosimetry_Photon/dosimetry_photon.py:82-98);
    here at 1 mm, then a CT projection of it against the oracle."""
    wd = str(tmp_path)
    configs.setup(wd)
    vc = GGEMSVolumeCreatorManager()
    vc.clean()
    vc.set_dimensions(120, 120, 160)
    vc.set_element_sizes(1.0, 1.0, 1.0, "mm")
    vc.set_output(os.path.join(wd, "phantom.mhd"))
    vc.set_range_output(os.path.join(wd, "range_phantom.txt"))
    vc.set_material("Air")
    vc.set_data_type("MET_INT")
    vc.initialize()
    tube = GGEMSTube(50.0, 50.0, 150.0, "mm")
    tube.set_position(0.0, 0.0, 0.0, "mm")
    tube.set_label_value(1)
    tube.set_material("Water")
    tube.initialize()
    tube.draw()
    vc.write()
    tube.delete()
    phantom = GGEMSVoxelizedPhantom("phantom")
    phantom.set_phantom(os.path.join(wd, "phantom.mhd"), os.path.join(wd, "range_phantom.txt"))
    phantom.set_rotation(0.0, 0.0, 0.0, "deg")
    phantom.set_position(0.0, 0.0, 0.0, "mm")
    ct = configs._detector(wd, n=64)
    configs._common_physics()
    src = configs._source(200_000, detector=64)
    src.set_monoenergy(60.0, "keV")
    g = GGEMS()
    g.initialize(55)
    g.run()
    proj, _ = authoring.read_mhd(os.path.join(wd, "projection.mhd"))
    lab, spacing = authoring.read_mhd(os.path.join(wd, "phantom.mhd"))
    sc = Scene()
    sc.add_phantom(lab.astype(np.uint8), ["Air", "Water"], tuple(spacing))
    sc.add_ct_flat("GOS", (1, 1), (64, 64, 1), (1.0, 1.0, 1.0), configs.SID, configs.SDD, 10e-3)
    sc.set_source((-configs.SID, 0.0, 0.0), math.degrees(math.atan(32 * math.sqrt(2.0) / configs.SDD)), [0.06],
                  mono=True)
    sc.build()
    ref = sc.run_oracle(200_000, 55)
    h_ref = ref["hist"][0]
    assert h_ref.sum() > 1000
    parity.l1("tube phantom projection", h_ref, proj[0], 1e-3, 10)
    g.delete()
    del phantom, ct, src
