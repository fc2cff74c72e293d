"""Input authoring (ggems_amd.authoring): MHD/raw and range files in the
formats GGEMS reads, the seeded phantoms and the filtered Kramers spectrum."""

import os

import numpy as np
import pytest

from ggems_amd import authoring

# NIST XCOM / Hubbell & Seltzer, aluminium, total attenuation with coherent
# scattering, cm^2/g; density 2.699 g/cm3
XCOM_AL = [(0.02, 3.441), (0.03, 1.128), (0.04, 0.5685), (0.05, 0.3681), (0.06, 0.2778), (0.08, 0.2018),
           (0.10, 0.1704), (0.15, 0.1378)]


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.float32, np.float64, np.int32])
def test_mhd_round_trip(tmp_path, dtype):
    vol = (np.arange(4 * 5 * 6) % 7).reshape(4, 5, 6).astype(dtype)
    path = authoring.write_mhd(str(tmp_path / "img.mhd"), vol, (0.5, 1.0, 2.5))
    back, spacing = authoring.read_mhd(path)
    assert back.dtype == vol.dtype and np.array_equal(back, vol)
    assert spacing == [0.5, 1.0, 2.5]


def test_mhd_header_keys(tmp_path):
    """GGEMSMHDImage::Read (src/io/GGEMSMHDImage.cc:129-216) reads DimSize
    (x y z), ElementSpacing, ElementType and ElementDataFile."""
    path = authoring.write_mhd(str(tmp_path / "img.mhd"), np.zeros((3, 4, 5), np.uint8), (1, 2, 3))
    keys = dict(l.split(" = ", 1) for l in open(path).read().splitlines())
    assert keys["DimSize"] == "5 4 3"
    assert keys["ElementSpacing"] == "1 2 3"
    assert keys["ElementType"] == "MET_UCHAR"
    assert keys["ElementDataFile"] == "img.raw"
    assert os.path.getsize(str(tmp_path / "img.raw")) == 60


def test_range_file(tmp_path):
    p = authoring.write_range(str(tmp_path / "range.txt"), ["Air", "Water"])
    assert open(p).read() == "0 0 Air\n1 1 Water\n"


def test_cylinder_phantom():
    lab = authoring.cylinder_phantom(100, 20.0, 1.0)
    assert lab.shape == (100, 100, 100)
    area = lab[0].sum()
    assert abs(area - np.pi * 20 ** 2) < 0.02 * np.pi * 20 ** 2
    assert (lab == lab[0]).all()


def test_ellipsoid_phantom_seeded_and_bounded():
    a = authoring.ellipsoid_body_phantom(20240, shape=(32, 64, 64), voxel=6.4)
    b = authoring.ellipsoid_body_phantom(20240, shape=(32, 64, 64), voxel=6.4)
    c = authoring.ellipsoid_body_phantom(20241, shape=(32, 64, 64), voxel=6.4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.max() <= 4
    # air exactly outside the elliptic body cylinder
    xs = (np.arange(64) + 0.5) * 6.4 - 64 * 6.4 / 2
    body = (xs[None, :] / 180.0) ** 2 + (xs[:, None] / 140.0) ** 2 <= 1.0
    assert (a[:, ~body] == 0).all() and (a[:, body] > 0).all()
    assert len(np.unique(a)) == 5


def test_aluminium_mu_against_xcom():
    mu = authoring.aluminium_mu()
    for e, mu_rho in XCOM_AL:
        got = mu(e) / 2.699
        assert abs(got / mu_rho - 1) < (0.06 if e < 0.03 else 0.04), (e, got, mu_rho)


def test_kramers_spectrum(tmp_path):
    e, w = authoring.kramers_spectrum(120.0, 20.0, al_mm=0.0)
    assert e[0] == pytest.approx(0.020) and e[-1] == pytest.approx(0.120) and len(e) == 101
    assert w[-1] == 0.0 and np.all(np.diff(w) < 0)
    mu = authoring.aluminium_mu()
    _, wf = authoring.kramers_spectrum(120.0, 20.0, al_mm=2.5, mu_al_per_cm=mu)
    # filtration hardens the beam: the mean energy rises
    assert (e * wf).sum() / wf.sum() > (e * w).sum() / w.sum() + 0.005
    p = authoring.write_spectrum(str(tmp_path / "s.dat"), e, wf)
    back = np.loadtxt(p)
    assert np.allclose(back[:, 0], e) and np.allclose(back[:, 1], wf, rtol=1e-8)
