"""The product's host side (libggems.so) on the CPU: physics tables built
through the GGEMS API are bit-identical to the oracle's, and the API keeps
GGEMS's argument meaning and error behaviour.  No GPU call is made."""

import numpy as np
import pytest

import ggems_oracle as O
from ggems_amd import configs
from ggems_amd.ggems import (GGEMSCrossSections, GGEMSError, GGEMSMaterials, GGEMSMaterialsDatabaseManager,
                             GGEMSProcessesManager, GGEMSVoxelizedPhantom, clean_safely)

F = np.float32
MATS = ["Water", "Air", "Lung", "Adipose", "SoftTissus_15", "RibBone", "GOS"]
PROCESSES = {"Compton": 0, "Photoelectric": 1, "Rayleigh": 2}


@pytest.fixture(scope="module")
def product_tables():
    clean_safely()
    GGEMSMaterialsDatabaseManager().set_materials(configs.MATERIALS_TXT)
    pm = GGEMSProcessesManager()
    pm.set_cross_section_table_number_of_bins(220)
    pm.set_cross_section_table_energy_min(1.0, "keV")
    pm.set_cross_section_table_energy_max(1.0, "MeV")
    m = GGEMSMaterials()
    for name in MATS:
        m.add_material(name)
    m.initialize()
    cs = GGEMSCrossSections(m)
    for p in PROCESSES:
        cs.add_process(p, "gamma")
    cs.initialize()
    out = {"ebins": cs.get_energy_bins(220),
           "cs": {(p, n): cs.get_table(p, n, 220) for p in PROCESSES for n in MATS},
           "density": {n: m.get_density(n) for n in MATS},
           "cut": {n: m.get_photon_cut(n) for n in MATS}}
    yield out
    cs.clean()
    m.clean()


@pytest.fixture(scope="module")
def oracle_tables():
    mt = O.material_tables(MATS, cut_distance=1e-3, cut_emin=F(1e-3))
    eb = O.energy_bins(220, F(1e-3), F(1.0))
    cs, _ = O.cs_tables(mt, 220, eb)
    return mt, eb, cs


def test_energy_grid_bit_exact(product_tables, oracle_tables):
    assert np.array_equal(product_tables["ebins"], oracle_tables[1])


def test_cross_sections_bit_exact(product_tables, oracle_tables):
    _, _, cs = oracle_tables
    for (p, name), row in product_tables["cs"].items():
        assert np.array_equal(row, cs[PROCESSES[p], MATS.index(name)]), (p, name)


def test_photon_cut_bit_exact(product_tables, oracle_tables):
    mt = oracle_tables[0]
    for i, name in enumerate(MATS):
        assert product_tables["cut"][name] == F(mt.photon_cut[i]), name


def test_density_in_g_cm3(product_tables):
    assert abs(product_tables["density"]["Water"] - 1.0) < 1e-6
    assert abs(product_tables["density"]["GOS"] - 7.44) < 1e-5


def test_unknown_material_raises():
    clean_safely()
    GGEMSMaterialsDatabaseManager().set_materials(configs.MATERIALS_TXT)
    m = GGEMSMaterials()
    m.add_material("Unobtainium")
    with pytest.raises(GGEMSError):
        m.initialize()
    m.clean()


def test_unknown_process_raises():
    """GGEMSProcessesManager::AddProcess forwards to every registered navigator
    (src/physics/GGEMSProcessesManager.cc:126-140), whose cross sections reject
    an unknown process."""
    clean_safely()
    pm = GGEMSProcessesManager()
    pm.add_process("Bremsstrahlung", "gamma", "all")  # no navigator yet: nothing to do, as in GGEMS
    phantom = GGEMSVoxelizedPhantom("phantom")
    with pytest.raises(GGEMSError):
        pm.add_process("Bremsstrahlung", "gamma", "all")
    pm.add_process("Compton", "gamma", "phantom")
    with pytest.raises(GGEMSError):
        pm.add_process("Compton", "gamma", "no_such_phantom")
    del phantom
    clean_safely()
