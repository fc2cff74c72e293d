"""Pins of the CPU oracle against what the mathematics and published data fix.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it:

* std::mt19937 and the slot seeding against numpy's independent MT19937;
* the exact KissUniform mapping used by the HIP engine against the
  reference's literal FP64 formula;
* the energy grid's closed form;
* total attenuation of water against the NIST XCOM tables, and the
  Compton cross section of hydrogen against the Klein-Nishina formula;
* atom densities against rho * N_A * w / A;
* the Klein-Nishina and Livermore Rayleigh samplers against their analytic
  densities (chi-square), and Compton kinematics sample by sample;
* Beer-Lambert transmission of a water slab with the left-bin cross section;
* voxel index sequences along axis-aligned and diagonal rays against exact
  rational ray/box enumeration;
* dose / uncertainty special cases (Ma et al. closed forms).
"""

import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import ggems_oracle as O
from scene import Scene

F = np.float32


# ---------------------------------------------------------------------------
# random numbers
# ---------------------------------------------------------------------------
def _numpy_mt(seed, n):
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    return bg.random_raw(n).astype(np.uint32)


def test_mt19937_known_answer_and_numpy():
    out = np.zeros(10000, np.uint32)
    O.lib().ggo_mt19937_stream(5489, 10000, O.ptr(out))
    assert int(out[-1]) == 4123659995  # ISO C++ [rand.predef]: 10000th output of default-seeded mt19937
    for seed in (1, 777, 20240, 4294967295):
        out = np.zeros(4096, np.uint32)
        O.lib().ggo_mt19937_stream(seed, 4096, O.ptr(out))
        assert np.array_equal(out, _numpy_mt(seed, 4096))


def test_slot_seeding_order():
    """GGEMSPseudoRandomGenerator::InitializeSeeds (src/randoms/GGEMSPseudoRandomGenerator.cc:140-167):
    slot j takes the draws 4j..4j+3 of mt19937(seed), fifth word 0."""
    st = O.seed_slots(777, 2000)
    ref = _numpy_mt(777, 8000).reshape(2000, 4)
    assert np.array_equal(st[:4].T, ref)
    assert not st[4].any()


def _kiss_literal(s):
    """The reference expression (GGEMSKissEngine.hh:73-74) with FP64 enabled."""
    f = np.float32(np.uint32(s))
    return np.float32((np.float64(f) / 4294967295.0) * np.float64(np.float32(1.0) - np.float32(2.0 ** -23)))


def _kiss_bits(s):
    """The division-free mapping of ggems_amd/csrc/engine/transport_device.hpp (uniform_from_sum)."""
    x = np.float32(np.float32(np.uint32(s)) * np.float32(2.0 ** -32))
    b = int(np.array([x], np.float32).view(np.uint32)[0])
    b -= 2 if ((b & 0x7FFFFF) - 1) % (2 ** 32) >= 4218928 else 1
    return np.float32(0.0) if s == 0 else np.array([b], np.uint32).view(np.float32)[0]


def test_kiss_uniform_mapping_matches_literal_formula():
    rng = np.random.default_rng(5)
    cases = list(rng.integers(0, 2 ** 32, 20000, dtype=np.uint64))
    # every binade edge and the mantissa threshold 12607537 in every binade
    for e in range(0, 32):
        for m in (0x800000, 0x800001, 12607535, 12607536, 12607537, 12607538, 0xFFFFFF):
            v = (m << e) >> 23 if e < 23 else m << (e - 23)
            if v < 2 ** 32:
                cases.extend([v, v - 1, v + 1])
    cases.extend([0, 1, 2, 3, 2 ** 32 - 1, 2 ** 32 - 128, 2 ** 32 - 129])
    for s in cases:
        s = int(s) % (2 ** 32)
        assert _kiss_literal(s) == _kiss_bits(s), s
    assert _kiss_literal(2 ** 32 - 1) < 1.0


def test_kiss_stream_against_literal_formula():
    st = O.seed_slots(777, 4)
    for slot in range(4):
        got, _ = O.kiss_stream(st[:, slot], 200)
        x, y, z, w, c = (int(v) for v in st[:, slot])
        for i in range(200):
            y ^= (y << 5) & 0xFFFFFFFF
            y ^= y >> 7
            y ^= (y << 22) & 0xFFFFFFFF
            t = (z + w + c) & 0xFFFFFFFF
            t_signed = t - 2 ** 32 if t >= 2 ** 31 else t
            z = w
            c = 1 if t_signed < 0 else 0
            w = t & 0x7FFFFFFF
            x = (x + 1411392427) & 0xFFFFFFFF
            assert got[i] == _kiss_literal((x + y + w) & 0xFFFFFFFF)


# ---------------------------------------------------------------------------
# tables
# ---------------------------------------------------------------------------
def test_energy_grid_closed_form():
    e = O.energy_bins(220, F(1e-3), F(1.0))
    assert e[0] == F(1e-3)
    assert abs(e[-1] - 1.0) <= 2 * np.spacing(F(1.0))
    assert np.all(np.diff(e) > 0)
    # e_i = Emin exp(ln(Emax/Emin) i/(n-1)), evaluated in FP32 (argument rounding ~ 5e-7 relative)
    i = np.arange(220)
    dev = np.log(e.astype(np.float64) / 1e-3) - math.log(1000.0) * i / 219
    assert np.abs(dev).max() < 5e-6


# NIST XCOM / Hubbell & Seltzer, water (liquid), total attenuation with
# coherent scattering, cm^2/g
XCOM_WATER = [(0.02, 0.8096), (0.03, 0.3756), (0.04, 0.2683), (0.05, 0.2269), (0.06, 0.2059), (0.08, 0.1837),
              (0.10, 0.1707), (0.15, 0.1505), (0.20, 0.1370), (0.30, 0.1186), (0.50, 0.09687), (1.00, 0.07072)]


@pytest.fixture(scope="module")
def water_tables():
    mt = O.material_tables(["Water", "GOS", "RibBone"], cut_distance=0.1, cut_emin=F(1e-3))
    eb = O.energy_bins(220, F(1e-3), F(1.0))
    cs, atom = O.cs_tables(mt, 220, eb)
    return mt, eb, cs, atom


def test_water_attenuation_against_xcom(water_tables):
    mt, eb, cs, _ = water_tables
    tot = cs[:, 0, :].sum(axis=0).astype(np.float64)  # mm^-1
    rho = 1.0
    for e, mu_rho in XCOM_WATER:
        ours = np.exp(np.interp(np.log(e), np.log(eb.astype(np.float64)), np.log(tot))) * 10.0 / rho
        tol = 0.06 if e < 0.03 else 0.03
        assert abs(ours / mu_rho - 1.0) < tol, (e, ours, mu_rho)


def _klein_nishina_barn(e_mev):
    k = e_mev / 0.510998910
    re2 = 0.0794077  # r_e^2 in barn
    l = math.log(1 + 2 * k)
    return 2 * math.pi * re2 * ((1 + k) / k ** 2 * (2 * (1 + k) / (1 + 2 * k) - l / k) + l / (2 * k)
                                - (1 + 3 * k) / (1 + 2 * k) ** 2)


def test_compton_hydrogen_against_klein_nishina():
    barn_mm2 = 1e-22
    for e in (0.1, 0.2, 0.5, 1.0):
        ours = O.lib().ggo_cs_compton_atom(F(e), 1) / barn_mm2
        kn = _klein_nishina_barn(e)
        assert abs(ours / kn - 1.0) < 0.03, (e, ours, kn)
    # Z electrons per atom at high energy
    assert abs(O.lib().ggo_cs_compton_atom(F(1.0), 8) / O.lib().ggo_cs_compton_atom(F(1.0), 1) - 8.0) < 0.3


def test_atom_density_closed_form(water_tables):
    mt = water_tables[0]
    a = O.mt_arrays(mt)
    na = 6.02214076e23
    # water: H 11.1 %, O 88.9 % by mass (data/materials.txt), rho = 1 g/cm3; atoms/mm3
    idx, n = a["elem_index"][0], a["n_elements"][0]
    z = list(a["z"][idx:idx + n])
    dens = dict(zip(z, a["atom_density"][idx:idx + n]))
    for zz, w, molar in ((1, 0.111, 1.00794), (8, 0.889, 15.9994)):
        expect = 1.0 * na * w / molar / 1e3  # per mm^3
        assert abs(dens[zz] / expect - 1) < 2e-3, (zz, dens[zz], expect)


def test_photon_cut_is_the_absorption_length(water_tables):
    """photon_energy_cut: 5 absorption lengths (PE + Compton) equal the 0.1 mm cut."""
    mt, _, _, _ = water_tables
    cut = O.lib().ggo_photon_cut(C.byref(mt), 0, F(0.1), F(1e-3))
    assert 1.0e-3 < cut < 1.3e-3
    # a longer distance cut needs a higher energy
    assert O.lib().ggo_photon_cut(C.byref(mt), 0, F(1.0), F(1e-3)) > cut
    assert O.lib().ggo_photon_cut(C.byref(mt), 0, F(10.0), F(1e-3)) > O.lib().ggo_photon_cut(
        C.byref(mt), 0, F(1.0), F(1e-3))


# ---------------------------------------------------------------------------
# samplers
# ---------------------------------------------------------------------------
def _chi2_p(obs, exp):
    from scipy.stats import chi2
    m = exp > 5
    x2 = float(np.sum((obs[m] - exp[m]) ** 2 / exp[m]))
    return chi2.sf(x2, int(m.sum()) - 1)


def test_klein_nishina_sampler():
    e0 = 0.06
    k = e0 / 0.510998910
    st = O.seed_slots(11, 1)[:, 0].copy()
    n = 60000
    cos = np.empty(n)
    for i in range(n):
        d = np.array([0, 0, 1], F)
        e1 = O.lib().ggo_sample_compton(F(e0), O.ptr(d), O.ptr(st))
        cos[i] = d[2]
        # Compton kinematics: E' = E / (1 + k (1 - cos))
        assert abs(e1 - e0 / (1 + k * (1 - d[2]))) < 2e-6 * e0
    edges = np.linspace(-1, 1, 41)
    cc = np.linspace(-1, 1, 40001)
    p = 1 / (1 + k * (1 - cc))
    pdf = p ** 2 * (p + 1 / p - (1 - cc ** 2))
    cdf = np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) / 2 * np.diff(cc))])
    exp = np.diff(np.interp(edges, cc, cdf)) / cdf[-1] * n
    obs, _ = np.histogram(cos, edges)
    assert _chi2_p(obs, exp) > 1e-3


@pytest.mark.parametrize("energy", [0.03, 0.08])
def test_rayleigh_sampler_against_form_factor_density(energy):
    sc = Scene()
    sc.add_phantom(np.zeros((2, 2, 2), np.uint8), ["Calcium"], (1.0, 1.0, 1.0))
    sc.set_source((-600.0, 0, 0), 1.0, [0.06], mono=True)
    sc.build()
    ph, _, _ = sc._oracle_structs({})
    z = 20
    pp = sc.pp
    kxx = sc.factor * energy * energy
    st = O.seed_slots(3, 1)[:, 0].copy()
    status = C.c_int(0)
    n = 40000
    cos = np.empty(n)
    for i in range(n):
        d = np.array([0, 0, 1], F)
        O.lib().ggo_sample_rayleigh(C.byref(ph), 0, F(energy), O.ptr(d), O.ptr(st), C.byref(status))
        cos[i] = d[2]
    cc = np.linspace(-1, 1, 40001)
    ff = sum(pp[i, z] * (1 + pp[3 + i, z] * kxx * (1 - cc)) ** (-pp[6 + i, z]) for i in range(3))
    pdf = (1 + cc ** 2) * ff
    cdf = np.concatenate([[0], np.cumsum((pdf[1:] + pdf[:-1]) / 2 * np.diff(cc))])
    edges = np.linspace(-1, 1, 41)
    exp = np.diff(np.interp(edges, cc, cdf)) / cdf[-1] * n
    obs, _ = np.histogram(cos, edges)
    assert _chi2_p(obs, exp) > 1e-3


def test_rayleigh_kills_below_250_ev():
    sc = Scene()
    sc.add_phantom(np.zeros((2, 2, 2), np.uint8), ["Water"], (1.0, 1.0, 1.0))
    sc.set_source((-600.0, 0, 0), 1.0, [0.06], mono=True)
    sc.build()
    ph, _, _ = sc._oracle_structs({})
    st = O.seed_slots(3, 1)[:, 0].copy()
    status = C.c_int(0)
    d = np.array([0, 0, 1], F)
    O.lib().ggo_sample_rayleigh(C.byref(ph), 0, F(200e-6), O.ptr(d), O.ptr(st), C.byref(status))
    assert status.value == 1


# ---------------------------------------------------------------------------
# transport
# ---------------------------------------------------------------------------
def test_beer_lambert_water_slab_left_bin():
    """Photoelectric only: a photon reaches the (thick, fully absorbing)
    detector iff it crossed the slab without interaction, so the detector
    count is Binomial(N, exp(-sigma_PE(E_left-bin) * L))."""
    L, e0, n = 50, 0.030, 100_000
    sc = Scene(processes=(1,))
    sc.add_phantom(np.zeros((20, 20, L), np.uint8), ["Water"], (1.0, 1.0, 1.0))
    sc.add_ct_flat("GOS", (1, 1), (8, 8, 1), (2.0, 2.0, 30.0), 600.0, 1000.0, 1e-3, store_scatter=False)
    sc.set_source((-600.0, 0.0, 0.0), 0.01, [e0], mono=True)
    sc.build()
    r = sc.run_oracle(n, 4242, threads=4)
    counts = int(r["hist"][0].sum())
    eid = int(np.searchsorted(sc.ebins, F(e0), side="right") - 1)
    sigma_left = float(sc.cs[1, 0, eid])
    sigma_interp = float(np.interp(e0, sc.ebins[eid:eid + 2], sc.cs[1, 0, eid:eid + 2]))
    p = math.exp(-sigma_left * L)
    sd = math.sqrt(n * p * (1 - p))
    assert abs(counts - n * p) < 4 * sd, (counts, n * p, sd)
    # the interpolated cross section gives a clearly different transmission
    assert abs(n * math.exp(-sigma_interp * L) - n * p) > 8 * sd


def _walk(scene, p, d, maxn=4096):
    """Voxels visited from p (inside the phantom) along d with no interaction."""
    ph, solids, _ = scene._oracle_structs({})
    out = np.zeros(maxn, np.int32)
    k = O.lib().ggo_voxel_walk(C.byref(solids[0]), O.ptr(np.array(p, F)), O.ptr(np.array(d, F)), O.ptr(out), maxn)
    return list(out[:k])


def _exact_walk(n, vs, p, d):
    """Voxels crossed by the ray p + t d, t >= 0 (exact rationals), in order."""
    p = [Fraction(float(v)) for v in p]
    d = [Fraction(float(v)) for v in d]
    size = Fraction(float(vs))
    lo = [-Fraction(n[i]) * size / 2 for i in range(3)]
    ts = {Fraction(0)}
    for i in range(3):
        if d[i] != 0:
            for k in range(n[i] + 1):
                t = (lo[i] + k * size - p[i]) / d[i]
                if t > 0:
                    ts.add(t)
    ts = sorted(ts)
    seq = []
    for a, b in zip(ts[:-1], ts[1:]):
        m = (a + b) / 2
        idx = [math.floor((p[i] + m * d[i] - lo[i]) / size) for i in range(3)]
        if all(0 <= idx[i] < n[i] for i in range(3)):
            v = idx[0] + idx[1] * n[0] + idx[2] * n[0] * n[1]
            if not seq or seq[-1] != v:
                seq.append(v)
        else:
            break
    return seq


def _box_scene(shape):
    sc = Scene()
    sc.add_phantom(np.zeros(shape, np.uint8), ["Water"], (1.0, 1.0, 1.0))
    sc.set_source((-600.0, 0, 0), 1.0, [0.06], mono=True)
    return sc.build()


def test_voxel_walk_axis_aligned_rays():
    sc = _box_scene((6, 5, 4))  # x in [-2, 2], y in [-2.5, 2.5], z in [-3, 3]
    # through voxel centres along +x, -y, +z, starting just inside a face
    assert _walk(sc, (-1.999, -2.0, -2.5), (1, 0, 0)) == [i for i in range(4)]
    assert _walk(sc, (0.5, 2.499, 2.5), (0, -1, 0)) == [2 + j * 4 + 5 * 20 for j in range(4, -1, -1)]
    assert _walk(sc, (-1.5, 2.0, -2.999), (0, 0, 1)) == [0 + 4 * 4 + k * 20 for k in range(6)]
    # a ray along a voxel face is nudged to alternate sides of it (safety tolerance)
    assert _walk(sc, (-1.5, 1.5, -2.999), (0, 0, 1)) == [16, 32, 56, 72, 96, 112]


def test_voxel_walk_diagonals_against_exact_enumeration():
    n = (4, 4, 4)
    sc = _box_scene((4, 4, 4))
    rng = np.random.default_rng(17)
    checked = 0
    for _ in range(60):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        d = d.astype(F)
        p = rng.uniform(-1.9, 1.9, 3).astype(F)
        got = _walk(sc, p, d)
        exact = _exact_walk(n, 1.0, p, d)
        # the reference nudges 5e-4 mm past every boundary, so a corner
        # voxel whose chord is shorter than the nudge can be skipped; every
        # visited voxel is on the exact path, in the exact order
        assert got[0] == exact[0]
        assert set(got) <= set(exact)
        assert [v for v in exact if v in got] == got
        assert len(exact) - len(got) <= 2
        checked += len(got) == len(exact)
    assert checked > 40


# ---------------------------------------------------------------------------
# dose
# ---------------------------------------------------------------------------
def test_dose_and_uncertainty_closed_forms():
    sc = Scene()
    sc.add_phantom(np.zeros((1, 1, 4), np.uint8), ["Water"], (1.0, 1.0, 1.0), dosimetry={})
    sc.set_source((-600.0, 0, 0), 1.0, [0.06], mono=True)
    sc.build()
    t = {}
    ph, solids, _ = sc._oracle_structs(t)
    # dosel 0: 3 deposits of 1 MeV; dosel 1: 0.2 and 0.6 MeV; dosel 2: one hit; dosel 3: none
    t["edep"].reshape(-1)[:] = [3.0, 0.8, 0.5, 0.0]
    t["edep2"].reshape(-1)[:] = [3.0, 0.04 + 0.36, 0.25, 0.0]
    t["hit"].reshape(-1)[:] = [3, 2, 1, 0]
    dens = (sc.arrays["density"]).astype(F)
    dose = np.zeros(4, F)
    unc = np.zeros(4, F)
    O.lib().ggo_compute_dose(C.byref(solids[0]), O.ptr(dens), F(1.0), 0, F(0.0), O.ptr(dose), O.ptr(unc))
    mev_j, kg = 1.602176634e-13, 1e-6  # 1 mm^3 of water
    assert abs(dose[0] / (3.0 * mev_j / kg) - 1) < 1e-5
    assert unc[0] == 0.0
    assert abs(unc[1] - 0.4 / 0.8) < 1e-6  # |e1-e2|/(e1+e2)
    assert unc[2] == 1.0 and unc[3] == 1.0 and dose[3] == 0.0
    # water reference and minimum density
    O.lib().ggo_compute_dose(C.byref(solids[0]), O.ptr(dens), F(2.0), 1, F(0.0), O.ptr(dose), O.ptr(unc))
    assert abs(dose[0] / (6.0 * mev_j / kg) - 1) < 1e-5
    O.lib().ggo_compute_dose(C.byref(solids[0]), O.ptr(dens), F(1.0), 0, F(2.0 * 6.241509704e21 / 1e3), O.ptr(dose),
                             O.ptr(unc))
    assert not dose.any()


def test_batch_layout_follows_the_float_ceiling_and_fits_the_slot_buffer():
    """GGEMSSource::OrganizeParticlesInBatch takes ceil(float(n) / 2^20)
    batches (src/sources/GGEMSSource.cc:403-425).  Wherever its largest batch
    fits the 2^20 slots the layout is the reference's; where float rounding
    leaves one batch too few (2^24 + 1 primaries gives 16 batches, the first
    of 2^20 + 1) the exact ceiling is taken."""
    slots = 1 << 20
    rng = np.random.default_rng(11)
    ns = [1, 3, slots, slots + 1, 1_000_000_000, (1 << 24) + 1, (1 << 25) + 7, (1 << 32) + 12_345, 10 ** 12]
    ns += [int(v) for v in rng.integers(1, 1 << 40, 2000)]
    ns += [int(k) * slots + int(d) for k, d in zip(rng.integers(16, 1 << 20, 500), rng.integers(1, 64, 500))]
    n_exact = 0
    for n in ns:
        nb, largest = O.batch_layout(n)
        ref = max(1, int(np.ceil(np.float32(n) / np.float32(slots))))
        assert nb * slots >= n and largest <= slots and largest == -(-n // nb), n
        if ref * slots >= n:
            assert nb == ref, n
        else:
            n_exact += 1
            assert nb == -(-n // slots), n
    assert O.batch_layout((1 << 24) + 1) == (17, -(-((1 << 24) + 1) // 17))
    assert n_exact > 100  # the rounding case is common above 2^24
