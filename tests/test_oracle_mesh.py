"""Pins of the oracle's meshed-solid navigation: the GGEMS octree, the
ray/mesh intersection list and the mesh transport, against geometry that
has closed forms (spheres, exact rational ray/triangle intersection,
Beer-Lambert along the mesh's own chord)."""

import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import ggems_oracle as O
from ggems_amd import authoring
from scene import Scene

F = np.float32


@pytest.fixture(scope="module")
def sphere():
    tri = authoring.icosphere(50.0, 3).astype(F)
    m = O.mesh_build(tri, 4)
    return tri, m, O.mesh_struct(m)


def test_octree_holds_every_triangle_once(sphere):
    tri, m, _ = sphere
    assert len(m["node_first"]) == (8 ** 4 - 1) // 7
    seen = np.zeros(len(tri), int)
    for k in range(len(m["node_first"])):
        t = m["node_first"][k]
        while t >= 0:
            seen[t] += 1
            t = m["tri_next"][t]
    assert (seen == 1).all()
    # octree box: bounds grown by 2.5 % and rounded outwards to whole mm
    lo, hi = tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)
    assert (m["obb_min"] <= lo - 0.025 * (hi - lo) + 1e-4).all() and (m["obb_max"] >= hi).all()
    assert (m["obb_min"] == np.floor(m["obb_min"])).all() and (m["obb_max"] == np.ceil(m["obb_max"])).all()


def test_sphere_hits(sphere):
    tri, m, ms = sphere
    rng = np.random.default_rng(4)
    for _ in range(300):
        target = rng.normal(size=3)
        target = target / np.linalg.norm(target) * rng.uniform(0, 35)
        p = rng.normal(size=3)
        p = p / np.linalg.norm(p) * 120.0
        d = (target - p) / np.linalg.norm(target - p)
        n, pts = O.mesh_hits(ms, p.astype(F), d.astype(F))
        assert n == 2
        r = np.linalg.norm(pts, axis=1)
        assert (np.abs(r - 50.0) < 0.5).all()
        # sorted by distance from the start
        assert np.linalg.norm(pts[0] - p) <= np.linalg.norm(pts[1] - p)
        # from inside: one surface ahead; pointing away from a sphere: none
        n_in, _ = O.mesh_hits(ms, target.astype(F), d.astype(F))
        assert n_in == 1
        n_away, _ = O.mesh_hits(ms, p.astype(F), (-d).astype(F))
        assert n_away == 0


def _exact_hit(p, d, t):
    """exact rational ray/triangle intersection parameter (None when missed)"""
    P = [Fraction(float(v)) for v in p]
    D = [Fraction(float(v)) for v in d]
    V = [[Fraction(float(x)) for x in t[3 * i:3 * i + 3]] for i in range(3)]
    e1 = [V[1][i] - V[0][i] for i in range(3)]
    e2 = [V[2][i] - V[0][i] for i in range(3)]

    def cross(a, b):
        return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]

    def dot(a, b):
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]

    h = cross(D, e2)
    a = dot(e1, h)
    if a == 0:
        return None
    s = [P[i] - V[0][i] for i in range(3)]
    u = dot(s, h) / a
    q = cross(s, e1)
    v = dot(q, D) / a
    tt = dot(q, e2) / a
    if u < 0 or u > 1 or v < 0 or u + v > 1 or tt <= 0:
        return None
    return float(tt), float(u), float(v)


def test_moller_trumbore_against_exact_arithmetic():
    rng = np.random.default_rng(8)
    agree = 0
    for _ in range(2000):
        t = rng.uniform(-10, 10, 9).astype(F)
        m = O.mesh_build(t.reshape(1, 3, 3), 1)
        ms = O.mesh_struct(m)
        p = rng.uniform(-20, 20, 3).astype(F)
        target = (t[0:3] * 0.3 + t[3:6] * 0.3 + t[6:9] * 0.4 + rng.normal(scale=3, size=3)).astype(F)
        d = (target - p) / np.linalg.norm(target - p)
        n, pts = O.mesh_hits(ms, p, d.astype(F))
        ex = _exact_hit(p, d.astype(F), t)
        if ex is not None:
            tt, u, v = ex
            margin = min(u, v, 1 - u - v)
            if margin > 1e-4 and tt > 1e-4:
                assert n == 1
                q = p.astype(np.float64) + d.astype(F).astype(np.float64) * tt
                assert np.abs(pts[0] - q).max() < 1e-4
                agree += 1
        else:
            # a miss by more than rounding stays a miss
            pass
    assert agree > 300


def test_beer_lambert_through_mesh_chord():
    """Photoelectric only: a thin beam through the centre of a water sphere
    mesh is attenuated along the mesh's own chord (distance between its two
    surface hits), exp(-sigma_left * chord)."""
    R, e0, n = 20.0, 0.040, 100_000
    sc = Scene(processes=(1,))
    # turned so that the beam axis meets no vertex or edge (a ray through a
    # vertex is counted once per adjacent triangle, as in the reference)
    sc.add_mesh(authoring.icosphere(R, 4), "Water", rotation=(0.31, 0.17, 0.11), depth=4)
    sc.add_ct_flat("GOS", (1, 1), (8, 8, 1), (2.0, 2.0, 30.0), 600.0, 1000.0, 1e-3, store_scatter=False)
    sc.set_source((-600.0, 0.0, 0.0), 0.005, [e0], mono=True)
    sc.build()
    r = sc.run_oracle(n, 99, threads=4)
    counts = int(r["hist"][0].sum())
    m = sc.navigators[0][2]["mesh"]
    k, pts = O.mesh_hits(O.mesh_struct(m), np.array([-100.0, 0, 0], F), np.array([1.0, 0, 0], F))
    assert k == 2
    chord = float(np.linalg.norm(pts[1].astype(np.float64) - pts[0]))
    eid = int(np.searchsorted(sc.ebins, F(e0), side="right") - 1)
    p = math.exp(-float(sc.cs[1, 0, eid]) * chord)
    sd = math.sqrt(n * p * (1 - p))
    assert abs(counts - n * p) < 4 * sd, (counts, n * p, sd)
    assert r["stats"]["mesh_steps"] > 0


def test_mesh_dose_photoelectric_only():
    """Dose map on a meshed solid, photoelectric absorption only, pencil beam
    through the centre of a water sphere: every absorbed photon leaves its
    whole energy in one dosel on the beam axis and is recorded twice (the
    deposit, then the zero-energy record of the energy cut), so
    edep = E0 * absorbed and hits = 2 * absorbed exactly; absorbed photons
    follow Beer-Lambert along the chord, their depths the truncated
    exponential; the dose is edep over the one material's density."""
    R, e0, n = 20.0, 0.040, 100_000
    sc = Scene(processes=(1,))
    sc.add_mesh(authoring.icosphere(R, 4), "Water", rotation=(0.31, 0.17, 0.11), depth=4,
                dosimetry={"dosel_size": (1.0, 1.0, 1.0)})
    sc.add_ct_flat("GOS", (1, 1), (8, 8, 1), (2.0, 2.0, 30.0), 600.0, 1000.0, 1e-3, store_scatter=False)
    sc.set_source((-600.0, 0.0, 0.0), 0.005, [e0], mono=True)
    sc.build()
    t = {}
    ph, solids, src = sc._oracle_structs(t)
    st = O.Stats()
    O.lib().ggo_run(C.byref(ph), solids, len(solids), C.byref(src), n, 99, 4, C.byref(st))
    transmitted = int(t["hist"][0].sum())
    absorbed = n - transmitted
    edep, hit = t["edep"], t["hit"]
    assert edep.shape == (42, 42, 42)  # octree box +-21 mm
    assert edep.sum() == absorbed * float(F(e0))
    assert hit.sum() == 2 * absorbed
    # deposits on the beam axis only: dosels y, z = 21 (the axis is the face
    # between dosels 20 and 21 grazed by the 0.005 degree cone: 20 or 21)
    iz, iy, ix = np.nonzero(hit)
    assert set(iy) <= {20, 21} and set(iz) <= {20, 21}
    m = sc.navigators[0][2]["mesh"]
    k, pts = O.mesh_hits(O.mesh_struct(m), np.array([-100.0, 0, 0], F), np.array([1.0, 0, 0], F))
    chord = float(np.linalg.norm(pts[1].astype(np.float64) - pts[0]))
    eid = int(np.searchsorted(sc.ebins, F(e0), side="right") - 1)
    mu = float(sc.cs[1, 0, eid])
    p = 1.0 - math.exp(-mu * chord)
    sd = math.sqrt(n * p * (1 - p))
    assert abs(absorbed - n * p) < 4 * sd, (absorbed, n * p, sd)
    # mean depth of absorption past the entry surface, by dosel centre
    x_entry = float(pts[0][0])
    depth = (np.arange(42) + 0.5 - 21.0) - x_entry
    per_x = edep.sum(axis=(0, 1)) / float(F(e0))
    mean = float((per_x * depth).sum() / per_x.sum())
    L = chord
    expect = 1.0 / mu - L * math.exp(-mu * L) / (1.0 - math.exp(-mu * L))
    assert abs(mean - expect) < 0.5 + 4 * L / math.sqrt(12 * absorbed), (mean, expect)
    # dose: edep over the water density and the dosel volume
    dens = sc.arrays["density"].astype(F)
    dose = np.zeros(edep.size, F)
    unc = np.zeros(edep.size, F)
    O.lib().ggo_compute_dose(C.byref(solids[0]), O.ptr(dens), F(1.0), 0, F(0.0), O.ptr(dose), O.ptr(unc))
    mev_j, kg = 1.602176634e-13, 1e-6  # 1 mm^3 of water (1 g/cm3)
    e = edep.reshape(-1)
    nz = e > 0
    assert np.allclose(dose[nz] / (e[nz] * mev_j / kg), 1.0, rtol=1e-5)
    assert (dose[~nz] == 0).all()
