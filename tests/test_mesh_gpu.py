"""Meshed-solid navigation on the GPU against the oracle: the ray/mesh
intersection list, and transport through nested icosphere meshes with a
curved CT detector, through the engine and through a GGEMS script."""

import math
import os

import numpy as np
import pytest

import ggems_oracle as O
import parity
from ggems_amd import authoring, configs
from scene import Scene

pytestmark = pytest.mark.gpu
F = np.float32


def _rays(rng, n, r_out, r_in):
    target = rng.normal(size=(n, 3))
    target *= (rng.uniform(0, r_in, n) / np.linalg.norm(target, axis=1))[:, None]
    p = rng.normal(size=(n, 3))
    p *= (r_out / np.linalg.norm(p, axis=1))[:, None]
    inside = rng.random(n) < 0.3
    p[inside] = target[inside] * 0.5
    d = rng.normal(size=(n, 3))
    d[~inside] = target[~inside] - p[~inside]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return p.astype(F), d.astype(F)


def test_mesh_hits_match_oracle():
    sc = Scene()
    sc.add_mesh(authoring.icosphere(50.0, 4), "Water", rotation=(0.31, 0.17, 0.11), depth=4)
    sc.add_ct_flat("GOS", (1, 1), (16, 16, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 1.0, [0.08], mono=True)
    sc.build()
    eng = sc.make_engine()
    rng = np.random.default_rng(21)
    p, d = _rays(rng, 20000, 120.0, 45.0)
    cnt, near = eng.debug_mesh_hits(0, p, d)
    ms = O.mesh_struct(sc.navigators[0][2]["mesh"])
    bad = 0
    for i in range(len(p)):
        n, pts = O.mesh_hits(ms, p[i], d[i])
        ok = n == cnt[i]
        if ok and n >= 1:
            ok = np.array_equal(pts[0], near[i, 0])
        if ok and n >= 2:
            ok = np.array_equal(pts[1], near[i, 1])
        bad += not ok
    assert bad == 0
    assert (cnt == 2).sum() > 10000 and (cnt == 1).sum() > 3000
    eng.close()


def _nested_scene(curved=True):
    sc = Scene()
    sc.add_mesh(authoring.icosphere(50.0, 4), "SoftTissus_15", rotation=(0.31, 0.17, 0.11), depth=4)
    sc.add_mesh(authoring.icosphere(15.0, 3), "RibBone", rotation=(0.11, 0.23, 0.05), depth=3)
    if curved:
        sc.add_ct_curved("GOS", (1, 11), (32, 8, 1), (1.0, 1.0, 1.0), 600.0, 1040.0, 10e-3)
    else:
        sc.add_ct_flat("GOS", (1, 1), (128, 128, 1), (1.0, 1.0, 1.0), 600.0, 1000.0, 10e-3)
    sc.set_source((-600.0, 0.0, 0.0), 4.0, [0.08], mono=True)
    return sc.build()


@pytest.mark.parametrize("curved", [False, True])
def test_mesh_transport_matches_oracle(curved):
    sc = _nested_scene(curved)
    n = 200_000
    ref = sc.run_oracle(n, 31, threads=0)
    got = sc.run_engine(n, 31)
    rs, gs = ref["stats"], got["stats"]
    assert gs["primaries"] == rs["primaries"] == n
    tag = "nested meshes curved=%s" % curved
    parity.counters(tag, rs, gs, ("mesh_steps", "mesh_interactions", "hops"), 1e-3)
    h_ref = np.concatenate([h.ravel() for h in ref["hist"]])
    h_got = np.concatenate([h.ravel() for h in got["hist"]])
    parity.l1(tag + " histogram", h_ref, h_got, 1e-3, 10)


def _mesh_ct_oracle(workdir, water_box, dosimetry):
    """The oracle's copy of configs.mesh_ct: navigators in the same order."""
    outer, inner, box, _ = configs.mesh_ct_inputs(workdir)
    sc = Scene()
    sc.add_mesh(authoring.read_stl(outer), "SoftTissus_15", depth=4,
                dosimetry={"dosel_size": (1.0, 1.0, 1.0)} if dosimetry else None)
    sc.add_mesh(authoring.read_stl(inner), "RibBone", depth=4)
    if water_box:
        lab, spacing = authoring.read_mhd(box)
        sc.add_phantom(lab, ["Water"], tuple(spacing))
    sc.add_ct_curved("GOS", configs.MESH_MODULES, configs.MESH_ELEMENTS, (1.0, 1.0, 1.0), configs.MESH_SID,
                     configs.MESH_SDD, 10e-3)
    sc.set_source((-configs.MESH_SID, 0.0, 0.0), 5.0, [0.08], mono=True)
    return sc.build()


@pytest.mark.parametrize("water_box", [False, True])
def test_mesh_ct_script_matches_oracle(tmp_path, water_box):
    """configs.mesh_ct through the GGEMS API: the curved projection (888
    columns x 64 rows), the run statistics and, without the box, the dose
    map on the outer mesh.  With the enclosing water box no photon reaches a
    mesh, in both implementations."""
    n = 300_000
    wd = str(tmp_path)
    keep = configs.mesh_ct(wd, n_particles=n, water_box=water_box, dosimetry=not water_box)
    g = configs.run(seed=4242)
    stats = g.engine_stats()
    proj, spacing = authoring.read_mhd(os.path.join(wd, "projection.mhd"))
    assert proj.shape == (1, 888, 64)
    ref = _mesh_ct_oracle(wd, water_box, not water_box).run_oracle(n, 4242)
    rs = ref["stats"]
    assert stats["primaries"] == rs["primaries"] == n
    if water_box:
        assert stats["mesh_steps"] == rs["mesh_steps"] == 0
        assert abs(stats["voxel_steps"] - rs["voxel_steps"]) <= 0.01 * rs["voxel_steps"]
    else:
        parity.counters("mesh CT script", rs, stats, ("mesh_steps", "mesh_interactions", "hops"), 1e-3)
    h_ref = np.concatenate(ref["hist"], axis=0)
    assert h_ref.shape == (888, 64)
    parity.l1("mesh CT script water_box=%s projection" % water_box, h_ref, proj[0], 5e-3, 10)
    if not water_box:
        edep, _ = authoring.read_mhd(os.path.join(wd, "dosimetry_edep.mhd"))
        hit, _ = authoring.read_mhd(os.path.join(wd, "dosimetry_hit.mhd"))
        dose, _ = authoring.read_mhd(os.path.join(wd, "dosimetry_dose.mhd"))
        assert edep.shape == ref["edep"].shape == (106, 106, 106)  # octree box +-53 mm, 1 mm dosels
        parity.total("mesh CT script edep", ref["edep"], edep, 1e-3)
        parity.l1("mesh CT script hit", ref["hit"], hit, 1e-3, 10)
        # one material: dose is edep over the soft-tissue density everywhere
        nz = edep > 0
        ratio = dose[nz] / edep[nz]
        assert np.allclose(ratio, ratio[0], rtol=1e-5)
    g.delete()
    del keep
