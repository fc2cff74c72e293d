"""Writes oracle_vectors.npz: fixed input/output vectors of the CPU oracle
for the tables, random streams, source and voxel walks of the standard
scenes.  They anchor the oracle against regressions (tests/test_golden.py)
and give the GPU tests fixed expected outputs.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, REPO)

import ctypes as C  # noqa: E402

import ggems_oracle as O  # noqa: E402
from scene import cfg1_scene  # noqa: E402

F = np.float32
MATS = ["Water", "Air", "Lung", "Adipose", "SoftTissus_15", "RibBone", "GOS"]


def rays(n, seed):
    """Rays starting inside the cfg1 phantom (100 mm cube)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = rng.uniform(-49.0, 49.0, size=(n, 3))
    # a few axis-aligned and voxel-face rays
    d[:6] = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]]
    p[:6] = [[-49.9, 0.5, 0.5], [0.0, -49.9, 0.5], [0.5, 0.0, -49.9], [49.9, 3.0, 3.5], [2.5, 49.9, 0.0],
             [-7.5, 7.0, 49.9]]
    return p.astype(F), d.astype(F)


def main():
    out = {}
    ebins = O.energy_bins(220, F(1e-3), F(1.0))
    mt = O.material_tables(MATS, cut_distance=0.1, cut_emin=F(1e-3))
    cs, atom = O.cs_tables(mt, 220, ebins)
    a = O.mt_arrays(mt)
    out.update(ebins=ebins, cs=cs, cs_atom_rayleigh_z=atom[[1, 6, 8, 20, 64]], photon_cut=a["photon_cut"],
               density=a["density"], n_elements=a["n_elements"], elem_index=a["elem_index"], z=a["z"],
               atom_density=a["atom_density"])
    mtw = O.material_tables(["Water"], cut_distance=F(1e-3), cut_emin=F(1e-3))
    out["photon_cut_water_1um"] = np.array([mtw.photon_cut[0]], F)
    mt_out = np.zeros(64, np.uint32)
    O.lib().ggo_mt19937_stream(777, 64, O.ptr(mt_out))
    out["mt19937_777"] = mt_out
    st = O.seed_slots(777, 8)
    out["slots_777"] = st
    out["kiss_777"] = np.stack([O.kiss_stream(st[:, s], 64)[0] for s in range(4)])
    sc = cfg1_scene(n_side=100, detector=256, dosimetry=False)
    ph, solids, src = sc._oracle_structs({})
    st = O.seed_slots(777, 64)
    pos = np.zeros((64, 3), F)
    dirs = np.zeros((64, 3), F)
    en = np.zeros(64, F)
    for s in range(64):
        r = st[:, s].copy()
        p = np.zeros(3, F)
        d = np.zeros(3, F)
        e = np.zeros(1, F)
        O.lib().ggo_sample_primary(C.byref(src), O.ptr(r), O.ptr(p), O.ptr(d), O.ptr(e))
        pos[s], dirs[s], en[s] = p, d, e[0]
    out.update(primary_pos=pos, primary_dir=dirs, primary_e=en)
    p, d = rays(24, 5)
    walks = np.full((24, 400), -1, np.int32)
    for k in range(24):
        buf = np.zeros(400, np.int32)
        n = O.lib().ggo_voxel_walk(C.byref(solids[0]), O.ptr(p[k]), O.ptr(d[k]), O.ptr(buf), 400)
        walks[k, :n] = buf[:n]
    out.update(walk_p=p, walk_d=d, walks=walks)
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
