"""Oracle restatement of the ray-traced primary transmission (TEST
INFRASTRUCTURE ONLY): per pixel of a CT module, the probability that a
source photon aimed at the pixel centre crosses the voxel phantoms without
interacting, sum_b w_b exp(-sum_m sigma_tot(m, b) L_m).  Path lengths come
from Siddon's parametric plane crossings in FP64; sigma_tot is the left-bin
total of the activated processes
(include/GGEMS/navigators/GGEMSPhotonNavigator.hh:57-93); w_b is the
probability of energy bin b under the source's sampling
(src/kernels/GetPrimariesGGEMSXRaySource.cl:103-125)."""

import math

import numpy as np

F = np.float32


def left_bin(ebins, e):
    """BinarySearch (include/GGEMS/maths/GGEMSMathAlgorithms.hh:46-60)"""
    low, high = 0, len(ebins) - 1
    while high - low > 1:
        mid = (high + low) // 2
        if ebins[mid] == e:
            return mid
        if ebins[mid] < e:
            low = mid
        else:
            high = mid
    return low


def energy_weights(ebins, energy, cdf, is_interp):
    """w[b] over the cross-section bins: line i has probability
    cdf[i+1] - cdf[i]; interpolated, it is spread uniformly over
    [E_i, E_i+1)."""
    nb = len(ebins)
    w = np.zeros(nb, np.float64)
    for i in range(len(cdf) - 1):
        p = float(cdf[i + 1]) - float(cdf[i])
        if p <= 0:
            continue
        a, b = float(energy[i]), float(energy[i + 1])
        if not is_interp or not b > a:
            w[left_bin(ebins, F(energy[i]))] += p
            continue
        for k in range(nb - 1):
            lo = -math.inf if k == 0 else float(ebins[k])
            hi = math.inf if k + 2 == nb else float(ebins[k + 1])
            ov = min(b, hi) - max(a, lo)
            if ov > 0:
                w[k] += p * ov / (b - a)
    return w


def siddon_lengths(p0, p1, bmin, bmax, vs, dims):
    """[(flat voxel index, length mm)] of the segment p0->p1 (local frame)
    through the grid bmin + i * vs, dims = (nx, ny, nz)."""
    p0, p1 = np.asarray(p0, np.float64), np.asarray(p1, np.float64)
    d = p1 - p0
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if abs(d[a]) < 1e-300:
            if p0[a] < bmin[a] or p0[a] > bmax[a]:
                return []
        else:
            ta, tb = (bmin[a] - p0[a]) / d[a], (bmax[a] - p0[a]) / d[a]
            t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    if not t0 < t1:
        return []
    ts = [t0, t1]
    for a in range(3):
        if abs(d[a]) < 1e-300:
            continue
        planes = bmin[a] + np.arange(dims[a] + 1) * vs[a]
        t = (planes - p0[a]) / d[a]
        ts.extend(t[(t > t0) & (t < t1)].tolist())
    ts = np.unique(np.array(ts))
    length = float(np.linalg.norm(d))
    out = []
    for ta, tb in zip(ts[:-1], ts[1:]):
        mid = p0 + d * (0.5 * (ta + tb))
        idx = [min(max(int(math.floor((mid[a] - bmin[a]) / vs[a])), 0), dims[a] - 1) for a in range(3)]
        out.append((idx[0] + idx[1] * dims[0] + idx[2] * dims[0] * dims[1], (tb - ta) * length))
    return out


def _apply(m, p):
    m = np.asarray(m, np.float64)
    return m[:3, :3] @ np.asarray(p, np.float64) + m[:3, 3]


def _inverse_apply(m, p):
    m = np.asarray(m, np.float64)
    return np.linalg.solve(m[:3, :3], np.asarray(p, np.float64) - m[:3, 3])


def drr(scene, module=0, src_m=None, box_m=None):
    """Transmission image [rows][cols] of CT module `module` of the scene."""
    mods = [mod for kind, _, pay in scene.navigators if kind == "ct" for mod in pay["modules"]]
    mod = mods[module]
    src_m = scene.source["m"] if src_m is None else src_m
    box_m = mod["m"] if box_m is None else box_m
    S = scene.source
    w = energy_weights(scene.ebins, S["energy"], S["cdf"], S["is_interp"])
    bins = np.nonzero(w)[0]
    sig = np.zeros((scene.cs.shape[1], scene.cs.shape[2]), np.float64)
    for p in scene.processes:
        sig += scene.cs[p].astype(np.float64)
    nx, ny = int(mod["ne"][0]), int(mod["ne"][1])
    esx, esy = float(F(mod["size"][0]) / F(nx)), float(F(mod["size"][1]) / F(ny))
    lo, hi = np.asarray(mod["bmin"], np.float64), np.asarray(mod["bmax"], np.float64)
    src = _apply(src_m, (0.0, 0.0, 0.0))
    phantoms = [(pay, off) for (kind, _, pay), off in zip(scene.navigators, scene.offsets) if kind == "voxel"]
    img = np.zeros((ny, nx), np.float64)
    for iy in range(ny):
        for ix in range(nx):
            pc = (lo[0] + (ix + 0.5) * esx, lo[1] + (iy + 0.5) * esy, 0.5 * (lo[2] + hi[2]))
            pe = _apply(box_m, pc)
            L = np.zeros(sig.shape[0], np.float64)
            for pay, off in phantoms:
                a, b = _inverse_apply(pay["m"], src), _inverse_apply(pay["m"], pe)
                nz_, ny_, nx_ = pay["labels"].shape
                lab = pay["labels"].reshape(-1)
                for v, ln in siddon_lengths(a, b, pay["bmin"], pay["bmax"], pay["vs"], (nx_, ny_, nz_)):
                    L[off + int(lab[v])] += ln
            mu = L @ sig[:, bins]
            img[iy, ix] = float(np.sum(w[bins] * np.exp(-mu)))
    return img
