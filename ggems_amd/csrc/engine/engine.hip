// engine.hip -- MI355X photon-transport engine behind the thin C ABI of
// include/ggems_engine.h.
//
// One launch transports a whole run.  Thread s owns particle slot s: its
// JKISS32 state stays in registers and it transports, in order, photon s of
// every batch (src/global/GGEMS.cc:303-363 runs batches of
// MAXIMUM_PARTICLES slots; slot s of every batch continues stream s).  Each
// photon walks the reference's hop loop (find solid -> project -> track)
// inside the thread, so the host never synchronises per hop.
//
// Divergence: the voxel step is the hot path.  A wave alternates a cold
// phase (every lane not inside a voxel solid advances: new primary, hop
// bookkeeping, detector transport) with a hot phase that steps voxels while
// enough lanes are inside a phantom; lanes that leave the phantom wait for
// the next cold phase instead of forcing the whole wave through the cold
// code every iteration.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../../include/ggems_engine.h"
#include "transport_device.hpp"

using namespace ggh;

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// device-side scene description
// ---------------------------------------------------------------------------
enum SolidType : int32_t { kVoxel = 0, kBoxSolid = 1 };

struct DevSolid {
  Mat34 m;           // local->global rows 0..2
  Box border;        // local borders
  int32_t type;
  int32_t mat_offset;
  // voxel solid
  f3 voxel_size;
  int32_t nv[3];
  const uint8_t* labels;
  // box solid (CT module)
  int32_t ne[3];
  f3 box_size;
  float threshold;
  int32_t* histogram;
  int32_t* scatter;
};

struct RunParams {
  const DevSolid* solids;
  int32_t n_solids;
  // physics
  int32_t nbins;
  int32_t n_mat;
  int32_t n_proc;
  int32_t proc[3];
  const float* ebins;
  const float4* cs4;          // [n_mat*nbins] {sigma_C, sigma_PE, sigma_R, photon_cut}
  const float* cs_atom_r;     // [101*nbins]
  const int32_t* n_elem;
  const int32_t* elem_idx;
  const uint8_t* z;
  const float* atom_dens;
  const float* pp;            // [9*101]
  float factor;
  // source
  Mat34 src_m;
  float aperture;
  f3 focal;
  int32_t n_energy_arg;       // number_of_energy_bins_ + 1, as passed to the reference kernel
  const float* spectrum;
  const float* cdf;
  int32_t is_interp;
  // dosimetry
  int32_t dose_solid;
  f3 dose_lo, dose_hi, inv_dosel;
  int32_t nd[3];
  double* edep;
  double* edep2;
  int32_t* hit;
  int32_t* ptrack;
  // particle slots
  uint32_t* rng;              // SoA [5][rng_stride]
  uint32_t rng_stride;
  uint32_t n_slots;
  uint64_t n_batches, base, rem;
  unsigned long long* stats;  // ggh_stats fields
  int32_t hot_threshold;
};

struct Photon {
  f3 p, d;
  float e;
  int32_t eid;
  int32_t status;
  int32_t scatter;
  int32_t solid;
  float dist;
};

struct Counters {
  uint32_t primaries, voxel_steps, interactions, detector_steps, detector_interactions, hops;
};

// ---------------------------------------------------------------------------
// physics (include/GGEMS/navigators/GGEMSPhotonNavigator.hh:57-131)
// ---------------------------------------------------------------------------
__device__ inline float sigma_of(const float4& s, int pid) { return pid == 0 ? s.x : (pid == 1 ? s.y : s.z); }

// GetPhotonNextInteraction: competing exponentials over the activated
// processes with the left-bin cross section (GGEMSPhotonNavigator.hh:57-93)
__device__ inline float next_interaction(const RunParams& P, const float4& sig, Rng& r, int& proc) {
  float next = kOutOfWorld;
  proc = kTransportation;  // NO_PROCESS == TRANSPORTATION == 99
  for (int i = 0; i < P.n_proc; ++i) {
    int pid = P.proc[i];
    float d = -logf(r.uniform()) / sigma_of(sig, pid);
    if (d < next) {
      next = d;
      proc = pid;
    }
  }
  return next;
}

// KleinNishinaComptonSampleSecondaries (GGEMSComptonScatteringModels.hh:47-132)
__device__ void compton(Photon& g, Rng& r) {
  float kE0 = g.e;
  float kE0_MeC2 = kE0 / kElectronMassC2;
  f3 dir0 = g.d;
  float kEps0 = 1.0f / (1.0f + 2.0f * kE0_MeC2);
  float kEps0Eps0 = kEps0 * kEps0;
  float kAlpha1 = -logf(kEps0);
  float kAlpha2 = kAlpha1 + 0.5f * (1.0f - kEps0Eps0);
  float epsilon, epsilonsq, onecost, sint2, greject;
  int nloop = 0;
  do {
    ++nloop;
    if (nloop > 1000) return;  // false interaction
    float r1 = r.uniform();
    float r2 = r.uniform();
    float r3 = r.uniform();
    if (kAlpha1 > kAlpha2 * r1) {
      epsilon = expf(-kAlpha1 * r2);
      epsilonsq = epsilon * epsilon;
    } else {
      epsilonsq = kEps0Eps0 + (1.0f - kEps0Eps0) * r2;
      epsilon = sqrtf(epsilonsq);
    }
    onecost = (1.0f - epsilon) / (epsilon * kE0_MeC2);
    sint2 = onecost * (2.0f - onecost);
    greject = 1.0f - epsilon * sint2 / (1.0f + epsilonsq);
    if (!(greject < r3)) break;
  } while (true);
  if (sint2 < 0.0f) sint2 = 0.0f;
  float costheta = 1.0f - onecost;
  float sintheta = sqrtf(sint2);
  float phi = r.uniform() * kTwoPi;
  f3 gd = mk(sintheta * cosf(phi), sintheta * sinf(phi), costheta);
  gd = normalize(rotate_unit_z(gd, dir0));
  g.e = kE0 * epsilon;
  g.d = gd;
}

// LivermoreRayleighSampleSecondaries (GGEMSRayleighScatteringModels.hh:169-316)
__device__ void rayleigh(const RunParams& P, Photon& g, int gmat, Rng& r) {
  float kE0 = g.e;
  if (kE0 <= 250.0e-6f) {
    g.status = kDead;
    return;
  }
  f3 dir0 = g.d;
  const int nb = P.nbins;
  const int nm1 = P.n_elem[gmat] - 1;
  const int mix = P.elem_idx[gmat];
  const int eid = g.eid;
  int selz = P.z[mix + nm1];
  if (nm1 > 0) {
    float e_lo = P.ebins[eid], e_hi = P.ebins[eid + 1];
    float kCS = lin_interp(e_lo, P.cs4[gmat * nb + eid].z, e_hi, P.cs4[gmat * nb + eid + 1].z, kE0);
    float x = r.uniform() * kCS;
    float cross = 0.0f;
    for (int i = 0; i < nm1; ++i) {
      int zz = P.z[mix + i];
      cross += P.atom_dens[mix + i] *
               lin_interp(e_lo, P.cs_atom_r[eid + nb * zz], e_hi, P.cs_atom_r[eid + 1 + nb * zz], kE0);
      if (x < cross) {
        selz = zz;
        break;
      }
    }
  }
  const float* pp = P.pp;
  float kXX = P.factor * kE0 * kE0;
  float kN0 = pp[6 * 101 + selz] - 1.0f, kN1 = pp[7 * 101 + selz] - 1.0f, kN2 = pp[8 * 101 + selz] - 1.0f;
  float kB0 = pp[3 * 101 + selz], kB1 = pp[4 * 101 + selz], kB2 = pp[5 * 101 + selz];
  float x = 2.0f * kXX * kB0;
  float kW0 = (x < 0.02f) ? kN0 * x * (1.0f - 0.5f * (kN0 - 1.0f) * x * (1.0f - (kN0 - 2.0f) * x / 3.0f))
                          : 1.0f - expf(-kN0 * logf(1.0f + x));
  x = 2.0f * kXX * kB1;
  float kW1 = (x < 0.02f) ? kN1 * x * (1.0f - 0.5f * (kN1 - 1.0f) * x * (1.0f - (kN1 - 2.0f) * x / 3.0f))
                          : 1.0f - expf(-kN1 * logf(1.0f + x));
  x = 2.0f * kXX * kB2;
  float kW2 = (x < 0.02f) ? kN2 * x * (1.0f - 0.5f * (kN2 - 1.0f) * x * (1.0f - (kN2 - 2.0f) * x / 3.0f))
                          : 1.0f - expf(-kN2 * logf(1.0f + x));
  float kX0 = kW0 * pp[0 * 101 + selz] / (kB0 * kN0);
  float kX1 = kW1 * pp[1 * 101 + selz] / (kB1 * kN1);
  float kX2 = kW2 * pp[2 * 101 + selz] / (kB2 * kN2);
  float costheta;
  for (;;) {
    float w = kW0, n = kN0, b = kB0;
    x = r.uniform() * (kX0 + kX1 + kX2);
    if (x > kX0) {
      x -= kX0;
      if (x <= kX1) {
        w = kW1; n = kN1; b = kB1;
      } else {
        w = kW2; n = kN2; b = kB2;
      }
    }
    n = 1.0f / n;
    float y = r.uniform() * w;
    if (y < 0.02f) x = y * n * (1.0f + 0.5f * (n + 1.0f) * y * (1.0f - (n + 2.0f) * y / 3.0f));
    else x = expf(-n * logf(1.0f - y)) - 1.0f;
    costheta = 1.0f - x / (b * kXX);
    float u = r.uniform();
    if (!(2.0f * u > 1.0f + costheta * costheta || costheta < -1.0f)) break;
  }
  float phi = kTwoPi * r.uniform();
  float sintheta = sqrtf((1.0f - costheta) * (1.0f + costheta));
  f3 gd = mk(sintheta * cosf(phi), sintheta * sinf(phi), costheta);
  g.d = normalize(rotate_unit_z(gd, dir0));
}

// PhotonDiscreteProcess (GGEMSPhotonNavigator.hh:109-131)
__device__ inline void discrete(const RunParams& P, Photon& g, int proc, int gmat, Rng& r) {
  if (proc == kCompton) {
    compton(g, r);
  } else if (proc == kPhotoelectric) {
    g.status = kDead;
    g.e = 0.0f;
  } else if (proc == kRayleigh) {
    rayleigh(P, g, gmat, r);
  }
}

// ---------------------------------------------------------------------------
// tallies (include/GGEMS/navigators/GGEMSDoseRecording.hh:65-118)
// ---------------------------------------------------------------------------
__device__ inline int dosel_index(const RunParams& P, f3 p) {
  if (p.x < P.dose_lo.x + kEps6 || p.x > P.dose_hi.x - kEps6) return -1;
  if (p.y < P.dose_lo.y + kEps6 || p.y > P.dose_hi.y - kEps6) return -1;
  if (p.z < P.dose_lo.z + kEps6 || p.z > P.dose_hi.z - kEps6) return -1;
  int ix = (int)((p.x - P.dose_lo.x) * P.inv_dosel.x);
  int iy = (int)((p.y - P.dose_lo.y) * P.inv_dosel.y);
  int iz = (int)((p.z - P.dose_lo.z) * P.inv_dosel.z);
  if (ix < 0 || ix >= P.nd[0]) return -1;
  if (iy < 0 || iy >= P.nd[1]) return -1;
  if (iz < 0 || iz >= P.nd[2]) return -1;
  return ix + iy * P.nd[0] + iz * P.nd[0] * P.nd[1];
}

__device__ inline void dose_record(const RunParams& P, float edep, f3 p) {
  int id = dosel_index(P, p);
  if (id < 0) return;
  if (P.hit) atomicAdd(&P.hit[id], 1);
  unsafeAtomicAdd(&P.edep[id], (double)edep);
  if (P.edep2) unsafeAtomicAdd(&P.edep2[id], (double)edep * (double)edep);
}

// ---------------------------------------------------------------------------
// voxel solid transport (src/kernels/TrackThroughGGEMSVoxelizedSolid.cl:76-323)
// ---------------------------------------------------------------------------
struct VoxelWalk {
  f3 lp;          // local position
  Ray ld;         // local direction + reciprocals
  Box border;
  f3 vs;
  int32_t nx, ny, nz;
  const uint8_t* labels;
  int32_t mat_offset;
  int32_t dose;   // this solid carries the dose grid
  int32_t solid;  // index of the solid being walked
};

__device__ inline void voxel_enter(VoxelWalk& v, const Photon& g, const DevSolid& s, int solid, int dose) {
  v.solid = solid;
  v.lp = g2l_pos(s.m, g.p);
  v.ld.set(g2l_dir(s.m, g.d));
  v.border = s.border;
  v.vs = s.voxel_size;
  v.nx = s.nv[0];
  v.ny = s.nv[1];
  v.nz = s.nv[2];
  v.labels = s.labels;
  v.mat_offset = s.mat_offset;
  v.dose = dose;
}

// One iteration of the reference do-while.  Returns true while the photon
// stays inside the solid and alive.  kWalk: no interaction (sigma == 0),
// records the visited voxel index instead.
template <bool kWalk>
__device__ inline bool voxel_step(const RunParams& P, const float4* __restrict__ cs, VoxelWalk& v, Photon& g, Rng& r,
                                  Counters& c, int32_t* walk_out, int& walk_n, int walk_max) {
  int vx = (int)((v.lp.x - v.border.lo.x) / v.vs.x);
  int vy = (int)((v.lp.y - v.border.lo.y) / v.vs.y);
  int vz = (int)((v.lp.z - v.border.lo.z) / v.vs.z);
  // The reference checks the upper bound only; a negative index would read
  // outside the label volume, so it ends the walk like an exit.
  if (vx >= v.nx || vy >= v.ny || vz >= v.nz || vx < 0 || vy < 0 || vz < 0) {
    g.dist = kOutOfWorld;
    g.solid = -1;
    return false;
  }
  int vid = vx + vy * v.nx + vz * v.nx * v.ny;
  int proc = kTransportation;
  float next = kOutOfWorld;
  int gmat = 0;
  float4 sig = make_float4(0.f, 0.f, 0.f, 0.f);
  if (kWalk) {
    if (walk_n < walk_max) walk_out[walk_n] = vid;
    ++walk_n;
  } else {
    gmat = v.mat_offset + v.labels[vid];
    sig = cs[gmat * P.nbins + g.eid];
    next = next_interaction(P, sig, r, proc);
    ++c.voxel_steps;
  }
  Box vox;
  vox.lo = mk(v.border.lo.x + (float)vx * v.vs.x, v.border.lo.y + (float)vy * v.vs.y,
              v.border.lo.z + (float)vz * v.vs.z);
  vox.hi = add(vox.lo, v.vs);
  safety_inside(v.lp, vox);
  float db = distance_to_aabb(v.lp, v.ld, vox);
  if (db <= next) {
    next = db + kTol;
    proc = kTransportation;
    if (!kWalk && v.dose && P.ptrack) {
      int id = dosel_index(P, v.lp);
      if (id >= 0) atomicAdd(&P.ptrack[id], 1);
    }
  }
  v.lp = add(v.lp, scl(v.ld.d, next));
  safety_outside(v.lp, vox);
  if (!in_aabb(v.lp, v.border)) {
    g.dist = kOutOfWorld;
    g.solid = -1;
    return false;
  }
  if (kWalk) return true;
  float e0 = g.e;
  if (proc != kTransportation) {
    ++c.interactions;
    g.d = v.ld.d;
    discrete(P, g, proc, gmat, r);
    if (proc == kCompton || proc == kRayleigh) g.scatter = 1;
    if (v.dose) dose_record(P, e0 - g.e, v.lp);
    if (proc != kPhotoelectric) {
      v.ld.set(g.d);
      if (g.e != e0) g.eid = binary_search(g.e, P.ebins, P.nbins);
    }
  }
  if (g.e <= sig.w) {  // photon_energy_cut_[material]
    if (v.dose) dose_record(P, g.e, v.lp);
    g.status = kDead;
  }
  return g.status == kAlive;
}

__device__ inline void voxel_leave(const VoxelWalk& v, Photon& g, const DevSolid& s) {
  g.p = l2g_pos(s.m, v.lp);
  g.d = l2g_dir(s.m, v.ld.d);
}

// ---------------------------------------------------------------------------
// CT module transport (src/kernels/TrackThroughGGEMSSolidBox.cl:56-261)
// ---------------------------------------------------------------------------
__device__ void track_box(const RunParams& P, const float4* __restrict__ cs, const DevSolid& s, Photon& g, Rng& r,
                          Counters& c) {
  f3 lp = g2l_pos(s.m, g.p);
  Ray ld;
  ld.set(g2l_dir(s.m, g.d));
  const Box b = s.border;
  const int gmat = s.mat_offset;
  do {
    int proc;
    float4 sig = cs[gmat * P.nbins + g.eid];
    float next = next_interaction(P, sig, r, proc);
    ++c.detector_steps;
    safety_inside(lp, b);
    float db = distance_to_aabb(lp, ld, b);
    if (db <= next) {
      next = db + kTol;
      proc = kTransportation;
    }
    lp = add(lp, scl(ld.d, next));
    safety_outside(lp, b);
    if (!in_aabb(lp, b)) {
      g.dist = kOutOfWorld;
      g.solid = -1;
      break;
    }
    if (g.e < s.threshold) g.status = kDead;
    if (proc != kTransportation) {
      ++c.detector_interactions;
      float e0 = g.e;
      g.d = ld.d;
      discrete(P, g, proc, gmat, r);
      if (proc != kPhotoelectric) {
        ld.set(g.d);
        if (g.e != e0) g.eid = binary_search(g.e, P.ebins, P.nbins);
      }
      if (proc == kPhotoelectric || proc == kCompton) {
        float esx = s.box_size.x / (float)s.ne[0];
        float esy = s.box_size.y / (float)s.ne[1];
        int ix = (int)((lp.x - b.lo.x) / esx);
        int iy = (int)((lp.y - b.lo.y) / esy);
        if (ix >= 0 && ix < s.ne[0] && iy >= 0 && iy < s.ne[1]) {
          int id = ix + iy * s.ne[0];
          atomicAdd(&s.histogram[id], 1);
          if (s.scatter && g.scatter) atomicAdd(&s.scatter[id], 1);
        }
      }
    }
  } while (g.status == kAlive);
  g.p = l2g_pos(s.m, lp);
  g.d = l2g_dir(s.m, ld.d);
}

// ---------------------------------------------------------------------------
// source (src/kernels/GetPrimariesGGEMSXRaySource.cl:52-143)
// ---------------------------------------------------------------------------
__device__ void sample_primary(const RunParams& P, Rng& r, Photon& g) {
  double phi = r.uniform();
  double theta = r.uniform();
  phi *= (double)kTwoPi;
  double new_aperture = 1.0 - cos((double)P.aperture);
  theta = acos(1.0 - new_aperture * theta);
  double st = sin(theta);
  f3 rotation = mk((float)(cos(phi) * st), (float)(sin(phi) * st), (float)cos(theta));
  f3 gp = l2g_pos(P.src_m, mk(0.0f, 0.0f, 0.0f));
  f3 d = normalize(mk(0.0f - gp.x, 0.0f - gp.y, 0.0f - gp.z));
  d = normalize(rotate_unit_z(rotation, d));
  gp.x = P.focal.x * (r.uniform() - 0.5f);
  gp.y = P.focal.y * (r.uniform() - 0.5f);
  gp.z = P.focal.z * (r.uniform() - 0.5f);
  gp = l2g_pos(P.src_m, gp);
  int idx = 0;
  float rnd = 0.0f;
  if (P.n_energy_arg != 1) {
    rnd = r.uniform();
    idx = binary_search(rnd, P.cdf, P.n_energy_arg + 1);
  }
  g.e = (!P.is_interp) ? P.spectrum[idx]
                       : lin_interp(P.cdf[idx], P.spectrum[idx], P.cdf[idx + 1], P.spectrum[idx + 1], rnd);
  g.p = gp;
  g.d = d;
  g.status = kAlive;
  g.scatter = 0;
  g.solid = -1;
  g.dist = kOutOfWorld;
  g.eid = binary_search(g.e, P.ebins, P.nbins);
}

// ---------------------------------------------------------------------------
// navigation between solids
// ---------------------------------------------------------------------------
// particle_solid_distance_* over every solid in registration order
// (src/kernels/ParticleSolidDistanceGGEMSVoxelizedSolid.cl:43-108) followed by
// project_to_* (src/kernels/ProjectToGGEMSVoxelizedSolid.cl:44-129).
// Returns false when the photon leaves the world.
__device__ inline bool find_and_project(const RunParams& P, Photon& g) {
  for (int k = 0; k < P.n_solids; ++k) {
    if (g.dist == 0.0f) break;
    const DevSolid& s = P.solids[k];
    f3 lp = g2l_pos(s.m, g.p);
    if (in_aabb(lp, s.border)) {
      g.dist = 0.0f;
      g.solid = k;
      break;
    }
    Ray lr;
    lr.set(g2l_dir(s.m, g.d));
    float d = distance_to_aabb(lp, lr, s.border);
    if (d < g.dist) {
      g.dist = d;
      g.solid = k;
    }
  }
  if (g.solid == -1 || g.dist == kOutOfWorld) {
    g.solid = -1;
    g.status = kDead;
    return false;
  }
  const DevSolid& s = P.solids[g.solid];
  g.p = add(g.p, scl(g.d, g.dist + kTol));
  f3 lp = g2l_pos(s.m, g.p);
  safety_inside(lp, s.border);
  g.p = l2g_pos(s.m, lp);
  g.dist = 0.0f;
  return true;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

enum Mode : int32_t { kNeedPrimary = 0, kHop = 1, kInVoxel = 2, kDone = 3 };

template <bool kLDS>
__global__ __launch_bounds__(kBlock) void transport_kernel(RunParams P) {
  extern __shared__ float4 s_cs[];
  const float4* cs = P.cs4;
  if (kLDS) {
    const int n = P.n_mat * P.nbins;
    for (int i = threadIdx.x; i < n; i += kBlock) s_cs[i] = P.cs4[i];
    __syncthreads();
    cs = s_cs;
  }
  const uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = slot < P.n_slots;
  Rng r = {0, 0, 0, 0, 0};
  uint64_t todo = 0;
  if (valid) {
    r.x = P.rng[0 * (size_t)P.rng_stride + slot];
    r.y = P.rng[1 * (size_t)P.rng_stride + slot];
    r.z = P.rng[2 * (size_t)P.rng_stride + slot];
    r.w = P.rng[3 * (size_t)P.rng_stride + slot];
    r.c = P.rng[4 * (size_t)P.rng_stride + slot];
    // batch b holds base + (b < rem) photons; slot s takes photon s of each
    todo = slot < P.base ? P.n_batches : (slot == P.base ? P.rem : 0);
  }
  Counters c = {0, 0, 0, 0, 0, 0};
  Photon g;
  g.status = kDead;
  VoxelWalk v;
  int hop = 0;
  int mode = todo ? kNeedPrimary : kDone;
  int32_t dummy_n = 0;

  for (;;) {
    // ---- cold phase: advance every lane to a voxel walk or to the end
    while (mode == kNeedPrimary || mode == kHop) {
      if (mode == kNeedPrimary) {
        if (todo == 0) {
          mode = kDone;
          break;
        }
        --todo;
        sample_primary(P, r, g);
        ++c.primaries;
        hop = 0;
        mode = kHop;
        continue;
      }
      if (hop >= GGH_MAX_HOPS) {
        mode = kNeedPrimary;
        continue;
      }
      ++hop;
      ++c.hops;
      if (!find_and_project(P, g)) {
        mode = kNeedPrimary;
        continue;
      }
      const DevSolid& s = P.solids[g.solid];
      if (s.type == kVoxel) {
        voxel_enter(v, g, s, g.solid, g.solid == P.dose_solid);
        mode = kInVoxel;
      } else {
        track_box(P, cs, s, g, r, c);
        mode = (g.status == kDead) ? kNeedPrimary : kHop;
      }
    }
    // ---- hot phase: voxel steps while enough lanes are inside a phantom
    bool all_done = false;
    for (;;) {
      const uint64_t act = __ballot(mode == kInVoxel);
      const uint64_t live = __ballot(mode != kDone);
      if (live == 0) {
        all_done = true;
        break;
      }
      const int na = __popcll(act), nl = __popcll(live);
      if (na == 0 || na < min(P.hot_threshold, nl)) break;
      if (mode == kInVoxel) {
        if (!voxel_step<false>(P, cs, v, g, r, c, nullptr, dummy_n, 0)) {
          voxel_leave(v, g, P.solids[v.solid]);
          mode = (g.status == kDead) ? kNeedPrimary : kHop;
        }
      }
    }
    if (all_done) break;
  }

  if (valid) {
    P.rng[0 * (size_t)P.rng_stride + slot] = r.x;
    P.rng[1 * (size_t)P.rng_stride + slot] = r.y;
    P.rng[2 * (size_t)P.rng_stride + slot] = r.z;
    P.rng[3 * (size_t)P.rng_stride + slot] = r.w;
    P.rng[4 * (size_t)P.rng_stride + slot] = r.c;
  }
  uint32_t sums[6] = {c.primaries, c.voxel_steps, c.interactions, c.detector_steps, c.detector_interactions, c.hops};
  for (int k = 0; k < 6; ++k) {
    uint32_t t = wave_sum(sums[k]);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(&P.stats[k], (unsigned long long)t);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// dose finalisation (src/kernels/ComputeDoseGGEMSVoxelizedSolid.cl:53-121)
// ---------------------------------------------------------------------------
namespace {

struct DoseParams {
  const double* edep;
  const double* edep2;
  const int32_t* hit;
  const uint8_t* labels;
  const float* density;     // by global material
  int32_t mat_offset;
  int32_t nd[3];
  f3 dosel_size;
  f3 vs;
  f3 vlo;
  int32_t nv[3];
  float scale;
  int32_t water_reference;
  float min_density;
  float* dose;
  float* unc;
};

__global__ __launch_bounds__(kBlock) void dose_kernel(DoseParams D) {
  const int64_t total = (int64_t)D.nd[0] * D.nd[1] * D.nd[2];
  const int slice = D.nd[0] * D.nd[1];
  const float kGy = 1.0e-12f;
  const float kWater = 1.0f * (6.241509704e+21f / 1.e3f);
  const float vol = D.dosel_size.x * D.dosel_size.y * D.dosel_size.z;
  for (int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x; gid < total; gid += (int64_t)gridDim.x * kBlock) {
    int z = (int)(gid / slice);
    int x = (int)((gid - (int64_t)z * slice) % D.nd[0]);
    int y = (int)((gid - (int64_t)z * slice) / D.nd[0]);
    float px = (float)x * D.dosel_size.x + (float)(D.nd[0] - 1) * (-0.5f * D.dosel_size.x);
    float py = (float)y * D.dosel_size.y + (float)(D.nd[1] - 1) * (-0.5f * D.dosel_size.y);
    float pz = (float)z * D.dosel_size.z + (float)(D.nd[2] - 1) * (-0.5f * D.dosel_size.z);
    int vx = min(max((int)((px - D.vlo.x) / D.vs.x), 0), D.nv[0] - 1);
    int vy = min(max((int)((py - D.vlo.y) / D.vs.y), 0), D.nv[1] - 1);
    int vz = min(max((int)((pz - D.vlo.z) / D.vs.z), 0), D.nv[2] - 1);
    int mat = D.labels[vx + vy * D.nv[0] + vz * D.nv[0] * D.nv[1]];
    float density = D.water_reference ? kWater : D.density[D.mat_offset + mat];
    double e = D.edep[gid];
    D.dose[gid] = density < D.min_density ? 0.0f : (float)((double)D.scale * e / (double)density / (double)vol / (double)kGy);
    if (D.unc) {
      int h = D.hit[gid];
      if (h > 1 && e != 0.0) {
        double s2 = e * e;
        D.unc[gid] = (float)sqrt(((double)h * D.edep2[gid] - s2) / ((double)(h - 1) * s2));
      } else {
        D.unc[gid] = 1.0f;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// test kernels
// ---------------------------------------------------------------------------
__global__ void walk_kernel(RunParams P, int solid, int n_rays, const float* pos, const float* dir, int max_steps,
                            int32_t* out, int32_t* count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rays) return;
  Photon g;
  g.p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  g.d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
  g.e = 1.0f;
  g.eid = 0;
  g.status = kAlive;
  g.scatter = 0;
  g.solid = solid;
  g.dist = 0.0f;
  VoxelWalk v;
  voxel_enter(v, g, P.solids[solid], solid, 0);
  Rng r = {0, 0, 0, 0, 0};
  Counters c = {0, 0, 0, 0, 0, 0};
  int n = 0;
  while (voxel_step<true>(P, P.cs4, v, g, r, c, out + (size_t)i * max_steps, n, max_steps)) {
    if (n > 16 * max_steps) break;
  }
  count[i] = n;
}

__global__ void primaries_kernel(RunParams P, int n, float* pos, float* dir, float* e) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  Rng r;
  r.x = P.rng[0 * (size_t)P.rng_stride + s];
  r.y = P.rng[1 * (size_t)P.rng_stride + s];
  r.z = P.rng[2 * (size_t)P.rng_stride + s];
  r.w = P.rng[3 * (size_t)P.rng_stride + s];
  r.c = P.rng[4 * (size_t)P.rng_stride + s];
  Photon g;
  sample_primary(P, r, g);
  pos[3 * s] = g.p.x; pos[3 * s + 1] = g.p.y; pos[3 * s + 2] = g.p.z;
  dir[3 * s] = g.d.x; dir[3 * s + 1] = g.d.y; dir[3 * s + 2] = g.d.z;
  e[s] = g.e;
  P.rng[0 * (size_t)P.rng_stride + s] = r.x;
  P.rng[1 * (size_t)P.rng_stride + s] = r.y;
  P.rng[2 * (size_t)P.rng_stride + s] = r.z;
  P.rng[3 * (size_t)P.rng_stride + s] = r.w;
  P.rng[4 * (size_t)P.rng_stride + s] = r.c;
}

__global__ void uniforms_kernel(RunParams P, int n_slots, int n, float* out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  Rng r;
  r.x = P.rng[0 * (size_t)P.rng_stride + s];
  r.y = P.rng[1 * (size_t)P.rng_stride + s];
  r.z = P.rng[2 * (size_t)P.rng_stride + s];
  r.w = P.rng[3 * (size_t)P.rng_stride + s];
  r.c = P.rng[4 * (size_t)P.rng_stride + s];
  for (int k = 0; k < n; ++k) out[(size_t)s * n + k] = r.uniform();
  P.rng[0 * (size_t)P.rng_stride + s] = r.x;
  P.rng[1 * (size_t)P.rng_stride + s] = r.y;
  P.rng[2 * (size_t)P.rng_stride + s] = r.z;
  P.rng[3 * (size_t)P.rng_stride + s] = r.w;
  P.rng[4 * (size_t)P.rng_stride + s] = r.c;
}

}  // namespace

// ===========================================================================
// host side: context and C ABI
// ===========================================================================
namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(-100 - (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

template <typename T>
int dev_alloc(DevBuf<T>& b, size_t n) {
  b.release();
  if (n == 0) return 0;
  HIP_TRY(hipMalloc(reinterpret_cast<void**>(&b.p), n * sizeof(T)));
  b.n = n;
  return 0;
}

template <typename T>
int dev_upload(DevBuf<T>& b, const T* host, size_t n, hipStream_t s) {
  int rc = dev_alloc(b, n);
  if (rc) return rc;
  if (n) HIP_TRY(hipMemcpyAsync(b.p, host, n * sizeof(T), hipMemcpyHostToDevice, s));
  return 0;
}

Mat34 to_mat34(const float* m16) {
  Mat34 m;
  for (int i = 0; i < 12; ++i) m.m[i] = m16[i];
  return m;
}

}  // namespace

struct ggh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
  bool have_physics = false, have_source = false, have_dose = false, seeded = false;
  // physics
  int nbins = 0, n_mat = 0, n_proc = 0, proc[3] = {0, 0, 0};
  float factor = 0.0f;
  std::vector<float> density_host;
  DevBuf<float> ebins, cs_atom_r, atom_dens, pp, density;
  DevBuf<float4> cs4;
  DevBuf<int32_t> n_elem, elem_idx;
  DevBuf<uint8_t> z;
  // solids
  std::vector<DevSolid> solids;
  std::vector<DevBuf<uint8_t>> labels;
  std::vector<DevBuf<int32_t>> hist, scat;
  std::vector<int> hist_elems;
  std::vector<int> max_material;  // largest global material index a solid can reach
  DevBuf<DevSolid> solids_dev;
  bool solids_dirty = true;
  // dosimetry
  ggh_dosimetry_desc dose_desc{};
  DevBuf<double> edep, edep2;
  DevBuf<int32_t> hit, ptrack;
  // source
  Mat34 src_m{};
  float aperture = 0.0f;
  f3 focal{0, 0, 0};
  int n_energy_arg = 1, is_interp = 0;
  DevBuf<float> spectrum, cdf;
  // slots
  DevBuf<uint32_t> rng;
  DevBuf<unsigned long long> stats;
  int hot_threshold = 40;
};

namespace {

int check_ctx(ggh_context* c) {
  if (!c) return fail(-1, "null context");
  HIP_TRY(hipSetDevice(c->device));
  return 0;
}

int sync_solids(ggh_context* c) {
  if (!c->solids_dirty) return 0;
  int rc = dev_upload(c->solids_dev, c->solids.data(), c->solids.size(), c->stream);
  if (rc) return rc;
  c->solids_dirty = false;
  return 0;
}

RunParams make_params(ggh_context* c) {
  RunParams P;
  std::memset(&P, 0, sizeof P);
  P.solids = c->solids_dev.p;
  P.n_solids = (int32_t)c->solids.size();
  P.nbins = c->nbins;
  P.n_mat = c->n_mat;
  P.n_proc = c->n_proc;
  for (int i = 0; i < 3; ++i) P.proc[i] = c->proc[i];
  P.ebins = c->ebins.p;
  P.cs4 = c->cs4.p;
  P.cs_atom_r = c->cs_atom_r.p;
  P.n_elem = c->n_elem.p;
  P.elem_idx = c->elem_idx.p;
  P.z = c->z.p;
  P.atom_dens = c->atom_dens.p;
  P.pp = c->pp.p;
  P.factor = c->factor;
  P.src_m = c->src_m;
  P.aperture = c->aperture;
  P.focal = c->focal;
  P.n_energy_arg = c->n_energy_arg;
  P.spectrum = c->spectrum.p;
  P.cdf = c->cdf.p;
  P.is_interp = c->is_interp;
  P.dose_solid = c->have_dose ? c->dose_desc.solid : -1;
  if (c->have_dose) {
    const ggh_dosimetry_desc& d = c->dose_desc;
    P.dose_lo = mk(d.border_min[0], d.border_min[1], d.border_min[2]);
    P.dose_hi = mk(d.border_max[0], d.border_max[1], d.border_max[2]);
    // inv_size_of_dosels_ = 1.0f / size (src/navigators/GGEMSDosimetryCalculator.cc:466-469)
    P.inv_dosel = mk(1.0f / d.dosel_size[0], 1.0f / d.dosel_size[1], 1.0f / d.dosel_size[2]);
    for (int i = 0; i < 3; ++i) P.nd[i] = d.n_dosels[i];
    P.edep = c->edep.p;
    P.edep2 = c->edep2.p;
    P.hit = c->hit.p;
    P.ptrack = c->ptrack.p;
  }
  P.rng = c->rng.p;
  P.rng_stride = GGH_MAX_SLOTS;
  P.stats = c->stats.p;
  P.hot_threshold = c->hot_threshold;
  return P;
}

uint32_t urandom_seed() {
  uint32_t s = 0;
  int fd = ::open("/dev/urandom", O_RDONLY);
  if (fd >= 0) {
    if (::read(fd, &s, sizeof s) != (ssize_t)sizeof s) s = 0;
    ::close(fd);
  }
  return s;
}

}  // namespace

extern "C" {

const char* ggh_last_error(void) { return g_error.c_str(); }

int ggh_device_count(int* count) {
  HIP_TRY(hipGetDeviceCount(count));
  return 0;
}

int ggh_context_create(int device, ggh_context** out) {
  if (!out) return fail(-1, "null output pointer");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(-2, "device index out of range");
  HIP_TRY(hipSetDevice(device));
  ggh_context* c = new (std::nothrow) ggh_context();
  if (!c) return fail(-3, "out of host memory");
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    delete c;
    return fail(-100 - (int)e, std::string("stream/event creation: ") + hipGetErrorString(e));
  }
  int rc = dev_alloc(c->stats, 8);
  if (!rc) rc = dev_alloc(c->rng, (size_t)5 * GGH_MAX_SLOTS);
  if (rc) {
    ggh_context_destroy(c);
    return rc;
  }
  (void)hipMemsetAsync(c->stats.p, 0, 8 * sizeof(unsigned long long), c->stream);
  *out = c;
  return 0;
}

int ggh_context_destroy(ggh_context* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->ebins.release(); c->cs_atom_r.release(); c->atom_dens.release(); c->pp.release(); c->density.release();
  c->cs4.release(); c->n_elem.release(); c->elem_idx.release(); c->z.release();
  for (auto& b : c->labels) b.release();
  for (auto& b : c->hist) b.release();
  for (auto& b : c->scat) b.release();
  c->solids_dev.release();
  c->edep.release(); c->edep2.release(); c->hit.release(); c->ptrack.release();
  c->spectrum.release(); c->cdf.release(); c->rng.release(); c->stats.release();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int ggh_set_physics(ggh_context* c, const ggh_physics_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || d->nbins < 2 || d->nbins > 2048 || d->n_materials < 1 || d->n_materials > 4096)
    return fail(-2, "invalid physics description");
  if (d->n_processes < 0 || d->n_processes > 3) return fail(-2, "invalid number of processes");
  for (int i = 0; i < d->n_processes; ++i)
    if (d->process_id[i] < 0 || d->process_id[i] > 2) return fail(-2, "invalid process id");
  const int nb = d->nbins, nm = d->n_materials;
  std::vector<float4> cs4((size_t)nb * nm);
  for (int m = 0; m < nm; ++m)
    for (int i = 0; i < nb; ++i) {
      float4 v;
      v.x = d->cross_sections[(size_t)0 * nm * nb + (size_t)m * nb + i];
      v.y = d->cross_sections[(size_t)1 * nm * nb + (size_t)m * nb + i];
      v.z = d->cross_sections[(size_t)2 * nm * nb + (size_t)m * nb + i];
      v.w = d->photon_cut[m];
      cs4[(size_t)m * nb + i] = v;
    }
  for (int m = 0; m < nm; ++m)
    if (d->element_index[m] < 0 || d->n_elements[m] < 1 ||
        d->element_index[m] + d->n_elements[m] > d->total_elements)
      return fail(-2, "invalid element table");
  for (int k = 0; k < d->total_elements; ++k)
    if (d->atomic_number[k] < 1 || d->atomic_number[k] > 100) return fail(-2, "atomic number out of range");
  c->nbins = nb;
  c->n_mat = nm;
  c->n_proc = d->n_processes;
  for (int i = 0; i < 3; ++i) c->proc[i] = i < d->n_processes ? d->process_id[i] : 0;
  c->factor = d->rayleigh_factor;
  c->density_host.assign(d->density, d->density + nm);
  if ((rc = dev_upload(c->ebins, d->energy_bins, (size_t)nb, c->stream))) return rc;
  if ((rc = dev_upload(c->cs4, cs4.data(), cs4.size(), c->stream))) return rc;
  if ((rc = dev_upload(c->cs_atom_r, d->cs_atom_rayleigh, (size_t)101 * nb, c->stream))) return rc;
  if ((rc = dev_upload(c->n_elem, d->n_elements, (size_t)nm, c->stream))) return rc;
  if ((rc = dev_upload(c->elem_idx, d->element_index, (size_t)nm, c->stream))) return rc;
  if ((rc = dev_upload(c->z, d->atomic_number, (size_t)d->total_elements, c->stream))) return rc;
  if ((rc = dev_upload(c->atom_dens, d->atom_density, (size_t)d->total_elements, c->stream))) return rc;
  if ((rc = dev_upload(c->pp, d->rayleigh_pp, (size_t)9 * 101, c->stream))) return rc;
  if ((rc = dev_upload(c->density, d->density, (size_t)nm, c->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_physics = true;
  return 0;
}

int ggh_add_voxel_solid(ggh_context* c, const ggh_voxel_solid_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || !d->labels) return fail(-2, "invalid voxel solid description");
  if (c->solids.size() >= GGH_MAX_SOLIDS) return fail(-2, "too many solids");
  size_t n = (size_t)d->n_voxels[0] * d->n_voxels[1] * d->n_voxels[2];
  if (d->n_voxels[0] <= 0 || d->n_voxels[1] <= 0 || d->n_voxels[2] <= 0 || n > (size_t)INT32_MAX)
    return fail(-2, "invalid voxel dimensions");
  DevSolid s;
  std::memset(&s, 0, sizeof s);
  s.m = to_mat34(d->obb.matrix);
  s.border.lo = mk(d->obb.border_min[0], d->obb.border_min[1], d->obb.border_min[2]);
  s.border.hi = mk(d->obb.border_max[0], d->obb.border_max[1], d->obb.border_max[2]);
  s.type = kVoxel;
  s.mat_offset = d->material_offset;
  s.voxel_size = mk(d->voxel_size[0], d->voxel_size[1], d->voxel_size[2]);
  for (int i = 0; i < 3; ++i) s.nv[i] = d->n_voxels[i];
  int max_label = 0;
  for (size_t i = 0; i < n; ++i) max_label = d->labels[i] > max_label ? d->labels[i] : max_label;
  DevBuf<uint8_t> lab;
  if ((rc = dev_upload(lab, d->labels, n, c->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  s.labels = lab.p;
  c->labels.push_back(lab);
  c->hist.push_back(DevBuf<int32_t>());
  c->scat.push_back(DevBuf<int32_t>());
  c->hist_elems.push_back(0);
  c->solids.push_back(s);
  c->max_material.push_back(d->material_offset + max_label);
  c->solids_dirty = true;
  return (int)c->solids.size() - 1;
}

int ggh_add_box_solid(ggh_context* c, const ggh_box_solid_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || d->n_elements[0] <= 0 || d->n_elements[1] <= 0 || d->n_elements[2] <= 0)
    return fail(-2, "invalid box solid description");
  if (c->solids.size() >= GGH_MAX_SOLIDS) return fail(-2, "too many solids");
  DevSolid s;
  std::memset(&s, 0, sizeof s);
  s.m = to_mat34(d->obb.matrix);
  s.border.lo = mk(d->obb.border_min[0], d->obb.border_min[1], d->obb.border_min[2]);
  s.border.hi = mk(d->obb.border_max[0], d->obb.border_max[1], d->obb.border_max[2]);
  s.type = kBoxSolid;
  s.mat_offset = d->material_offset;
  for (int i = 0; i < 3; ++i) s.ne[i] = d->n_elements[i];
  s.box_size = mk(d->box_size[0], d->box_size[1], d->box_size[2]);
  s.threshold = d->threshold;
  const int ne = d->n_elements[0] * d->n_elements[1];
  DevBuf<int32_t> h, sc;
  if ((rc = dev_alloc(h, (size_t)ne))) return rc;
  HIP_TRY(hipMemsetAsync(h.p, 0, (size_t)ne * sizeof(int32_t), c->stream));
  if (d->store_scatter) {
    if ((rc = dev_alloc(sc, (size_t)ne))) return rc;
    HIP_TRY(hipMemsetAsync(sc.p, 0, (size_t)ne * sizeof(int32_t), c->stream));
  }
  s.histogram = h.p;
  s.scatter = sc.p;
  c->labels.push_back(DevBuf<uint8_t>());
  c->hist.push_back(h);
  c->scat.push_back(sc);
  c->hist_elems.push_back(ne);
  c->solids.push_back(s);
  c->max_material.push_back(d->material_offset);
  c->solids_dirty = true;
  return (int)c->solids.size() - 1;
}

int ggh_set_dosimetry(ggh_context* c, const ggh_dosimetry_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || d->solid < 0 || d->solid >= (int)c->solids.size() || c->solids[d->solid].type != kVoxel)
    return fail(-2, "dosimetry needs a voxel solid");
  size_t n = (size_t)d->n_dosels[0] * d->n_dosels[1] * d->n_dosels[2];
  if (d->n_dosels[0] <= 0 || d->n_dosels[1] <= 0 || d->n_dosels[2] <= 0 || n > (size_t)INT32_MAX)
    return fail(-2, "invalid dosel grid");
  c->dose_desc = *d;
  if ((rc = dev_alloc(c->edep, n))) return rc;
  c->edep2.release();
  c->hit.release();
  c->ptrack.release();
  if (d->edep_squared && (rc = dev_alloc(c->edep2, n))) return rc;
  if (d->hit && (rc = dev_alloc(c->hit, n))) return rc;
  if (d->photon_tracking && (rc = dev_alloc(c->ptrack, n))) return rc;
  c->have_dose = true;
  return ggh_reset_tallies(c);
}

int ggh_set_xray_source(ggh_context* c, const ggh_xray_source_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || d->n_energy < 1 || !d->energy || !d->cdf) return fail(-2, "invalid source description");
  c->src_m = to_mat34(d->matrix);
  c->aperture = d->aperture;
  c->focal = mk(d->focal_spot[0], d->focal_spot[1], d->focal_spot[2]);
  c->n_energy_arg = d->n_energy + 1;
  c->is_interp = d->is_interp;
  // the reference binary search reads cdf[0..n_energy] (n_energy+1 entries);
  // one padding entry keeps the interpolation's idx+1 read in bounds
  std::vector<float> e(d->energy, d->energy + d->n_energy + 1), f(d->cdf, d->cdf + d->n_energy + 1);
  e.push_back(e.back());
  f.push_back(1.0f);
  if ((rc = dev_upload(c->spectrum, e.data(), e.size(), c->stream))) return rc;
  if ((rc = dev_upload(c->cdf, f.data(), f.size(), c->stream))) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->have_source = true;
  return 0;
}

int ggh_seed(ggh_context* c, uint32_t seed) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (seed == 0) seed = urandom_seed();
  std::mt19937 mt(seed);
  std::vector<uint32_t> st((size_t)5 * GGH_MAX_SLOTS);
  const size_t n = GGH_MAX_SLOTS;
  for (size_t j = 0; j < n; ++j) {
    st[0 * n + j] = (uint32_t)mt();
    st[1 * n + j] = (uint32_t)mt();
    st[2 * n + j] = (uint32_t)mt();
    st[3 * n + j] = (uint32_t)mt();
    st[4 * n + j] = 0;
  }
  HIP_TRY(hipMemcpyAsync(c->rng.p, st.data(), st.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->seeded = true;
  return 0;
}

int ggh_reset_tallies(ggh_context* c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  for (size_t i = 0; i < c->solids.size(); ++i) {
    if (c->hist[i].p) HIP_TRY(hipMemsetAsync(c->hist[i].p, 0, c->hist[i].n * sizeof(int32_t), c->stream));
    if (c->scat[i].p) HIP_TRY(hipMemsetAsync(c->scat[i].p, 0, c->scat[i].n * sizeof(int32_t), c->stream));
  }
  if (c->edep.p) HIP_TRY(hipMemsetAsync(c->edep.p, 0, c->edep.n * sizeof(double), c->stream));
  if (c->edep2.p) HIP_TRY(hipMemsetAsync(c->edep2.p, 0, c->edep2.n * sizeof(double), c->stream));
  if (c->hit.p) HIP_TRY(hipMemsetAsync(c->hit.p, 0, c->hit.n * sizeof(int32_t), c->stream));
  if (c->ptrack.p) HIP_TRY(hipMemsetAsync(c->ptrack.p, 0, c->ptrack.n * sizeof(int32_t), c->stream));
  HIP_TRY(hipMemsetAsync(c->stats.p, 0, 8 * sizeof(unsigned long long), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int ggh_run(ggh_context* c, uint64_t n) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_physics) return fail(-3, "physics tables not set");
  if (!c->have_source) return fail(-3, "source not set");
  if (!c->seeded) return fail(-3, "particle slots not seeded");
  if (c->solids.empty()) return fail(-3, "no solid registered");
  for (size_t i = 0; i < c->solids.size(); ++i)
    if (c->solids[i].mat_offset < 0 || c->max_material[i] >= c->n_mat)
      return fail(-3, "solid labels reach a material without physics tables");
  if (n == 0) return 0;
  if ((rc = sync_solids(c))) return rc;
  RunParams P = make_params(c);
  // batch layout of GGEMSSource::OrganizeParticlesInBatch (src/sources/GGEMSSource.cc:403-425)
  uint64_t nb = (uint64_t)std::ceil((float)n / (float)GGH_MAX_SLOTS);
  if (nb == 0) nb = 1;
  P.n_batches = nb;
  P.base = nb == 1 ? n : n / nb;
  P.rem = nb == 1 ? 0 : n % nb;
  uint64_t slots = P.base + (P.rem ? 1 : 0);
  if (slots > GGH_MAX_SLOTS) return fail(-3, "batch larger than the slot buffer");
  P.n_slots = (uint32_t)slots;
  const size_t lds = (size_t)c->n_mat * c->nbins * sizeof(float4);
  const unsigned grid = (unsigned)((slots + kBlock - 1) / kBlock);
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  if (lds <= 48 * 1024) {
    hipLaunchKernelGGL(transport_kernel<true>, dim3(grid), dim3(kBlock), lds, c->stream, P);
  } else {
    hipLaunchKernelGGL(transport_kernel<false>, dim3(grid), dim3(kBlock), 0, c->stream, P);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  HIP_TRY(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  return 0;
}

int ggh_synchronize(ggh_context* c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int ggh_last_kernel_ms(ggh_context* c, float* ms) {
  if (!c || !ms) return fail(-1, "null argument");
  *ms = c->last_ms;
  return 0;
}

int ggh_get_stats(ggh_context* c, ggh_stats* out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  unsigned long long h[8];
  HIP_TRY(hipMemcpyAsync(h, c->stats.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  out->primaries = h[0];
  out->voxel_steps = h[1];
  out->interactions = h[2];
  out->detector_steps = h[3];
  out->detector_interactions = h[4];
  out->hops = h[5];
  return 0;
}

int ggh_get_histogram(ggh_context* c, int solid, int32_t* counts, int32_t* scatter) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (solid < 0 || solid >= (int)c->solids.size() || c->solids[solid].type != kBoxSolid)
    return fail(-2, "not a box solid");
  const size_t n = (size_t)c->hist_elems[solid];
  if (counts) HIP_TRY(hipMemcpyAsync(counts, c->hist[solid].p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  if (scatter) {
    if (!c->scat[solid].p) return fail(-2, "scatter histogram not stored for this solid");
    HIP_TRY(hipMemcpyAsync(scatter, c->scat[solid].p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int ggh_get_dose_tallies(ggh_context* c, double* edep, double* edep2, int32_t* hit, int32_t* ptrack) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_dose) return fail(-2, "dosimetry not set");
  const size_t n = c->edep.n;
  if (edep) HIP_TRY(hipMemcpyAsync(edep, c->edep.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (edep2) {
    if (!c->edep2.p) return fail(-2, "edep squared not stored");
    HIP_TRY(hipMemcpyAsync(edep2, c->edep2.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (hit) {
    if (!c->hit.p) return fail(-2, "hits not stored");
    HIP_TRY(hipMemcpyAsync(hit, c->hit.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  }
  if (ptrack) {
    if (!c->ptrack.p) return fail(-2, "photon tracking not stored");
    HIP_TRY(hipMemcpyAsync(ptrack, c->ptrack.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int ggh_compute_dose(ggh_context* c, float scale, int32_t water_ref, float min_density, float* dose, float* unc) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_dose) return fail(-2, "dosimetry not set");
  if (unc && (!c->hit.p || !c->edep2.p)) return fail(-2, "uncertainty needs hit and edep squared tallies");
  const size_t n = c->edep.n;
  const DevSolid& s = c->solids[c->dose_desc.solid];
  DevBuf<float> d_dose, d_unc;
  if ((rc = dev_alloc(d_dose, n))) return rc;
  if (unc && (rc = dev_alloc(d_unc, n))) {
    d_dose.release();
    return rc;
  }
  DoseParams D;
  std::memset(&D, 0, sizeof D);
  D.edep = c->edep.p;
  D.edep2 = c->edep2.p;
  D.hit = c->hit.p;
  D.labels = s.labels;
  D.density = c->density.p;
  D.mat_offset = s.mat_offset;
  for (int i = 0; i < 3; ++i) {
    D.nd[i] = c->dose_desc.n_dosels[i];
    D.nv[i] = s.nv[i];
  }
  D.dosel_size = mk(c->dose_desc.dosel_size[0], c->dose_desc.dosel_size[1], c->dose_desc.dosel_size[2]);
  D.vs = s.voxel_size;
  D.vlo = s.border.lo;
  D.scale = scale;
  D.water_reference = water_ref;
  D.min_density = min_density;
  D.dose = d_dose.p;
  D.unc = unc ? d_unc.p : nullptr;
  unsigned grid = (unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, 65536);
  hipLaunchKernelGGL(dose_kernel, dim3(grid), dim3(kBlock), 0, c->stream, D);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(dose, d_dose.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && unc) e = hipMemcpyAsync(unc, d_unc.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  d_dose.release();
  d_unc.release();
  if (e != hipSuccess) return fail(-100 - (int)e, std::string("compute dose: ") + hipGetErrorString(e));
  return 0;
}

int ggh_debug_voxel_walk(ggh_context* c, int solid, int n_rays, const float* pos, const float* dir, int max_steps,
                         int32_t* out_index, int32_t* out_count) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (solid < 0 || solid >= (int)c->solids.size() || c->solids[solid].type != kVoxel)
    return fail(-2, "not a voxel solid");
  if (n_rays <= 0 || max_steps <= 0) return fail(-2, "invalid ray count");
  if ((rc = sync_solids(c))) return rc;
  RunParams P = make_params(c);
  DevBuf<float> dp, dd;
  DevBuf<int32_t> dout, dcnt;
  if ((rc = dev_upload(dp, pos, (size_t)3 * n_rays, c->stream))) return rc;
  if ((rc = dev_upload(dd, dir, (size_t)3 * n_rays, c->stream))) return rc;
  if ((rc = dev_alloc(dout, (size_t)n_rays * max_steps))) return rc;
  if ((rc = dev_alloc(dcnt, (size_t)n_rays))) return rc;
  hipLaunchKernelGGL(walk_kernel, dim3((n_rays + 63) / 64), dim3(64), 0, c->stream, P, solid, n_rays, dp.p, dd.p,
                     max_steps, dout.p, dcnt.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out_index, dout.p, (size_t)n_rays * max_steps * sizeof(int32_t), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(out_count, dcnt.p, (size_t)n_rays * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  dp.release(); dd.release(); dout.release(); dcnt.release();
  return 0;
}

int ggh_debug_primaries(ggh_context* c, int n, float* pos, float* dir, float* energy) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_source || !c->have_physics || !c->seeded) return fail(-3, "source, physics and seed required");
  if (n <= 0 || n > GGH_MAX_SLOTS) return fail(-2, "invalid count");
  RunParams P = make_params(c);
  DevBuf<float> dp, dd, de;
  if ((rc = dev_alloc(dp, (size_t)3 * n)) || (rc = dev_alloc(dd, (size_t)3 * n)) || (rc = dev_alloc(de, (size_t)n)))
    return rc;
  hipLaunchKernelGGL(primaries_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, P, n, dp.p, dd.p, de.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(pos, dp.p, (size_t)3 * n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(dir, dd.p, (size_t)3 * n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(energy, de.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  dp.release(); dd.release(); de.release();
  return 0;
}

int ggh_debug_uniforms(ggh_context* c, int n_slots, int n, float* out) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->seeded) return fail(-3, "particle slots not seeded");
  if (n_slots <= 0 || n_slots > GGH_MAX_SLOTS || n <= 0) return fail(-2, "invalid count");
  RunParams P = make_params(c);
  DevBuf<float> d;
  if ((rc = dev_alloc(d, (size_t)n_slots * n))) return rc;
  hipLaunchKernelGGL(uniforms_kernel, dim3((n_slots + 255) / 256), dim3(256), 0, c->stream, P, n_slots, n, d.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, d.p, (size_t)n_slots * n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  d.release();
  return 0;
}

void ggh_host_uniform_from_sum(const uint32_t* sums, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = uniform_from_sum(sums[i]);
}

}  // extern "C"
