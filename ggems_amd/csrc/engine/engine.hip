// engine.hip -- MI355X photon-transport engine behind the thin C ABI of
// include/ggems_engine.h.
//
// One persistent launch transports a whole run.  A particle slot is the
// unit of work: slot s owns JKISS32 stream s and transports, in order,
// photon s of every batch (src/global/GGEMS.cc:303-363 runs batches of
// MAXIMUM_PARTICLES slots and slot s of every batch continues stream s).
// A lane takes a slot from a device-wide queue, keeps its stream in
// registers while it transports the slot's photons, writes the stream back
// and takes the next slot, so lanes stay busy until the queue drains.  Each
// photon walks the reference's hop loop (find solid -> project -> track)
// inside the lane; the host never synchronises per hop.
//
// Divergence.  The voxel step is the hot path (hundreds of steps per
// primary, about one interaction per hundred steps).  A lane's work is a
// small state machine with two kinds of work:
//   * voxel steps while the photon is inside a phantom.  A step that ends in
//     an interaction, an exit or the energy cut only records why the lane
//     stopped, so the step itself stays short and branch-light;
//   * navigation for a stopped lane: it resolves its interaction, leaves the
//     phantom, starts its next primary, crosses to the next solid or tracks
//     through a detector module, until it is back inside a phantom.
// A block works in rounds: a few navigation passes, voxel steps, then the
// block regroups its photons through LDS so that the photons inside a
// phantom fill whole waves and the stopped ones gather in a few others
// (transport_kernel).  Regrouping and the state machine only reorder work
// between lanes; each photon consumes its slot's stream in exactly the
// reference order.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/ggems_engine.h"
#include "transport_device.hpp"

using namespace ggh;

namespace {

constexpr int kBlock = 512;            // 8 waves; 2 blocks per CU at 4 waves per SIMD
#ifndef GGH_WAVES_PER_EU
#define GGH_WAVES_PER_EU 4
#endif
constexpr int kWavesPerEU = GGH_WAVES_PER_EU;  // register budget: 512 / waves VGPRs per lane
constexpr int kLdsBudget = 64 * 1024;  // step tables + energy bins staged per block

// ---------------------------------------------------------------------------
// device-side scene description
// ---------------------------------------------------------------------------
enum SolidType : int32_t { kVoxel = 0, kBoxSolid = 1, kMesh = 2 };

// Everything the voxel step reads about its phantom.  rvs = RN(1/voxel_size)
// (exact division, see div_rn); in_lo/in_hi = border -/+ the geometry
// tolerance, the bounds IsParticleInAABB compares against.  labels is the
// MHD order (x fastest); bricks holds the same labels in 128-byte bricks of
// 8 x 4 x 4 voxels (kBrick*), nbx bricks along x and nbxy per brick layer.
struct VoxGeom {
  f3 lo, vs, rvs, in_lo, in_hi;
  int32_t nx, ny, nz, nxy;
  const uint8_t* labels;
  const uint8_t* bricks;
  int32_t mat_offset;
  int32_t solid;
  int32_t nbx, nbxy;
};

// 128-byte label bricks: one L2 line holds an 8 x 4 x 4 neighbourhood
// instead of 128 voxels of one row, so a walk in any direction touches fewer
// lines.  The tau-sampling voxel step reads them (measured faster: its loop
// waits on the gather); the reference step keeps the row-major labels (its
// loop is bound by issue, and the brick index costs it more than it saves).
constexpr int kBrickX = 8, kBrickY = 4, kBrickZ = 4;
static_assert(kBrickX * kBrickY * kBrickZ == 128, "brick_offset's shifts assume 8 x 4 x 4");

__host__ __device__ inline uint32_t brick_offset(int nbx, int nbxy, int vx, int vy, int vz) {
  const uint32_t b = (uint32_t)((vz >> 2) * nbxy + (vy >> 2) * nbx + (vx >> 3));
  return (b << 7) | ((uint32_t)(vz & 3) << 5) | ((uint32_t)(vy & 3) << 3) | (uint32_t)(vx & 7);
}

struct DevSolid {
  Mat34 m;           // local->global rows 0..2
  Box border;        // local borders
  int32_t type;
  int32_t mat_offset;
  VoxGeom vox;       // voxel solid
  // box solid (CT module)
  int32_t ne[3];
  f3 box_size;
  float threshold;
  int32_t* histogram;
  int32_t* scatter;
  // meshed solid: BVH nodes (two float4 each: {lo, skip}, {hi, first << 4 | count})
  // and triangles (three float4 each: v0, e1 = v1 - v0, e2 = v2 - v0)
  const float4* bvh;
  const float4* tris;
  int32_t n_bvh;
  int32_t box_index;  // rank among the box solids (per-projection matrices), -1 otherwise
  int32_t pad_[2];
};

static_assert(sizeof(DevSolid) % 16 == 0, "solids are staged in LDS after 16-byte table entries");

// One dosel's tallies in one aligned 32-byte record (GGEMSDoseRecording's
// edep, edep_squared and hit buffers, include/GGEMS/navigators/
// GGEMSDoseRecording.hh:44-52).  The three atomics of a deposit then touch
// one sector, so the dosel's line is fetched and written back once instead of
// three times; ggh_get_dose_tallies splits the records into the three arrays.
// edep^2 and hit are tallied whether or not they were requested: they ride in
// the same sector, and the readback refuses the ones not requested.
struct alignas(32) DoseTally {
  double edep;
  double edep2;
  int32_t hit;
  int32_t pad_[3];
};

static_assert(sizeof(DoseTally) == 32, "one dosel record per 32-byte sector");

struct RunParams {
  const DevSolid* solids;
  int32_t n_solids;
  // physics
  int32_t nbins;
  int32_t n_mat;
  int32_t n_proc;
  int32_t proc[3];            // activation order
  const float* ebins;
  float ebin_l2min, ebin_inv;  // energy_bin's guess: log2 of the first bin, (nbins - 1) / log2 span
  const float4* tab_sig;      // [n_mat*nbins] {sigma of proc[0], proc[1], proc[2], photon cut}
  const float4* tab_rcp;      // [n_mat*nbins] {RN(1/sigma) of proc[0..2], RN(1/sigma_tot)}
  const float* cs_r;          // [n_mat*nbins] Rayleigh sigma of the material (element selection)
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
  double new_aperture;        // 1 - cos(aperture) in FP64, computed on the host
  f3 focal;
  int32_t n_energy_arg;       // number_of_energy_bins_ + 1, as passed to the reference kernel
  const float* spectrum;
  const float* cdf;
  int32_t is_interp;
  // dosimetry
  int32_t dose_solid;
  f3 dose_lo, dose_hi, inv_dosel;
  int32_t nd[3];
  DoseTally* tally;
  int32_t* ptrack;
  // particle slots
  uint32_t* rng;              // SoA [5][rng_stride]
  uint32_t rng_stride;
  uint32_t n_slots;
  uint32_t n_batches, base, rem;
  uint32_t* queue;            // next slot to hand out
  unsigned long long* stats;  // ggh_stats fields
  int32_t round_steps;        // voxel steps a wave makes at least per round
  int32_t cold_passes;        // navigation passes per round
  int32_t interact_batch;
  uint32_t xchg_offset;       // byte offset of the lane-state exchange area in dynamic LDS
  VoxGeom vox0;               // the only voxel solid, when there is exactly one
  // CT acquisition: projection p places the source at acq_src[p] and box
  // solid j at acq_box[p * n_box + j]; its primaries are laid out in
  // batches as one run of that many primaries ({n_batches, base, rem}); the
  // histograms of box solids hold n_proj consecutive images
  const Mat34* acq_src;
  const Mat34* acq_box;
  const uint4* acq_layout;
  const int32_t* acq_phantom;  // [n_proj] the one non-box solid projection p sees (-1: all), or null
  int32_t n_proj;
  int32_t n_box;
};

// tables as seen by a block: LDS copies or the global arrays
struct Tabs {
  const float4* sig;
  const float4* rcp;
  const float* ebins;
  const DevSolid* solids;
  const float* spectrum;
  const float* cdf;
};

// ---------------------------------------------------------------------------
// lane state
// ---------------------------------------------------------------------------
enum Mode : int32_t {
  kInVoxel = 0,    // stepping through a voxel solid (hot)
  kInteract = 1,   // a voxel step ended on a discrete process (pend = proc | gmat << 2)
  kCut = 2,        // a voxel step left the photon at or below the material's energy cut
  kExit = 3,       // left the voxel solid
  kNeedPrimary = 4,
  kHop = 5,
  kNeedSlot = 6,
  kDone = 7
};

// p and ray hold the local frame of the solid while mode is kInVoxel ..
// kExit, the global frame otherwise.
struct Lane {
  Rng r;
  f3 p;
  Ray ray;
  float e;
  int32_t eid;
  int32_t mode;
  int32_t solid;
  int32_t hop;
  int32_t scatter;
  int32_t pend;
  uint32_t todo;
  uint32_t slot;
  uint32_t proj;  // current projection of an acquisition
  float tau;      // tau sampling: mean free paths left in the current flight
};

// Run counters.  Which form is faster depends on the kernel variant's
// register allocation (measured on the voxel-CT workload), so each variant
// takes the faster one; both count exactly the same events.
enum { kCntPrimaries = 0, kCntInteractions, kCntDetectorSteps, kCntDetectorInteractions, kCntHops, kNumCounters };

// 32-bit counters per lane, summed per wave in 64 bits at the end (the
// reference-sampling variants).  A lane wraps only past 2^32 counted events
// in one launch: with runs capped at 2^48 primaries, spread over 2^18 lanes
// by the slot queue, a lane sees about 2^30 primaries of a few events each.
struct LaneCounters {
  uint32_t v[kNumCounters];
  __device__ inline void add(int k) { ++v[k]; }
};

// 64-bit counters per wave in LDS (the tau-sampling variants): a counting
// site adds the number of its active lanes through the first of them.  Each
// wave owns its counters and a wave's LDS operations complete in program
// order, so no update is lost.
struct WaveCounters {
  unsigned long long* w;  // this wave's kNumCounters counters in LDS
  __device__ inline void add(int k) {
    const uint64_t m = __ballot(1);
    if (__lane_id() == __ffsll((unsigned long long)m) - 1) w[k] += (unsigned long long)__popcll(m);
  }
};

// ---------------------------------------------------------------------------
// physics (include/GGEMS/navigators/GGEMSPhotonNavigator.hh:57-131)
// ---------------------------------------------------------------------------
// -logf(u) for the uniform u of sum s, with the arithmetic of the device
// logf for normal arguments (log2 by v_log_f32, then times ln 2 in two
// pieces).  u is 0 or at least 2^-32, never subnormal, so logf's subnormal
// rescaling is left out.  For s == 0 (u = 0, where the reference's
// -log(0)/sigma is +inf) the result is NaN, which every caller compares
// with `<` exactly as +inf.
__device__ inline float neg_log_uniform(uint32_t s) {
  const float x = (float)s * 0x1p-32f;
  uint32_t b = __float_as_uint(x);
  b -= ((b & 0x7fffffu) - 1u) >= 4218928u ? 2u : 1u;  // uniform_from_sum; s == 0 leaves a NaN pattern
  const float r = __builtin_amdgcn_logf(__uint_as_float(b));
  const float c = __uint_as_float(0x3f317217u), cc = __uint_as_float(0x3377d1cfu);
  const float y = r * c;
  float t = __builtin_fmaf(r, c, -y);
  t = __builtin_fmaf(cc, r, t);
  return -(y + t);
}

// GetPhotonNextInteraction: competing exponentials over the activated
// processes, in activation order, with the left-bin cross section
// (GGEMSPhotonNavigator.hh:57-93).  The table entry holds the cross sections
// in activation order and their correctly rounded reciprocals, so that
// -log(u)/sigma is the IEEE quotient without a division.  NP is the number
// of activated processes, or -1 to read it from the run parameters.
template <int NP>
__device__ inline float next_interaction(const RunParams& P, const float4& s, const float4& y, const float* lg,
                                         int& proc) {
  float next = kOutOfWorld;
  proc = kTransportation;  // NO_PROCESS == TRANSPORTATION == 99
  const int np = NP >= 0 ? NP : P.n_proc;
  if (np > 0) {
    const float d = div_rn(lg[0], s.x, y.x);
    const bool w = d < next;
    next = w ? d : next;
    proc = w ? P.proc[0] : proc;
  }
  if (np > 1) {
    const float d = div_rn(lg[1], s.y, y.y);
    const bool w = d < next;
    next = w ? d : next;
    proc = w ? P.proc[1] : proc;
  }
  if (np > 2) {
    const float d = div_rn(lg[2], s.z, y.z);
    const bool w = d < next;
    next = w ? d : next;
    proc = w ? P.proc[2] : proc;
  }
  return next;
}

// the -log(u) of every activated process, drawn in activation order
template <int NP>
__device__ inline void draw_logs(const RunParams& P, Rng& r, float* lg) {
  const int np = NP >= 0 ? NP : P.n_proc;
  if (np > 0) lg[0] = neg_log_uniform(r.next_sum());
  if (np > 1) lg[1] = neg_log_uniform(r.next_sum());
  if (np > 2) lg[2] = neg_log_uniform(r.next_sum());
}

// Tau sampling (opt-in; no counterpart in GGEMS, which draws three competing
// exponentials per voxel step, GGEMSPhotonNavigator.hh:57-93): one number of
// mean free paths -log(u) per flight, drawn when the photon enters a voxel
// solid and after each interaction that leaves it alive.  u = 0 gives +inf
// (the photon never interacts), as -logf(0) does.
__device__ inline float draw_tau(Rng& r) {
  const uint32_t s = r.next_sum();
  const float t = neg_log_uniform(s);
  return s == 0u ? __builtin_inff() : t;
}

// KleinNishinaComptonSampleSecondaries (GGEMSComptonScatteringModels.hh:47-132)
__device__ void compton(float& e, f3& dir, Rng& r) {
  float kE0 = e;
  float kE0_MeC2 = kE0 / kElectronMassC2;
  f3 dir0 = dir;
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
  dir = normalize(rotate_unit_z(gd, dir0));
  e = kE0 * epsilon;
}

// LivermoreRayleighSampleSecondaries (GGEMSRayleighScatteringModels.hh:169-316).
// Returns false when the photon is killed (E <= 250 eV).
__device__ bool rayleigh(const RunParams& P, const float* ebins, float kE0, int eid, f3& dir, int gmat,
                                      Rng& r) {
  if (kE0 <= 250.0e-6f) return false;
  f3 dir0 = dir;
  const int nb = P.nbins;
  const int nm1 = P.n_elem[gmat] - 1;
  const int mix = P.elem_idx[gmat];
  int selz = P.z[mix + nm1];
  if (nm1 > 0) {
    float e_lo = ebins[eid], e_hi = ebins[eid + 1];
    float kCS = lin_interp(e_lo, P.cs_r[gmat * nb + eid], e_hi, P.cs_r[gmat * nb + eid + 1], kE0);
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
  dir = normalize(rotate_unit_z(gd, dir0));
  return true;
}

// PhotonDiscreteProcess (GGEMSPhotonNavigator.hh:109-131).  Returns false
// when the photon dies.
__device__ inline bool discrete(const RunParams& P, const float* ebins, int proc, int gmat, float& e, int eid, f3& dir,
                                Rng& r) {
  if (proc == kCompton) {
    compton(e, dir, r);
    return true;
  }
  if (proc == kPhotoelectric) {
    e = 0.0f;
    return false;
  }
  if (proc == kRayleigh) return rayleigh(P, ebins, e, eid, dir, gmat, r);
  return true;
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
  DoseTally* t = P.tally + id;
  atomicAdd(&t->hit, 1);
  unsafeAtomicAdd(&t->edep, (double)edep);
  unsafeAtomicAdd(&t->edep2, (double)edep * (double)edep);
}

// ---------------------------------------------------------------------------
// voxel solid transport (src/kernels/TrackThroughGGEMSVoxelizedSolid.cl:133-310)
// ---------------------------------------------------------------------------
__device__ inline bool in_bounds(f3 p, const VoxGeom& G) {
  // IsParticleInAABB with its tolerance folded into the bounds
  return !((p.x < G.in_lo.x) | (p.x > G.in_hi.x) | (p.y < G.in_lo.y) | (p.y > G.in_hi.y) | (p.z < G.in_lo.z) |
           (p.z > G.in_hi.z));
}

// Index of the voxel holding p: (int)((p - border_min) / voxel_size) with the
// IEEE quotient.  False when outside the label volume (the reference checks
// the upper bound only; a negative index would read outside the volume, so
// it ends the walk like an exit).
__device__ inline bool voxel_of(f3 p, const VoxGeom& G, int& vx, int& vy, int& vz) {
  vx = (int)div_rn(p.x - G.lo.x, G.vs.x, G.rvs.x);
  vy = (int)div_rn(p.y - G.lo.y, G.vs.y, G.rvs.y);
  vz = (int)div_rn(p.z - G.lo.z, G.vs.z, G.rvs.z);
  return ((uint32_t)vx < (uint32_t)G.nx) & ((uint32_t)vy < (uint32_t)G.ny) & ((uint32_t)vz < (uint32_t)G.nz);
}

__device__ inline Box voxel_box(const VoxGeom& G, int vx, int vy, int vz) {
  Box b;
  b.lo = mk(G.lo.x + (float)vx * G.vs.x, G.lo.y + (float)vy * G.vs.y, G.lo.z + (float)vz * G.vs.z);
  b.hi = add(b.lo, G.vs);
  return b;
}

// One iteration of the reference do-while, up to the point where the step's
// outcome is known.  Returns true when the step counts as a voxel step (the
// voxel index was inside the volume).  The label gather is issued first and
// the random draws of the step, which do not depend on it, cover its latency.
template <int NP, bool kTrack>
__device__ inline bool hot_step(const RunParams& P, const Tabs& T, const VoxGeom& G, Lane& L) {
  int vx, vy, vz;
  if (!voxel_of(L.p, G, vx, vy, vz)) {
    L.mode = kExit;
    return false;
  }
  const uint32_t label = G.labels[vx + vy * G.nx + vz * G.nxy];
  float lg[3];
  draw_logs<NP>(P, L.r, lg);
  const Box vox = voxel_box(G, vx, vy, vz);
  safety_inside(L.p, vox);
  const float db = distance_to_aabb(L.p, L.ray, vox);
  const int gmat = G.mat_offset + (int)label;
  const int ti = gmat * P.nbins + L.eid;
  const float4 sig = T.sig[ti];
  const float4 rcp = T.rcp[ti];
  int proc;
  float next = next_interaction<NP>(P, sig, rcp, lg, proc);
  if (db <= next) {
    next = db + kTol;
    proc = kTransportation;
    if (kTrack && P.ptrack && G.solid == P.dose_solid) {
      int id = dosel_index(P, L.p);
      if (id >= 0) atomicAdd(&P.ptrack[id], 1);
    }
  }
  L.p = add(L.p, scl(L.ray.d, next));
  safety_outside(L.p, vox);
  // outcome by selects: stores to different lane fields on different paths
  // would be merged by the compiler into one store through a selected
  // address, which keeps the lane state out of registers
  const bool out = !in_bounds(L.p, G);
  const bool inter = !out & (proc != kTransportation);
  const bool cut = !out & !inter & (L.e <= sig.w);  // photon_energy_cut_[material]
  L.mode = out ? kExit : (inter ? kInteract : (cut ? kCut : kInVoxel));
  L.pend = inter ? (proc | (gmat << 2)) : L.pend;
  return true;
}

// The voxel step of tau sampling: the walk, tolerances and outcomes of
// hot_step, with the interaction decided by the flight's tau instead of
// three fresh exponentials.  A step whose boundary distance db costs
// a = sigma_tot * db <= tau crosses the voxel and spends a; otherwise the
// photon interacts at tau / sigma_tot (the IEEE quotient) and the process
// is chosen when the interaction is resolved.  sigma_tot = (sigma_0 +
// sigma_1) + sigma_2 in activation order, its reciprocal in the w of the
// second table entry.
template <bool kTrack>
__device__ inline bool hot_step_tau(const RunParams& P, const Tabs& T, const VoxGeom& G, Lane& L) {
  int vx, vy, vz;
  if (!voxel_of(L.p, G, vx, vy, vz)) {
    L.mode = kExit;
    return false;
  }
  const uint32_t label = G.bricks[brick_offset(G.nbx, G.nbxy, vx, vy, vz)];
  const Box vox = voxel_box(G, vx, vy, vz);
  safety_inside(L.p, vox);
  const float db = distance_to_aabb(L.p, L.ray, vox);
  const int gmat = G.mat_offset + (int)label;
  const int ti = gmat * P.nbins + L.eid;
  const float4 sig = T.sig[ti];
  const float rtot = T.rcp[ti].w;
  const float tot = (sig.x + sig.y) + sig.z;
  const float a = tot * db;
  const bool cross = !(L.tau < a);
  const float next = cross ? db + kTol : div_rn(L.tau, tot, rtot);
  L.tau = cross ? L.tau - a : L.tau;
  if (kTrack && cross && P.ptrack && G.solid == P.dose_solid) {
    int id = dosel_index(P, L.p);
    if (id >= 0) atomicAdd(&P.ptrack[id], 1);
  }
  L.p = add(L.p, scl(L.ray.d, next));
  safety_outside(L.p, vox);
  const bool out = !in_bounds(L.p, G);
  const bool inter = !out & !cross;
  const bool cut = !out & !inter & (L.e <= sig.w);  // photon_energy_cut_[material]
  L.mode = out ? kExit : (inter ? kInteract : (cut ? kCut : kInVoxel));
  L.pend = inter ? (gmat << 2) : L.pend;
  return true;
}

// The same walk with no interaction (sigma == 0): records the voxel indices.
__device__ inline bool walk_step(const VoxGeom& G, f3& p, const Ray& ray, int32_t* out, int& n, int max_out) {
  int vx, vy, vz;
  if (!voxel_of(p, G, vx, vy, vz)) return false;
  if (n < max_out) out[n] = vx + vy * G.nx + vz * G.nxy;
  ++n;
  const Box vox = voxel_box(G, vx, vy, vz);
  safety_inside(p, vox);
  const float db = distance_to_aabb(p, ray, vox);
  float next = kOutOfWorld;
  if (db <= next) next = db + kTol;
  p = add(p, scl(ray.d, next));
  safety_outside(p, vox);
  return in_bounds(p, G);
}

// ---------------------------------------------------------------------------
// CT module transport (src/kernels/TrackThroughGGEMSSolidBox.cl:56-261)
// ---------------------------------------------------------------------------
// kAcq: the module sits where the lane's projection puts it and scores into
// that projection's image.
template <bool kAcq, class Counters>
__device__ void track_box(const RunParams& P, const Tabs& T, const DevSolid& s, Lane& L, Counters& c) {
  const Mat34& m = kAcq ? P.acq_box[L.proj * P.n_box + s.box_index] : s.m;
  f3 lp = g2l_pos(m, L.p);
  Ray ld;
  ld.set(g2l_dir(m, L.ray.d));
  const Box b = s.border;
  const int gmat = s.mat_offset;
  bool alive = true;
  do {
    int proc;
    const int ti = gmat * P.nbins + L.eid;
    float lg[3];
    draw_logs<-1>(P, L.r, lg);
    float next = next_interaction<-1>(P, T.sig[ti], T.rcp[ti], lg, proc);
    c.add(kCntDetectorSteps);
    safety_inside(lp, b);
    float db = distance_to_aabb(lp, ld, b);
    if (db <= next) {
      next = db + kTol;
      proc = kTransportation;
    }
    lp = add(lp, scl(ld.d, next));
    safety_outside(lp, b);
    if (!in_aabb(lp, b)) break;  // leaves the module: back to navigation
    if (L.e < s.threshold) alive = false;
    if (proc != kTransportation) {
      c.add(kCntDetectorInteractions);
      float e0 = L.e;
      f3 dir = ld.d;
      if (!discrete(P, T.ebins, proc, gmat, L.e, L.eid, dir, L.r)) alive = false;
      if (proc != kPhotoelectric) {
        ld.set(dir);
        if (L.e != e0) L.eid = energy_bin(L.e, T.ebins, P.nbins, P.ebin_l2min, P.ebin_inv);
      }
      if (proc == kPhotoelectric || proc == kCompton) {
        float esx = s.box_size.x / (float)s.ne[0];
        float esy = s.box_size.y / (float)s.ne[1];
        int ix = (int)((lp.x - b.lo.x) / esx);
        int iy = (int)((lp.y - b.lo.y) / esy);
        if (ix >= 0 && ix < s.ne[0] && iy >= 0 && iy < s.ne[1]) {
          // image of the lane's projection (acquisitions stack them)
          const size_t id = (size_t)(ix + iy * s.ne[0]) + (kAcq ? (size_t)L.proj * s.ne[0] * s.ne[1] : 0);
          atomicAdd(&s.histogram[id], 1);
          if (s.scatter && L.scatter) atomicAdd(&s.scatter[id], 1);
        }
      }
    }
  } while (alive);
  L.p = l2g_pos(m, lp);
  L.ray.d = l2g_dir(m, ld.d);
  L.mode = alive ? kHop : kNeedPrimary;
}

// ---------------------------------------------------------------------------
// meshed solid transport (src/kernels/TrackThroughGGEMSMeshSolid.cl)
// ---------------------------------------------------------------------------
// MollerTrumboreRayTriangle (TrackThroughGGEMSMeshSolid.cl:119-176), with the
// edges precomputed on the host (the same float subtractions); dot products
// accumulate left to right as DotMesh does.
__device__ inline bool moller_trumbore(f3 pi, f3 d, const float4& v0, const float4& e1, const float4& e2, f3& po) {
  const f3 h = mk((d.y * e2.z) - (d.z * e2.y), (d.z * e2.x) - (d.x * e2.z), (d.x * e2.y) - (d.y * e2.x));
  float a = e1.x * h.x;
  a += e1.y * h.y;
  a += e1.z * h.z;
  if (a > -kEps6 && a < kEps6) return false;
  const float f = 1.0f / a;
  const f3 sv = mk(pi.x - v0.x, pi.y - v0.y, pi.z - v0.z);
  float sh = sv.x * h.x;
  sh += sv.y * h.y;
  sh += sv.z * h.z;
  const float u = f * sh;
  if (u < 0.0f || u > 1.0f) return false;
  const f3 q = mk((sv.y * e1.z) - (sv.z * e1.y), (sv.z * e1.x) - (sv.x * e1.z), (sv.x * e1.y) - (sv.y * e1.x));
  float qd = q.x * d.x;
  qd += q.y * d.y;
  qd += q.z * d.z;
  const float v = f * qd;
  if (v < 0.0f || u + v > 1.0f) return false;
  float qe = q.x * e2.x;
  qe += q.y * e2.y;
  qe += q.z * e2.z;
  const float t = f * qe;
  if (!(t > kEps6)) return false;
  po = mk(pi.x + d.x * t, pi.y + d.y * t, pi.z + d.z * t);
  return true;
}

// GetMeshIOPositions (TrackThroughGGEMSMeshSolid.cl:253-314): every
// intersection of the ray with the mesh, counted, and the two nearest in the
// order of the reference's stable insertion sort.  The reference scans all
// octree nodes; here a BVH in depth-first order with skip links is walked
// without a stack.  Its boxes are widened on the host so that pruning never
// drops a triangle the exact test would hit.
__device__ inline int mesh_hits(const DevSolid& s, f3 p, f3 d, f3& q0, f3& q1) {
  const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
  int n = 0;
  float d0 = 0.0f, d1 = 0.0f;
  int i = 0;
  while (i < s.n_bvh) {
    const float4 a = s.bvh[2 * i], b = s.bvh[2 * i + 1];
    float t1 = (a.x - p.x) * inv.x, t2 = (b.x - p.x) * inv.x;
    float tmin = fmaxf(0.0f, fminf(t1, t2)), tmax = fmaxf(t1, t2);
    t1 = (a.y - p.y) * inv.y;
    t2 = (b.y - p.y) * inv.y;
    tmin = fmaxf(tmin, fminf(t1, t2));
    tmax = fminf(tmax, fmaxf(t1, t2));
    t1 = (a.z - p.z) * inv.z;
    t2 = (b.z - p.z) * inv.z;
    tmin = fmaxf(tmin, fminf(t1, t2));
    tmax = fminf(tmax, fmaxf(t1, t2));
    if (tmin <= tmax) {
      const int info = __float_as_int(b.w);
      const int count = info & 15, first = info >> 4;
      for (int k = 0; k < count; ++k) {
        const int t = first + k;
        f3 q;
        if (!moller_trumbore(p, d, s.tris[3 * t], s.tris[3 * t + 1], s.tris[3 * t + 2], q)) continue;
        // DistanceSquare (:101-104)
        const float d2 = (q.x - p.x) * (q.x - p.x) + (q.y - p.y) * (q.y - p.y) + (q.z - p.z) * (q.z - p.z);
        ++n;
        if (n == 1) {
          q0 = q;
          d0 = d2;
        } else if (d2 < d0) {
          q1 = q0;
          d1 = d0;
          q0 = q;
          d0 = d2;
        } else if (n == 2 || d2 < d1) {
          q1 = q;
          d1 = d2;
        }
      }
      ++i;
    } else {
      i = __float_as_int(a.w);
    }
  }
  return n;
}

// track_through_ggems_meshed_solid (TrackThroughGGEMSMeshSolid.cl:322-534):
// one material (the navigator's first), global coordinates.  A photon
// outside the mesh is moved onto its first surface; when no surface is
// ahead it is projected out of the octree box.
__device__ void track_mesh(const RunParams& P, const Tabs& T, const DevSolid& s, Lane& L) {
  const int gmat = s.mat_offset;
  const bool dose = L.solid == P.dose_solid;
  f3 lp = L.p, ld = L.ray.d;
  bool alive = true;
  uint32_t steps = 0, inter = 0;
  for (;;) {
    f3 q0 = lp, q1 = lp;
    const int n = mesh_hits(s, lp, ld, q0, q1);
    if (n == 0) break;
    f3 next_tri = q0;
    if ((n & 1) == 0) {
      lp = q0;
      next_tri = q1;
    }
    float lg[3];
    draw_logs<-1>(P, L.r, lg);
    int proc;
    const int ti = gmat * P.nbins + L.eid;
    float next = next_interaction<-1>(P, T.sig[ti], T.rcp[ti], lg, proc);
    ++steps;
    const float dx = (next_tri.x - lp.x), dy = (next_tri.y - lp.y), dz = (next_tri.z - lp.z);
    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    if (dist <= next) {
      next = dist + kTol;
      proc = kTransportation;
      if (dose && P.ptrack) {  // dose_photon_tracking at the point of departure
        const int id = dosel_index(P, lp);
        if (id >= 0) atomicAdd(&P.ptrack[id], 1);
      }
    }
    lp = add(lp, scl(ld, next));
    if (proc != kTransportation) {
      ++inter;
      const float e0 = L.e;
      f3 dir = ld;
      if (!discrete(P, T.ebins, proc, gmat, L.e, L.eid, dir, L.r)) alive = false;
      if (proc == kCompton || proc == kRayleigh) L.scatter = 1;
      if (dose) dose_record(P, e0 - L.e, lp);
      if (proc != kPhotoelectric) {
        ld = dir;
        if (L.e != e0) L.eid = energy_bin(L.e, T.ebins, P.nbins, P.ebin_l2min, P.ebin_inv);
      }
    }
    if (L.e <= T.sig[gmat * P.nbins].w) {  // photon_energy_cut_[0]
      if (dose) dose_record(P, L.e, lp);
      alive = false;
    }
    if (!alive) break;
  }
  // ComputeDistanceToOBB, then the projection out of the box
  Ray r;
  r.set(g2l_dir(s.m, ld));
  const float dist = distance_to_aabb(g2l_pos(s.m, lp), r, s.border);
  L.p = add(lp, scl(ld, dist + kTol));
  L.ray.d = ld;
  L.mode = alive ? kHop : kNeedPrimary;
  atomicAdd(&P.stats[6], (unsigned long long)steps);
  atomicAdd(&P.stats[7], (unsigned long long)inter);
}

// ---------------------------------------------------------------------------
// source (src/kernels/GetPrimariesGGEMSXRaySource.cl:52-143)
// ---------------------------------------------------------------------------
template <bool kAcq, bool kTau>
__device__ void sample_primary(const RunParams& P, const Tabs& T, const Mat34& src_m, Rng& r, f3& pos, f3& dir,
                               float& energy, int& eid) {
  double phi = r.uniform();
  double theta = r.uniform();
  phi *= (double)kTwoPi;
  // The acquisition and tau-sampling variants read 1 - cos(aperture) from
  // the host and take sin and cos of phi from one sincos (the same values:
  // one range reduction for both); the plain reference-sampling variant
  // recomputes the first and calls sin and cos, which measured faster there
  // (register allocation).  The host's and the device's FP64 cos agree on
  // every aperture the tests use.
  const double new_aperture = (kAcq || kTau) ? P.new_aperture : 1.0 - cos((double)P.aperture);
  theta = acos(1.0 - new_aperture * theta);
  double st = sin(theta);
  double sp, cp;
  if (kTau) {
    sincos(phi, &sp, &cp);
  } else {
    sp = sin(phi);
    cp = cos(phi);
  }
  f3 rotation = mk((float)(cp * st), (float)(sp * st), (float)cos(theta));
  f3 gp = l2g_pos(src_m, mk(0.0f, 0.0f, 0.0f));
  f3 d = normalize(mk(0.0f - gp.x, 0.0f - gp.y, 0.0f - gp.z));
  d = normalize(rotate_unit_z(rotation, d));
  gp.x = P.focal.x * (r.uniform() - 0.5f);
  gp.y = P.focal.y * (r.uniform() - 0.5f);
  gp.z = P.focal.z * (r.uniform() - 0.5f);
  gp = l2g_pos(src_m, gp);
  int idx = 0;
  float rnd = 0.0f;
  if (P.n_energy_arg != 1) {
    rnd = r.uniform();
    idx = binary_search(rnd, T.cdf, P.n_energy_arg + 1);
  }
  energy = (!P.is_interp) ? T.spectrum[idx]
                          : lin_interp(T.cdf[idx], T.spectrum[idx], T.cdf[idx + 1], T.spectrum[idx + 1], rnd);
  pos = gp;
  dir = d;
  eid = energy_bin(energy, T.ebins, P.nbins, P.ebin_l2min, P.ebin_inv);
}

// ---------------------------------------------------------------------------
// navigation between solids
// ---------------------------------------------------------------------------
// particle_solid_distance_* over every solid in registration order
// (src/kernels/ParticleSolidDistanceGGEMSVoxelizedSolid.cl:43-108) followed by
// project_to_* (src/kernels/ProjectToGGEMSVoxelizedSolid.cl:44-129).  Returns
// the solid the photon now sits in, or -1 when it leaves the world.
template <bool kAcq>
__device__ inline const Mat34& solid_matrix(const RunParams& P, const DevSolid& s, uint32_t proj) {
  return (kAcq && s.type == kBoxSolid) ? P.acq_box[proj * P.n_box + s.box_index] : s.m;
}

template <bool kAcq>
__device__ int find_and_project(const RunParams& P, const Tabs& T, f3& pos, f3 dir, uint32_t proj) {
  float dist = kOutOfWorld;
  int solid = -1;
  // an acquisition over several phantoms: the other phantoms are absent
  const int only = (kAcq && P.acq_phantom) ? P.acq_phantom[proj] : -1;
  for (int k = 0; k < P.n_solids; ++k) {
    const DevSolid& s = T.solids[k];
    if (kAcq && only >= 0 && k != only && s.type != kBoxSolid) continue;
    const Mat34& m = solid_matrix<kAcq>(P, s, proj);
    f3 lp = g2l_pos(m, pos);
    if (in_aabb(lp, s.border)) {
      dist = 0.0f;
      solid = k;
      break;
    }
    Ray lr;
    lr.set(g2l_dir(m, dir));
    float d = distance_to_aabb(lp, lr, s.border);
    if (d < dist) {
      dist = d;
      solid = k;
    }
  }
  if (solid == -1 || dist == kOutOfWorld) return -1;
  const DevSolid& s = T.solids[solid];
  const Mat34& m = solid_matrix<kAcq>(P, s, proj);
  pos = add(pos, scl(dir, dist + kTol));
  f3 lp = g2l_pos(m, pos);
  safety_inside(lp, s.border);
  pos = l2g_pos(m, lp);
  return solid;
}

// a wave's counters into the run statistics (ggh_stats order: primaries,
// voxel steps, interactions, detector steps, detector interactions, hops)
__device__ inline int stat_slot(int k) { return k == kCntPrimaries ? 0 : k + 1; }

__device__ inline void flush_counters(const RunParams& P, const LaneCounters& c) {
  for (int k = 0; k < kNumCounters; ++k) {
    unsigned long long t = c.v[k];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (__lane_id() == 0 && t) atomicAdd(&P.stats[stat_slot(k)], t);
  }
}

__device__ inline void flush_counters(const RunParams& P, const WaveCounters& c) {
  const int k = __lane_id();
  if (k < kNumCounters) {
    const unsigned long long v = c.w[k];
    if (v) atomicAdd(&P.stats[stat_slot(k)], v);
  }
}

__device__ inline void load_stream(const RunParams& P, uint32_t slot, Rng& r) {
  r.x = P.rng[0 * (size_t)P.rng_stride + slot];
  r.y = P.rng[1 * (size_t)P.rng_stride + slot];
  r.z = P.rng[2 * (size_t)P.rng_stride + slot];
  r.w = P.rng[3 * (size_t)P.rng_stride + slot];
  r.c = P.rng[4 * (size_t)P.rng_stride + slot];
}

__device__ inline void store_stream(const RunParams& P, uint32_t slot, const Rng& r) {
  P.rng[0 * (size_t)P.rng_stride + slot] = r.x;
  P.rng[1 * (size_t)P.rng_stride + slot] = r.y;
  P.rng[2 * (size_t)P.rng_stride + slot] = r.z;
  P.rng[3 * (size_t)P.rng_stride + slot] = r.w;
  P.rng[4 * (size_t)P.rng_stride + slot] = r.c;
}

// The discrete process a voxel step stopped on, then the energy cut
// (TrackThroughGGEMSVoxelizedSolid.cl:238-310).
template <int NP, bool kTau, class Counters>
__device__ inline void resolve_interaction(const RunParams& P, const Tabs& T, Lane& L, Counters& c) {
  const int gmat = L.pend >> 2;
  int proc = L.pend & 3;
  if (kTau) {
    // the process, in proportion to its left-bin cross section: the first
    // whose running sum in activation order exceeds u * sigma_tot
    const int np = NP >= 0 ? NP : P.n_proc;
    const float4 s = T.sig[gmat * P.nbins + L.eid];
    const float x = L.r.uniform() * ((s.x + s.y) + s.z);
    proc = P.proc[np > 0 ? np - 1 : 0];
    if (np > 1) {
      const float c0 = s.x;
      if (x < c0) proc = P.proc[0];
      else if (np > 2 && x < c0 + s.y) proc = P.proc[1];
    }
  }
  c.add(kCntInteractions);
  const float e0 = L.e;
  f3 dir = L.ray.d;
  bool alive = discrete(P, T.ebins, proc, gmat, L.e, L.eid, dir, L.r);
  if (proc == kCompton || proc == kRayleigh) L.scatter = 1;
  const bool dose = L.solid == P.dose_solid;
  if (dose) dose_record(P, e0 - L.e, L.p);
  if (proc != kPhotoelectric) {
    L.ray.set(dir);
    if (L.e != e0) L.eid = energy_bin(L.e, T.ebins, P.nbins, P.ebin_l2min, P.ebin_inv);
  }
  if (L.e <= T.sig[gmat * P.nbins].w) {  // photon_energy_cut_[material]
    if (dose) dose_record(P, L.e, L.p);
    alive = false;
  }
  L.mode = alive ? kInVoxel : kNeedPrimary;
  if (kTau && alive) L.tau = draw_tau(L.r);
}

// primaries of particle slot s in a run laid out as {n_batches, base, rem}:
// batch b holds base + (b < rem) photons and slot s takes photon s of each
__device__ inline uint32_t slot_todo(uint32_t n_batches, uint32_t base, uint32_t rem, uint32_t s) {
  return s < base ? n_batches : (s == base ? rem : 0u);
}

// an acquisition slot runs its primaries of projection 0, then of 1, ...:
// its random stream continues exactly as over consecutive runs
__device__ inline void next_projection(const RunParams& P, Lane& L) {
  while (L.todo == 0 && L.proj + 1 < (uint32_t)P.n_proj) {
    ++L.proj;
    const uint4 lay = P.acq_layout[L.proj];
    L.todo = slot_todo(lay.x, lay.y, lay.z, L.slot);
  }
}

// the cold work of one lane, one state at a time in an order that lets a
// lane pass through several states in one call
template <int NP, bool kSingle, bool kAcq, bool kTau, class Counters>
__device__ inline void cold_step(const RunParams& P, const Tabs& T, Lane& L, Counters& c) {
  if (L.mode == kInteract) resolve_interaction<NP, kTau>(P, T, L, c);
  if (L.mode == kCut) {
    if (L.solid == P.dose_solid) dose_record(P, L.e, L.p);
    L.mode = kNeedPrimary;
  }
  if (L.mode == kExit) {
    const DevSolid& s = T.solids[L.solid];
    L.p = l2g_pos(s.m, L.p);
    L.ray.d = l2g_dir(s.m, L.ray.d);
    L.mode = kHop;
  }
  if (kAcq && L.mode == kNeedPrimary) next_projection(P, L);
  if (L.mode == kNeedPrimary && L.todo == 0) {
    if (L.slot != 0xffffffffu) store_stream(P, L.slot, L.r);
    L.mode = kNeedSlot;
  }
  if (L.mode == kNeedSlot) {
    // one queue update per wave: the first lane that needs a slot takes a
    // contiguous range for all of them
    const uint64_t need = __ballot(1);
    const int leader = __ffsll((unsigned long long)need) - 1;
    const int lane = __lane_id();
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(P.queue, (uint32_t)__popcll(need));
    base = __shfl(base, leader, 64);
    const uint32_t my = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
    if (my < P.n_slots) {
      L.slot = my;
      load_stream(P, my, L.r);
      if (kAcq) {
        const uint4 lay = P.acq_layout[0];
        L.proj = 0;
        L.todo = slot_todo(lay.x, lay.y, lay.z, my);
        next_projection(P, L);
      } else {
        L.todo = slot_todo(P.n_batches, P.base, P.rem, my);
      }
      L.mode = kNeedPrimary;
    } else {
      L.slot = 0xffffffffu;
      L.mode = kDone;
    }
  }
  if (L.mode == kNeedPrimary) {
    if (L.todo == 0) {
      L.mode = kNeedSlot;  // empty slot: take another one on the next pass
    } else {
      --L.todo;
      sample_primary<kAcq, kTau>(P, T, kAcq ? P.acq_src[L.proj] : P.src_m, L.r, L.p, L.ray.d, L.e, L.eid);
      c.add(kCntPrimaries);
      L.scatter = 0;
      L.hop = 0;
      L.mode = kHop;
    }
  }
  if (L.mode == kHop) {
    if (L.hop >= GGH_MAX_HOPS) {
      L.mode = kNeedPrimary;
      return;
    }
    ++L.hop;
    c.add(kCntHops);
    const int k = find_and_project<kAcq>(P, T, L.p, L.ray.d, L.proj);
    if (k < 0) {
      L.mode = kNeedPrimary;
      return;
    }
    L.solid = k;
    const DevSolid& s = T.solids[k];
    if (s.type == kVoxel) {
      L.p = g2l_pos(s.m, L.p);
      L.ray.set(g2l_dir(s.m, L.ray.d));
      L.mode = kInVoxel;
      if (kTau) L.tau = draw_tau(L.r);
    } else if (!kSingle && s.type == kMesh) {
      // ggh_run launches the single-voxel-solid variant only on scenes
      // without meshes, which keeps the mesh loop (and even this test) out
      // of its register allocation
      track_mesh(P, T, s, L);
    } else {
      track_box<kAcq>(P, T, s, L, c);
    }
  }
}

// Block-level regrouping of photons between waves.  A lane's whole state
// moves through LDS to the position `pos` of the block, in two passes of 12
// words; each pass reads back exactly the words it wrote, so the second
// pass still writes this lane's own remaining words.  The permutation does
// not change any history: a photon carries its slot, its random stream and
// its counters of work left.
constexpr int kWaves = kBlock / 64;
constexpr size_t kXchgBytes = (size_t)kBlock * 12 * sizeof(uint32_t);

template <bool kTau>
__device__ inline void exchange_lane(Lane& L, uint4* s_x, uint32_t pos) {
  const uint4* src = s_x + (size_t)threadIdx.x * 3;
  uint4* dst = s_x + (size_t)pos * 3;
  dst[0] = make_uint4(L.r.x, L.r.y, L.r.z, L.r.w);
  dst[1] = make_uint4(L.r.c, __float_as_uint(L.p.x), __float_as_uint(L.p.y), __float_as_uint(L.p.z));
  dst[2] = make_uint4(__float_as_uint(L.ray.d.x), __float_as_uint(L.ray.d.y), __float_as_uint(L.ray.d.z),
                      __float_as_uint(L.ray.inv.x));
  __syncthreads();
  uint4 a = src[0], b = src[1], d = src[2];
  L.r = Rng{a.x, a.y, a.z, a.w, b.x};
  L.p = mk(__uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w));
  L.ray.d = mk(__uint_as_float(d.x), __uint_as_float(d.y), __uint_as_float(d.z));
  L.ray.inv.x = __uint_as_float(d.w);
  __syncthreads();
  dst[0] = make_uint4(__float_as_uint(L.ray.inv.y), __float_as_uint(L.ray.inv.z), __float_as_uint(L.e),
                      (uint32_t)L.eid);
  if (kTau) {
    // mode (3 bits), scatter (1 bit) and hop (<= GGH_MAX_HOPS) share a word to make room for tau
    dst[1] = make_uint4((uint32_t)L.mode | ((uint32_t)L.scatter << 3) | ((uint32_t)L.hop << 4), (uint32_t)L.solid,
                        (uint32_t)L.pend, L.todo);
    dst[2] = make_uint4(L.slot, L.proj, __float_as_uint(L.tau), 0u);
  } else {
    dst[1] = make_uint4((uint32_t)L.mode, (uint32_t)L.solid, (uint32_t)L.hop, (uint32_t)L.scatter);
    dst[2] = make_uint4((uint32_t)L.pend, L.todo, L.slot, L.proj);
  }
  __syncthreads();
  a = src[0];
  b = src[1];
  d = src[2];
  L.ray.inv.y = __uint_as_float(a.x);
  L.ray.inv.z = __uint_as_float(a.y);
  L.e = __uint_as_float(a.z);
  L.eid = (int32_t)a.w;
  if (kTau) {
    L.mode = (int32_t)(b.x & 7u);
    L.scatter = (int32_t)((b.x >> 3) & 1u);
    L.hop = (int32_t)(b.x >> 4);
    L.solid = (int32_t)b.y;
    L.pend = (int32_t)b.z;
    L.todo = b.w;
    L.slot = d.x;
    L.proj = d.y;
    L.tau = __uint_as_float(d.z);
  } else {
    L.mode = (int32_t)b.x;
    L.solid = (int32_t)b.y;
    L.hop = (int32_t)b.z;
    L.scatter = (int32_t)b.w;
    L.pend = (int32_t)d.x;
    L.todo = d.y;
    L.slot = d.z;
    L.proj = d.w;
  }
  __syncthreads();
}

// The kernel's RunParams argument, read through a pointer the compiler cannot
// follow past the empty asm: fields the navigation code reads through it are
// loaded (scalar loads from the kernel argument segment, RunParams being the
// first argument) where that code runs, instead of being hoisted to the
// kernel's start and kept in scalar registers across the voxel loop.
__device__ inline const RunParams& cold_params() {
  auto k = __builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(k));
  return *(const RunParams*)k;
}

__device__ inline uint32_t lds_load_relaxed(uint32_t* p) {
  return __builtin_amdgcn_readfirstlane(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
}

// NP: activated processes (-1: read at run time); kSingle: the scene has
// exactly one voxel solid, whose geometry the voxel step reads from the
// kernel arguments, and no meshed solid; kLDS: step tables staged in LDS; kTrack: photon
// fluence tracking may be on; kAcq: a CT acquisition (per-projection source
// and module placement); kRegroup: the scene has a voxel solid, and blocks
// work in rounds with regrouping (below).  Without a voxel solid every
// photon is navigated and tracked by its lane's state machine alone.
template <int NP, bool kSingle, bool kLDS, bool kTrack, bool kAcq, bool kRegroup, bool kTau = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWavesPerEU)))
void transport_kernel(RunParams P) {
  extern __shared__ float4 s_tab[];
  __shared__ unsigned long long s_counts[kBlock / 64][kNumCounters];
  if (kTau) {
    if (threadIdx.x < (kBlock / 64) * kNumCounters) (&s_counts[0][0])[threadIdx.x] = 0ull;
    __syncthreads();
  }
  Tabs T = {P.tab_sig, P.tab_rcp, P.ebins, P.solids, P.spectrum, P.cdf};
  if (kLDS) {
    // LDS layout (the host sizes it with the same arithmetic, lds_layout):
    // step tables, solids, energy bins, spectrum, cdf
    const int n = P.n_mat * P.nbins;
    float4* s_sig = s_tab;
    float4* s_rcp = s_tab + n;
    DevSolid* s_sol = reinterpret_cast<DevSolid*>(s_tab + 2 * n);
    float* s_eb = reinterpret_cast<float*>(s_sol + P.n_solids);
    const int ne = P.n_energy_arg + 1;  // spectrum and cdf entries, with the padding one
    float* s_spec = s_eb + P.nbins;
    float* s_cdf = s_spec + ne;
    for (int i = threadIdx.x; i < n; i += kBlock) {
      s_sig[i] = P.tab_sig[i];
      s_rcp[i] = P.tab_rcp[i];
    }
    const int sw = P.n_solids * (int)(sizeof(DevSolid) / 4);
    const uint32_t* gsol = reinterpret_cast<const uint32_t*>(P.solids);
    uint32_t* lsol = reinterpret_cast<uint32_t*>(s_sol);
    for (int i = threadIdx.x; i < sw; i += kBlock) lsol[i] = gsol[i];
    for (int i = threadIdx.x; i < P.nbins; i += kBlock) s_eb[i] = P.ebins[i];
    for (int i = threadIdx.x; i < ne; i += kBlock) {
      s_spec[i] = P.spectrum[i];
      s_cdf[i] = P.cdf[i];
    }
    __syncthreads();
    T.sig = s_sig;
    T.rcp = s_rcp;
    T.solids = s_sol;
    T.ebins = s_eb;
    T.spectrum = s_spec;
    T.cdf = s_cdf;
  }
  Lane L;
  L.r = Rng{0, 0, 0, 0, 0};
  L.p = mk(0.f, 0.f, 0.f);
  L.ray.d = mk(0.f, 0.f, 1.f);
  L.ray.inv = mk(0.f, 0.f, 1.f);
  L.e = 0.0f;
  L.eid = 0;
  L.mode = kNeedSlot;
  L.solid = -1;
  L.hop = 0;
  L.scatter = 0;
  L.pend = 0;
  L.todo = 0;
  L.slot = 0xffffffffu;
  L.proj = 0;
  L.tau = 0.0f;
  typename std::conditional<kTau, WaveCounters, LaneCounters>::type c;
  if constexpr (kTau) c.w = s_counts[threadIdx.x >> 6];
  else for (int k = 0; k < kNumCounters; ++k) c.v[k] = 0u;
  unsigned long long steps = 0;  // wave-uniform

  if (!kRegroup) {
    // no voxel solid, so no lane ever steps voxels: the loop ends when no
    // lane has navigation or tracking left
    for (;;) {
      const bool cold = L.mode != kInVoxel && L.mode != kDone;
      if (!__any(cold)) break;
      if (cold) cold_step<NP, kSingle, kAcq, kTau>(cold_params(), T, L, c);
    }
  } else {
    // The block works in rounds.  In each round every wave
    //   1. advances its stopped lanes (interaction, exit, detector, new
    //      primary or slot) by at most cold_passes navigation passes;
    //   2. makes voxel steps until every wave of the block is ready to stop:
    //      a wave is ready once all waves have finished step 1 and it has made
    //      round_steps steps (or has no lane inside a phantom); ready waves
    //      keep stepping until all are, so no wave idles at the barrier
    //      longer than about one step;
    //   3. regroups the block's photons: those inside a phantom first, then
    //      the stopped ones, then lanes with no work left.
    // Step 3 fills waves with photons in the same phase, so that voxel steps
    // run on full waves and the navigation code runs once for many photons.
    // Every wave passes the same barriers in every round, and every loop
    // above is bounded, so the block cannot deadlock.
    uint4* s_x = reinterpret_cast<uint4*>(reinterpret_cast<char*>(s_tab) + P.xchg_offset);
    __shared__ uint32_t s_cnt[kWaves][kTau ? 3 : 2];
    __shared__ uint32_t s_cold_done[2];
    __shared__ uint32_t s_ready[2];
    const int wave = (int)(threadIdx.x >> 6), lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    if (threadIdx.x < 2) {
      s_cold_done[threadIdx.x] = 0;
      s_ready[threadIdx.x] = 0;
    }
    __syncthreads();
    for (uint32_t round = 0;; ++round) {
      const uint32_t rb = round & 1u;
      for (int pass = 0; pass < P.cold_passes; ++pass) {
        const bool cold = L.mode != kInVoxel && L.mode != kDone;
        if (!__any(cold)) break;
        if (cold) cold_step<NP, kSingle, kAcq, kTau>(cold_params(), T, L, c);
      }
      if (lane == 0) atomicAdd(&s_cold_done[rb], 1u);
      bool ready = false;
      for (int k = 0;; k += kTau ? 2 : 1) {  // k counts voxel steps
        const uint64_t act = __ballot(L.mode == kInVoxel);
        if (!ready && (act == 0 || (k >= P.round_steps && lds_load_relaxed(&s_cold_done[rb]) == (uint32_t)kWaves))) {
          ready = true;
          if (lane == 0) atomicAdd(&s_ready[rb], 1u);
        }
        if (act == 0 || (ready && lds_load_relaxed(&s_ready[rb]) == (uint32_t)kWaves)) break;
        bool stepped = false, stepped2 = false;
        // the phantom's geometry: kernel arguments for the only voxel
        // solid, else the lane's solid in the (LDS-staged) solid table
        if (L.mode == kInVoxel) {
          const VoxGeom& G = kSingle ? P.vox0 : T.solids[L.solid].vox;
          stepped = kTau ? hot_step_tau<kTrack>(P, T, G, L) : hot_step<NP, kTrack>(P, T, G, L);
          // tau sampling makes two steps per pass: its step is short, so the
          // pass's ballots and round checks weigh more (measured faster; three
          // gained little more, four was slower)
          if (kTau && L.mode == kInVoxel) stepped2 = hot_step_tau<kTrack>(P, T, G, L);
        }
        steps += __popcll(__ballot(stepped)) + (kTau ? __popcll(__ballot(stepped2)) : 0);
        // pending interactions resolved in batches, so that scattered photons
        // keep stepping without waiting for the next round.  Tau sampling
        // leaves them to the navigation passes: its voxel loop runs faster
        // without the interaction code (measured), and its steps are cheap
        // enough that the waiting lanes cost less.
        if (!kTau && __popcll(__ballot(L.mode == kInteract)) >= P.interact_batch) {
          if (L.mode == kInteract) resolve_interaction<NP, kTau>(cold_params(), T, L, c);
        }
      }
      if constexpr (kTau) {
        // regrouping by event: rank 0 inside a phantom, 1 a pending
        // interaction, 2 other navigation (exit, cut, new primary or slot),
        // 3 no work left, so that the next navigation passes run one kind of
        // event per wave (measured faster for tau sampling, whose
        // interactions all wait for these passes; slower for the reference
        // sampling, which resolves most of them inside the voxel loop)
        const int key = L.mode == kInVoxel ? 0 : (L.mode == kInteract ? 1 : (L.mode != kDone ? 2 : 3));
        const uint64_t b0 = __ballot(key == 0), b1 = __ballot(key == 1), b2 = __ballot(key == 2);
        if (lane == 0) {
          s_cnt[wave][0] = (uint32_t)__popcll(b0);
          s_cnt[wave][1] = (uint32_t)__popcll(b1);
          s_cnt[wave][2] = (uint32_t)__popcll(b2);
        }
        __syncthreads();
        uint32_t n0 = 0, n1 = 0, n2 = 0, before0 = 0, before1 = 0, before2 = 0, before3 = 0;
        for (int w = 0; w < kWaves; ++w) {
          const uint32_t c0 = s_cnt[w][0], c1 = s_cnt[w][1], c2 = s_cnt[w][2];
          n0 += c0;
          n1 += c1;
          n2 += c2;
          if (w < wave) {
            before0 += c0;
            before1 += c1;
            before2 += c2;
            before3 += 64u - c0 - c1 - c2;
          }
        }
        if (n0 + n1 + n2 == 0) break;  // read from LDS after a barrier: the same for the whole block
        if (threadIdx.x == 0) {  // the next round's counters; nobody reads them before the barriers below
          s_cold_done[rb ^ 1u] = 0;
          s_ready[rb ^ 1u] = 0;
        }
        const uint64_t mine = key == 0 ? b0 : (key == 1 ? b1 : (key == 2 ? b2 : ~(b0 | b1 | b2)));
        const uint32_t pos = (key == 0 ? before0 : (key == 1 ? n0 + before1 : (key == 2 ? n0 + n1 + before2
                                                                                 : n0 + n1 + n2 + before3))) +
                             (uint32_t)__popcll(mine & below);
        exchange_lane<kTau>(L, s_x, pos);
      } else {
        // regrouping: rank 0 inside a phantom, 1 stopped, 2 no work left
        const int key = L.mode == kInVoxel ? 0 : (L.mode != kDone ? 1 : 2);
        const uint64_t b0 = __ballot(key == 0), b1 = __ballot(key == 1);
        if (lane == 0) {
          s_cnt[wave][0] = (uint32_t)__popcll(b0);
          s_cnt[wave][1] = (uint32_t)__popcll(b1);
        }
        __syncthreads();
        uint32_t n_hot = 0, n_cold = 0, before0 = 0, before1 = 0, before2 = 0;
        for (int w = 0; w < kWaves; ++w) {
          const uint32_t h = s_cnt[w][0], cl = s_cnt[w][1];
          n_hot += h;
          n_cold += cl;
          if (w < wave) {
            before0 += h;
            before1 += cl;
            before2 += 64u - h - cl;
          }
        }
        if (n_hot + n_cold == 0) break;  // read from LDS after a barrier: the same for the whole block
        if (threadIdx.x == 0) {  // the next round's counters; nobody reads them before the barriers below
          s_cold_done[rb ^ 1u] = 0;
          s_ready[rb ^ 1u] = 0;
        }
        const uint64_t mine = key == 0 ? b0 : (key == 1 ? b1 : ~(b0 | b1));
        const uint32_t pos = (key == 0 ? before0 : (key == 1 ? n_hot + before1 : n_hot + n_cold + before2)) +
                             (uint32_t)__popcll(mine & below);
        exchange_lane<kTau>(L, s_x, pos);
      }
    }
  }

  flush_counters(P, c);
  if (__lane_id() == 0 && steps) atomicAdd(&P.stats[1], steps);
}

}  // namespace

// ---------------------------------------------------------------------------
// dose finalisation (src/kernels/ComputeDoseGGEMSVoxelizedSolid.cl:53-121, and
// ComputeDoseGGEMSMeshSolid.cl:53-114 for a meshed solid: its one material)
// ---------------------------------------------------------------------------
namespace {

constexpr int kDoseBlock = 256;

struct DoseParams {
  const DoseTally* tally;
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

// field 0, 1 or 2 (edep, edep^2, hit) of every record into a plain array
__global__ __launch_bounds__(kDoseBlock) void tally_field_kernel(const DoseTally* t, int64_t n, int field, void* out) {
  for (int64_t i = blockIdx.x * (int64_t)kDoseBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kDoseBlock) {
    if (field == 0)
      static_cast<double*>(out)[i] = t[i].edep;
    else if (field == 1)
      static_cast<double*>(out)[i] = t[i].edep2;
    else
      static_cast<int32_t*>(out)[i] = t[i].hit;
  }
}

__global__ __launch_bounds__(kDoseBlock) void dose_kernel(DoseParams D) {
  const int64_t total = (int64_t)D.nd[0] * D.nd[1] * D.nd[2];
  const int slice = D.nd[0] * D.nd[1];
  const float kGy = 1.0e-12f;
  const float kWater = 1.0f * (6.241509704e+21f / 1.e3f);
  const float vol = D.dosel_size.x * D.dosel_size.y * D.dosel_size.z;
  for (int64_t gid = blockIdx.x * (int64_t)kDoseBlock + threadIdx.x; gid < total;
       gid += (int64_t)gridDim.x * kDoseBlock) {
    int z = (int)(gid / slice);
    int x = (int)((gid - (int64_t)z * slice) % D.nd[0]);
    int y = (int)((gid - (int64_t)z * slice) / D.nd[0]);
    float px = (float)x * D.dosel_size.x + (float)(D.nd[0] - 1) * (-0.5f * D.dosel_size.x);
    float py = (float)y * D.dosel_size.y + (float)(D.nd[1] - 1) * (-0.5f * D.dosel_size.y);
    float pz = (float)z * D.dosel_size.z + (float)(D.nd[2] - 1) * (-0.5f * D.dosel_size.z);
    int mat = 0;  // a meshed solid (no labels) holds one material
    if (D.labels) {
      int vx = min(max((int)((px - D.vlo.x) / D.vs.x), 0), D.nv[0] - 1);
      int vy = min(max((int)((py - D.vlo.y) / D.vs.y), 0), D.nv[1] - 1);
      int vz = min(max((int)((pz - D.vlo.z) / D.vs.z), 0), D.nv[2] - 1);
      mat = D.labels[vx + vy * D.nv[0] + vz * D.nv[0] * D.nv[1]];
    }
    float density = D.water_reference ? kWater : D.density[D.mat_offset + mat];
    const DoseTally& t = D.tally[gid];
    double e = t.edep;
    D.dose[gid] = density < D.min_density ? 0.0f : (float)((double)D.scale * e / (double)density / (double)vol / (double)kGy);
    if (D.unc) {
      int h = t.hit;
      if (h > 1 && e != 0.0) {
        double s2 = e * e;
        D.unc[gid] = (float)sqrt(((double)h * t.edep2 - s2) / ((double)(h - 1) * s2));
      } else {
        D.unc[gid] = 1.0f;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// ray-traced primary transmission (DRR)
// ---------------------------------------------------------------------------
// One thread per detector pixel: the segment from the focal point to the
// pixel centre crosses every voxel solid; the exact length inside each
// voxel (a 3D DDA in FP64) is added to the voxel's global material in LDS.
// The transmission is then sum_b w_b exp(-sum_m sigma_tot(m, b) L_m) over
// the energy bins b the source spectrum reaches, sigma_tot being the
// left-bin total of the activated processes the transport uses for its
// distances (GGEMSPhotonNavigator.hh:57-93).
constexpr int kDrrBlock = 128;
constexpr int kDrrMaxMat = 32;

struct DrrParams {
  const DevSolid* solids;
  int32_t n_solids;
  int32_t box;
  Mat34 src_m, box_m;
  const float4* sig;
  int32_t nbins, n_mat;
  const int32_t* w_bin;
  const float* w;
  int32_t n_w;
  const uint32_t* n_w_dev;  // weight count on the device (the allocation-free entry point), else null
  int32_t only;  // the one voxel solid present (an acquisition projection over several phantoms), or -1
  float* image;
};

__device__ void drr_voxel_solid(const DevSolid& s, f3 gs, f3 ge, double* L) {
  const VoxGeom& G = s.vox;
  const f3 ls = g2l_pos(s.m, gs), le = g2l_pos(s.m, ge);
  const double o[3] = {ls.x, ls.y, ls.z};
  const double d[3] = {(double)le.x - ls.x, (double)le.y - ls.y, (double)le.z - ls.z};
  const double lo[3] = {s.border.lo.x, s.border.lo.y, s.border.lo.z};
  const double hi[3] = {s.border.hi.x, s.border.hi.y, s.border.hi.z};
  const double vs[3] = {G.vs.x, G.vs.y, G.vs.z};
  const int n[3] = {G.nx, G.ny, G.nz};
  double t0 = 0.0, t1 = 1.0;
  for (int a = 0; a < 3; ++a) {
    if (fabs(d[a]) < 1e-300) {
      if (o[a] < lo[a] || o[a] > hi[a]) return;
    } else {
      const double ta = (lo[a] - o[a]) / d[a], tb = (hi[a] - o[a]) / d[a];
      t0 = fmax(t0, fmin(ta, tb));
      t1 = fmin(t1, fmax(ta, tb));
    }
  }
  if (!(t0 < t1)) return;
  const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  int i[3], step[3];
  double tmax[3], tdel[3];
  const double tm = 0.5 * (t0 + fmin(t1, t0 + 1e-9));
  for (int a = 0; a < 3; ++a) {
    const double p = o[a] + d[a] * tm;
    int k = (int)floor((p - lo[a]) / vs[a]);
    i[a] = k < 0 ? 0 : (k >= n[a] ? n[a] - 1 : k);
    if (d[a] > 0.0) {
      step[a] = 1;
      tmax[a] = (lo[a] + (i[a] + 1) * vs[a] - o[a]) / d[a];
      tdel[a] = vs[a] / d[a];
    } else if (d[a] < 0.0) {
      step[a] = -1;
      tmax[a] = (lo[a] + i[a] * vs[a] - o[a]) / d[a];
      tdel[a] = -vs[a] / d[a];
    } else {
      step[a] = 0;
      tmax[a] = INFINITY;
      tdel[a] = INFINITY;
    }
  }
  double t = t0;
  const int max_iter = n[0] + n[1] + n[2] + 3;
  for (int it = 0; it < max_iter; ++it) {
    const int a = tmax[0] < tmax[1] ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
    const double tn = fmin(tmax[a], t1);
    if (tn > t) {
      const int gm = G.mat_offset + G.labels[i[0] + i[1] * G.nx + i[2] * G.nxy];
      L[gm] += (tn - t) * len;
    }
    if (tn >= t1) break;
    t = tn;
    i[a] += step[a];
    if (i[a] < 0 || i[a] >= n[a]) break;
    tmax[a] += tdel[a];
  }
}

__global__ __launch_bounds__(kDrrBlock) void drr_kernel(DrrParams D) {
  __shared__ double sL[kDrrBlock * kDrrMaxMat];
  double* L = sL + threadIdx.x * kDrrMaxMat;
  for (int m = 0; m < kDrrMaxMat; ++m) L[m] = 0.0;
  const DevSolid& box = D.solids[D.box];
  const int npx = box.ne[0] * box.ne[1];
  const int px = blockIdx.x * kDrrBlock + threadIdx.x;
  if (px >= npx) return;
  const int ix = px % box.ne[0], iy = px / box.ne[0];
  const float esx = box.box_size.x / (float)box.ne[0], esy = box.box_size.y / (float)box.ne[1];
  const f3 pc = mk(box.border.lo.x + ((float)ix + 0.5f) * esx, box.border.lo.y + ((float)iy + 0.5f) * esy,
                   0.5f * (box.border.lo.z + box.border.hi.z));
  const f3 ge = l2g_pos(D.box_m, pc);
  const f3 gs = l2g_pos(D.src_m, mk(0.0f, 0.0f, 0.0f));
  for (int k = 0; k < D.n_solids; ++k)
    if (D.solids[k].type == kVoxel && (D.only < 0 || k == D.only)) drr_voxel_solid(D.solids[k], gs, ge, L);
  double t = 0.0;
  const int n_w = D.n_w_dev ? (int)*D.n_w_dev : D.n_w;
  for (int j = 0; j < n_w; ++j) {
    const int b = D.w_bin[j];
    double mu = 0.0;
    for (int m = 0; m < D.n_mat; ++m) {
      const float4 sg = D.sig[m * D.nbins + b];
      mu += ((double)sg.x + (double)sg.y + (double)sg.z) * L[m];
    }
    t += (double)D.w[j] * exp(-mu);
  }
  D.image[px] = (float)t;
}

// ---------------------------------------------------------------------------
// the ray tracer without allocations (ggh_drr_noalloc): the spectrum weights
// are built on the device into a file-scope staging array instead of buffers
// allocated per call.  Layout in 32-bit words: [0] the weight count, [64 ..
// 64 + kNoAllocMaxBins) the energy bins, then the FP32 weights.  Every call
// on one GPU shares the array.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxBins = 1024;
constexpr int kScratchBins = 64, kScratchWeights = kScratchBins + kNoAllocMaxBins;
constexpr int kScratchWords = kScratchWeights + kNoAllocMaxBins;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

// BinarySearch over the whole energy grid (GGEMSMathAlgorithms.hh:46-60)
__device__ inline int grid_bin(const float* a, int size, float key) {
  int low = 0, high = size - 1;
  while (high - low > 1) {
    const int mid = (high + low) / 2;
    if (a[mid] == key) return mid;
    if (a[mid] < key) low = mid;
    else high = mid;
  }
  return low;
}

// The probability w_b of each energy bin under the source's sampling, as the
// host builds it for ggh_drr: thread k sums bin k's share of every spectrum
// interval in interval order (the host's order of additions, so the FP64
// sums are identical), then one lane keeps the bins with weight, in order.
constexpr int kWeightsBlock = 256;
__global__ __launch_bounds__(kWeightsBlock) void drr_weights_kernel(const float* eb, int nb, const float* e, int n_e,
                                                                   const float* f, int n_f, int is_interp,
                                                                   uint32_t* stage) {
  __shared__ double W[kNoAllocMaxBins];
  for (int k = threadIdx.x; k < nb; k += kWeightsBlock) {
    double acc = 0.0;
    for (int idx = 0; idx + 1 < n_f && idx + 1 < n_e; ++idx) {
      const double p = (double)f[idx + 1] - (double)f[idx];
      if (!(p > 0.0)) continue;
      const double a = e[idx], b = e[idx + 1];
      if (!is_interp || !(b > a)) {
        if (grid_bin(eb, nb, e[idx]) == k) acc += p;
        continue;
      }
      if (k + 1 < nb) {
        const double lo = k == 0 ? -INFINITY : (double)eb[k];
        const double hi = k + 2 == nb ? INFINITY : (double)eb[k + 1];
        const double ov = (hi < b ? hi : b) - (a < lo ? lo : a);  // std::min(b, hi) - std::max(a, lo)
        if (ov > 0.0) acc += p * ov / (b - a);
      }
    }
    W[k] = acc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t* bins = reinterpret_cast<int32_t*>(stage + kScratchBins);
    float* w = reinterpret_cast<float*>(stage + kScratchWeights);
    uint32_t n = 0;
    for (int k = 0; k < nb; ++k)
      if (W[k] > 0.0) {
        bins[n] = k;
        w[n] = (float)W[k];
        ++n;
      }
    stage[0] = n;
  }
}

// ---------------------------------------------------------------------------
// test kernels
// ---------------------------------------------------------------------------
__global__ void walk_kernel(RunParams P, int solid, int n_rays, const float* pos, const float* dir, int max_steps,
                            int32_t* out, int32_t* count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rays) return;
  const DevSolid& s = P.solids[solid];
  f3 p = g2l_pos(s.m, mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
  Ray ray;
  ray.set(g2l_dir(s.m, mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2])));
  const VoxGeom G = s.vox;
  int n = 0;
  int32_t* o = out + (size_t)i * max_steps;
  while (walk_step(G, p, ray, o, n, max_steps)) {
    if (n > 16 * max_steps) break;
  }
  count[i] = n;
}

__global__ void mesh_hits_kernel(RunParams P, int solid, int n_rays, const float* pos, const float* dir,
                                 int32_t* count, float* nearest) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rays) return;
  const f3 p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  const f3 d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
  f3 q0 = mk(0.f, 0.f, 0.f), q1 = q0;
  count[i] = mesh_hits(P.solids[solid], p, d, q0, q1);
  float* o = nearest + 6 * (size_t)i;
  o[0] = q0.x; o[1] = q0.y; o[2] = q0.z;
  o[3] = q1.x; o[4] = q1.y; o[5] = q1.z;
}

// slot-queue reset as a kernel node: captured graphs hold no memset nodes
// row-major labels -> 128-byte bricks (brick_offset); voxels of the padding
// bricks past the volume's edges hold 0 and are never read
__global__ void brick_labels_kernel(const uint8_t* lin, uint8_t* br, int nx, int ny, int nz, int nbx, int nbxy,
                                    int64_t n_out) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < n_out; o += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = o >> 7;
    const int w = (int)(o & 127);
    const int bz = (int)(b / nbxy);
    const int r = (int)(b - (int64_t)bz * nbxy);
    const int by = r / nbx, bx = r - by * nbx;
    const int x = bx * kBrickX + (w & 7), y = by * kBrickY + ((w >> 3) & 3), z = bz * kBrickZ + (w >> 5);
    br[o] = (x < nx && y < ny && z < nz) ? lin[x + (int64_t)y * nx + (int64_t)z * nx * ny] : (uint8_t)0;
  }
}

__global__ void reset_queue_kernel(uint32_t* queue) {
  if (threadIdx.x == 0) *queue = 0u;
}

__global__ void primaries_kernel(RunParams P, int n, float* pos, float* dir, float* e) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  Rng r;
  load_stream(P, s, r);
  f3 gp, gd;
  float energy;
  int eid;
  const Tabs T = {P.tab_sig, P.tab_rcp, P.ebins, P.solids, P.spectrum, P.cdf};
  sample_primary<false, false>(P, T, P.src_m, r, gp, gd, energy, eid);
  pos[3 * s] = gp.x; pos[3 * s + 1] = gp.y; pos[3 * s + 2] = gp.z;
  dir[3 * s] = gd.x; dir[3 * s + 1] = gd.y; dir[3 * s + 2] = gd.z;
  e[s] = energy;
  store_stream(P, s, r);
}

__global__ void neglog_kernel(int n, const uint32_t* sums, float* fast, float* ref) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fast[i] = neg_log_uniform(sums[i]);
  ref[i] = -logf(uniform_from_sum(sums[i]));
}

__global__ void uniforms_kernel(RunParams P, int n_slots, int n, float* out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_slots) return;
  Rng r;
  load_stream(P, s, r);
  for (int k = 0; k < n; ++k) out[(size_t)s * n + k] = r.uniform();
  store_stream(P, s, r);
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
  int n_cu = 0;
  int max_lds = 0;  // bytes of LDS one block may allocate
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.0f;
  bool have_physics = false, have_source = false, have_dose = false, seeded = false;
  // physics
  int nbins = 0, n_mat = 0, n_proc = 0, proc[3] = {0, 0, 0};
  float factor = 0.0f;
  std::vector<float> density_host;
  std::vector<float> ebins_host;      // energy grid and step tables, for the ray tracer's weights
  std::vector<float4> sig_host;
  DevBuf<float> ebins, cs_r, cs_atom_r, atom_dens, pp, density;
  DevBuf<float4> tab_sig, tab_rcp;
  DevBuf<int32_t> n_elem, elem_idx;
  DevBuf<uint8_t> z;
  // solids
  std::vector<DevSolid> solids;
  std::vector<DevBuf<uint8_t>> labels, bricks;
  std::vector<DevBuf<int32_t>> hist, scat;
  std::vector<DevBuf<float4>> mesh_bvh, mesh_tris;
  std::vector<int> hist_elems;
  std::vector<int> max_material;  // largest global material index a solid can reach
  DevBuf<DevSolid> solids_dev;
  bool solids_dirty = true;
  // dosimetry
  ggh_dosimetry_desc dose_desc{};
  DevBuf<DoseTally> tally;
  DevBuf<int32_t> ptrack;
  // source
  Mat34 src_m{};
  float aperture = 0.0f;
  f3 focal{0, 0, 0};
  int n_energy_arg = 1, is_interp = 0;
  DevBuf<float> spectrum, cdf;
  std::vector<float> spectrum_host, cdf_host;
  // slots
  DevBuf<uint32_t> rng;
  DevBuf<uint32_t> queue;
  DevBuf<unsigned long long> stats;
  // launch shape
  int round_steps = 0;  // 0: 16 with tau sampling, 20 with the reference sampling (measured best)
  int cold_passes = 0;  // 0: 1 with tau sampling, 2 with the reference sampling (measured best)
  int interact_batch = 16;
  int blocks_per_cu = 0;  // 0: as many as fit
  bool force_general = false;
  int tau_mode = 0;  // voxel steps sample interactions by tau (one exponential per flight)
  unsigned last_grid = 0;
  int last_variant = -1;
  // CT acquisition
  int acq_n_proj = 0;
  uint32_t acq_slots = 0;
  DevBuf<Mat34> acq_src, acq_box;
  DevBuf<uint4> acq_layout;
  DevBuf<int32_t> acq_phantom;                      // per projection: the one non-box solid present, or -1
  std::vector<DevBuf<int32_t>> acq_hist, acq_scat;  // per solid: n_proj images (box solids)
  std::vector<Mat34> acq_src_host, acq_box_host;
  std::vector<int32_t> acq_phantom_host;
  DevBuf<DevSolid> acq_solids;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // this device's address of the file-scope staging array `scratch`
  uint32_t* scratch = nullptr;
};

namespace {

int check_ctx(ggh_context* c) {
  if (!c) return fail(-1, "null context");
  HIP_TRY(hipSetDevice(c->device));
  return 0;
}

// The device address of `scratch` on `device`, which must be the current
// device: each GPU holds its own copy of the array.  Resolved once per
// device; host code passes this address, never the array's name, to kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kMaxDevices] = {};

int scratch_address(int device, uint32_t** out) {
  if (device < 0 || device >= kMaxDevices) return fail(-2, "device index beyond the staging-array table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(-100 - (int)e, std::string("staging array address: ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch[device];
  return 0;
}

int sync_solids(ggh_context* c) {
  if (!c->solids_dirty) return 0;
  int rc = dev_upload(c->solids_dev, c->solids.data(), c->solids.size(), c->stream);
  if (rc) return rc;
  c->solids_dirty = false;
  return 0;
}

// RN(1/v) on the host, for div_rn; 1/0 = +inf keeps sigma == 0 a never-winning process
float host_rcp(float v) {
  volatile float one = 1.0f;
  return one / v;
}

VoxGeom make_vox(const ggh_voxel_solid_desc* d, int solid) {
  VoxGeom G;
  std::memset(&G, 0, sizeof G);
  G.lo = mk(d->obb.border_min[0], d->obb.border_min[1], d->obb.border_min[2]);
  G.vs = mk(d->voxel_size[0], d->voxel_size[1], d->voxel_size[2]);
  G.rvs = mk(host_rcp(G.vs.x), host_rcp(G.vs.y), host_rcp(G.vs.z));
  G.in_lo = mk(d->obb.border_min[0] + kTol, d->obb.border_min[1] + kTol, d->obb.border_min[2] + kTol);
  G.in_hi = mk(d->obb.border_max[0] - kTol, d->obb.border_max[1] - kTol, d->obb.border_max[2] - kTol);
  G.nx = d->n_voxels[0];
  G.ny = d->n_voxels[1];
  G.nz = d->n_voxels[2];
  G.nxy = d->n_voxels[0] * d->n_voxels[1];
  G.nbx = (G.nx + kBrickX - 1) / kBrickX;
  G.nbxy = G.nbx * ((G.ny + kBrickY - 1) / kBrickY);
  G.mat_offset = d->material_offset;
  G.solid = solid;
  return G;
}

int n_solids_of(const ggh_context* c, int32_t type) {
  int n = 0;
  for (const DevSolid& s : c->solids) n += s.type == type;
  return n;
}

// the guess of energy_bin for a grid (any values keep energy_bin exact)
void energy_bin_guess(const std::vector<float>& eb, float& l2min, float& inv) {
  l2min = 0.0f;
  inv = 0.0f;
  if (eb.size() < 2 || !(eb.front() > 0.0f) || !(eb.back() > eb.front())) return;
  l2min = std::log2(eb.front());
  inv = (float)(eb.size() - 1) / (std::log2(eb.back()) - l2min);
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
  energy_bin_guess(c->ebins_host, P.ebin_l2min, P.ebin_inv);
  P.tab_sig = c->tab_sig.p;
  P.tab_rcp = c->tab_rcp.p;
  P.cs_r = c->cs_r.p;
  P.cs_atom_r = c->cs_atom_r.p;
  P.n_elem = c->n_elem.p;
  P.elem_idx = c->elem_idx.p;
  P.z = c->z.p;
  P.atom_dens = c->atom_dens.p;
  P.pp = c->pp.p;
  P.factor = c->factor;
  P.src_m = c->src_m;
  P.aperture = c->aperture;
  P.new_aperture = 1.0 - std::cos((double)c->aperture);
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
    P.tally = c->tally.p;
    P.ptrack = c->ptrack.p;
  }
  P.rng = c->rng.p;
  P.rng_stride = GGH_MAX_SLOTS;
  P.queue = c->queue.p;
  P.stats = c->stats.p;
  P.round_steps = c->round_steps ? c->round_steps : (c->tau_mode ? 16 : 20);
  P.cold_passes = c->cold_passes ? c->cold_passes : (c->tau_mode ? 1 : 2);
  P.interact_batch = c->interact_batch;
  P.xchg_offset = 0;
  for (const DevSolid& s : c->solids)
    if (s.type == kVoxel) {
      P.vox0 = s.vox;
      break;
    }
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

void drop_graph(ggh_context* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

void drop_acquisition(ggh_context* c) {
  drop_graph(c);
  c->acq_src.release();
  c->acq_box.release();
  c->acq_layout.release();
  c->acq_phantom.release();
  c->acq_phantom_host.clear();
  c->acq_solids.release();
  for (auto& b : c->acq_hist) b.release();
  for (auto& b : c->acq_scat) b.release();
  c->acq_hist.clear();
  c->acq_scat.clear();
  c->acq_n_proj = 0;
  c->acq_slots = 0;
  c->acq_src_host.clear();
  c->acq_box_host.clear();
}

// bytes of the LDS image the transport kernel stages per block: step tables,
// solids, energy bins, spectrum and cdf (n_energy_arg + 1 entries each, as
// ggh_set_xray_source stores them)
size_t tables_lds_bytes(const ggh_context* c) {
  return (size_t)2 * c->n_mat * c->nbins * sizeof(float4) + c->solids.size() * sizeof(DevSolid) +
         (size_t)c->nbins * sizeof(float) + (size_t)2 * (c->n_energy_arg + 1) * sizeof(float);
}

// everything a transport launch needs, checked; uploads the solids
int check_ready(ggh_context* c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_physics) return fail(-3, "physics tables not set");
  if (!c->have_source) return fail(-3, "source not set");
  if (!c->seeded) return fail(-3, "particle slots not seeded");
  if (c->solids.empty()) return fail(-3, "no solid registered");
  for (size_t i = 0; i < c->solids.size(); ++i)
    if (c->solids[i].mat_offset < 0 || c->max_material[i] >= c->n_mat)
      return fail(-3, "solid labels reach a material without physics tables");
  return sync_solids(c);
}

// GGEMSSource::OrganizeParticlesInBatch (src/sources/GGEMSSource.cc:403-425):
// ceil(float(n) / slots) batches of n / batches photons, the first n % batches
// batches one more; slots = the largest batch.  Above 2^24 primaries the
// float quotient can round n down by enough to give one batch too few (2^24 + 1
// primaries: 16 batches, the first of 2^20 + 1 photons), whose largest batch
// would overrun the particle buffer; such n take the exact ceiling instead.
int batch_layout(uint64_t n, uint32_t& n_batches, uint32_t& base, uint32_t& rem, uint32_t& slots) {
  // the bound under which no 32-bit lane run counter can wrap (LaneCounters)
  if (n > (1ull << 48)) return fail(-2, "a run transports at most 2^48 primaries");
  uint64_t nb = (uint64_t)std::ceil((float)n / (float)GGH_MAX_SLOTS);
  if (nb == 0) nb = 1;
  if (nb * GGH_MAX_SLOTS < n) nb = (n + GGH_MAX_SLOTS - 1) / GGH_MAX_SLOTS;
  const uint64_t b = nb == 1 ? n : n / nb;
  const uint64_t r = nb == 1 ? 0 : n % nb;
  const uint64_t sl = b + (r ? 1 : 0);
  if (sl > GGH_MAX_SLOTS) return fail(-3, "batch larger than the slot buffer");
  if (nb > 0xffffffffull) return fail(-3, "too many batches");
  n_batches = (uint32_t)nb;
  base = (uint32_t)b;
  rem = (uint32_t)r;
  slots = (uint32_t)sl;
  return 0;
}

struct Launch {
  void (*fn)(RunParams);
  unsigned grid;
  size_t shmem;
  uint32_t xchg_offset;  // RunParams::xchg_offset of this launch
};

// kernel variant and persistent grid: the specialised variant (three
// activated processes, one voxel solid and no mesh, tables in LDS, no fluence
// tracking) or the general one, each with or without acquisition placement;
// block regrouping whenever the scene has a voxel solid to step through
int pick_launch(ggh_context* c, uint64_t slots, bool acq, Launch& L) {
  const bool single = n_solids_of(c, kVoxel) == 1 && n_solids_of(c, kMesh) == 0;
  const bool regroup = n_solids_of(c, kVoxel) > 0;
  // dynamic LDS: the staged tables (when they fit), then the lane-state
  // exchange area of the block's regrouping
  const size_t xchg = regroup ? kXchgBytes : 0;
  const size_t lds = (tables_lds_bytes(c) + 15) & ~(size_t)15;
  const bool use_lds = lds <= (size_t)kLdsBudget && lds + xchg <= (size_t)c->max_lds;
  if (xchg > (size_t)c->max_lds) return fail(-3, "the block's lane-state exchange area does not fit in LDS");
  // three activated processes, tables in LDS, no fluence tracking: the
  // specialised variant for one voxel solid, else its several-solid form
  const bool three = c->n_proc == 3 && use_lds && !c->ptrack.p && !c->force_general;
  const bool special = three && single;
  L.xchg_offset = use_lds ? (uint32_t)lds : 0u;
  L.shmem = L.xchg_offset + xchg;
  // tau sampling changes only the voxel step: scenes without a voxel solid
  // run the plain variants
  const bool tau = c->tau_mode != 0 && regroup;
  if (special && tau)
    L.fn = acq ? transport_kernel<3, true, true, false, true, true, true>
               : transport_kernel<3, true, true, false, false, true, true>;
  else if (special)
    L.fn = acq ? transport_kernel<3, true, true, false, true, true>
               : transport_kernel<3, true, true, false, false, true>;
  else if (three && regroup && tau)
    L.fn = acq ? transport_kernel<3, false, true, false, true, true, true>
               : transport_kernel<3, false, true, false, false, true, true>;
  else if (three && regroup)
    L.fn = acq ? transport_kernel<3, false, true, false, true, true>
               : transport_kernel<3, false, true, false, false, true>;
  else if (tau)
    L.fn = acq ? (use_lds ? transport_kernel<-1, false, true, true, true, true, true>
                          : transport_kernel<-1, false, false, true, true, true, true>)
               : (use_lds ? transport_kernel<-1, false, true, true, false, true, true>
                          : transport_kernel<-1, false, false, true, false, true, true>);
  else if (acq)
    L.fn = use_lds ? (regroup ? transport_kernel<-1, false, true, true, true, true>
                              : transport_kernel<-1, false, true, true, true, false>)
                   : (regroup ? transport_kernel<-1, false, false, true, true, true>
                              : transport_kernel<-1, false, false, true, true, false>);
  else
    L.fn = use_lds ? (regroup ? transport_kernel<-1, false, true, true, false, true>
                              : transport_kernel<-1, false, true, true, false, false>)
                   : (regroup ? transport_kernel<-1, false, false, true, false, true>
                              : transport_kernel<-1, false, false, true, false, false>);
  c->last_variant = special ? 0 : (three && regroup ? 2 : 1);
  int per_cu = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(L.fn), kBlock,
                                                       L.shmem));
  if (per_cu < 1) return fail(-3, "transport kernel does not fit on a compute unit");
  if (c->blocks_per_cu > 0) per_cu = std::min(per_cu, c->blocks_per_cu);
  // persistent blocks: no more than the slots can keep busy
  const uint64_t max_blocks = (slots + kBlock - 1) / kBlock;
  L.grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)per_cu * c->n_cu, max_blocks));
  c->last_grid = L.grid;
  return 0;
}

// parameters and launch of the set acquisition: the solids with their
// image stacks in place of the single histograms
int acquisition_launch(ggh_context* c, RunParams& P, Launch& L) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->acq_n_proj) return fail(-3, "no acquisition set");
  if (c->acq_hist.size() != c->solids.size()) return fail(-3, "solids changed since the acquisition was set");
  std::vector<DevSolid> sol = c->solids;
  for (size_t k = 0; k < sol.size(); ++k)
    if (sol[k].type == kBoxSolid) {
      sol[k].histogram = c->acq_hist[k].p;
      sol[k].scatter = c->acq_scat[k].p;
    }
  // the buffer is reused, not reallocated: a captured graph holds its address
  if (c->acq_solids.n != sol.size() && (rc = dev_alloc(c->acq_solids, sol.size()))) return rc;
  HIP_TRY(hipMemcpyAsync(c->acq_solids.p, sol.data(), sol.size() * sizeof(DevSolid), hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  P = make_params(c);
  P.solids = c->acq_solids.p;
  P.acq_src = c->acq_src.p;
  P.acq_box = c->acq_box.p;
  P.acq_layout = c->acq_layout.p;
  P.acq_phantom = c->acq_phantom.p;
  P.n_proj = c->acq_n_proj;
  P.n_box = n_solids_of(c, kBoxSolid);
  P.n_slots = c->acq_slots;
  if ((rc = pick_launch(c, c->acq_slots, true, L))) return rc;
  P.xchg_offset = L.xchg_offset;
  return 0;
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
  hipError_t e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    delete c;
    return fail(-100 - (int)e, std::string("device query / stream / event creation: ") + hipGetErrorString(e));
  }
  int rc = scratch_address(device, &c->scratch);
  if (!rc) rc = dev_alloc(c->stats, 8);
  if (!rc) rc = dev_alloc(c->rng, (size_t)5 * GGH_MAX_SLOTS);
  if (!rc) rc = dev_alloc(c->queue, 1);
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
  c->ebins.release(); c->cs_r.release(); c->cs_atom_r.release(); c->atom_dens.release(); c->pp.release();
  c->density.release(); c->tab_sig.release(); c->tab_rcp.release(); c->n_elem.release(); c->elem_idx.release();
  c->z.release();
  for (auto& b : c->labels) b.release();
  for (auto& b : c->bricks) b.release();
  for (auto& b : c->hist) b.release();
  for (auto& b : c->scat) b.release();
  for (auto& b : c->mesh_bvh) b.release();
  for (auto& b : c->mesh_tris) b.release();
  c->solids_dev.release();
  drop_acquisition(c);
  c->tally.release(); c->ptrack.release();
  c->spectrum.release(); c->cdf.release(); c->rng.release(); c->queue.release(); c->stats.release();
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
  for (int i = 0; i < d->n_processes; ++i) {
    if (d->process_id[i] < 0 || d->process_id[i] > 2) return fail(-2, "invalid process id");
    for (int j = 0; j < i; ++j)
      if (d->process_id[j] == d->process_id[i]) return fail(-2, "process activated twice");
  }
  const int nb = d->nbins, nm = d->n_materials;
  for (int m = 0; m < nm; ++m)
    if (d->element_index[m] < 0 || d->n_elements[m] < 1 ||
        d->element_index[m] + d->n_elements[m] > d->total_elements)
      return fail(-2, "invalid element table");
  for (int k = 0; k < d->total_elements; ++k)
    if (d->atomic_number[k] < 1 || d->atomic_number[k] > 100) return fail(-2, "atomic number out of range");
  // step tables: cross sections in activation order, their reciprocals, the cut
  std::vector<float4> sig((size_t)nb * nm), rcp((size_t)nb * nm);
  std::vector<float> cs_r((size_t)nb * nm);
  for (int m = 0; m < nm; ++m)
    for (int i = 0; i < nb; ++i) {
      float s[3] = {0.f, 0.f, 0.f}, y[3] = {0.f, 0.f, 0.f};
      for (int k = 0; k < d->n_processes; ++k) {
        s[k] = d->cross_sections[(size_t)d->process_id[k] * nm * nb + (size_t)m * nb + i];
        y[k] = host_rcp(s[k]);
      }
      const size_t at = (size_t)m * nb + i;
      sig[at] = make_float4(s[0], s[1], s[2], d->photon_cut[m]);
      rcp[at] = make_float4(y[0], y[1], y[2], host_rcp((s[0] + s[1]) + s[2]));
      cs_r[at] = d->cross_sections[(size_t)2 * nm * nb + (size_t)m * nb + i];
    }
  c->nbins = nb;
  c->n_mat = nm;
  c->n_proc = d->n_processes;
  for (int i = 0; i < 3; ++i) c->proc[i] = i < d->n_processes ? d->process_id[i] : 0;
  // the captured acquisition holds the table pointers reallocated below
  drop_graph(c);
  c->factor = d->rayleigh_factor;
  c->density_host.assign(d->density, d->density + nm);
  c->ebins_host.assign(d->energy_bins, d->energy_bins + nb);
  c->sig_host = sig;
  if ((rc = dev_upload(c->ebins, d->energy_bins, (size_t)nb, c->stream))) return rc;
  if ((rc = dev_upload(c->tab_sig, sig.data(), sig.size(), c->stream))) return rc;
  if ((rc = dev_upload(c->tab_rcp, rcp.data(), rcp.size(), c->stream))) return rc;
  if ((rc = dev_upload(c->cs_r, cs_r.data(), cs_r.size(), c->stream))) return rc;
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
  for (int i = 0; i < 3; ++i)
    if (!(d->voxel_size[i] > 0.0f) || !std::isfinite(d->voxel_size[i])) return fail(-2, "invalid voxel size");
  DevSolid s;
  std::memset(&s, 0, sizeof s);
  s.m = to_mat34(d->obb.matrix);
  s.border.lo = mk(d->obb.border_min[0], d->obb.border_min[1], d->obb.border_min[2]);
  s.border.hi = mk(d->obb.border_max[0], d->obb.border_max[1], d->obb.border_max[2]);
  s.type = kVoxel;
  s.box_index = -1;
  s.mat_offset = d->material_offset;
  int max_label = 0;
  for (size_t i = 0; i < n; ++i) max_label = d->labels[i] > max_label ? d->labels[i] : max_label;
  DevBuf<uint8_t> lab, br;
  s.vox = make_vox(d, (int)c->solids.size());
  const size_t nb = (size_t)s.vox.nbxy * (size_t)((s.vox.nz + kBrickZ - 1) / kBrickZ) * 128;
  if (nb > (size_t)UINT32_MAX) return fail(-2, "label volume too large for 32-bit brick offsets");
  if ((rc = dev_upload(lab, d->labels, n, c->stream))) return rc;
  if ((rc = dev_alloc(br, nb))) {
    lab.release();
    return rc;
  }
  {
    const unsigned grid = (unsigned)std::min<size_t>((nb + 255) / 256, 65536);
    hipLaunchKernelGGL(brick_labels_kernel, dim3(grid), dim3(256), 0, c->stream, (const uint8_t*)lab.p, br.p,
                       s.vox.nx, s.vox.ny, s.vox.nz, s.vox.nbx, s.vox.nbxy, (int64_t)nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      lab.release();
      br.release();
      return fail(-100 - (int)e, std::string("voxel solid: ") + hipGetErrorString(e));
    }
  }
  s.vox.labels = lab.p;
  s.vox.bricks = br.p;
  c->labels.push_back(lab);
  c->bricks.push_back(br);
  c->hist.push_back(DevBuf<int32_t>());
  c->scat.push_back(DevBuf<int32_t>());
  c->mesh_bvh.push_back(DevBuf<float4>());
  c->mesh_tris.push_back(DevBuf<float4>());
  c->hist_elems.push_back(0);
  drop_graph(c);  // a captured acquisition was recorded for the previous solid list
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
  s.box_index = n_solids_of(c, kBoxSolid);
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
  c->bricks.push_back(DevBuf<uint8_t>());
  c->hist.push_back(h);
  c->scat.push_back(sc);
  c->mesh_bvh.push_back(DevBuf<float4>());
  c->mesh_tris.push_back(DevBuf<float4>());
  c->hist_elems.push_back(ne);
  drop_graph(c);  // a captured acquisition was recorded for the previous solid list
  c->solids.push_back(s);
  c->max_material.push_back(d->material_offset);
  c->solids_dirty = true;
  return (int)c->solids.size() - 1;
}

}  // extern "C"

namespace {

// BVH over a triangle mesh, emitted depth first: a node is followed by its
// first child; `skip` is the index after its subtree.  Leaves hold up to 4
// consecutive triangles.  Boxes are widened by 1e-3 mm plus 1e-5 of their
// coordinates so that rounding in the device's slab test cannot prune a box
// whose triangles the exact intersection test would hit.
// Triangles per BVH leaf.  One measured fastest on the mesh-CT scene (with
// median splits: 164 M primaries/s, against 157 with two, 152 with four and
// 137 with eight): a Moller-Trumbore test costs several node tests, so it
// pays to run it only for triangles whose own box the ray crosses.  The
// binned surface-area splits below took the scene from 164 to 185 M/s.
constexpr int kBvhLeaf = 1;

struct BvhBuild {
  const float* tri;
  std::vector<int> order;
  std::vector<float4> nodes;
  void bounds(int b, int e, float lo[3], float hi[3]) const {
    for (int a = 0; a < 3; ++a) {
      lo[a] = INFINITY;
      hi[a] = -INFINITY;
    }
    for (int k = b; k < e; ++k)
      for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 3; ++a) {
          float x = tri[9 * order[k] + 3 * v + a];
          lo[a] = std::min(lo[a], x);
          hi[a] = std::max(hi[a], x);
        }
  }
  float centroid(int t, int a) const { return tri[9 * t + a] + tri[9 * t + 3 + a] + tri[9 * t + 6 + a]; }
  int build(int b, int e) {
    const int me = (int)nodes.size() / 2;
    nodes.push_back(make_float4(0, 0, 0, 0));
    nodes.push_back(make_float4(0, 0, 0, 0));
    float lo[3], hi[3];
    bounds(b, e, lo, hi);
    for (int a = 0; a < 3; ++a) {
      const float w = 1e-3f + 1e-5f * std::max(std::fabs(lo[a]), std::fabs(hi[a]));
      lo[a] -= w;
      hi[a] += w;
    }
    int info = 0;
    if (e - b <= kBvhLeaf) {
      info = (b << 4) | (e - b);
    } else {
      float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int k = b; k < e; ++k)
        for (int a = 0; a < 3; ++a) {
          clo[a] = std::min(clo[a], centroid(order[k], a));
          chi[a] = std::max(chi[a], centroid(order[k], a));
        }
      // binned surface-area heuristic over the three axes: the walk visits
      // every node whose box the ray line crosses, so the expected visits are
      // proportional to the children's surface areas
      constexpr int kBins = 16;
      int best_axis = -1, best_bin = 0;
      double best_cost = INFINITY;
      for (int a = 0; a < 3; ++a) {
        const float ext = chi[a] - clo[a];
        if (!(ext > 0.0f)) continue;
        int cnt[kBins] = {0};
        float blo[kBins][3], bhi[kBins][3];
        for (int k = 0; k < kBins; ++k)
          for (int q = 0; q < 3; ++q) {
            blo[k][q] = INFINITY;
            bhi[k][q] = -INFINITY;
          }
        for (int k = b; k < e; ++k) {
          const int t = order[k];
          int bin = (int)((centroid(t, a) - clo[a]) / ext * kBins);
          bin = bin < 0 ? 0 : (bin >= kBins ? kBins - 1 : bin);
          ++cnt[bin];
          for (int v = 0; v < 3; ++v)
            for (int q = 0; q < 3; ++q) {
              const float x = tri[9 * t + 3 * v + q];
              blo[bin][q] = std::min(blo[bin][q], x);
              bhi[bin][q] = std::max(bhi[bin][q], x);
            }
        }
        auto area = [](const float* lo, const float* hi) {
          const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
          return dx * dy + dy * dz + dz * dx;
        };
        // prefix (left) and suffix (right) bounds
        float llo[kBins][3], lhi[kBins][3];
        int lcnt[kBins];
        for (int k = 0; k < kBins; ++k) {
          for (int q = 0; q < 3; ++q) {
            llo[k][q] = k ? std::min(llo[k - 1][q], blo[k][q]) : blo[k][q];
            lhi[k][q] = k ? std::max(lhi[k - 1][q], bhi[k][q]) : bhi[k][q];
          }
          lcnt[k] = (k ? lcnt[k - 1] : 0) + cnt[k];
        }
        float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        int rcnt = 0;
        for (int k = kBins - 1; k >= 1; --k) {  // split between bin k - 1 and bin k
          for (int q = 0; q < 3; ++q) {
            rlo[q] = std::min(rlo[q], blo[k][q]);
            rhi[q] = std::max(rhi[q], bhi[k][q]);
          }
          rcnt += cnt[k];
          const int nl = lcnt[k - 1];
          if (nl == 0 || rcnt == 0) continue;
          const double cost = nl * area(llo[k - 1], lhi[k - 1]) + rcnt * area(rlo, rhi);
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = a;
            best_bin = k;
          }
        }
      }
      int m = (b + e) / 2;
      if (best_axis >= 0) {
        const int a = best_axis;
        const float ext = chi[a] - clo[a];
        auto mid = std::partition(order.begin() + b, order.begin() + e, [&](int t) {
          int bin = (int)((centroid(t, a) - clo[a]) / ext * kBins);
          bin = bin < 0 ? 0 : (bin >= kBins ? kBins - 1 : bin);
          return bin < best_bin;
        });
        m = (int)(mid - order.begin());
      }
      if (m <= b || m >= e) {  // degenerate: fall back to the median on the longest axis
        int axis = 0;
        for (int a = 1; a < 3; ++a)
          if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
        m = (b + e) / 2;
        std::nth_element(order.begin() + b, order.begin() + m, order.begin() + e,
                         [&](int x, int y) { return centroid(x, axis) < centroid(y, axis); });
      }
      build(b, m);
      build(m, e);
    }
    const int skip = (int)nodes.size() / 2;
    float sk, in;
    std::memcpy(&sk, &skip, 4);
    std::memcpy(&in, &info, 4);
    nodes[2 * me] = make_float4(lo[0], lo[1], lo[2], sk);
    nodes[2 * me + 1] = make_float4(hi[0], hi[1], hi[2], in);
    return me;
  }
};

}  // namespace

extern "C" {

int ggh_add_mesh_solid(ggh_context* c, const ggh_mesh_solid_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || !d->triangles || d->n_triangles <= 0 || d->n_triangles >= (1 << 26))
    return fail(-2, "invalid mesh solid description");
  if (c->solids.size() >= GGH_MAX_SOLIDS) return fail(-2, "too many solids");
  const int n = d->n_triangles;
  BvhBuild B;
  B.tri = d->triangles;
  B.order.resize(n);
  for (int i = 0; i < n; ++i) B.order[i] = i;
  B.build(0, n);
  std::vector<float4> tris((size_t)3 * n);
  for (int k = 0; k < n; ++k) {
    const float* t = d->triangles + 9 * (size_t)B.order[k];
    tris[3 * k] = make_float4(t[0], t[1], t[2], 0.f);
    tris[3 * k + 1] = make_float4(t[3] - t[0], t[4] - t[1], t[5] - t[2], 0.f);
    tris[3 * k + 2] = make_float4(t[6] - t[0], t[7] - t[1], t[8] - t[2], 0.f);
  }
  DevBuf<float4> bvh, tb;
  if ((rc = dev_upload(bvh, B.nodes.data(), B.nodes.size(), c->stream))) return rc;
  if ((rc = dev_upload(tb, tris.data(), tris.size(), c->stream))) {
    bvh.release();
    return rc;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  DevSolid s;
  std::memset(&s, 0, sizeof s);
  s.m = to_mat34(d->obb.matrix);
  s.border.lo = mk(d->obb.border_min[0], d->obb.border_min[1], d->obb.border_min[2]);
  s.border.hi = mk(d->obb.border_max[0], d->obb.border_max[1], d->obb.border_max[2]);
  s.type = kMesh;
  s.box_index = -1;
  s.mat_offset = d->material_offset;
  s.bvh = bvh.p;
  s.tris = tb.p;
  s.n_bvh = (int32_t)(B.nodes.size() / 2);
  c->labels.push_back(DevBuf<uint8_t>());
  c->bricks.push_back(DevBuf<uint8_t>());
  c->hist.push_back(DevBuf<int32_t>());
  c->scat.push_back(DevBuf<int32_t>());
  c->mesh_bvh.push_back(bvh);
  c->mesh_tris.push_back(tb);
  c->hist_elems.push_back(0);
  drop_graph(c);  // a captured acquisition was recorded for the previous solid list
  c->solids.push_back(s);
  c->max_material.push_back(d->material_offset);
  c->solids_dirty = true;
  return (int)c->solids.size() - 1;
}

int ggh_debug_mesh_hits(ggh_context* c, int solid, int n_rays, const float* pos, const float* dir, int32_t* count,
                        float* nearest) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (solid < 0 || solid >= (int)c->solids.size() || c->solids[solid].type != kMesh)
    return fail(-2, "not a meshed solid");
  if (n_rays <= 0) return fail(-2, "invalid ray count");
  if ((rc = sync_solids(c))) return rc;
  RunParams P = make_params(c);
  DevBuf<float> dp, dd, dn;
  DevBuf<int32_t> dc;
  if ((rc = dev_upload(dp, pos, (size_t)3 * n_rays, c->stream)) || (rc = dev_upload(dd, dir, (size_t)3 * n_rays, c->stream)) ||
      (rc = dev_alloc(dc, (size_t)n_rays)) || (rc = dev_alloc(dn, (size_t)6 * n_rays)))
    return rc;
  hipLaunchKernelGGL(mesh_hits_kernel, dim3((n_rays + 63) / 64), dim3(64), 0, c->stream, P, solid, n_rays, dp.p, dd.p,
                     dc.p, dn.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(count, dc.p, (size_t)n_rays * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(nearest, dn.p, (size_t)6 * n_rays * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  dp.release(); dd.release(); dn.release(); dc.release();
  return 0;
}

int ggh_set_dosimetry(ggh_context* c, const ggh_dosimetry_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || d->solid < 0 || d->solid >= (int)c->solids.size() || c->solids[d->solid].type == kBoxSolid)
    return fail(-2, "dosimetry needs a voxel or meshed solid");
  size_t n = (size_t)d->n_dosels[0] * d->n_dosels[1] * d->n_dosels[2];
  if (d->n_dosels[0] <= 0 || d->n_dosels[1] <= 0 || d->n_dosels[2] <= 0 || n > (size_t)INT32_MAX)
    return fail(-2, "invalid dosel grid");
  drop_graph(c);  // the captured acquisition holds the tally pointers reallocated below
  c->dose_desc = *d;
  if ((rc = dev_alloc(c->tally, n))) return rc;
  c->ptrack.release();
  if (d->photon_tracking && (rc = dev_alloc(c->ptrack, n))) return rc;
  c->have_dose = true;
  return ggh_reset_tallies(c);
}

int ggh_set_xray_source(ggh_context* c, const ggh_xray_source_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || d->n_energy < 1 || !d->energy || !d->cdf) return fail(-2, "invalid source description");
  drop_graph(c);  // the captured acquisition holds the spectrum and cdf reallocated below
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
  c->spectrum_host = e;
  c->cdf_host = f;
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
  if (c->tally.p) HIP_TRY(hipMemsetAsync(c->tally.p, 0, c->tally.n * sizeof(DoseTally), c->stream));
  if (c->ptrack.p) HIP_TRY(hipMemsetAsync(c->ptrack.p, 0, c->ptrack.n * sizeof(int32_t), c->stream));
  for (auto* v : {&c->acq_hist, &c->acq_scat})
    for (auto& b : *v)
      if (b.p) HIP_TRY(hipMemsetAsync(b.p, 0, b.n * sizeof(int32_t), c->stream));
  HIP_TRY(hipMemsetAsync(c->stats.p, 0, 8 * sizeof(unsigned long long), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int ggh_set_sampling(ggh_context* c, int32_t mode) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (mode != GGH_SAMPLING_REFERENCE && mode != GGH_SAMPLING_TAU) return fail(-2, "unknown sampling mode");
  drop_graph(c);  // a captured acquisition holds the kernel of the previous mode
  c->tau_mode = mode == GGH_SAMPLING_TAU;
  return 0;
}

int ggh_set_launch(ggh_context* c, const ggh_launch_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!d || d->round_steps < 0 || d->round_steps > 1024 || d->cold_passes < 0 || d->cold_passes > 16 ||
      d->interact_batch < 0 || d->interact_batch > 64 || d->blocks_per_cu < 0 || d->blocks_per_cu > 64)
    return fail(-2, "invalid launch parameters");
  c->round_steps = d->round_steps;
  c->cold_passes = d->cold_passes;
  c->interact_batch = d->interact_batch ? d->interact_batch : 16;
  c->blocks_per_cu = d->blocks_per_cu;
  c->force_general = d->general_kernel != 0;
  return 0;
}

int ggh_run(ggh_context* c, uint64_t n) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n == 0) return 0;
  RunParams P = make_params(c);
  // batch layout of GGEMSSource::OrganizeParticlesInBatch (src/sources/GGEMSSource.cc:403-425)
  uint32_t nb, base, rem, slots;
  if ((rc = batch_layout(n, nb, base, rem, slots))) return rc;
  P.n_batches = nb;
  P.base = base;
  P.rem = rem;
  P.n_slots = slots;
  Launch L;
  if ((rc = pick_launch(c, slots, false, L))) return rc;
  P.xchg_offset = L.xchg_offset;
  HIP_TRY(hipMemsetAsync(c->queue.p, 0, sizeof(uint32_t), c->stream));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(L.fn, dim3(L.grid), dim3(kBlock), L.shmem, c->stream, P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  HIP_TRY(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  return 0;
}

int ggh_set_acquisition(ggh_context* c, const ggh_acquisition_desc* d) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const int n_box = n_solids_of(c, kBoxSolid);
  if (!d || d->n_projections < 1 || d->n_projections > GGH_MAX_PROJECTIONS || !d->source_matrices ||
      !d->n_particles || (n_box > 0 && !d->box_matrices))
    return fail(-2, "invalid acquisition description");
  const int np = d->n_projections;
  std::vector<uint4> lay((size_t)np);
  uint32_t max_slots = 0;
  for (int p = 0; p < np; ++p) {
    uint32_t nb, base, rem, slots;
    if ((rc = batch_layout(d->n_particles[p], nb, base, rem, slots))) return rc;
    lay[p] = make_uint4(nb, base, rem, 0u);
    max_slots = std::max(max_slots, slots);
  }
  std::vector<int32_t> only;
  if (d->phantoms) {
    only.assign(d->phantoms, d->phantoms + np);
    for (int32_t k : only)
      if (k < -1 || k >= (int)c->solids.size() || (k >= 0 && c->solids[k].type == kBoxSolid))
        return fail(-2, "a projection's phantom must be -1 or the index of a voxel or meshed solid");
  }
  std::vector<Mat34> src((size_t)np), box((size_t)np * n_box);
  for (int p = 0; p < np; ++p) src[p] = to_mat34(d->source_matrices + 16 * (size_t)p);
  for (size_t k = 0; k < box.size(); ++k) box[k] = to_mat34(d->box_matrices + 16 * k);
  drop_acquisition(c);
  c->acq_src_host = src;
  c->acq_box_host = box;
  c->acq_phantom_host = only;
  if ((rc = dev_upload(c->acq_src, src.data(), src.size(), c->stream)) ||
      (rc = dev_upload(c->acq_box, box.data(), box.size(), c->stream)) ||
      (rc = dev_upload(c->acq_layout, lay.data(), lay.size(), c->stream)) ||
      (!only.empty() && (rc = dev_upload(c->acq_phantom, only.data(), only.size(), c->stream))))
    return rc;
  c->acq_hist.resize(c->solids.size());
  c->acq_scat.resize(c->solids.size());
  for (size_t k = 0; k < c->solids.size(); ++k) {
    if (c->solids[k].type != kBoxSolid) continue;
    const size_t n_img = (size_t)np * c->hist_elems[k];
    if ((rc = dev_alloc(c->acq_hist[k], n_img))) return rc;
    HIP_TRY(hipMemsetAsync(c->acq_hist[k].p, 0, n_img * sizeof(int32_t), c->stream));
    if (c->scat[k].p) {
      if ((rc = dev_alloc(c->acq_scat[k], n_img))) return rc;
      HIP_TRY(hipMemsetAsync(c->acq_scat[k].p, 0, n_img * sizeof(int32_t), c->stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->acq_n_proj = np;
  c->acq_slots = max_slots;
  return 0;
}

int ggh_run_acquisition(ggh_context* c) {
  RunParams P;
  Launch L;
  int rc = acquisition_launch(c, P, L);
  if (rc) return rc;
  if (P.n_slots == 0) return 0;
  HIP_TRY(hipMemsetAsync(c->queue.p, 0, sizeof(uint32_t), c->stream));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(L.fn, dim3(L.grid), dim3(kBlock), L.shmem, c->stream, P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  HIP_TRY(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  return 0;
}

int ggh_capture_acquisition(ggh_context* c) {
  RunParams P;
  Launch L;
  int rc = acquisition_launch(c, P, L);
  if (rc) return rc;
  if (P.n_slots == 0) return fail(-3, "the acquisition has no primaries");
  drop_graph(c);
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, c->stream, c->queue.p);
  hipLaunchKernelGGL(L.fn, dim3(L.grid), dim3(kBlock), L.shmem, c->stream, P);
  const hipError_t le = hipGetLastError();
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(c->stream, &g);
  if (le != hipSuccess || ce != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    const hipError_t e = le != hipSuccess ? le : ce;
    return fail(-100 - (int)e, std::string("acquisition capture: ") + hipGetErrorString(e));
  }
  c->graph = g;
  const hipError_t ie = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    drop_graph(c);
    return fail(-100 - (int)ie, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
  }
  return 0;
}

int ggh_replay_acquisition(ggh_context* c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->graph_exec) return fail(-3, "no captured acquisition");
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  HIP_TRY(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
  return 0;
}

}  // extern "C"

namespace {

// BinarySearch over the host energy grid, as the device does it
int host_bin(const std::vector<float>& a, float key) {
  int low = 0, high = (int)a.size() - 1;
  while (high - low > 1) {
    const int mid = (high + low) / 2;
    if (a[mid] == key) return mid;
    if (a[mid] < key) low = mid;
    else high = mid;
  }
  return low;
}

// Probability of each energy bin of the cross-section tables under the
// source's sampling (GetPrimariesGGEMSXRaySource.cl:103-125): line idx has
// probability cdf[idx+1] - cdf[idx]; an interpolated spectrum spreads it
// uniformly over [E_idx, E_idx+1), split at the bin edges.  W holds nbins
// (<= 2048) doubles of work space; bins and w receive the bins of nonzero
// weight, whose count is returned.  Allocates nothing.
int drr_weights(const ggh_context* c, double* W, int32_t* bins, float* w) {
  const std::vector<float>& eb = c->ebins_host;
  const int nb = (int)eb.size();
  for (int k = 0; k < nb; ++k) W[k] = 0.0;
  const std::vector<float>& e = c->spectrum_host;
  const std::vector<float>& f = c->cdf_host;
  for (int idx = 0; idx + 1 < (int)f.size() && idx + 1 < (int)e.size(); ++idx) {
    const double p = (double)f[idx + 1] - (double)f[idx];
    if (!(p > 0.0)) continue;
    const double a = e[idx], b = e[idx + 1];
    if (!c->is_interp || !(b > a)) {
      W[host_bin(eb, e[idx])] += p;
      continue;
    }
    for (int k = 0; k + 1 < nb; ++k) {
      const double lo = k == 0 ? -INFINITY : (double)eb[k];
      const double hi = k + 2 == nb ? INFINITY : (double)eb[k + 1];
      const double ov = std::min(b, hi) - std::max(a, lo);
      if (ov > 0.0) W[k] += p * ov / (b - a);
    }
  }
  int n = 0;
  for (int k = 0; k < nb; ++k)
    if (W[k] > 0.0) {
      bins[n] = k;
      w[n] = (float)W[k];
      ++n;
    }
  return n;
}

// Checks shared by both ray-tracer entry points; fills the launch
// parameters except the weights and the image.  Allocates nothing when the
// solid table is already on the device (allow_upload false).
int drr_setup(ggh_context* c, int solid, int projection, const float* image, bool allow_upload, DrrParams& D) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!image) return fail(-1, "null image");
  if (!c->have_physics || !c->have_source) return fail(-3, "physics tables and source must be set");
  if (solid < 0 || solid >= (int)c->solids.size() || c->solids[solid].type != kBoxSolid)
    return fail(-2, "not a box solid");
  if (c->n_mat > kDrrMaxMat) return fail(-2, "the ray tracer handles at most 32 materials");
  for (size_t i = 0; i < c->solids.size(); ++i)
    if (c->solids[i].type == kVoxel && c->max_material[i] >= c->n_mat)
      return fail(-3, "solid labels reach a material without physics tables");
  if (c->solids_dirty) {
    if (!allow_upload)
      return fail(-3, "the solid table is not on the device yet: a run or ggh_drr uploads it after a solid is added");
    if ((rc = sync_solids(c))) return rc;
  }
  std::memset(&D, 0, sizeof D);
  const DevSolid& box = c->solids[solid];
  D.only = -1;
  if (projection < 0) {
    D.src_m = c->src_m;
    D.box_m = box.m;
  } else {
    const int n_box = n_solids_of(c, kBoxSolid);
    if (projection >= c->acq_n_proj) return fail(-2, "no such projection");
    D.src_m = c->acq_src_host[projection];
    D.box_m = c->acq_box_host[(size_t)projection * n_box + box.box_index];
    if (!c->acq_phantom_host.empty()) D.only = c->acq_phantom_host[projection];
  }
  D.solids = c->solids_dev.p;
  D.n_solids = (int32_t)c->solids.size();
  D.box = solid;
  D.sig = c->tab_sig.p;
  D.nbins = c->nbins;
  D.n_mat = c->n_mat;
  return 0;
}


}  // namespace

extern "C" {

int ggh_drr(ggh_context* c, int solid, int projection, float* image) {
  DrrParams D;
  int rc = drr_setup(c, solid, projection, image, true, D);
  if (rc) return rc;
  const DevSolid& box = c->solids[solid];
  std::vector<double> W((size_t)c->nbins);
  std::vector<int32_t> bins((size_t)c->nbins);
  std::vector<float> w((size_t)c->nbins);
  const int n_w = drr_weights(c, W.data(), bins.data(), w.data());
  const size_t npx = (size_t)box.ne[0] * box.ne[1];
  DevBuf<int32_t> db;
  DevBuf<float> dw, dimg;
  if ((rc = dev_upload(db, bins.data(), (size_t)n_w, c->stream)) || (rc = dev_upload(dw, w.data(), (size_t)n_w, c->stream)) ||
      (rc = dev_alloc(dimg, npx)))
    return rc;
  D.w_bin = db.p;
  D.w = dw.p;
  D.n_w = n_w;
  D.image = dimg.p;
  hipLaunchKernelGGL(drr_kernel, dim3((unsigned)((npx + kDrrBlock - 1) / kDrrBlock)), dim3(kDrrBlock), 0, c->stream, D);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(image, dimg.p, npx * sizeof(float), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  db.release();
  dw.release();
  dimg.release();
  if (e != hipSuccess) return fail(-100 - (int)e, std::string("ray tracer: ") + hipGetErrorString(e));
  return 0;
}

int ggh_drr_noalloc(ggh_context* c, int solid, int projection, float* image, void* stream) {
  DrrParams D;
  int rc = drr_setup(c, solid, projection, image, false, D);
  if (rc) return rc;
  if (c->nbins > kNoAllocMaxBins)
    return fail(-2, "ggh_drr_noalloc stages the weights of at most 1024 energy bins (the physics tables have " +
                        std::to_string(c->nbins) + "); ggh_drr has no such limit");
  if (!c->scratch) return fail(-3, "the staging array was not resolved for this device");
  const hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  uint32_t* stage = c->scratch;
  hipLaunchKernelGGL(drr_weights_kernel, dim3(1), dim3(kWeightsBlock), 0, s, c->ebins.p, c->nbins, c->spectrum.p,
                     (int)c->spectrum_host.size(), c->cdf.p, (int)c->cdf_host.size(), c->is_interp, stage);
  const DevSolid& box = c->solids[solid];
  const size_t npx = (size_t)box.ne[0] * box.ne[1];
  D.w_bin = reinterpret_cast<const int32_t*>(stage + kScratchBins);
  D.w = reinterpret_cast<const float*>(stage + kScratchWeights);
  D.n_w_dev = stage;
  D.image = image;
  hipLaunchKernelGGL(drr_kernel, dim3((unsigned)((npx + kDrrBlock - 1) / kDrrBlock)), dim3(kDrrBlock), 0, s, D);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(-100 - (int)e, std::string("ray tracer (no allocation): ") + hipGetErrorString(e));
  return 0;
}

int ggh_get_projection(ggh_context* c, int solid, int projection, int32_t* counts, int32_t* scatter) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (solid < 0 || solid >= (int)c->solids.size() || c->solids[solid].type != kBoxSolid)
    return fail(-2, "not a box solid");
  if (!c->acq_n_proj || projection < 0 || projection >= c->acq_n_proj) return fail(-2, "no such projection");
  const size_t n = (size_t)c->hist_elems[solid], off = (size_t)projection * n;
  if (counts)
    HIP_TRY(hipMemcpyAsync(counts, c->acq_hist[solid].p + off, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
  if (scatter) {
    if (!c->acq_scat[solid].p) return fail(-2, "scatter histogram not stored for this solid");
    HIP_TRY(hipMemcpyAsync(scatter, c->acq_scat[solid].p + off, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int ggh_autotune(ggh_context* c, uint64_t n_trial, ggh_launch_desc* best) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (n_trial == 0) return fail(-2, "empty trial");
  DevBuf<uint32_t> saved;
  if ((rc = dev_alloc(saved, c->rng.n))) return rc;
  HIP_TRY(hipMemcpyAsync(saved.p, c->rng.p, c->rng.n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
  const int keep_round = c->round_steps, keep_cold = c->cold_passes, keep_batch = c->interact_batch,
            keep_blocks = c->blocks_per_cu;
  float best_ms = INFINITY;
  ggh_launch_desc b{};
  // With the reference sampling the grid is round steps x interaction
  // batch.  Tau sampling leaves its interactions to the navigation passes,
  // so the batch does not apply, and it keeps its passes per round (one
  // pass beat two at full size by less than a short trial's timing noise):
  // its grid is round steps only, centred on its measured best.
  const bool tau = c->tau_mode != 0;
  const int rounds[2][4] = {{12, 20, 32, 0}, {12, 16, 20, 24}};
  const int batches[3] = {8, 16, 32};
  for (int blocks : {1, 2}) {
    for (int round : rounds[tau]) {
      if (round == 0) continue;
      for (int bi = 0; bi < (tau ? 1 : 3); ++bi) {  // tau sampling: one trial per round length
        c->blocks_per_cu = blocks;
        c->round_steps = round;
        c->interact_batch = tau ? keep_batch : batches[bi];
        c->cold_passes = keep_cold;
        HIP_TRY(hipMemcpyAsync(c->rng.p, saved.p, saved.n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
        if ((rc = ggh_run(c, n_trial))) break;
        if (c->last_ms < best_ms) {
          best_ms = c->last_ms;
          b.round_steps = round;
          b.cold_passes = c->cold_passes;
          b.interact_batch = c->interact_batch;
          b.blocks_per_cu = blocks;
          b.general_kernel = c->force_general ? 1 : 0;
        }
      }
      if (rc) break;  // a failed trial ends the search with its error
    }
    if (rc) break;
  }
  const hipError_t e = hipMemcpyAsync(c->rng.p, saved.p, saved.n * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                      c->stream);
  (void)hipStreamSynchronize(c->stream);
  saved.release();
  if (rc || e != hipSuccess) {
    c->round_steps = keep_round;
    c->cold_passes = keep_cold;
    c->interact_batch = keep_batch;
    c->blocks_per_cu = keep_blocks;
    return rc ? rc : fail(-100 - (int)e, "autotune: restoring the random streams failed");
  }
  c->round_steps = b.round_steps;
  c->cold_passes = b.cold_passes;
  c->interact_batch = b.interact_batch;
  c->blocks_per_cu = b.blocks_per_cu;
  if (best) *best = b;
  return ggh_reset_tallies(c);
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
  out->mesh_steps = h[6];
  out->mesh_interactions = h[7];
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
  if (edep2 && !c->dose_desc.edep_squared) return fail(-2, "edep squared not stored");
  if (hit && !c->dose_desc.hit) return fail(-2, "hits not stored");
  if (ptrack && !c->ptrack.p) return fail(-2, "photon tracking not stored");
  const size_t n = c->tally.n;
  if (edep || edep2 || hit) {
    // one field of the records at a time through a device staging array;
    // the stream orders each split after the previous copy out of it
    DevBuf<double> stage;
    if ((rc = dev_alloc(stage, n))) return rc;
    void* out[3] = {edep, edep2, hit};
    const size_t bytes[3] = {sizeof(double), sizeof(double), sizeof(int32_t)};
    const unsigned grid = (unsigned)std::min<size_t>((n + kDoseBlock - 1) / kDoseBlock, 65536);
    hipError_t e = hipSuccess;
    for (int f = 0; f < 3 && e == hipSuccess; ++f) {
      if (!out[f]) continue;
      hipLaunchKernelGGL(tally_field_kernel, dim3(grid), dim3(kDoseBlock), 0, c->stream, c->tally.p, (int64_t)n, f,
                         (void*)stage.p);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(out[f], stage.p, n * bytes[f], hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    stage.release();
    if (e != hipSuccess) return fail(-100 - (int)e, std::string("dose tallies: ") + hipGetErrorString(e));
  }
  if (ptrack) HIP_TRY(hipMemcpyAsync(ptrack, c->ptrack.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

int ggh_compute_dose(ggh_context* c, float scale, int32_t water_ref, float min_density, float* dose, float* unc) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->have_dose) return fail(-2, "dosimetry not set");
  if (unc && (!c->dose_desc.hit || !c->dose_desc.edep_squared))
    return fail(-2, "uncertainty needs hit and edep squared tallies");
  const size_t n = c->tally.n;
  const DevSolid& s = c->solids[c->dose_desc.solid];
  DevBuf<float> d_dose, d_unc;
  if ((rc = dev_alloc(d_dose, n))) return rc;
  if (unc && (rc = dev_alloc(d_unc, n))) {
    d_dose.release();
    return rc;
  }
  DoseParams D;
  std::memset(&D, 0, sizeof D);
  D.tally = c->tally.p;
  D.labels = s.type == kVoxel ? s.vox.labels : nullptr;
  D.density = c->density.p;
  D.mat_offset = s.mat_offset;
  D.nd[0] = c->dose_desc.n_dosels[0];
  D.nd[1] = c->dose_desc.n_dosels[1];
  D.nd[2] = c->dose_desc.n_dosels[2];
  D.nv[0] = s.vox.nx;
  D.nv[1] = s.vox.ny;
  D.nv[2] = s.vox.nz;
  D.dosel_size = mk(c->dose_desc.dosel_size[0], c->dose_desc.dosel_size[1], c->dose_desc.dosel_size[2]);
  D.vs = s.vox.vs;
  D.vlo = s.border.lo;
  D.scale = scale;
  D.water_reference = water_ref;
  D.min_density = min_density;
  D.dose = d_dose.p;
  D.unc = unc ? d_unc.p : nullptr;
  unsigned grid = (unsigned)std::min<size_t>((n + kDoseBlock - 1) / kDoseBlock, 65536);
  hipLaunchKernelGGL(dose_kernel, dim3(grid), dim3(kDoseBlock), 0, c->stream, D);
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

int ggh_debug_neg_log(ggh_context* c, int n, const uint32_t* sums, float* fast, float* ref) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n <= 0) return fail(-2, "invalid count");
  DevBuf<uint32_t> ds;
  DevBuf<float> df, dr;
  if ((rc = dev_upload(ds, sums, (size_t)n, c->stream)) || (rc = dev_alloc(df, (size_t)n)) ||
      (rc = dev_alloc(dr, (size_t)n)))
    return rc;
  hipLaunchKernelGGL(neglog_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, ds.p, df.p, dr.p);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(fast, df.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(ref, dr.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  ds.release(); df.release(); dr.release();
  return 0;
}

int ggh_last_grid(ggh_context* c, int32_t* blocks, int32_t* block_size, int32_t* variant) {
  if (!c || !blocks || !block_size || !variant) return fail(-1, "null argument");
  *blocks = (int32_t)c->last_grid;
  *block_size = kBlock;
  *variant = c->last_variant;
  return 0;
}

void ggh_host_uniform_from_sum(const uint32_t* sums, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = uniform_from_sum(sums[i]);
}

void ggh_host_div_rn(const float* a, const float* b, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    volatile float one = 1.0f;
    const float y = one / b[i];
    out[i] = div_rn(a[i], b[i], y);
  }
}

void ggh_host_energy_bin(const float* e, size_t n, const float* ebins, int32_t nbins, int32_t* out) {
  const std::vector<float> eb(ebins, ebins + nbins);
  float l2min, inv;
  energy_bin_guess(eb, l2min, inv);
  for (size_t i = 0; i < n; ++i) out[i] = energy_bin(e[i], ebins, nbins, l2min, inv);
}

void ggh_host_distance_to_aabb(const float* p, const float* d, const float* box, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    Ray r;
    r.set(mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
    Box b;
    b.lo = mk(box[6 * i], box[6 * i + 1], box[6 * i + 2]);
    b.hi = mk(box[6 * i + 3], box[6 * i + 4], box[6 * i + 5]);
    out[i] = distance_to_aabb(mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]), r, b);
  }
}

}  // extern "C"
