/*
 * ggo_transport.c -- oracle restatement of the GGEMS device-side hot path.
 * TEST INFRASTRUCTURE ONLY (see ggems_oracle.h).
 *
 * Follows, in the reference's operation order:
 *   include/GGEMS/randoms/GGEMSKissEngine.hh:50-75            (JKISS32 uniform)
 *   src/randoms/GGEMSPseudoRandomGenerator.cc:140-167          (mt19937 slot seeding)
 *   src/kernels/GetPrimariesGGEMSXRaySource.cl:52-143          (cone-beam primary)
 *   src/kernels/ParticleSolidDistanceGGEMS*.cl:43-108          (nearest solid)
 *   src/kernels/ProjectToGGEMS*.cl:44-129                      (projection)
 *   src/kernels/TrackThroughGGEMSVoxelizedSolid.cl:58-323      (voxel transport)
 *   src/kernels/TrackThroughGGEMSSolidBox.cl:56-261            (CT module transport)
 *   src/kernels/TrackThroughGGEMSMeshSolid.cl:322-534          (meshed solid transport)
 *   include/GGEMS/navigators/GGEMSPhotonNavigator.hh:57-131    (next interaction)
 *   include/GGEMS/physics/GGEMSComptonScatteringModels.hh:47-132
 *   include/GGEMS/physics/GGEMSPhotoElectricEffectModels.hh:46-53
 *   include/GGEMS/physics/GGEMSRayleighScatteringModels.hh:169-316
 *   include/GGEMS/geometries/GGEMSRayTracing.hh:60-326
 *   include/GGEMS/maths/GGEMSReferentialTransformation.hh:49-141
 *   include/GGEMS/maths/GGEMSMatrixOperations.hh:51-140
 *   include/GGEMS/maths/GGEMSMathAlgorithms.hh:46-77
 *   include/GGEMS/navigators/GGEMSDoseRecording.hh:65-118
 *   src/kernels/ComputeDoseGGEMSVoxelizedSolid.cl:53-121
 *   src/global/GGEMS.cc:303-363, src/sources/GGEMSSource.cc:371-426 (batch/hop loop)
 *
 * Device math built-ins (log, exp, sqrt, cos, sin, acos) are taken from the
 * host C library; the OpenCL driver's own implementations are not part of
 * the reference tree, so per-history agreement with any GPU is statistical
 * (ulp-level differences), while geometry uses only IEEE +,-,*,/ and sqrt.
 */
#include "ggems_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define TOL 5.0e-04f        /* GEOMETRY_TOLERANCE (GGEMSGeometryConstants.hh) */
#define EPS6 1.0e-06f       /* EPSILON6 */
#define OUT_OF_WORLD FLT_MAX
#define TWO_PI_F 6.283185307179586f
#define E_MASS 0.510998910f
#define MAX_SLOTS 1048576ULL /* MAXIMUM_PARTICLES (CMakeLists.txt:128-132) */
#define MAX_HOPS 100         /* max_loop (src/global/GGEMS.cc:330) */

enum { COMPTON = 0, PHOTOELECTRIC = 1, RAYLEIGH = 2, TRANSPORTATION = 99 };
enum { ALIVE = 0, DEAD = 1 };

/* ------------------------------------------------------------------------ */
/* RNG                                                                      */
/* ------------------------------------------------------------------------ */
/* std::mt19937 (ISO C++ [rand.eng.mers]) */
typedef struct { uint32_t mt[624]; int i; } mt19937;

static void mt_seed(mt19937* g, uint32_t s) {
  g->mt[0] = s;
  for (int i = 1; i < 624; ++i) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->i = 624;
}

static uint32_t mt_next(mt19937* g) {
  if (g->i >= 624) {
    for (int k = 0; k < 624; ++k) {
      uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
      g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    g->i = 0;
  }
  uint32_t y = g->mt[g->i++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

void ggo_mt19937_stream(uint32_t seed, uint64_t n, uint32_t* out) {
  mt19937 g;
  mt_seed(&g, seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = mt_next(&g);
}

/* GGEMSPseudoRandomGenerator::InitializeSeeds (:140-167): per slot j, in
 * order, state1..state4 = mt(), state5 = 0.  Output is SoA [5][n_slots]. */
void ggo_seed_slots(uint32_t seed, uint64_t n, uint32_t* st) {
  mt19937 g;
  mt_seed(&g, seed);
  for (uint64_t j = 0; j < n; ++j) {
    st[0 * n + j] = mt_next(&g);
    st[1 * n + j] = mt_next(&g);
    st[2 * n + j] = mt_next(&g);
    st[3 * n + j] = mt_next(&g);
    st[4 * n + j] = 0;
  }
}

/* KissUniform (GGEMSKissEngine.hh:50-75); s = {x, y, z, w, c} */
float ggo_kiss_uniform(uint32_t* s) {
  s[1] ^= s[1] << 5;
  s[1] ^= s[1] >> 7;
  s[1] ^= s[1] << 22;
  int32_t t = (int32_t)(s[2] + s[3] + s[4]);
  s[2] = s[3];
  s[4] = t < 0;
  s[3] = (uint32_t)t & 2147483647u;
  s[0] += 1411392427u;
  return (float)(((double)(float)(s[0] + s[1] + s[3]) / 4294967295.0) * (double)(1.0f - 1.0f / (1 << 23)));
}

/* ------------------------------------------------------------------------ */
/* small vector helpers (OpenCL float3 semantics, left-to-right dot)        */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } f3;

static inline f3 v3(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* normalize(v): v * (1/sqrt(dot(v,v))) */
static inline f3 normalize3(f3 v) {
  float inv = 1.0f / sqrtf(dot3(v, v));
  return v3(v.x * inv, v.y * inv, v.z * inv);
}

/* GGfloat44MultGGfloat3 (GGEMSMatrixOperations.hh:51-75): dot(row, (p,1)) */
static inline f3 m44_mul_p(const float m[4][4], f3 p) {
  return v3(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3] * 1.0f,
            m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3] * 1.0f,
            m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3] * 1.0f);
}
/* GGfloat33MultGGfloat3 (:85-107) */
static inline f3 m33_mul(const float m[4][4], f3 p) {
  return v3(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z,
            m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z,
            m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z);
}
/* GGfloat33TransposeMultGGfloat3 (:117-140) */
static inline f3 m33t_mul(const float m[4][4], f3 p) {
  return v3(m[0][0] * p.x + m[1][0] * p.y + m[2][0] * p.z,
            m[0][1] * p.x + m[1][1] * p.y + m[2][1] * p.z,
            m[0][2] * p.x + m[1][2] * p.y + m[2][2] * p.z);
}
/* GGEMSReferentialTransformation.hh:99-141 */
static inline f3 g2l_pos(const float m[4][4], f3 p) {
  return m33t_mul(m, v3(p.x - m[0][3], p.y - m[1][3], p.z - m[2][3]));
}
static inline f3 l2g_pos(const float m[4][4], f3 p) { return m44_mul_p(m, p); }
static inline f3 g2l_dir(const float m[4][4], f3 d) { return normalize3(m33t_mul(m, d)); }
static inline f3 l2g_dir(const float m[4][4], f3 d) { return normalize3(m33_mul(m, d)); }

/* RotateUnitZ (GGEMSReferentialTransformation.hh:49-70) */
static inline f3 rotate_unit_z(f3 v, f3 u) {
  float u1 = u.x, u2 = u.y, u3 = u.z;
  float up = u1 * u1 + u2 * u2;
  if (up > 0) {
    up = sqrtf(up);
    float px = v.x, py = v.y, pz = v.z;
    v.x = (u1 * u3 * px - u2 * py) / up + u1 * pz;
    v.y = (u2 * u3 * px + u1 * py) / up + u2 * pz;
    v.z = -up * px + u3 * pz;
  } else if (u3 < 0.) {
    v.x = -v.x;
    v.z = -v.z;
  }
  return v;
}

/* BinarySearch (GGEMSMathAlgorithms.hh:46-60) */
static inline int bsearch_f(float key, const float* a, int size) {
  int low = 0, mid = 0, high = size - 1;
  while (high - low > 1) {
    mid = (high + low) / 2;
    if (a[mid] == key) return mid;
    if (a[mid] < key) low = mid;
    else high = mid;
  }
  return low;
}

/* LinearInterpolation (GGEMSMathAlgorithms.hh:67-77) */
static inline float lin_interp(float xa, float ya, float xb, float yb, float x) {
  if (xa > xb) return yb;
  if (xa >= x) return ya;
  if (xb <= x) return yb;
  return ya + (x - xa) * (yb - ya) / (xb - xa);
}

/* ------------------------------------------------------------------------ */
/* ray tracing (GGEMSRayTracing.hh)                                         */
/* ------------------------------------------------------------------------ */
static inline void safety_inside(f3* p, const float* lo, const float* hi) {
  float smin = fabsf(p->x - lo[0]), smax = fabsf(p->x - hi[0]);
  p->x = (smin < TOL) ? lo[0] + TOL : p->x;
  p->x = (smax < TOL) ? hi[0] - TOL : p->x;
  smin = fabsf(p->y - lo[1]);
  smax = fabsf(p->y - hi[1]);
  p->y = (smin < TOL) ? lo[1] + TOL : p->y;
  p->y = (smax < TOL) ? hi[1] - TOL : p->y;
  smin = fabsf(p->z - lo[2]);
  smax = fabsf(p->z - hi[2]);
  p->z = (smin < TOL) ? lo[2] + TOL : p->z;
  p->z = (smax < TOL) ? hi[2] - TOL : p->z;
}

static inline void safety_outside(f3* p, const float* lo, const float* hi) {
  float smin = fabsf(p->x - lo[0]), smax = fabsf(p->x - hi[0]);
  p->x = (smin < TOL) ? lo[0] - TOL : p->x;
  p->x = (smax < TOL) ? hi[0] + TOL : p->x;
  smin = fabsf(p->y - lo[1]);
  smax = fabsf(p->y - hi[1]);
  p->y = (smin < TOL) ? lo[1] - TOL : p->y;
  p->y = (smax < TOL) ? hi[1] + TOL : p->y;
  smin = fabsf(p->z - lo[2]);
  smax = fabsf(p->z - hi[2]);
  p->z = (smin < TOL) ? lo[2] - TOL : p->z;
  p->z = (smax < TOL) ? hi[2] + TOL : p->z;
}

static inline int in_aabb(f3 p, const float* lo, const float* hi) {
  if (p.x < (lo[0] + TOL) || p.x > (hi[0] - TOL)) return 0;
  if (p.y < (lo[1] + TOL) || p.y > (hi[1] - TOL)) return 0;
  if (p.z < (lo[2] + TOL) || p.z > (hi[2] - TOL)) return 0;
  return 1;
}

/* ComputeDistanceToAABB (GGEMSRayTracing.hh:223-298) */
static float dist_aabb(f3 p, f3 d, const float* lo, const float* hi) {
  float idx, idy, idz, tmp;
  float tmin = FLT_MIN, tmax = FLT_MAX;
  float tymin, tymax, tzmin, tzmax;
  if (fabsf(d.x) < EPS6) {
    if (p.x < lo[0] || p.x > hi[0]) return OUT_OF_WORLD;
  } else {
    idx = 1.0f / d.x;
    tmin = (lo[0] - p.x) * idx;
    tmax = (hi[0] - p.x) * idx;
    if (tmin > tmax) { tmp = tmin; tmin = tmax; tmax = tmp; }
    if (tmin > tmax) return OUT_OF_WORLD;
  }
  if (fabsf(d.y) < EPS6) {
    if (p.y < lo[1] || p.y > hi[1]) return OUT_OF_WORLD;
  } else {
    idy = 1.0f / d.y;
    tymin = (lo[1] - p.y) * idy;
    tymax = (hi[1] - p.y) * idy;
    if (tymin > tymax) { tmp = tymin; tymin = tymax; tymax = tmp; }
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    if (tmin > tmax) return OUT_OF_WORLD;
  }
  if (fabsf(d.z) < EPS6) {
    if (p.z < lo[2] || p.z > hi[2]) return OUT_OF_WORLD;
  } else {
    idz = 1.0f / d.z;
    tzmin = (lo[2] - p.z) * idz;
    tzmax = (hi[2] - p.z) * idz;
    if (tzmin > tzmax) { tmp = tzmin; tzmin = tzmax; tzmax = tmp; }
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    if (tmin > tmax) return OUT_OF_WORLD;
  }
  if (tmin < 0.0f && (tmax < 0.0f || tmax == 0.0f)) return OUT_OF_WORLD;
  if ((double)(tmax - tmin) < (2.0 * (double)TOL)) return OUT_OF_WORLD;
  if (tmin <= EPS6) return tmax;
  return tmin;
}

float ggo_distance_to_aabb(const float* p, const float* d, const float* lo, const float* hi) {
  return dist_aabb(v3(p[0], p[1], p[2]), v3(d[0], d[1], d[2]), lo, hi);
}

/* ------------------------------------------------------------------------ */
/* photon state and physics                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  f3 p, d;
  float e;
  int status;
  int scatter;
  int solid;
  float dist;
  int e_index;
} photon;

/* GetPhotonNextInteraction (GGEMSPhotonNavigator.hh:57-93) */
static inline float next_interaction(const ggo_physics* ph, photon* g, int gmat, uint32_t* rng, int* proc) {
  int eid = bsearch_f(g->e, ph->energy_bins, ph->nbins);
  float next = OUT_OF_WORLD;
  int np = TRANSPORTATION; /* NO_PROCESS == TRANSPORTATION == 99 */
  for (int i = 0; i < ph->n_processes; ++i) {
    int pid = ph->process_id[i];
    float d = -logf(ggo_kiss_uniform(rng)) / ph->cs[pid][eid + ph->nbins * gmat];
    if (d < next) {
      next = d;
      np = pid;
    }
  }
  g->e_index = eid;
  *proc = np;
  return next;
}

/* KleinNishinaComptonSampleSecondaries (GGEMSComptonScatteringModels.hh:47-132) */
static void compton(photon* g, uint32_t* rng) {
  float kE0 = g->e;
  float kE0_MeC2 = kE0 / E_MASS;
  f3 dir0 = g->d;
  float kEps0 = 1.0f / (1.0f + 2.0f * kE0_MeC2);
  float kEps0Eps0 = kEps0 * kEps0;
  float kAlpha1 = -logf(kEps0);
  float kAlpha2 = kAlpha1 + 0.5f * (1.0f - kEps0Eps0);
  float epsilon, epsilonsq, onecost, sint2, greject, r1, r2, r3;
  int nloop = 0;
  do {
    ++nloop;
    if (nloop > 1000) return;
    r1 = ggo_kiss_uniform(rng);
    r2 = ggo_kiss_uniform(rng);
    r3 = ggo_kiss_uniform(rng);
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
  } while (greject < r3);
  if (sint2 < 0.0f) sint2 = 0.0f;
  float costheta = 1.0f - onecost;
  float sintheta = sqrtf(sint2);
  float phi = ggo_kiss_uniform(rng) * TWO_PI_F;
  f3 gd = v3(sintheta * cosf(phi), sintheta * sinf(phi), costheta);
  gd = rotate_unit_z(gd, dir0);
  gd = normalize3(gd);
  g->e = kE0 * epsilon;
  g->d = gd;
}

/* LivermoreRayleighSampleSecondaries (GGEMSRayleighScatteringModels.hh:169-316) */
static const float* pp[9];
static float factor;

/* PP0..PP8 and FACTOR (GGEMSRayleighScatteringModels.hh:36-153) from the data blob */
static int rayleigh_init(void) {
  char nm[16] = "rayleigh_pp0";
  for (int k = 0; k < 9; ++k) {
    nm[11] = (char)('0' + k);
    pp[k] = (const float*)ggo_data(nm, NULL);
    if (!pp[k]) return -1;
  }
  const float* f = (const float*)ggo_data("rayleigh_factor", NULL);
  if (!f) return -1;
  factor = *f;
  return 0;
}

static void rayleigh(const ggo_physics* ph, photon* g, int gmat, uint32_t* rng) {
  float kE0 = g->e;
  if (kE0 <= 250.0e-6f) {
    g->status = DEAD;
    return;
  }
  f3 dir0 = g->d;
  int nb = ph->nbins;
  int nm1 = ph->n_elements[gmat] - 1;
  int mix = ph->elem_index[gmat];
  int eid = g->e_index;
  int selz = ph->z[mix + nm1];
  if (nm1 > 0) {
    const float* csr = ph->cs[RAYLEIGH];
    float kCS = lin_interp(ph->energy_bins[eid], csr[eid + nb * gmat], ph->energy_bins[eid + 1],
                           csr[eid + 1 + nb * gmat], kE0);
    float x = ggo_kiss_uniform(rng) * kCS;
    float cross = 0.0f;
    for (int i = 0; i < nm1; ++i) {
      int z = ph->z[mix + i];
      cross += ph->atom_density[mix + i] *
               lin_interp(ph->energy_bins[eid], ph->cs_atom_rayleigh[eid + nb * z], ph->energy_bins[eid + 1],
                          ph->cs_atom_rayleigh[eid + 1 + nb * z], kE0);
      if (x < cross) {
        selz = z;
        break;
      }
    }
  }
  float kXX = factor * kE0 * kE0;
  float kN0 = pp[6][selz] - 1.0f, kN1 = pp[7][selz] - 1.0f, kN2 = pp[8][selz] - 1.0f;
  float kB0 = pp[3][selz], kB1 = pp[4][selz], kB2 = pp[5][selz];
  float x = 2.0f * kXX * kB0;
  float kW0 = (x < 0.02f) ? kN0 * x * (1.0f - 0.5f * (kN0 - 1.0f) * x * (1.0f - (kN0 - 2.0f) * x / 3.0f))
                          : 1.0f - expf(-kN0 * logf(1.0f + x));
  x = 2.0f * kXX * kB1;
  float kW1 = (x < 0.02f) ? kN1 * x * (1.0f - 0.5f * (kN1 - 1.0f) * x * (1.0f - (kN1 - 2.0f) * x / 3.0f))
                          : 1.0f - expf(-kN1 * logf(1.0f + x));
  x = 2.0f * kXX * kB2;
  float kW2 = (x < 0.02f) ? kN2 * x * (1.0f - 0.5f * (kN2 - 1.0f) * x * (1.0f - (kN2 - 2.0f) * x / 3.0f))
                          : 1.0f - expf(-kN2 * logf(1.0f + x));
  float kX0 = kW0 * pp[0][selz] / (kB0 * kN0);
  float kX1 = kW1 * pp[1][selz] / (kB1 * kN1);
  float kX2 = kW2 * pp[2][selz] / (kB2 * kN2);
  float costheta;
  do {
    float w = kW0, n = kN0, b = kB0;
    x = ggo_kiss_uniform(rng) * (kX0 + kX1 + kX2);
    if (x > kX0) {
      x -= kX0;
      if (x <= kX1) { w = kW1; n = kN1; b = kB1; }
      else { w = kW2; n = kN2; b = kB2; }
    }
    n = 1.0f / n;
    float y = ggo_kiss_uniform(rng) * w;
    if (y < 0.02f) x = y * n * (1.0f + 0.5f * (n + 1.0f) * y * (1.0f - (n + 2.0f) * y / 3.0f));
    else x = expf(-n * logf(1.0f - y)) - 1.0f;
    costheta = 1.0f - x / (b * kXX);
  } while (2.0f * ggo_kiss_uniform(rng) > 1.0f + costheta * costheta || costheta < -1.0f);
  float phi = TWO_PI_F * ggo_kiss_uniform(rng);
  float sintheta = sqrtf((1.0f - costheta) * (1.0f + costheta));
  f3 gd = v3(sintheta * cosf(phi), sintheta * sinf(phi), costheta);
  gd = rotate_unit_z(gd, dir0);
  gd = normalize3(gd);
  g->d = gd;
}

/* PhotonDiscreteProcess (GGEMSPhotonNavigator.hh:109-131) */
static inline void discrete(const ggo_physics* ph, photon* g, int proc, int gmat, uint32_t* rng) {
  if (proc == COMPTON) compton(g, rng);
  else if (proc == PHOTOELECTRIC) {
    g->status = DEAD;
    g->e = 0.0f;
  } else if (proc == RAYLEIGH) rayleigh(ph, g, gmat, rng);
}

/* ------------------------------------------------------------------------ */
/* tallies                                                                  */
/* ------------------------------------------------------------------------ */
static inline void atomic_add_f64(double* a, double v) {
  uint64_t* u = (uint64_t*)a;
  uint64_t old = __atomic_load_n(u, __ATOMIC_RELAXED);
  for (;;) {
    double cur;
    memcpy(&cur, &old, 8);
    double nv = cur + v;
    uint64_t nu;
    memcpy(&nu, &nv, 8);
    if (__atomic_compare_exchange_n(u, &old, nu, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
  }
}

static inline int dosel_index(const ggo_solid* s, f3 p) {
  if (p.x < s->dose_bmin[0] + EPS6 || p.x > s->dose_bmax[0] - EPS6) return -1;
  if (p.y < s->dose_bmin[1] + EPS6 || p.y > s->dose_bmax[1] - EPS6) return -1;
  if (p.z < s->dose_bmin[2] + EPS6 || p.z > s->dose_bmax[2] - EPS6) return -1;
  int ix = (int)((p.x - s->dose_bmin[0]) * s->inv_dosel[0]);
  int iy = (int)((p.y - s->dose_bmin[1]) * s->inv_dosel[1]);
  int iz = (int)((p.z - s->dose_bmin[2]) * s->inv_dosel[2]);
  if (ix < 0 || ix >= s->ndosel[0]) return -1;
  if (iy < 0 || iy >= s->ndosel[1]) return -1;
  if (iz < 0 || iz >= s->ndosel[2]) return -1;
  return ix + iy * s->ndosel[0] + iz * s->ndosel[0] * s->ndosel[1];
}

/* dose_record_standard (GGEMSDoseRecording.hh:94-118) */
static inline void dose_record(const ggo_solid* s, float edep, f3 p) {
  int id = dosel_index(s, p);
  if (id < 0) return;
  if (s->hit) __atomic_fetch_add(&s->hit[id], 1, __ATOMIC_RELAXED);
  atomic_add_f64(&s->edep[id], (double)edep);
  if (s->edep2) atomic_add_f64(&s->edep2[id], (double)edep * (double)edep);
}

/* ------------------------------------------------------------------------ */
/* navigation                                                               */
/* ------------------------------------------------------------------------ */
static inline f3 add3(f3 a, f3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 scale3(f3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }

/* Opt-in sampling with one number of mean free paths per flight (tau_mode;
 * no counterpart in the reference, which draws three competing exponentials
 * in every voxel step, GGEMSPhotonNavigator.hh:57-93).  tau = -log(u) is
 * drawn when the photon enters a voxel solid and after each interaction that
 * leaves it alive.  A voxel step whose boundary distance db costs
 * a = sigma_tot * db <= tau crosses the voxel and spends a; otherwise the
 * photon interacts at tau / sigma_tot, the process chosen by one uniform in
 * proportion to its left-bin cross section.  Per voxel this gives the
 * reference's probabilities: crossing exp(-sigma_tot db), process i with
 * sigma_i / sigma_tot, and an exponential depth truncated at db.  The
 * voxel walk, the tolerances and the interactions are unchanged. */
static inline float total_cs(const ggo_physics* ph, int gmat, int eid) {
  float t = 0.0f;
  for (int i = 0; i < ph->n_processes; ++i) t += ph->cs[ph->process_id[i]][eid + ph->nbins * gmat];
  return t;
}

static inline int select_process(const ggo_physics* ph, int gmat, int eid, float tot, uint32_t* rng) {
  const float x = ggo_kiss_uniform(rng) * tot;
  float c = 0.0f;
  for (int i = 0; i + 1 < ph->n_processes; ++i) {
    c += ph->cs[ph->process_id[i]][eid + ph->nbins * gmat];
    if (x < c) return ph->process_id[i];
  }
  return ph->process_id[ph->n_processes - 1];
}

#define INTERACTION_PENDING (-1) /* tau mode: the process is chosen once the step stays inside */

/* track_through_ggems_voxelized_solid (TrackThroughGGEMSVoxelizedSolid.cl:76-323) */
static void track_voxel(const ggo_physics* ph, const ggo_solid* s, photon* g, uint32_t* rng,
                        ggo_stats* st, int32_t* walk, int walk_max, int* walk_n) {
  const float(*m)[4] = s->obb.m;
  f3 lp = g2l_pos(m, g->p);
  f3 ld = g2l_dir(m, g->d);
  const float* bmin = s->obb.bmin;
  const float* bmax = s->obb.bmax;
  const float* vs = s->voxel_size;
  const int tau_mode = ph->tau_mode && !walk;
  float tau = tau_mode ? -logf(ggo_kiss_uniform(rng)) : 0.0f;
  do {
    int vx = (int)((lp.x - bmin[0]) / vs[0]);
    int vy = (int)((lp.y - bmin[1]) / vs[1]);
    int vz = (int)((lp.z - bmin[2]) / vs[2]);
    /* The reference checks the upper bound only (TrackThroughGGEMSVoxelizedSolid.cl:137);
     * a negative index would read outside the label volume, so it ends the walk. */
    if (vx >= s->nvox[0] || vy >= s->nvox[1] || vz >= s->nvox[2] || vx < 0 || vy < 0 || vz < 0) {
      g->dist = OUT_OF_WORLD;
      g->solid = -1;
      break;
    }
    int vid = vx + vy * s->nvox[0] + vz * s->nvox[0] * s->nvox[1];
    if (walk) {
      if (*walk_n < walk_max) walk[*walk_n] = vid;
      ++*walk_n;
    }
    int gmat = s->mat_offset + s->labels[vid];
    if (st) st->voxel_steps++;
    int proc = TRANSPORTATION;
    float next = OUT_OF_WORLD, tot = 0.0f;
    int eid = 0;
    if (tau_mode) {
      eid = bsearch_f(g->e, ph->energy_bins, ph->nbins);
      g->e_index = eid;
      tot = total_cs(ph, gmat, eid);
    } else if (!walk) {
      next = next_interaction(ph, g, gmat, rng, &proc);
    }
    float vlo[3] = {bmin[0] + (float)vx * vs[0], bmin[1] + (float)vy * vs[1], bmin[2] + (float)vz * vs[2]};
    float vhi[3] = {vlo[0] + vs[0], vlo[1] + vs[1], vlo[2] + vs[2]};
    safety_inside(&lp, vlo, vhi);
    float db = dist_aabb(lp, ld, vlo, vhi);
    int cross;
    if (tau_mode) {
      const float a = tot * db;
      cross = !(tau < a);
      if (cross) {
        tau = tau - a;
      } else {
        next = tau / tot;
        proc = INTERACTION_PENDING;
      }
    } else {
      cross = db <= next;
    }
    if (cross) {
      next = db + TOL;
      proc = TRANSPORTATION;
      if (s->dosimetry && s->photon_tracking) {
        int id = dosel_index(s, lp);
        if (id >= 0) __atomic_fetch_add(&s->photon_tracking[id], 1, __ATOMIC_RELAXED);
      }
    }
    lp = add3(lp, scale3(ld, next));
    safety_outside(&lp, vlo, vhi);
    if (!in_aabb(lp, bmin, bmax)) {
      g->dist = OUT_OF_WORLD;
      g->solid = -1;
      break;
    }
    float e0 = g->e;
    if (proc != TRANSPORTATION) {
      if (st) st->interactions++;
      if (proc == INTERACTION_PENDING) proc = select_process(ph, gmat, eid, tot, rng);
      g->d = ld;
      discrete(ph, g, proc, gmat, rng);
      if (proc == COMPTON || proc == RAYLEIGH) g->scatter = 1;
      if (s->dosimetry) dose_record(s, e0 - g->e, lp);
      ld = g->d;
    }
    if (g->e <= ph->photon_cut[gmat]) {
      if (s->dosimetry) dose_record(s, g->e, lp);
      g->status = DEAD;
    }
    if (tau_mode && proc != TRANSPORTATION && g->status == ALIVE) tau = -logf(ggo_kiss_uniform(rng));
  } while (g->status == ALIVE);
  g->p = l2g_pos(m, lp);
  g->d = l2g_dir(m, ld);
}

/* track_through_ggems_solid_box (TrackThroughGGEMSSolidBox.cl:56-261) */
static void track_box(const ggo_physics* ph, const ggo_solid* s, photon* g, uint32_t* rng, ggo_stats* st) {
  const float(*m)[4] = s->obb.m;
  f3 lp = g2l_pos(m, g->p);
  f3 ld = g2l_dir(m, g->d);
  const float* bmin = s->obb.bmin;
  const float* bmax = s->obb.bmax;
  int gmat = s->mat_offset;
  do {
    int proc;
    float next = next_interaction(ph, g, gmat, rng, &proc);
    if (st) st->detector_steps++;
    safety_inside(&lp, bmin, bmax);
    float db = dist_aabb(lp, ld, bmin, bmax);
    if (db <= next) {
      next = db + TOL;
      proc = TRANSPORTATION;
    }
    lp = add3(lp, scale3(ld, next));
    safety_outside(&lp, bmin, bmax);
    if (!in_aabb(lp, bmin, bmax)) {
      g->dist = OUT_OF_WORLD;
      g->solid = -1;
      break;
    }
    if (g->e < s->threshold) g->status = DEAD;
    if (proc != TRANSPORTATION) {
      if (st) st->detector_interactions++;
      g->d = ld;
      discrete(ph, g, proc, gmat, rng);
      ld = g->d;
      if (proc == PHOTOELECTRIC || proc == COMPTON) {
        float esx = s->box_size[0] / (float)s->nelem[0];
        float esy = s->box_size[1] / (float)s->nelem[1];
        int ix = (int)((lp.x - bmin[0]) / esx);
        int iy = (int)((lp.y - bmin[1]) / esy);
        if (ix >= 0 && ix < s->nelem[0] && iy >= 0 && iy < s->nelem[1]) {
          int id = ix + iy * s->nelem[0];
          __atomic_fetch_add(&s->histogram[id], 1, __ATOMIC_RELAXED);
          if (s->scatter && g->scatter) __atomic_fetch_add(&s->scatter[id], 1, __ATOMIC_RELAXED);
        }
      }
    }
  } while (g->status == ALIVE);
  g->p = l2g_pos(m, lp);
  g->d = l2g_dir(m, ld);
}

static inline float dist_obb(const ggo_obb* o, f3 p, f3 d);

/* track_through_ggems_meshed_solid (TrackThroughGGEMSMeshSolid.cl:322-534): a
 * single material (the navigator's material 0); global coordinates; the
 * photon is teleported to the first surface when it is outside the mesh,
 * and projected out of the octree box (the OBB) when no surface is ahead */
static void track_mesh(const ggo_physics* ph, const ggo_solid* s, photon* g, uint32_t* rng, ggo_stats* st) {
  const int gmat = s->mat_offset;
  f3 lp = g->p, ld = g->d;
  float pts[16 * 3];
  do {
    float pa[3] = {lp.x, lp.y, lp.z}, da[3] = {ld.x, ld.y, ld.z};
    int n = ggo_mesh_hits(&s->mesh, pa, da, pts);
    if (n == 0) {
      g->dist = OUT_OF_WORLD;
      g->solid = -1;
      break;
    }
    f3 next_tri;
    if (n % 2 == 0) {
      lp = v3(pts[0], pts[1], pts[2]);
      next_tri = v3(pts[3], pts[4], pts[5]);
    } else {
      next_tri = v3(pts[0], pts[1], pts[2]);
    }
    int proc;
    float next = next_interaction(ph, g, gmat, rng, &proc);
    if (st) st->mesh_steps++;
    float dx = (next_tri.x - lp.x), dy = (next_tri.y - lp.y), dz = (next_tri.z - lp.z);
    float dtri = sqrtf(dx * dx + dy * dy + dz * dz);
    if (dtri <= next) {
      next = dtri + TOL;
      proc = TRANSPORTATION;
      if (s->dosimetry && s->photon_tracking) {
        int id = dosel_index(s, lp);
        if (id >= 0) __atomic_fetch_add(&s->photon_tracking[id], 1, __ATOMIC_RELAXED);
      }
    }
    lp = add3(lp, scale3(ld, next));
    if (proc != TRANSPORTATION) {
      if (st) st->mesh_interactions++;
      float e0 = g->e;
      g->d = ld;
      discrete(ph, g, proc, gmat, rng);
      if (proc == COMPTON || proc == RAYLEIGH) g->scatter = 1;
      if (s->dosimetry) dose_record(s, e0 - g->e, lp);
      ld = g->d;
    }
    if (g->e <= ph->photon_cut[gmat]) {
      if (s->dosimetry) dose_record(s, g->e, lp);
      g->status = DEAD;
    }
  } while (g->status == ALIVE);
  float dist = dist_obb(&s->obb, lp, ld);
  g->p = add3(lp, scale3(ld, dist + TOL));
  g->d = ld;
}

/* get_primaries_ggems_xray_source (GetPrimariesGGEMSXRaySource.cl:52-143) */
void ggo_sample_primary(const ggo_source* src, uint32_t* rng, float* pos, float* dir, float* e) {
  double phi = ggo_kiss_uniform(rng);
  double theta = ggo_kiss_uniform(rng);
  phi *= (double)TWO_PI_F;
  double new_aperture = 1.0 - cos((double)src->aperture);
  theta = acos(1.0 - new_aperture * theta);
  f3 rotation = v3((float)(cos(phi) * sin(theta)), (float)(sin(phi) * sin(theta)), (float)cos(theta));
  f3 gp = l2g_pos(src->m, v3(0.0f, 0.0f, 0.0f));
  f3 d = normalize3(v3(0.0f - gp.x, 0.0f - gp.y, 0.0f - gp.z));
  d = rotate_unit_z(rotation, d);
  d = normalize3(d);
  gp.x = src->focal[0] * (ggo_kiss_uniform(rng) - 0.5f);
  gp.y = src->focal[1] * (ggo_kiss_uniform(rng) - 0.5f);
  gp.z = src->focal[2] * (ggo_kiss_uniform(rng) - 0.5f);
  gp = l2g_pos(src->m, gp);
  int nb_arg = src->n_energy + 1;
  int idx = 0;
  float rnd = 0.0f;
  if (nb_arg != 1) {
    rnd = ggo_kiss_uniform(rng);
    idx = bsearch_f(rnd, src->cdf, nb_arg + 1);
  }
  *e = (!src->is_interp) ? src->energy[idx]
                         : lin_interp(src->cdf[idx], src->energy[idx], src->cdf[idx + 1], src->energy[idx + 1], rnd);
  pos[0] = gp.x; pos[1] = gp.y; pos[2] = gp.z;
  dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

/* IsParticleInOBB / ComputeDistanceToOBB / TransportGetSafetyInsideOBB (GGEMSRayTracing.hh:94-203,312-326) */
static inline int in_obb(const ggo_obb* o, f3 p) { return in_aabb(g2l_pos(o->m, p), o->bmin, o->bmax); }

static inline float dist_obb(const ggo_obb* o, f3 p, f3 d) {
  return dist_aabb(g2l_pos(o->m, p), g2l_dir(o->m, d), o->bmin, o->bmax);
}

static inline f3 safety_inside_obb(const ggo_obb* o, f3 p) {
  f3 lp = g2l_pos(o->m, p);
  safety_inside(&lp, o->bmin, o->bmax);
  return l2g_pos(o->m, lp);
}

/* One primary through the hop loop (src/global/GGEMS.cc:323-345). */
static void transport_primary(const ggo_physics* ph, const ggo_solid* solids, int ns, const ggo_source* src,
                              uint32_t* rng, ggo_stats* st) {
  photon g;
  float pos[3], dir[3];
  ggo_sample_primary(src, rng, pos, dir, &g.e);
  g.p = v3(pos[0], pos[1], pos[2]);
  g.d = v3(dir[0], dir[1], dir[2]);
  g.status = ALIVE;
  g.scatter = 0;
  g.solid = -1;
  g.dist = OUT_OF_WORLD;
  g.e_index = 0;
  if (st) st->primaries++;
  for (int hop = 0; hop < MAX_HOPS; ++hop) {
    if (st) st->hops++;
    /* FindSolid: particle_solid_distance_* over every solid, in order */
    for (int k = 0; k < ns; ++k) {
      if (g.status == DEAD || g.dist == 0.0f) break;
      const ggo_obb* o = &solids[k].obb;
      if (in_obb(o, g.p)) {
        g.dist = 0.0f;
        g.solid = k;
        break;
      }
      float d = dist_obb(o, g.p, g.d);
      if (d < g.dist) {
        g.dist = d;
        g.solid = k;
      }
    }
    /* ProjectToSolid */
    if (g.solid == -1) g.status = DEAD;
    if (g.dist == OUT_OF_WORLD) {
      g.solid = -1;
      g.status = DEAD;
    }
    if (g.status == DEAD) break;
    const ggo_solid* s = &solids[g.solid];
    g.p = add3(g.p, scale3(g.d, g.dist + TOL));
    g.p = safety_inside_obb(&s->obb, g.p);
    g.dist = 0.0f;
    /* TrackThroughSolid */
    if (s->type == GGO_SOLID_VOXEL) track_voxel(ph, s, &g, rng, st, NULL, 0, NULL);
    else if (s->type == GGO_SOLID_MESH) track_mesh(ph, s, &g, rng, st);
    else track_box(ph, s, &g, rng, st);
    if (g.status == DEAD) break;
  }
}

/* Worker threads when the caller passes 0: OMP_NUM_THREADS when set (the
 * machine's CPU share), else the online processors, at most 16. */
static int default_threads(void) {
  const char* e = getenv("OMP_NUM_THREADS");
  int t = e ? atoi(e) : 0;
  if (t <= 0) {
    t = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (t > 16) t = 16;
  }
  return t < 1 ? 1 : t;
}

/* GGEMSSource::OrganizeParticlesInBatch (src/sources/GGEMSSource.cc:403-425):
 * ceil(float(n) / slots) batches.  Where the float quotient gives one batch
 * too few (n above 2^24 that rounds down, e.g. 2^24 + 1), the largest batch
 * would exceed the slot buffer, and the exact ceiling is taken instead. */
void ggo_batch_layout(uint64_t n, uint64_t* nb, uint64_t* ns) {
  uint64_t b = (uint64_t)ceilf((float)n / (float)MAX_SLOTS);
  if (b == 0) b = 1;
  if (b * MAX_SLOTS < n) b = (n + MAX_SLOTS - 1) / MAX_SLOTS;
  *nb = b;
  *ns = b == 1 ? n : n / b + (n % b ? 1 : 0);
}

typedef struct {
  const ggo_physics* ph;
  const ggo_solid* solids;
  int ns;
  const ggo_source* src;
  uint64_t n, n_batches, base, rem;
  uint32_t* states;
  uint64_t n_slots;
  uint64_t s0, s1;
  ggo_stats st;
} worker;

static void* run_worker(void* arg) {
  worker* w = (worker*)arg;
  memset(&w->st, 0, sizeof w->st);
  for (uint64_t s = w->s0; s < w->s1; ++s) {
    uint32_t rng[5];
    for (int k = 0; k < 5; ++k) rng[k] = w->states[k * w->n_slots + s];
    for (uint64_t b = 0; b < w->n_batches; ++b) {
      uint64_t size = w->n_batches == 1 ? w->n : w->base + (b < w->rem ? 1 : 0);
      if (s < size) transport_primary(w->ph, w->solids, w->ns, w->src, rng, &w->st);
    }
    for (int k = 0; k < 5; ++k) w->states[k * w->n_slots + s] = rng[k];
  }
  return NULL;
}

uint64_t ggo_run_slots(const ggo_physics* ph, const ggo_solid* solids, int ns, const ggo_source* src,
                       uint64_t n, uint32_t seed, uint64_t slot_begin, uint64_t slot_end, int threads,
                       ggo_stats* stats) {
  uint64_t nb, nslots;
  if (rayleigh_init()) return 0;
  ggo_batch_layout(n, &nb, &nslots);
  uint64_t seeded = MAX_SLOTS; /* the reference seeds every slot of the buffer */
  uint32_t* states = (uint32_t*)malloc(5 * seeded * sizeof(uint32_t));
  if (!states) return 0;
  ggo_seed_slots(seed, seeded, states);
  if (slot_end > nslots) slot_end = nslots;
  if (slot_begin > slot_end) slot_begin = slot_end;
  if (threads <= 0) threads = default_threads();
  if (threads < 1) threads = 1;
  worker* ws = (worker*)calloc((size_t)threads, sizeof(worker));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  uint64_t span = slot_end - slot_begin;
  for (int t = 0; t < threads; ++t) {
    worker* w = &ws[t];
    w->ph = ph;
    w->solids = solids;
    w->ns = ns;
    w->src = src;
    w->n = n;
    w->n_batches = nb;
    w->base = n / nb;
    w->rem = n % nb;
    w->states = states;
    w->n_slots = seeded;
    w->s0 = slot_begin + span * (uint64_t)t / (uint64_t)threads;
    w->s1 = slot_begin + span * (uint64_t)(t + 1) / (uint64_t)threads;
    pthread_create(&th[t], NULL, run_worker, w);
  }
  ggo_stats tot;
  memset(&tot, 0, sizeof tot);
  for (int t = 0; t < threads; ++t) {
    pthread_join(th[t], NULL);
    tot.primaries += ws[t].st.primaries;
    tot.voxel_steps += ws[t].st.voxel_steps;
    tot.interactions += ws[t].st.interactions;
    tot.detector_steps += ws[t].st.detector_steps;
    tot.detector_interactions += ws[t].st.detector_interactions;
    tot.hops += ws[t].st.hops;
    tot.mesh_steps += ws[t].st.mesh_steps;
    tot.mesh_interactions += ws[t].st.mesh_interactions;
  }
  if (stats) *stats = tot;
  free(ws);
  free(th);
  free(states);
  return tot.primaries;
}

int ggo_run(const ggo_physics* ph, const ggo_solid* solids, int ns, const ggo_source* src, uint64_t n,
            uint32_t seed, int threads, ggo_stats* stats) {
  ggo_run_slots(ph, solids, ns, src, n, seed, 0, MAX_SLOTS, threads, stats);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CT acquisition: consecutive runs, one per gantry position                */
/* ------------------------------------------------------------------------ */
typedef struct {
  const ggo_physics* ph;
  const ggo_solid* solids;
  int ns;
  const ggo_source* src;
  int np;
  const float* src_mats;
  const float* box_mats;
  const int32_t* phantoms; /* [np] the one non-box solid present, -1 all; or NULL */
  int n_box;
  const uint64_t* lay; /* [np][3]: n_batches, base, rem */
  uint32_t* states;
  uint64_t n_slots;
  uint64_t s0, s1;
  ggo_stats st;
} acq_worker;

static void* acq_run(void* arg) {
  acq_worker* w = (acq_worker*)arg;
  memset(&w->st, 0, sizeof w->st);
  ggo_solid* sol = (ggo_solid*)malloc((size_t)w->ns * sizeof(ggo_solid));
  ggo_solid* present = (ggo_solid*)malloc((size_t)w->ns * sizeof(ggo_solid));
  if (!sol || !present) {
    free(sol);
    free(present);
    return NULL;
  }
  memcpy(sol, w->solids, (size_t)w->ns * sizeof(ggo_solid));
  ggo_source src = *w->src;
  for (uint64_t s = w->s0; s < w->s1; ++s) {
    uint32_t rng[5];
    for (int k = 0; k < 5; ++k) rng[k] = w->states[k * w->n_slots + s];
    for (int p = 0; p < w->np; ++p) {
      const uint64_t nb = w->lay[3 * p], base = w->lay[3 * p + 1], rem = w->lay[3 * p + 2];
      /* photon s of each batch: batch b holds base + (b < rem) photons */
      uint64_t todo = s < base ? nb : (s == base ? rem : 0);
      if (!todo) continue;
      memcpy(src.m, w->src_mats + 16 * (size_t)p, sizeof src.m);
      for (int k = 0, j = 0; k < w->ns; ++k) {
        if (w->solids[k].type != GGO_SOLID_BOX) continue;
        const size_t img = (size_t)p * w->solids[k].nelem[0] * w->solids[k].nelem[1];
        memcpy(sol[k].obb.m, w->box_mats + 16 * ((size_t)p * w->n_box + j), sizeof sol[k].obb.m);
        sol[k].histogram = w->solids[k].histogram + img;
        sol[k].scatter = w->solids[k].scatter ? w->solids[k].scatter + img : NULL;
        ++j;
      }
      /* several phantoms in one acquisition: projection p is a run in a
       * world of the box solids and its one phantom, in registration order */
      const int only = w->phantoms ? w->phantoms[p] : -1;
      if (only < 0) {
        while (todo--) transport_primary(w->ph, sol, w->ns, &src, rng, &w->st);
      } else {
        int n = 0;
        for (int k = 0; k < w->ns; ++k)
          if (k == only || sol[k].type == GGO_SOLID_BOX) present[n++] = sol[k];
        while (todo--) transport_primary(w->ph, present, n, &src, rng, &w->st);
      }
    }
    for (int k = 0; k < 5; ++k) w->states[k * w->n_slots + s] = rng[k];
  }
  free(sol);
  free(present);
  return NULL;
}

int ggo_run_acquisition(const ggo_physics* ph, const ggo_solid* solids, int ns, const ggo_source* src, int np,
                        const float* src_mats, const float* box_mats, const uint64_t* n_particles,
                        const int32_t* phantoms, uint32_t seed, int threads, ggo_stats* stats) {
  if (rayleigh_init() || np < 1) return -1;
  if (phantoms)
    for (int p = 0; p < np; ++p)
      if (phantoms[p] < -1 || phantoms[p] >= ns || (phantoms[p] >= 0 && solids[phantoms[p]].type == GGO_SOLID_BOX))
        return -1;
  uint64_t* lay = (uint64_t*)malloc(3 * (size_t)np * sizeof(uint64_t));
  uint32_t* states = (uint32_t*)malloc(5 * MAX_SLOTS * sizeof(uint32_t));
  if (!lay || !states) {
    free(lay);
    free(states);
    return -1;
  }
  uint64_t max_slots = 0;
  for (int p = 0; p < np; ++p) {
    uint64_t nb, nslots, n = n_particles[p];
    ggo_batch_layout(n, &nb, &nslots);
    lay[3 * p] = nb;
    lay[3 * p + 1] = nb == 1 ? n : n / nb;
    lay[3 * p + 2] = nb == 1 ? 0 : n % nb;
    if (nslots > max_slots) max_slots = nslots;
  }
  int n_box = 0;
  for (int k = 0; k < ns; ++k) n_box += solids[k].type == GGO_SOLID_BOX;
  ggo_seed_slots(seed, MAX_SLOTS, states);
  if (threads <= 0) threads = default_threads();
  if (threads < 1) threads = 1;
  acq_worker* ws = (acq_worker*)calloc((size_t)threads, sizeof(acq_worker));
  pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; ++t) {
    acq_worker* w = &ws[t];
    w->ph = ph;
    w->solids = solids;
    w->ns = ns;
    w->src = src;
    w->np = np;
    w->src_mats = src_mats;
    w->box_mats = box_mats;
    w->phantoms = phantoms;
    w->n_box = n_box;
    w->lay = lay;
    w->states = states;
    w->n_slots = MAX_SLOTS;
    w->s0 = max_slots * (uint64_t)t / (uint64_t)threads;
    w->s1 = max_slots * (uint64_t)(t + 1) / (uint64_t)threads;
    pthread_create(&th[t], NULL, acq_run, w);
  }
  ggo_stats tot;
  memset(&tot, 0, sizeof tot);
  for (int t = 0; t < threads; ++t) {
    pthread_join(th[t], NULL);
    tot.primaries += ws[t].st.primaries;
    tot.voxel_steps += ws[t].st.voxel_steps;
    tot.interactions += ws[t].st.interactions;
    tot.detector_steps += ws[t].st.detector_steps;
    tot.detector_interactions += ws[t].st.detector_interactions;
    tot.hops += ws[t].st.hops;
    tot.mesh_steps += ws[t].st.mesh_steps;
    tot.mesh_interactions += ws[t].st.mesh_interactions;
  }
  if (stats) *stats = tot;
  free(ws);
  free(th);
  free(states);
  free(lay);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* test entry points                                                        */
/* ------------------------------------------------------------------------ */
int ggo_voxel_walk(const ggo_solid* s, const float* p, const float* d, int32_t* out, int max_out) {
  photon g;
  memset(&g, 0, sizeof g);
  /* p, d are global; track_voxel converts them to the solid frame */
  g.p = v3(p[0], p[1], p[2]);
  g.d = v3(d[0], d[1], d[2]);
  g.e = 1.0f;
  g.status = ALIVE;
  int n = 0;
  ggo_physics ph;
  memset(&ph, 0, sizeof ph);
  float cut[256];
  memset(cut, 0, sizeof cut);
  ph.photon_cut = cut;
  track_voxel(&ph, s, &g, NULL, NULL, out, max_out, &n);
  return n;
}

float ggo_sample_compton(float e, float* dir, uint32_t* rng) {
  photon g;
  memset(&g, 0, sizeof g);
  g.e = e;
  g.d = v3(dir[0], dir[1], dir[2]);
  compton(&g, rng);
  dir[0] = g.d.x; dir[1] = g.d.y; dir[2] = g.d.z;
  return g.e;
}

float ggo_sample_rayleigh(const ggo_physics* ph, int gmat, float e, float* dir, uint32_t* rng, int* status) {
  photon g;
  memset(&g, 0, sizeof g);
  g.e = e;
  g.d = v3(dir[0], dir[1], dir[2]);
  g.e_index = bsearch_f(e, ph->energy_bins, ph->nbins);
  if (rayleigh_init()) return -1.0f;
  rayleigh(ph, &g, gmat, rng);
  dir[0] = g.d.x; dir[1] = g.d.y; dir[2] = g.d.z;
  if (status) *status = g.status;
  return g.e;
}

/* compute_dose_ggems_voxelized_solid (ComputeDoseGGEMSVoxelizedSolid.cl:53-121);
 * a meshed solid takes its one material (compute_dose_ggems_meshed_solid,
 * ComputeDoseGGEMSMeshSolid.cl:53-114) */
void ggo_compute_dose(const ggo_solid* s, const float* density_by_gmat, float scale, int water_ref,
                      float min_density, float* dose, float* unc) {
  int nx = s->ndosel[0], ny = s->ndosel[1], nz = s->ndosel[2];
  int slice = nx * ny;
  long total = (long)slice * nz;
  const float U_G_CM3 = 6.241509704e+21f / 1.e3f;
  const float GY = 1.0e-12f;
  float vol = s->dosel_size[0] * s->dosel_size[1] * s->dosel_size[2];
  for (long gid = 0; gid < total; ++gid) {
    int z = (int)(gid / slice);
    int x = (int)((gid - (long)z * slice) % nx);
    int y = (int)((gid - (long)z * slice) / nx);
    float px = (float)x * s->dosel_size[0] + (float)(nx - 1) * (-0.5f * s->dosel_size[0]);
    float py = (float)y * s->dosel_size[1] + (float)(ny - 1) * (-0.5f * s->dosel_size[1]);
    float pz = (float)z * s->dosel_size[2] + (float)(nz - 1) * (-0.5f * s->dosel_size[2]);
    int mat = 0;
    if (s->type == GGO_SOLID_VOXEL) {
      int vx = (int)((px - s->obb.bmin[0]) / s->voxel_size[0]);
      int vy = (int)((py - s->obb.bmin[1]) / s->voxel_size[1]);
      int vz = (int)((pz - s->obb.bmin[2]) / s->voxel_size[2]);
      mat = s->labels[vx + vy * s->nvox[0] + vz * s->nvox[0] * s->nvox[1]];
    }
    float density = water_ref ? 1.0f * U_G_CM3 : density_by_gmat[s->mat_offset + mat];
    dose[gid] = density < min_density ? 0.0f : (float)((double)scale * s->edep[gid] / (double)density / (double)vol / (double)GY);
    if (unc) {
      if (s->hit[gid] > 1 && s->edep[gid] != 0.0) {
        double s2 = s->edep[gid] * s->edep[gid];
        unc[gid] = (float)sqrt(((double)s->hit[gid] * s->edep2[gid] - s2) / ((double)(s->hit[gid] - 1) * s2));
      } else {
        unc[gid] = 1.0f;
      }
    }
  }
}
