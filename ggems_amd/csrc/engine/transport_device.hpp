// transport_device.hpp -- device-side building blocks of the MI355X photon
// transport engine: JKISS32 streams, frame changes, slab ray tracing,
// Compton / photoelectric / Rayleigh sampling and tallies.
//
// Every function keeps the operation order of the GGEMS kernel code it
// replaces (file:line given per function) and the engine is compiled with
// -ffp-contract=off, so geometry and tables agree bit for bit with a host
// restatement of the same arithmetic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ggh {

constexpr float kTol = 5.0e-04f;          // GEOMETRY_TOLERANCE (GGEMSGeometryConstants.hh)
constexpr float kEps6 = 1.0e-06f;         // EPSILON6
constexpr float kOutOfWorld = 3.402823466e+38f;  // FLT_MAX (GGEMSParticleConstants.hh)
constexpr float kFltMin = 1.175494351e-38f;      // FLT_MIN
constexpr float kTwoPi = 6.283185307179586f;
constexpr float kElectronMassC2 = 0.510998910f;
constexpr int kCompton = 0, kPhotoelectric = 1, kRayleigh = 2, kTransportation = 99;
constexpr int kAlive = 0, kDead = 1;

struct f3 {
  float x, y, z;
};
__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 normalize(f3 v) {
  float inv = 1.0f / sqrtf(dot(v, v));
  return mk(v.x * inv, v.y * inv, v.z * inv);
}

// ---------------------------------------------------------------------------
// JKISS32 (include/GGEMS/randoms/GGEMSKissEngine.hh:50-75)
// ---------------------------------------------------------------------------
// The reference maps the 32-bit sum s to
//   (float)(((double)(float)s / 4294967295.0) * (double)(1.0f - 2^-23)).
// With f = (float)s = M * 2^e (24-bit mantissa M), that value is exactly the
// float k ulps below f * 2^-32, where k = 2 when M == 2^23 or
// 511 * M > 3 * 2^31 (M >= 12607537) and k = 1 otherwise: the FP64 division
// and product only decide which neighbour the final rounding lands on.  The
// identity is checked exhaustively over every float reachable from a uint32.
// Branch-free form: with t = mantissa bits of f * 2^-32, the two-ulp cases
// M == 2^23 and M >= 12607537 are exactly (t - 1) >= 12607537 - 2^23 - 1
// in unsigned arithmetic (t == 0 wraps around).
__host__ __device__ inline float uniform_from_sum(uint32_t s) {
  float x = (float)s * 0x1p-32f;
  uint32_t b;
  __builtin_memcpy(&b, &x, 4);
  b -= ((b & 0x7fffffu) - 1u) >= 4218928u ? 2u : 1u;
  float u;
  __builtin_memcpy(&u, &b, 4);
  return s == 0u ? 0.0f : u;
}

// a / b rounded to nearest, from y = RN(1/b) computed once on the host:
// q0 = RN(a y) is within 1.5 ulp of a/b, one correction with the exact
// remainder r = a - q b (an FMA) makes it faithful, and a second one makes
// it the correctly rounded quotient (Markstein, IBM J. Res. Dev. 34(1), 1990),
// for every a/b whose remainders stay normal (|a| >= 2^-100 and a normal
// quotient here).  When a/b is +inf (a = +inf, or b = 0 with y = +inf) the
// result is NaN; every caller compares it with `<`, which is false for
// NaN exactly as for +inf.
__host__ __device__ inline float div_rn(float a, float b, float y) {
  float q = a * y;
  float r = __builtin_fmaf(-q, b, a);
  q = __builtin_fmaf(r, y, q);
  r = __builtin_fmaf(-q, b, a);
  return __builtin_fmaf(r, y, q);
}

struct Rng {
  uint32_t x, y, z, w, c;
  // advances the state and returns the 32-bit sum KissUniform maps to [0, 1)
  __device__ inline uint32_t next_sum() {
    y ^= y << 5;
    y ^= y >> 7;
    y ^= y << 22;
    int32_t t = (int32_t)(z + w + c);
    z = w;
    c = t < 0;
    w = (uint32_t)t & 2147483647u;
    x += 1411392427u;
    return x + y + w;
  }
  __device__ inline float uniform() { return uniform_from_sum(next_sum()); }
};

// ---------------------------------------------------------------------------
// frames (include/GGEMS/maths/GGEMSMatrixOperations.hh:51-140,
//         include/GGEMS/maths/GGEMSReferentialTransformation.hh:49-141)
// m holds rows 0..2 of the 4x4 local->global matrix.
// ---------------------------------------------------------------------------
struct Mat34 {
  float m[12];
};
__device__ inline f3 l2g_pos(const Mat34& M, f3 p) {
  const float* m = M.m;
  return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3] * 1.0f, m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7] * 1.0f,
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11] * 1.0f);
}
__device__ inline f3 m33_mul(const Mat34& M, f3 p) {
  const float* m = M.m;
  return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z,
            m[8] * p.x + m[9] * p.y + m[10] * p.z);
}
__device__ inline f3 m33t_mul(const Mat34& M, f3 p) {
  const float* m = M.m;
  return mk(m[0] * p.x + m[4] * p.y + m[8] * p.z, m[1] * p.x + m[5] * p.y + m[9] * p.z,
            m[2] * p.x + m[6] * p.y + m[10] * p.z);
}
__device__ inline f3 g2l_pos(const Mat34& M, f3 p) {
  return m33t_mul(M, mk(p.x - M.m[3], p.y - M.m[7], p.z - M.m[11]));
}
__device__ inline f3 g2l_dir(const Mat34& M, f3 d) { return normalize(m33t_mul(M, d)); }
__device__ inline f3 l2g_dir(const Mat34& M, f3 d) { return normalize(m33_mul(M, d)); }

// RotateUnitZ (GGEMSReferentialTransformation.hh:49-70)
__device__ inline f3 rotate_unit_z(f3 v, f3 u) {
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

// BinarySearch (include/GGEMS/maths/GGEMSMathAlgorithms.hh:46-60)
__device__ inline int binary_search(float key, const float* a, int size) {
  int low = 0, mid = 0, high = size - 1;
  while (high - low > 1) {
    mid = (high + low) / 2;
    float v = a[mid];
    if (v == key) return mid;
    if (v < key) low = mid;
    else high = mid;
  }
  return low;
}

// The energy bin BinarySearch returns on the cross-section grid, found from
// a guess instead of a search.  On a strictly increasing grid the search
// returns the largest i <= size - 2 with a[i] <= key, or 0 when there is
// none (key below the grid or NaN).  The grid is log-spaced
// (GGEMSCrossSections.cc:204-206: E_i = E_min exp(slope i / (n - 1))), so
// (log2 E - l2min) * inv, with l2min = log2 a[0] and inv = (n - 1) /
// (log2 a[n-1] - log2 a[0]), lands on or next to that bin; the two loops
// then step to it by comparisons against the grid itself, which makes the
// result exact whatever the guess.  Two reads instead of eight dependent
// ones.
__host__ __device__ inline int energy_bin(float key, const float* a, int size, float l2min, float inv) {
  float g = (log2f(key) - l2min) * inv;
  g = fminf(fmaxf(g, 0.0f), (float)(size - 2));  // NaN -> 0
  int i = (int)g;
  while (i < size - 2 && a[i + 1] <= key) ++i;
  while (i > 0 && a[i] > key) --i;
  return i;
}

// LinearInterpolation (GGEMSMathAlgorithms.hh:67-77)
__device__ inline float lin_interp(float xa, float ya, float xb, float yb, float x) {
  if (xa > xb) return yb;
  if (xa >= x) return ya;
  if (xb <= x) return yb;
  return ya + (x - xa) * (yb - ya) / (xb - xa);
}

// ---------------------------------------------------------------------------
// slab ray tracing (include/GGEMS/geometries/GGEMSRayTracing.hh:60-298)
// ---------------------------------------------------------------------------
struct Box {
  f3 lo, hi;
};

// TransportGetSafetyInsideAABB (GGEMSRayTracing.hh:130-153)
__host__ __device__ inline void safety_inside(f3& p, const Box& b) {
  float smin = fabsf(p.x - b.lo.x), smax = fabsf(p.x - b.hi.x);
  p.x = (smin < kTol) ? b.lo.x + kTol : p.x;
  p.x = (smax < kTol) ? b.hi.x - kTol : p.x;
  smin = fabsf(p.y - b.lo.y);
  smax = fabsf(p.y - b.hi.y);
  p.y = (smin < kTol) ? b.lo.y + kTol : p.y;
  p.y = (smax < kTol) ? b.hi.y - kTol : p.y;
  smin = fabsf(p.z - b.lo.z);
  smax = fabsf(p.z - b.hi.z);
  p.z = (smin < kTol) ? b.lo.z + kTol : p.z;
  p.z = (smax < kTol) ? b.hi.z - kTol : p.z;
}

// TransportGetSafetyOutsideAABB (GGEMSRayTracing.hh:156-178)
__host__ __device__ inline void safety_outside(f3& p, const Box& b) {
  float smin = fabsf(p.x - b.lo.x), smax = fabsf(p.x - b.hi.x);
  p.x = (smin < kTol) ? b.lo.x - kTol : p.x;
  p.x = (smax < kTol) ? b.hi.x + kTol : p.x;
  smin = fabsf(p.y - b.lo.y);
  smax = fabsf(p.y - b.hi.y);
  p.y = (smin < kTol) ? b.lo.y - kTol : p.y;
  p.y = (smax < kTol) ? b.hi.y + kTol : p.y;
  smin = fabsf(p.z - b.lo.z);
  smax = fabsf(p.z - b.hi.z);
  p.z = (smin < kTol) ? b.lo.z - kTol : p.z;
  p.z = (smax < kTol) ? b.hi.z + kTol : p.z;
}

// IsParticleInAABB (GGEMSRayTracing.hh:60-82): the tolerance argument is
// ignored there in favour of GEOMETRY_TOLERANCE
__host__ __device__ inline bool in_aabb(f3 p, const Box& b) {
  if (p.x < (b.lo.x + kTol) || p.x > (b.hi.x - kTol)) return false;
  if (p.y < (b.lo.y + kTol) || p.y > (b.hi.y - kTol)) return false;
  if (p.z < (b.lo.z + kTol) || p.z > (b.hi.z - kTol)) return false;
  return true;
}

// Direction with its per-axis reciprocals (1.0f/d, the value the reference
// recomputes on every call) cached until the direction changes.
struct Ray {
  f3 d, inv;
  __host__ __device__ inline void set(f3 dir) {
    d = dir;
    inv.x = 1.0f / dir.x;
    inv.y = 1.0f / dir.y;
    inv.z = 1.0f / dir.z;
  }
};

// ComputeDistanceToAABB (GGEMSRayTracing.hh:223-298) without branches.
// The reference returns OUT_OF_WORLD at the first failing test; here every
// test is evaluated and OR-ed, which returns the same value because no test
// has side effects.  Per axis:
//   |d| < 1e-6: parallel, out when p lies outside the slab;
//   otherwise the slab interval [min(t1,t2), max(t1,t2)] (the reference's
//   compare-and-swap) replaces [tmin, tmax] on x and narrows it on y and z,
//   and an empty interval is out.
// The products of a parallel axis can be NaN (0 * inf); they are never selected.
// (tmax - tmin) < 2*tol is evaluated in FP64 by the reference; 2*kTol is
// exactly representable in FP32, so the FP32 comparison is identical.
__host__ __device__ inline float distance_to_aabb(f3 p, const Ray& r, const Box& b) {
  // Each axis gives the interval of t inside its slab.  An axis the ray is
  // parallel to (|d| < EPSILON6) keeps the running interval when the point
  // lies within the slab and empties it otherwise - the reference's early
  // return.  Its "tmin > tmax" returns after the second and third axes
  // reduce to one test at the end, since tmin only grows and tmax only
  // shrinks.  All of it is selects on floats: no branches, no lane masks
  // kept in registers.
  const float inf = __builtin_inff();
  bool par = fabsf(r.d.x) < kEps6;
  bool in = (p.x >= b.lo.x) & (p.x <= b.hi.x);
  float t1 = (b.lo.x - p.x) * r.inv.x, t2 = (b.hi.x - p.x) * r.inv.x;
  float tmin = par ? (in ? kFltMin : inf) : (t1 > t2 ? t2 : t1);
  float tmax = par ? (in ? kOutOfWorld : -inf) : (t1 > t2 ? t1 : t2);

  par = fabsf(r.d.y) < kEps6;
  in = (p.y >= b.lo.y) & (p.y <= b.hi.y);
  t1 = (b.lo.y - p.y) * r.inv.y;
  t2 = (b.hi.y - p.y) * r.inv.y;
  float lo = par ? (in ? -inf : inf) : (t1 > t2 ? t2 : t1);
  float hi = par ? (in ? inf : -inf) : (t1 > t2 ? t1 : t2);
  tmin = lo > tmin ? lo : tmin;
  tmax = hi < tmax ? hi : tmax;

  par = fabsf(r.d.z) < kEps6;
  in = (p.z >= b.lo.z) & (p.z <= b.hi.z);
  t1 = (b.lo.z - p.z) * r.inv.z;
  t2 = (b.hi.z - p.z) * r.inv.z;
  lo = par ? (in ? -inf : inf) : (t1 > t2 ? t2 : t1);
  hi = par ? (in ? inf : -inf) : (t1 > t2 ? t1 : t2);
  tmin = lo > tmin ? lo : tmin;
  tmax = hi < tmax ? hi : tmax;

  bool out = tmin > tmax;
  out |= (tmin < 0.0f) & (tmax <= 0.0f);   // tmax < 0 || tmax == 0
  out |= (tmax - tmin) < 2.0f * kTol;
  const float d = tmin <= kEps6 ? tmax : tmin;
  return out ? kOutOfWorld : d;
}

}  // namespace ggh
