/* ggo_mesh.c -- CPU oracle (TEST INFRASTRUCTURE ONLY): restatement of the
 * GGEMS meshed solid: triangle bounding spheres, the octree and its
 * insertion, and the ray/mesh intersection list.
 *
 * src/io/GGEMSSTLReader.cc:111-226       GGEMSMeshTriangle bounding sphere (Ritter)
 * src/geometries/GGEMSMeshedSolid.cc     ComputeOctreeCenter / ComputeHalfWidthCenter (:340-458)
 * src/geometries/GGEMSOctree.cc          node layout (:40-80,140-210), InsertTrianglesOnDevice (:240-316)
 * src/kernels/TrackThroughGGEMSMeshSolid.cl
 *                                        MollerTrumboreRayTriangle (:119-176), TestLineNode (:190-240),
 *                                        GetMeshIOPositions (:253-314)
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ggems_oracle.h"

#define EPS6 1.0e-06f
#define MAX_TRIANGLE_INTERACTION 16

typedef struct { float x, y, z; } p3;

static inline float dotp(p3 a, p3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* MostSeparatedPointsOnAABB + SphereFromDistantPoints + SphereOfSphereAndPoint */
static void triangle_sphere(const float* t, float* center, float* radius) {
  p3 pts[3] = {{t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]}};
  int minx = 0, maxx = 0, miny = 0, maxy = 0, minz = 0, maxz = 0;
  for (int i = 1; i < 3; ++i) {
    if (pts[i].x < pts[minx].x) minx = i;
    if (pts[i].x > pts[maxx].x) maxx = i;
    if (pts[i].y < pts[miny].y) miny = i;
    if (pts[i].y > pts[maxy].y) maxy = i;
    if (pts[i].z < pts[minz].z) minz = i;
    if (pts[i].z > pts[maxz].z) maxz = i;
  }
  p3 px = {pts[maxx].x - pts[minx].x, pts[maxx].y - pts[minx].y, pts[maxx].z - pts[minx].z};
  p3 py = {pts[maxy].x - pts[miny].x, pts[maxy].y - pts[miny].y, pts[maxy].z - pts[miny].z};
  p3 pz = {pts[maxz].x - pts[minz].x, pts[maxz].y - pts[minz].y, pts[maxz].z - pts[minz].z};
  float d2x = dotp(px, px), d2y = dotp(py, py), d2z = dotp(pz, pz);
  int mn = minx, mx = maxx;
  if (d2y > d2x && d2y > d2z) {
    mn = miny;
    mx = maxy;
  } else if (d2z > d2y && d2z > d2x) {
    mn = minz;
    mx = maxz;
  }
  p3 c = {(pts[mn].x + pts[mx].x) * 0.5f, (pts[mn].y + pts[mx].y) * 0.5f, (pts[mn].z + pts[mx].z) * 0.5f};
  p3 dd = {pts[mx].x - c.x, pts[mx].y - c.y, pts[mx].z - c.z};
  float r = dotp(dd, dd);
  r = sqrtf(r);
  for (int i = 0; i < 3; ++i) {
    p3 ps = {pts[i].x - c.x, pts[i].y - c.y, pts[i].z - c.z};
    float dist2 = dotp(ps, ps);
    if (dist2 > r * r) {
      float dist = sqrtf(dist2);
      float nr = (r + dist) * 0.5f;
      float k = (nr - r) / dist;
      r = nr;
      c.x += ps.x * k;
      c.y += ps.y * k;
      c.z += ps.z * k;
    }
  }
  center[0] = c.x;
  center[1] = c.y;
  center[2] = c.z;
  *radius = r;
}

/* ComputeOctreeCenter / ComputeHalfWidthCenter: bounds (hi starts at FLT_MIN,
 * as in the reference), expanded by 2.5 %, rounded outwards to whole mm */
static void octree_box(const float* tri, int n, float* center, float* half) {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {FLT_MIN, FLT_MIN, FLT_MIN};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 3; ++j)
      for (int a = 0; a < 3; ++a) {
        float v = tri[9 * i + 3 * j + a];
        if (v < lo[a]) lo[a] = v;
        if (v > hi[a]) hi[a] = v;
      }
  for (int a = 0; a < 3; ++a) {
    float e = (hi[a] - lo[a]) * 0.025f;
    lo[a] -= e;
    hi[a] += e;
    lo[a] = lo[a] < 0.0f ? floorf(lo[a]) : ceilf(lo[a]);
    hi[a] = hi[a] < 0.0f ? floorf(hi[a]) : ceilf(hi[a]);
    center[a] = (hi[a] + lo[a]) * 0.5f;
    half[a] = (hi[a] - lo[a]) * 0.5f;
  }
}

int ggo_mesh_build(const float* tri, int n, int depth, float* node_box, int* node_first, int* tri_next,
                   float* obb_min, float* obb_max) {
  if (depth < 1 || depth > 8 || n < 1) return -1;
  const int total = (int)(powf(8.0f, (float)depth) - 1) / 7;
  float center[3], half[3];
  octree_box(tri, n, center, half);
  int* first_child = (int*)calloc((size_t)total, sizeof(int));
  if (!first_child) return -1;
  for (int j = 0; j < total; ++j) {
    node_first[j] = -1;
    for (int a = 0; a < 6; ++a) node_box[6 * j + a] = 0.0f;
  }
  /* BuildOctreeOnDevice */
  for (int a = 0; a < 3; ++a) {
    node_box[a] = center[a];
    node_box[3 + a] = half[a];
  }
  first_child[0] = 1;
  int current = 1;
  for (int i = 2; i <= depth; ++i) {
    int through = ((int)pow(8, i) - 1) / 7;
    int in_level = through - ((int)pow(8.0, i - 1) - 1) / 7;
    int cluster = 0;
    for (int j = 0; j < in_level; ++j) {
      int mother = (((current + 1) + 6) >> 3) - 1;
      float step[3];
      for (int a = 0; a < 3; ++a) step[a] = (float)(half[a] / powf(2.0f, (float)i - 1.0f));
      float ox = ((cluster % 8) & 1) ? step[0] : -step[0];
      float oy = ((cluster % 8) & 2) ? step[1] : -step[1];
      float oz = ((cluster % 8) & 4) ? step[2] : -step[2];
      node_box[6 * current + 0] = node_box[6 * mother + 0] + ox;
      node_box[6 * current + 1] = node_box[6 * mother + 1] + oy;
      node_box[6 * current + 2] = node_box[6 * mother + 2] + oz;
      node_box[6 * current + 3] = step[0];
      node_box[6 * current + 4] = step[1];
      node_box[6 * current + 5] = step[2];
      first_child[current] = 8 * (current + 1) - 7;
      ++current;
      ++cluster;
    }
  }
  /* InsertTrianglesOnDevice: each triangle goes to the deepest node its
   * bounding sphere does not straddle; lists are built by prepending */
  for (int t = 0; t < n; ++t) {
    float c[3], r;
    triangle_sphere(tri + 9 * t, c, &r);
    int cur = 0, next = 0;
    for (int i = 0; i < depth; ++i) {
      int index = 0, straddle = 0;
      for (int j = 0; j < 3; ++j) {
        float delta = c[j] - node_box[6 * cur + j];
        if (fabsf(delta) <= r) {
          straddle = 1;
          break;
        }
        if (delta > 0.0f) index |= (1 << j);
      }
      if (!straddle && i != depth - 1) {
        next = first_child[cur] + index;
      } else {
        tri_next[t] = node_first[cur];
        node_first[cur] = t;
        break;
      }
      cur = next;
    }
  }
  free(first_child);
  for (int a = 0; a < 3; ++a) {
    obb_min[a] = center[a] - half[a];
    obb_max[a] = center[a] + half[a];
  }
  return total;
}

/* MollerTrumboreRayTriangle */
static int moller_trumbore(p3 pi, p3 d, const float* t, p3* po) {
  p3 e1 = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
  p3 e2 = {t[6] - t[0], t[7] - t[1], t[8] - t[2]};
  p3 h = {(d.y * e2.z) - (d.z * e2.y), (d.z * e2.x) - (d.x * e2.z), (d.x * e2.y) - (d.y * e2.x)};
  float a = e1.x * h.x;
  a += e1.y * h.y;
  a += e1.z * h.z;
  if (a > -EPS6 && a < EPS6) return 0;
  float f = 1.0f / a;
  p3 s = {pi.x - t[0], pi.y - t[1], pi.z - t[2]};
  float sh = s.x * h.x;
  sh += s.y * h.y;
  sh += s.z * h.z;
  float u = f * sh;
  if (u < 0.0f || u > 1.0f) return 0;
  p3 q = {(s.y * e1.z) - (s.z * e1.y), (s.z * e1.x) - (s.x * e1.z), (s.x * e1.y) - (s.y * e1.x)};
  float qd = q.x * d.x;
  qd += q.y * d.y;
  qd += q.z * d.z;
  float v = f * qd;
  if (v < 0.0f || u + v > 1.0f) return 0;
  float qe = q.x * e2.x;
  qe += q.y * e2.y;
  qe += q.z * e2.z;
  float tt = f * qe;
  if (tt > EPS6) {
    po->x = pi.x + d.x * tt;
    po->y = pi.y + d.y * tt;
    po->z = pi.z + d.z * tt;
    return 1;
  }
  return 0;
}

/* TestLineNode */
static int line_node(p3 p, p3 d, const float* box) {
  float mn[3] = {box[0] - box[3], box[1] - box[4], box[2] - box[5]};
  float mx[3] = {box[0] + box[3], box[1] + box[4], box[2] + box[5]};
  float tmin = 0.0f, tmax = FLT_MAX;
  float dir[3] = {d.x, d.y, d.z}, pos[3] = {p.x, p.y, p.z};
  for (int i = 0; i < 3; ++i) {
    if (fabsf(dir[i]) < EPS6) {
      if (pos[i] < mn[i] || pos[i] > mx[i]) return 0;
    } else {
      float ood = 1.0f / dir[i];
      float t1 = (mn[i] - pos[i]) * ood;
      float t2 = (mx[i] - pos[i]) * ood;
      if (t1 > t2) {
        float tmp = t2;
        t2 = t1;
        t1 = tmp;
      }
      if (t1 > tmin) tmin = t1;
      if (t2 < tmax) tmax = t2;
      if (tmin > tmax) return 0;
    }
  }
  return 1;
}

int ggo_mesh_hits(const ggo_mesh* m, const float* pp, const float* dd, float* out) {
  p3 p = {pp[0], pp[1], pp[2]}, d = {dd[0], dd[1], dd[2]};
  p3 pos[MAX_TRIANGLE_INTERACTION];
  float near[MAX_TRIANGLE_INTERACTION];
  int n = 0;
  for (int k = 0; k < m->n_nodes; ++k) {
    if (m->node_first[k] < 0) continue;
    if (!line_node(p, d, m->node_box + 6 * k)) continue;
    for (int t = m->node_first[k]; t >= 0; t = m->tri_next[t]) {
      p3 q;
      if (!moller_trumbore(p, d, m->tri + 9 * t, &q)) continue;
      float d2 = (q.x - p.x) * (q.x - p.x) + (q.y - p.y) * (q.y - p.y) + (q.z - p.z) * (q.z - p.z);
      int shift = 0;
      if (n == 0) {
        pos[0] = q;
        near[0] = d2;
      } else {
        for (int i = n - 1; i >= 0; --i) {
          if (d2 < near[i]) shift++;
          else i = 0;
        }
        for (int j = 0; j < shift; ++j) {
          pos[n - j] = pos[n - j - 1];
          near[n - j] = near[n - j - 1];
        }
        pos[n - shift] = q;
        near[n - shift] = d2;
      }
      n++;
      /* The reference leaves only the triangle loop at 16 hits and then
       * writes past its 16-entry array in later nodes; the oracle stops. */
      if (n == MAX_TRIANGLE_INTERACTION) goto done;
    }
  }
done:
  if (out)
    for (int i = 0; i < n; ++i) {
      out[3 * i] = pos[i].x;
      out[3 * i + 1] = pos[i].y;
      out[3 * i + 2] = pos[i].z;
    }
  return n;
}
