/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * Oracle (test infrastructure only): reference src/mesh.cpp restated in C for the path's inputs: the bounding-box
 * centre every coordinate is recentred by, sensor points/normals by the physical rules, and the occluder triangles of
 * tri / quad faces.
 */
#include "orc.h"
#include <math.h>
#include <stdlib.h>

typedef struct { double x, y, z; } v3;
static v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 vadd(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static v3 vmul(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static v3 vcross(v3 a, v3 b) { v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; return r; }
static double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static double vlen(v3 a) { return sqrt(vdot(a, a)); }
static v3 vunit(v3 a) { const double l = vlen(a); v3 z = {0, 0, 0}; return l > 0 ? vmul(a, 1 / l) : z; }
static v3 pt(const orc_mesh* m, int i) { v3 r = {m->points[3 * i], m->points[3 * i + 1], m->points[3 * i + 2]}; return r; }

/* mesh.cpp:107-118: per axis (lo + hi) / 2 over every point of every mesh */
void orc_bbox_centre(const orc_mesh* m, int nMesh, double c[3]) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < nMesh; ++k)
    for (long i = 0; i < 3 * m[k].nPoints; ++i) {
      lo[i % 3] = fmin(lo[i % 3], m[k].points[i]);
      hi[i % 3] = fmax(hi[i % 3], m[k].points[i]);
    }
  for (int a = 0; a < 3; ++a) c[a] = lo[a] <= hi[a] ? (lo[a] + hi[a]) / 2 : 0;
}

/* Newell's normal of a polygon: length twice the area, direction by the winding (mesh.cpp:35-43) */
static v3 newell(const orc_mesh* m, const int* idx, int n) {
  v3 N = {0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const v3 a = pt(m, idx[i]), b = pt(m, idx[(i + 1) % n]);
    N.x += (a.y - b.y) * (a.z + b.z);
    N.y += (a.z - b.z) * (a.x + b.x);
    N.z += (a.x - b.x) * (a.y + b.y);
  }
  return N;
}

/* face_geom, physical rules (mesh.cpp:45-71): unit Newell normal (0 when the area is under 1e-12), the signed-area
 * weighted centroid of the fan from corner 0 (the corner mean when that area vanishes) */
static void face_geom(const orc_mesh* m, const int* idx, int n, int ladybug, v3* nrm, double* area, v3* cen) {
  v3 zero = {0, 0, 0};
  *nrm = zero; *area = 0; *cen = zero;
  if (n < 3) return;
  const v3 P0 = pt(m, idx[0]);
  if (ladybug) {   /* mesh.cpp:49-55: the corners' mean; a triangle's cross product; a quad's (0 1 2) + (2 3 1) normals */
    v3 c = zero, N = zero;
    for (int i = 0; i < n; ++i) c = vadd(c, pt(m, idx[i]));
    *cen = vmul(c, 1.0 / n);
    if (n == 3) { N = vcross(vsub(pt(m, idx[1]), P0), vsub(pt(m, idx[2]), P0)); *area = vlen(N) / 2; }
    else if (n == 4) {
      const v3 P1 = pt(m, idx[1]), P2 = pt(m, idx[2]), P3 = pt(m, idx[3]);
      const v3 n1 = vcross(vsub(P1, P0), vsub(P2, P0)), n2 = vcross(vsub(P3, P2), vsub(P1, P2));
      N = vadd(n1, n2);
      *area = (vlen(n1) + vlen(n2)) / 2;
    }
    *nrm = *area < 1e-12 ? zero : vunit(N);
    return;
  }
  const v3 N = newell(m, idx, n), Nu = vunit(N);
  *area = vlen(N) / 2;
  double sum = 0;
  v3 c = zero;
  for (int i = 1; i + 1 < n; ++i) {
    const v3 a = pt(m, idx[i]), b = pt(m, idx[i + 1]);
    const double w = vdot(vcross(vsub(a, P0), vsub(b, P0)), Nu) / 2;
    c = vadd(c, vmul(vadd(vadd(P0, a), b), w / 3));
    sum += w;
  }
  if (sum > 1e-12) c = vmul(c, 1 / sum);
  else {
    c = zero;
    for (int i = 0; i < n; ++i) c = vadd(c, pt(m, idx[i]));
    c = vmul(c, 1.0 / n);
  }
  *cen = c;
  *nrm = *area < 1e-12 ? zero : vunit(N);
}

static void put(float* p, float* n, long* j, v3 q, v3 nr, int sides, double offset, v3 C) {
  if (sides != 1) {   /* front */
    const v3 a = vsub(vadd(q, vmul(nr, offset)), C);
    p[3 * *j] = (float)a.x; p[3 * *j + 1] = (float)a.y; p[3 * *j + 2] = (float)a.z;
    n[3 * *j] = (float)nr.x; n[3 * *j + 1] = (float)nr.y; n[3 * *j + 2] = (float)nr.z;
    ++*j;
  }
  if (sides != 0) {   /* back */
    const v3 a = vsub(vsub(q, vmul(nr, offset)), C), nb = vmul(nr, -1);
    p[3 * *j] = (float)a.x; p[3 * *j + 1] = (float)a.y; p[3 * *j + 2] = (float)a.z;
    n[3 * *j] = (float)nb.x; n[3 * *j + 1] = (float)nb.y; n[3 * *j + 2] = (float)nb.z;
    ++*j;
  }
}

long orc_sensors(const orc_mesh* m, int onVertices, int sides, double offset, const double c[3], float* p, float* n) {
  return orc_sensors_mode(m, onVertices, sides, offset, 0, c, p, n);
}

/* make_sensors_into (mesh.cpp:315-349) by the mode's rules; ladybug: a face of more than 4 corners makes nothing (-1) */
long orc_sensors_mode(const orc_mesh* m, int onVertices, int sides, double offset, int ladybug, const double c[3], float* p, float* n) {
  for (long f = 0; ladybug && f < m->nFaces; ++f) if (m->counts[f] > 4) return -1;
  const v3 C = {c[0], c[1], c[2]};
  long j = 0, off = 0;
  v3* vn = onVertices ? calloc((size_t)m->nPoints, sizeof(v3)) : NULL;
  for (long f = 0; f < m->nFaces; ++f) {
    const int k = m->counts[f];
    const int* idx = m->indices + off;
    off += k;
    v3 nrm, cen;
    double area;
    face_geom(m, idx, k, ladybug, &nrm, &area, &cen);
    if (onVertices) for (int q = 0; q < k; ++q) vn[idx[q]] = vadd(vn[idx[q]], vmul(nrm, area));
    else put(p, n, &j, cen, nrm, sides, offset, C);
  }
  if (onVertices) {
    for (long v = 0; v < m->nPoints; ++v) {
      v3 a = vn[v];
      if (vlen(a) == 0) { a.x = 0; a.y = 0; a.z = 1; }
      put(p, n, &j, pt(m, (int)v), vunit(a), sides, offset, C);
    }
    free(vn);
    return m->nPoints;
  }
  return m->nFaces;
}

/* triangulate_into (mesh.cpp:137-168) for tris and quads: points recentred then float; quad (0,1,2) (0,2,3) */
long orc_triangulate(const orc_mesh* m, const double c[3], float* v, uint32_t* tri, uint32_t base, int* skipped) {
  const v3 C = {c[0], c[1], c[2]};
  for (long i = 0; i < m->nPoints; ++i) {
    const v3 a = vsub(pt(m, (int)i), C);
    v[3 * i] = (float)a.x; v[3 * i + 1] = (float)a.y; v[3 * i + 2] = (float)a.z;
  }
  long made = 0, off = 0;
  for (long f = 0; f < m->nFaces; ++f) {
    const int k = m->counts[f];
    const int* idx = m->indices + off;
    off += k;
    if (k == 3 || k == 4) {
      for (int t = 0; t < k - 2; ++t) {
        tri[3 * made] = base + (uint32_t)idx[0];
        tri[3 * made + 1] = base + (uint32_t)idx[t + 1];
        tri[3 * made + 2] = base + (uint32_t)idx[t + 2];
        ++made;
      }
    } else ++*skipped;
  }
  return made;
}
