/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * Oracle (test infrastructure only): the ray semantics of reference src/trace_dev.cu (__raygen__bits, __raygen__view)
 * and src/params.h restated on the CPU, over a host LBVH (Karras 2012) whose build is the one the GPU build must
 * reproduce bit for bit.
 *
 * What is restated, and how it differs from the reference where the reference cannot be reproduced off OptiX:
 *  - offset_ray (trace_dev.cu:16-25), tmin 1e-5, the hemisphere cull (:101), the own-plane rule (:53-88) with its
 *    constants (2^-22 x the triangle's extent, the 1/32 m gate, at most 3 own planes passed), views' frame, cells,
 *    jitter hash, virtual ground and closest-hit loop (:177-232) with labels and panes.
 *  - The triangle test is the watertight test of Woop, Benthin & Wald (JCGT 2(1), 2013), not OptiX's hardware test; a
 *    ray through an edge or a vertex goes to one of the triangles that share it by a top-left rule, as hardware
 *    tests do (OptiX's own choice of triangle is not documented).
 *  - The first hit of a visibility ray: OptiX reports any hit; here a ray is open when every hit on [1e-5, inf) is an
 *    own-plane hit (gate, plane, facing) and they lie at 3 distinct distances at most. That is the reference's walk for
 *    every order OptiX may report hits in, except exact ties of an own-plane hit with a blocker.
 *  - Closest hits break exact ties of t by the lower triangle index.
 *  - rsqrtf / sincospif (CUDA fast math) are 1 / sqrtf and a polynomial sincospi (Taylor to degree 9/10 in float fma),
 *    the same operations the HIP kernels run; logs of view distances are double log.
 * All float arithmetic here is compiled without contraction (-ffp-contract=off); fmaf is used only where named.
 */
#include "orc.h"
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- small helpers */
static float f_as(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_as(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* offset_ray (trace_dev.cu:16-25): 256 ulps along n per axis, or 2^-16 n within 1/32 m of the origin */
void orc_offset_ray(const float p[3], const float n[3], float out[3]) {
  for (int a = 0; a < 3; ++a) {
    const int o = (int)(256.f * n[a]);
    const float stepped = f_as((uint32_t)((int32_t)u_as(p[a]) + (p[a] < 0.f ? -o : o)));
    out[a] = fabsf(p[a]) < (1.f / 32.f) ? p[a] + (1.f / 65536.f) * n[a] : stepped;
  }
}

/* pcg3d (Jarzynski & Olano 2020), params.h:64-70 */
void orc_pcg3d(uint32_t x, uint32_t y, uint32_t z, uint32_t out[3]) {
  x = x * 1664525u + 1013904223u; y = y * 1664525u + 1013904223u; z = z * 1664525u + 1013904223u;
  x += y * z; y += z * x; z += x * y;
  x ^= x >> 16; y ^= y >> 16; z ^= z >> 16;
  x += y * z; y += z * x; z += x * y;
  out[0] = x; out[1] = y; out[2] = z;
}
static float u01(uint32_t v) { return (float)(v >> 8) * (1.f / 16777216.f); }

/* sin(pi x), cos(pi x) in float: reduce to |r| <= 1/4 by quarter turns, Taylor polynomials (degree 9 / 10) by fmaf */
static void sincospi_f(float x, float* s, float* c) {
  const float q = rintf(x * 2.f);          /* nearest quarter turn (x * 2 is exact) */
  const float r = x - q * 0.5f;            /* exact: q / 2 is a nearby multiple of 1/2 */
  const int k = (int)((long long)q & 3);
  const float r2 = r * r;
  float ps = 8.21458866e-2f;               /* pi^9 / 9! */
  ps = fmaf(ps, r2, -5.99264528e-1f);      /* -pi^7 / 7! */
  ps = fmaf(ps, r2, 2.55016404e+0f);       /* pi^5 / 5! */
  ps = fmaf(ps, r2, -5.16771278e+0f);      /* -pi^3 / 3! */
  ps = fmaf(ps, r2, 3.14159274e+0f);       /* pi */
  const float sr = ps * r;
  float pc = -2.58068914e-2f;              /* -pi^10 / 10! */
  pc = fmaf(pc, r2, 2.35330630e-1f);       /* pi^8 / 8! */
  pc = fmaf(pc, r2, -1.33526277e+0f);      /* -pi^6 / 6! */
  pc = fmaf(pc, r2, 4.05871213e+0f);       /* pi^4 / 4! */
  pc = fmaf(pc, r2, -4.93480220e+0f);      /* -pi^2 / 2! */
  const float cr = fmaf(pc, r2, 1.f);
  switch (k) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

/* view_cell (params.h:77-85): the cell's point in the field frame (x right, y forward, z up) */
void orc_view_cell(unsigned d, unsigned col, float rsLo, float rsHi, float az0, float dAz, int jitter, unsigned element,
                   unsigned side, float out[3]) {
  float a = 0.5f, s = 0.5f * (rsLo + rsHi);
  if (jitter) {
    uint32_t h[3];
    orc_pcg3d(element, d, side, h);
    a = u01(h[0]);
    s = rsLo + u01(h[1]) * (rsHi - rsLo);
  }
  float sa, ca;
  sincospi_f((az0 + ((float)col + a) * dAz) * (1.f / 180.f), &sa, &ca);
  const float cc = sqrtf(fmaxf(0.f, 1.f - s * s));
  out[0] = cc * sa; out[1] = cc * ca; out[2] = s;
}

/* ---------------------------------------------------------------- the watertight triangle test */
typedef struct {
  float o[3], d[3];
  int kx, ky, kz;
  float Sx, Sy, Sz;
  float inv[3], oNear[3], oFar[3];   /* box test: per axis the padded plane offsets, near / far by the sign of d */
} ray_t;

static void ray_setup(ray_t* r, const float o[3], const float d[3], float pad) {
  memcpy(r->o, o, 12); memcpy(r->d, d, 12);
  const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
  r->kz = ax >= ay ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
  r->kx = r->kz == 2 ? 0 : r->kz + 1;
  r->ky = r->kx == 2 ? 0 : r->kx + 1;
  if (d[r->kz] < 0.f) { const int t = r->kx; r->kx = r->ky; r->ky = t; }
  r->Sx = d[r->kx] / d[r->kz];
  r->Sy = d[r->ky] / d[r->kz];
  r->Sz = 1.f / d[r->kz];
  for (int a = 0; a < 3; ++a) {
    const float di = d[a] != 0.f ? d[a] : (signbit(d[a]) ? -1e-30f : 1e-30f);
    r->inv[a] = 1.f / di;
    /* t of the padded near plane = (lo - pad - o) inv for inv > 0: store o + pad / o - pad per side */
    if (r->inv[a] >= 0.f) { r->oNear[a] = o[a] + pad; r->oFar[a] = o[a] - pad; }
    else { r->oNear[a] = o[a] - pad; r->oFar[a] = o[a] + pad; }
  }
}

/* Hit with t in [tmin, tmax]: 1 and t, front (the ray meets the side the winding's normal faces: det > 0). */
static int tri_hit_bc(const ray_t* r, const float* v0, const float* v1, const float* v2, float tmin, float tmax, int cull, float* tOut,
                      int* front, float* bc) {
  const float A[3] = {v0[0] - r->o[0], v0[1] - r->o[1], v0[2] - r->o[2]};
  const float B[3] = {v1[0] - r->o[0], v1[1] - r->o[1], v1[2] - r->o[2]};
  const float C[3] = {v2[0] - r->o[0], v2[1] - r->o[1], v2[2] - r->o[2]};
  const float Ax = A[r->kx] - r->Sx * A[r->kz], Ay = A[r->ky] - r->Sy * A[r->kz];
  const float Bx = B[r->kx] - r->Sx * B[r->kz], By = B[r->ky] - r->Sy * B[r->kz];
  const float Cx = C[r->kx] - r->Sx * C[r->kz], Cy = C[r->ky] - r->Sy * C[r->kz];
  float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  int edge = 0;
  if (U == 0.f || V == 0.f || W == 0.f) {   /* on an edge in float: decide in double (exact products) */
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    edge = U == 0.f || V == 0.f || W == 0.f;
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return 0;
  const float det = U + V + W;
  if (det == 0.f) return 0;
  if (cull && det < 0.f) return 0;
  /* on an edge or a vertex: the top-left rule of each zero edge function (the edge's direction in the winding, times
   * the facing's sign: y > 0, or y == 0 and x > 0) gives it to one of the triangles that share it */
  if (edge) {
    const float sg = det > 0.f ? 1.f : -1.f;
    const float ex[3] = {sg * (Cx - Bx), sg * (Ax - Cx), sg * (Bx - Ax)}, ey[3] = {sg * (Cy - By), sg * (Ay - Cy), sg * (By - Ay)};
    const float e[3] = {U, V, W};
    for (int k = 0; k < 3; ++k)
      if (e[k] == 0.f && !(ey[k] > 0.f || (ey[k] == 0.f && ex[k] > 0.f))) return 0;
  }
  const float Az = r->Sz * A[r->kz], Bz = r->Sz * B[r->kz], Cz = r->Sz * C[r->kz];
  const float T = U * Az + V * Bz + W * Cz;
  const float t = T / det;
  if (!(t >= tmin && t <= tmax)) return 0;
  *tOut = t;
  *front = det > 0.f;
  if (bc) { bc[0] = V / det; bc[1] = W / det; }
  return 1;
}
static int tri_hit(const ray_t* r, const float* v0, const float* v1, const float* v2, float tmin, float tmax, int cull, float* tOut,
                   int* front) {
  return tri_hit_bc(r, v0, v1, v2, tmin, tmax, cull, tOut, front, NULL);
}

/* own_hit (trace_dev.cu:60-72): under the 1/32 m gate, non-degenerate, its plane within 2^-22 x its extent of the
 * origin and facing away from the sensor's normal n */
static int own_plane(const float d[3], const float n[3], float t, int front, const float* v0, const float* v1, const float* v2);
static int own_hit(const float d[3], const float n[3], float t, int front, const float* v0, const float* v1, const float* v2) {
  return t * dot3(n, d) < 1.f / 32.f && own_plane(d, n, t, front, v0, v1, v2);
}
/* own_plane (trace_dev.cu:63-66): non-degenerate, its plane within 2^-22 x its extent of the origin, facing away from n */
static int own_plane(const float d[3], const float n[3], float t, int front, const float* v0, const float* v1, const float* v2) {
  const float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  const float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  const float ng[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  const float l2 = dot3(ng, ng);
  if (!(l2 > 0.f)) return 0;
  const float s = front ? 1.f / sqrtf(l2) : -(1.f / sqrtf(l2));
  const float hn[3] = {s * ng[0], s * ng[1], s * ng[2]};
  const float ext = fmaxf(fmaxf(fmaxf(fabsf(e1[0]), fabsf(e1[1])), fmaxf(fabsf(e1[2]), fabsf(e2[0]))), fmaxf(fabsf(e2[1]), fabsf(e2[2])));
  return t * fabsf(dot3(d, hn)) < ext * 2.3841858e-7f && dot3(n, hn) <= 0.f;
}

/* ---------------------------------------------------------------- host LBVH (Karras 2012) */
struct orc_scene {
  long nt;
  float pad;            /* box padding: 2^-18 x (largest |coordinate| + 1) */
  float* tri;           /* [nt][12]: v0 xyz, v1 xyz, v2 xyz in sorted (leaf) order, 3 floats unused (prim as bits) */
  uint32_t* prim;       /* original index per leaf */
  float* node;          /* [nt - 1 or 1][16]: left box lo xyz hi xyz, right box lo xyz hi xyz, left, right (as bits), 2 unused */
  long nNode;
};

static uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
static int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
static int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
typedef struct { uint64_t key; uint32_t idx; } kv_t;
static int kv_cmp(const void* a, const void* b) {
  const kv_t *x = a, *y = b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : x->idx > y->idx;
}
static int delta(const kv_t* k, long n, long i, long j) {
  if (j < 0 || j >= n) return -1;
  if (k[i].key == k[j].key) return 64 + clz32((uint32_t)i ^ (uint32_t)j);
  return clz64(k[i].key ^ k[j].key);
}

/* Morton code of each triangle's centroid ((v0 + v1) + v2) / 3 in the centroid box, 21 bits per axis, x highest. */
void orc_lbvh_keys(const float* V, const uint32_t* I, long nt, uint64_t* keys) {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float* cen = malloc(sizeof(float) * 3 * (size_t)(nt ? nt : 1));
  for (long t = 0; t < nt; ++t)
    for (int a = 0; a < 3; ++a) {
      const float c = ((V[3 * I[3 * t] + a] + V[3 * I[3 * t + 1] + a]) + V[3 * I[3 * t + 2] + a]) * (1.f / 3.f);
      cen[3 * t + a] = c;
      lo[a] = fminf(lo[a], c);
      hi[a] = fmaxf(hi[a], c);
    }
  float sc[3];
  for (int a = 0; a < 3; ++a) { const float e = hi[a] - lo[a]; sc[a] = e > 0.f ? 1.f / e : 0.f; }
  for (long t = 0; t < nt; ++t) {
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
      const float u = (cen[3 * t + a] - lo[a]) * sc[a];
      q[a] = (uint32_t)fminf(u * 2097152.f, 2097151.f);
    }
    keys[t] = spread21(q[0]) << 2 | spread21(q[1]) << 1 | spread21(q[2]);
  }
  free(cen);
}

/* Karras 2012: node i's children from the sorted keys. child >= 0x80000000: leaf (child & 0x7fffffff). */
void orc_lbvh_topology(const uint64_t* sortedKeys, long n, uint32_t* left, uint32_t* right) {
  kv_t* k = malloc(sizeof(kv_t) * (size_t)n);
  for (long i = 0; i < n; ++i) { k[i].key = sortedKeys[i]; k[i].idx = (uint32_t)i; }
  for (long i = 0; i + 1 < n; ++i) {
    const int d = delta(k, n, i, i + 1) - delta(k, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    long lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    long l = 0;
    for (long t = lmax / 2; t >= 1; t /= 2)
      if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const long j = i + l * d;
    const int dnode = delta(k, n, i, j);
    long s = 0;
    for (long div = 2;; div *= 2) {
      const long t = (l + div - 1) / div;
      if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
      if (t <= 1) break;
    }
    const long g = i + s * d + (d < 0 ? -1 : 0);
    const long mn = i < j ? i : j, mx = i < j ? j : i;
    left[i] = mn == g ? 0x80000000u | (uint32_t)g : (uint32_t)g;
    right[i] = mx == g + 1 ? 0x80000000u | (uint32_t)(g + 1) : (uint32_t)(g + 1);
  }
  free(k);
}

static void box_of_child(const orc_scene* s, uint32_t c, float b[6], const float* boxes) {
  if (c & 0x80000000u) {
    const float* t = s->tri + 12 * (c & 0x7fffffffu);
    for (int a = 0; a < 3; ++a) {
      b[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
      b[3 + a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
    }
  } else memcpy(b, boxes + 6 * c, 24);
}

orc_scene* orc_scene_new(const float* V, long nv, const uint32_t* I, long nt) {
  orc_scene* s = calloc(1, sizeof *s);
  s->nt = nt;
  float mx = 0.f;
  for (long i = 0; i < 3 * nv; ++i) mx = fmaxf(mx, fabsf(V[i]));
  s->pad = (mx + 1.f) * 3.81469727e-6f;   /* 2^-18 */
  if (nt == 0) return s;
  uint64_t* keys = malloc(sizeof(uint64_t) * (size_t)nt);
  orc_lbvh_keys(V, I, nt, keys);
  kv_t* kv = malloc(sizeof(kv_t) * (size_t)nt);
  for (long t = 0; t < nt; ++t) { kv[t].key = keys[t]; kv[t].idx = (uint32_t)t; }
  qsort(kv, (size_t)nt, sizeof(kv_t), kv_cmp);   /* (key, index): the order a stable radix sort gives */
  s->tri = calloc((size_t)nt * 12, sizeof(float));
  s->prim = malloc(sizeof(uint32_t) * (size_t)nt);
  for (long k = 0; k < nt; ++k) {
    const uint32_t t = kv[k].idx;
    s->prim[k] = t;
    for (int c = 0; c < 3; ++c) memcpy(s->tri + 12 * k + 3 * c, V + 3 * I[3 * t + c], 12);
    const float pb = f_as(t);
    s->tri[12 * k + 9] = pb;
  }
  for (long k = 0; k < nt; ++k) keys[k] = kv[k].key;
  const long nn = nt > 1 ? nt - 1 : 1;
  uint32_t* L = malloc(sizeof(uint32_t) * (size_t)nn);
  uint32_t* R = malloc(sizeof(uint32_t) * (size_t)nn);
  if (nt == 1) { L[0] = R[0] = 0x80000000u; }
  else orc_lbvh_topology(keys, nt, L, R);
  /* boxes bottom-up: post-order over the tree from the root */
  float* boxes = malloc(sizeof(float) * 6 * (size_t)nn);
  s->node = calloc((size_t)nn * 16, sizeof(float));
  s->nNode = nn;
  uint32_t* stack = malloc(sizeof(uint32_t) * 2 * (size_t)(nn + 2));
  long sp = 0;
  stack[sp++] = 0;   /* node, visited flag in bit 31 */
  while (sp) {
    const uint32_t e = stack[--sp], i = e & 0x7fffffffu;
    if (!(e & 0x80000000u)) {
      stack[sp++] = i | 0x80000000u;
      if (!(R[i] & 0x80000000u)) stack[sp++] = R[i];
      if (!(L[i] & 0x80000000u)) stack[sp++] = L[i];
      continue;
    }
    float bl[6], br[6];
    box_of_child(s, L[i], bl, boxes);
    box_of_child(s, R[i], br, boxes);
    float* nd = s->node + 16 * i;
    memcpy(nd, bl, 24);
    memcpy(nd + 6, br, 24);
    nd[12] = f_as(L[i]);
    nd[13] = f_as(R[i]);
    for (int a = 0; a < 3; ++a) { boxes[6 * i + a] = fminf(bl[a], br[a]); boxes[6 * i + 3 + a] = fmaxf(bl[3 + a], br[3 + a]); }
  }
  free(stack); free(boxes); free(L); free(R); free(kv); free(keys);
  return s;
}

/* ---- the binned-SAH tree (the MI355X build's default, kit-sim_daylight_amd/csrc/hip/sah.hip), restated on the host.
 * No reference counterpart: OptiX builds its own GAS (reference src/gpu.cu:386-457); any tree gives the same ray results.
 * Rules: a triangle's box by ordered min / max of its vertices, centroid (lo + hi) * 0.5; a range's 16 bins per axis over
 * its centroid bounds [l, h] (ext = h - l, scale = 16 / ext, axis used when ext > 0 and scale finite, bin = min(15,
 * int((c - l) scale))); candidates (a, s), s = 0..14, valid when both sides hold triangles, cost A(L) n(L) + A(R) n(R)
 * with A = (dx dy + dy dz) + dz dx, the least wins, ties to the lower a * 15 + s; none valid: halve (left = n / 2); the
 * partition is stable; an inner node over leaf positions [b, e) split at m has id m - 1, and the root's id swaps with 0;
 * leaf k is the k-th triangle of the final order. Starts from the Morton order of orc_scene_new. */
static uint32_t okey(float f) { const uint32_t u = u_as(f); return (u & 0x80000000u) ? ~u : (u | 0x80000000u); }
static float okf(uint32_t k) { return f_as((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }
static float omin(float a, float b) { return okey(b) < okey(a) ? b : a; }
static float omax(float a, float b) { return okey(b) > okey(a) ? b : a; }

typedef struct { long b, e; uint32_t slot; int depth; } sah_job;

int orc_scene_sah(orc_scene* s) {
  const long nt = s->nt;
  if (nt < 2) return 0;
  const long nn = nt - 1;
  float* pb = malloc(sizeof(float) * 6 * (size_t)nt);
  for (long k = 0; k < nt; ++k) {
    const float* t = s->tri + 12 * k;
    for (int a = 0; a < 3; ++a) {
      pb[6 * k + a] = omin(omin(t[a], t[3 + a]), t[6 + a]);
      pb[6 * k + 3 + a] = omax(omax(t[a], t[3 + a]), t[6 + a]);
    }
  }
  uint32_t* ref = malloc(sizeof(uint32_t) * (size_t)nt);
  uint32_t* tmp = malloc(sizeof(uint32_t) * (size_t)nt);
  for (long k = 0; k < nt; ++k) ref[k] = (uint32_t)k;
  float* raw = calloc((size_t)nn * 16, sizeof(float));
  sah_job* st = malloc(sizeof(sah_job) * (size_t)(nt + 1));
  long sp = 0;
  st[sp++] = (sah_job){0, nt, 0xffffffffu, 1};
  uint32_t root = 0;
  int height = 0;
  while (sp) {
    const sah_job J = st[--sp];
    const long n = J.e - J.b;
    uint32_t cb[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0};
    for (long k = J.b; k < J.e; ++k)
      for (int a = 0; a < 3; ++a) {
        const uint32_t c = okey((pb[6 * ref[k] + a] + pb[6 * ref[k] + 3 + a]) * 0.5f);
        if (c < cb[a]) cb[a] = c;
        if (c > cb[3 + a]) cb[3 + a] = c;
      }
    float lo[3], sc[3];
    int use[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = okf(cb[a]);
      const float ext = okf(cb[3 + a]) - lo[a];
      sc[a] = 16.f / ext;
      use[a] = ext > 0.f && sc[a] < INFINITY;
    }
    /* bins[a][q]: lo xyz, hi xyz keys, count */
    uint32_t bins[3][16][7];
    for (int a = 0; a < 3; ++a)
      for (int q = 0; q < 16; ++q)
        for (int j = 0; j < 7; ++j) bins[a][q][j] = j < 3 ? 0xffffffffu : 0u;
    for (long k = J.b; k < J.e; ++k) {
      const float* b = pb + 6 * ref[k];
      for (int a = 0; a < 3; ++a) {
        if (!use[a]) continue;
        int q = (int)(((b[a] + b[3 + a]) * 0.5f - lo[a]) * sc[a]);
        if (q > 15) q = 15;
        uint32_t* B = bins[a][q];
        for (int j = 0; j < 3; ++j) {
          if (okey(b[j]) < B[j]) B[j] = okey(b[j]);
          if (okey(b[3 + j]) > B[3 + j]) B[3 + j] = okey(b[3 + j]);
        }
        B[6] += 1;
      }
    }
    int bestIdx = -1;
    uint32_t bestCost = 0;
    for (int idx = 0; idx < 45; ++idx) {
      const int a = idx / 15, sIdx = idx % 15;
      if (!use[a]) continue;
      uint32_t L[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0}, R[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0};
      uint32_t nl = 0, nr = 0;
      for (int q = 0; q < 16; ++q) {
        uint32_t* X = q <= sIdx ? L : R;
        for (int j = 0; j < 3; ++j) {
          if (bins[a][q][j] < X[j]) X[j] = bins[a][q][j];
          if (bins[a][q][3 + j] > X[3 + j]) X[3 + j] = bins[a][q][3 + j];
        }
        if (q <= sIdx) nl += bins[a][q][6]; else nr += bins[a][q][6];
      }
      if (!nl || !nr) continue;
      const float ldx = okf(L[3]) - okf(L[0]), ldy = okf(L[4]) - okf(L[1]), ldz = okf(L[5]) - okf(L[2]);
      const float rdx = okf(R[3]) - okf(R[0]), rdy = okf(R[4]) - okf(R[1]), rdz = okf(R[5]) - okf(R[2]);
      const float al = (ldx * ldy + ldy * ldz) + ldz * ldx, ar = (rdx * rdy + rdy * rdz) + rdz * rdx;
      const float cost = al * (float)nl + ar * (float)nr;
      if (bestIdx < 0 || u_as(cost) < bestCost) { bestIdx = idx; bestCost = u_as(cost); }
    }
    long nL = 0;
    if (bestIdx < 0) {
      nL = n / 2;
    } else {
      const int a = bestIdx / 15, sIdx = bestIdx % 15;
      long l = J.b, r = 0;
      for (long k = J.b; k < J.e; ++k) {
        const float* b = pb + 6 * ref[k];
        int q = (int)(((b[a] + b[3 + a]) * 0.5f - lo[a]) * sc[a]);
        if (q > 15) q = 15;
        if (q <= sIdx) ref[l++] = ref[k];
        else tmp[r++] = ref[k];
      }
      nL = l - J.b;
      memcpy(ref + l, tmp, sizeof(uint32_t) * (size_t)r);
    }
    const uint32_t id = (uint32_t)(J.b + nL - 1);
    if (J.slot == 0xffffffffu) root = id;
    else raw[16 * (size_t)(J.slot >> 1) + 12 + (J.slot & 1u)] = f_as(id);
    if (J.depth > height) height = J.depth;
    float* nd = raw + 16 * (size_t)id;
    for (int side = 0; side < 2; ++side) {
      const long b = side ? J.b + nL : J.b, e = side ? J.e : J.b + nL;
      float bl[6] = {0};
      for (long k = b; k < e; ++k)
        for (int a = 0; a < 3; ++a) {
          const float x = pb[6 * ref[k] + a], y = pb[6 * ref[k] + 3 + a];
          bl[a] = k == b ? x : omin(bl[a], x);
          bl[3 + a] = k == b ? y : omax(bl[3 + a], y);
        }
      memcpy(nd + 6 * side, bl, 24);
      if (e - b == 1) nd[12 + side] = f_as(0x80000000u | (uint32_t)b);
      else st[sp++] = (sah_job){b, e, id << 1 | (uint32_t)side, J.depth + 1};
    }
  }
  /* ids to slots: the root and 0 swap */
#define MAPID(x) ((x) == root ? 0u : (x) == 0u ? root : (x))
  for (long i = 0; i < nn; ++i) {
    float* dst = s->node + 16 * (size_t)MAPID((uint32_t)i);
    memcpy(dst, raw + 16 * (size_t)i, 64);
    for (int side = 0; side < 2; ++side) {
      const uint32_t c = u_as(dst[12 + side]);
      if (!(c & 0x80000000u)) dst[12 + side] = f_as(MAPID(c));
    }
    dst[14] = 0.f;
    dst[15] = 0.f;
  }
#undef MAPID
  float* tri = malloc(sizeof(float) * 12 * (size_t)nt);
  uint32_t* prim = malloc(sizeof(uint32_t) * (size_t)nt);
  for (long k = 0; k < nt; ++k) {
    memcpy(tri + 12 * k, s->tri + 12 * (size_t)ref[k], 48);
    prim[k] = s->prim[ref[k]];
  }
  free(s->tri); free(s->prim);
  s->tri = tri; s->prim = prim;
  free(pb); free(ref); free(tmp); free(raw); free(st);
  return height;
}

/* the tree's height: inner nodes on the longest root-to-leaf path */
int orc_scene_height(const orc_scene* s) {
  if (s->nt < 2) return s->nt == 1 ? 1 : 0;
  uint32_t* st = malloc(sizeof(uint32_t) * 2 * (size_t)(s->nNode + 2));
  long sp = 0;
  int h = 0;
  st[sp++] = 0; st[sp++] = 1;
  while (sp) {
    const uint32_t d = st[--sp], n = st[--sp];
    if ((int)d > h) h = (int)d;
    for (int side = 0; side < 2; ++side) {
      const uint32_t c = u_as(s->node[16 * (size_t)n + 12 + side]);
      if (!(c & 0x80000000u)) { st[sp++] = c; st[sp++] = d + 1; }
    }
  }
  free(st);
  return h;
}

void orc_scene_free(orc_scene* s) {
  if (!s) return;
  free(s->tri); free(s->prim); free(s->node); free(s);
}

/* The scene's arrays as the GPU build must produce them (tests compare them bit for bit). */
long orc_scene_nodes(const orc_scene* s, const float** node, const float** tri, const uint32_t** prim) {
  *node = s->node; *tri = s->tri; *prim = s->prim;
  return s->nNode;
}
float orc_scene_pad(const orc_scene* s) { return s->pad; }

/* slab test against the padded box; 1 with the entry distance */
static int box_hit(const ray_t* r, const float* b, float tmin, float tmax, float* tEnter) {
  float t0 = tmin, t1 = tmax;
  for (int a = 0; a < 3; ++a) {
    const float lo = r->inv[a] >= 0.f ? b[a] : b[3 + a], hi = r->inv[a] >= 0.f ? b[3 + a] : b[a];
    const float tn = (lo - r->oNear[a]) * r->inv[a], tf = (hi - r->oFar[a]) * r->inv[a];
    t0 = fmaxf(t0, tn);
    t1 = fminf(t1, tf);
  }
  *tEnter = t0;
  return t0 <= t1;
}

/* Every hit on [tmin, tmax]: f(leaf) returns 1 to stop. Depth-first, near child first. */
typedef int (*leaf_fn)(void* ctx, long leaf);
static void traverse(const orc_scene* s, const ray_t* r, float tmin, const float* tmax, leaf_fn f, void* ctx) {
  if (s->nt == 0) return;
  uint32_t stack[128];
  int sp = 0;
  uint32_t cur = 0;
  for (;;) {
    const float* nd = s->node + 16 * cur;
    float tl, tr;
    const int hl = box_hit(r, nd, tmin, *tmax, &tl), hr = box_hit(r, nd + 6, tmin, *tmax, &tr);
    const uint32_t cl = u_as(nd[12]), cr = u_as(nd[13]);
    uint32_t next[2];
    int nn = 0;
    /* leaves first (their hits shrink tmax), then the nearer inner child */
    if (hl && (cl & 0x80000000u)) { if (f(ctx, cl & 0x7fffffffu)) return; }
    if (hr && (cr & 0x80000000u)) { if (f(ctx, cr & 0x7fffffffu)) return; }
    const int il = hl && !(cl & 0x80000000u), ir = hr && !(cr & 0x80000000u);
    if (il && ir) {
      if (tl <= tr) { next[0] = cl; next[1] = cr; } else { next[0] = cr; next[1] = cl; }
      nn = 2;
    } else if (il) { next[0] = cl; nn = 1; }
    else if (ir) { next[0] = cr; nn = 1; }
    if (nn == 2) stack[sp++] = next[1];
    if (nn >= 1) { cur = next[0]; continue; }
    if (!sp) return;
    cur = stack[--sp];
  }
}

/* ---------------------------------------------------------------- visibility (bits) */
typedef struct {
  const orc_scene* s;
  const ray_t* r;
  const float* n;
  int cull, blocked, nt;
  float ts[3];
} vis_ctx;
static int vis_leaf(void* c, long leaf) {
  vis_ctx* v = c;
  const float* T = v->s->tri + 12 * leaf;
  float t;
  int front;
  if (!tri_hit(v->r, T, T + 3, T + 6, 1e-5f, FLT_MAX, v->cull, &t, &front)) return 0;
  if (!own_hit(v->r->d, v->n, t, front, T, T + 3, T + 6)) { v->blocked = 1; return 1; }
  for (int k = 0; k < v->nt; ++k) if (v->ts[k] == t) return 0;
  if (v->nt == 3) { v->blocked = 1; return 1; }   /* a fourth own plane: the walk gives up (trace_dev.cu:79-87) */
  v->ts[v->nt++] = t;
  return 0;
}

static int visible_from(const orc_scene* s, const float o[3], const float n[3], const float d[3], int cull) {
  if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] <= 0.f) return 0;   /* hemisphere cull */
  ray_t r;
  ray_setup(&r, o, d, s->pad);
  vis_ctx v = {s, &r, n, cull, 0, 0, {0, 0, 0}};
  const float tmax = FLT_MAX;
  traverse(s, &r, 1e-5f, &tmax, vis_leaf, &v);
  return !v.blocked;
}

int orc_visible(const orc_scene* s, const float p[3], const float n[3], const float d[3], int cull) {
  float o[3];
  orc_offset_ray(p, n, o);
  return visible_from(s, o, n, d, cull);
}

typedef struct {
  const orc_scene* s;
  const float *P, *N, *D;
  long nc;
  int nDir, cull;
  uint32_t* bits;
  long next;
  pthread_mutex_t mu;
  /* views */
  const orc_view* v;
  unsigned i0;
  uint8_t* cls;
  float* dist;
} job_t;

static long take(job_t* j, long chunk) {
  pthread_mutex_lock(&j->mu);
  const long i = j->next;
  j->next += chunk;
  pthread_mutex_unlock(&j->mu);
  return i;
}

static void* bits_worker(void* arg) {
  job_t* j = arg;
  const int nWords = (j->nDir + 31) / 32;
  for (;;) {
    const long i0 = take(j, 16);
    if (i0 >= j->nc) break;
    const long i1 = i0 + 16 < j->nc ? i0 + 16 : j->nc;
    for (long i = i0; i < i1; ++i) {
      const float* n = j->N + 3 * i;
      float o[3];
      orc_offset_ray(j->P + 3 * i, n, o);
      for (int w = 0; w < nWords; ++w) {
        uint32_t word = 0;
        for (int k = 0; k < 32 && w * 32 + k < j->nDir; ++k)
          if (visible_from(j->s, o, n, j->D + 3 * (w * 32 + k), j->cull)) word |= 1u << k;
        j->bits[(size_t)w * j->nc + i] = word;
      }
    }
  }
  return NULL;
}

static void run_workers(job_t* j, void* (*fn)(void*), int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  pthread_mutex_init(&j->mu, NULL);
  for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, fn, j);
  fn(j);
  for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&j->mu);
}

void orc_trace_bits(const orc_scene* s, const float* P, const float* N, long nc, const float* D, int nDir, int cull, uint32_t* bits,
                    int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.s = s; j.P = P; j.N = N; j.nc = nc; j.D = D; j.nDir = nDir; j.cull = cull; j.bits = bits;
  run_workers(&j, bits_worker, nthreads);
}

/* ---------------------------------------------------------------- byte mode (transmission) */
/* tn_from_Tn (glass.h:11-13, Radiance m_glass.c): glass transmissivity from the transmittance at normal incidence */
double orc_tn_from_Tn(double Tn) {
  return fmax(0., (sqrt(.8402528435 + .0072522239 * Tn * Tn) - .9166530661) / .0036261119 / Tn);
}
/* glass_TR's T (glass.h:17-25): the pane at incidence cosine c, n = 1.52, both polarisations' two-face series, in double,
 * rounded to float */
void orc_glass_TR(double c, double tn, float* T, float* R) {
  const double n = 1.52;
  const double ci = sqrt((1. - 1. / (n * n)) + c * c / (n * n));   /* cosine inside the glass (Snell) */
  const double a = pow(tn, 1. / ci), a2 = a * a;                   /* the slanted path's transmissivity */
  const double s0 = (c - n * ci) / (c + n * ci), p0 = (ci - n * c) / (ci + n * c);
  const double rs = s0 * s0, rp = p0 * p0;
  *T = (float)(.5 * (1. - rs) * (1. - rs) * a / (1. - rs * rs * a2) + .5 * (1. - rp) * (1. - rp) * a / (1. - rp * rp * a2));
  /* reflected: the first face plus what the internal reflections send back */
  *R = (float)(.5 * rs * (1. + (1. - 2. * rs) * a2) / (1. - rs * rs * a2) + .5 * rp * (1. + (1. - 2. * rp) * a2) / (1. - rp * rp * a2));
}
float orc_glass_T(double c, double tn) {
  float T, R;
  orc_glass_TR(c, tn, &T, &R);
  return T;
}
/* __anyhit__tr's factor (trace_dev.cu:149-165): a thin solid's exit face 1; glass by the angle of the hit triangle's
 * geometric normal (float edges, double products) to d; else tau. slot: rho, tau, tn, flags (as bits). */
static float pass_factor(const float* slot, int front, const float* v0, const float* v1, const float* v2, const float d[3]) {
  uint32_t flags;
  memcpy(&flags, slot + 3, 4);
  if ((flags & 2u) && front == ((flags & 4u) != 0)) return 1.f;
  if (!(flags & 1u)) return slot[1];
  const float e1f[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]}, e2f[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  const double ax = e1f[0], ay = e1f[1], az = e1f[2], bx = e2f[0], by = e2f[1], bz = e2f[2];
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double l = sqrt(nx * nx + ny * ny + nz * nz);
  if (!(l > 0.)) return 0.f;
  return orc_glass_T(fabs(nx * d[0] + ny * d[1] + nz * d[2]) / l, slot[2]);
}

typedef struct { float t; uint32_t prim; long leaf; int front; } hit_t;
typedef struct { const orc_scene* s; const ray_t* r; int cull; hit_t* h; long n, cap; } all_ctx;
static int all_leaf(void* c, long leaf) {
  all_ctx* a = c;
  const float* T = a->s->tri + 12 * leaf;
  float t;
  int front;
  if (!tri_hit(a->r, T, T + 3, T + 6, 1e-5f, FLT_MAX, a->cull, &t, &front)) return 0;
  if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 16; a->h = realloc(a->h, (size_t)a->cap * sizeof(hit_t)); }
  a->h[a->n].t = t; a->h[a->n].prim = a->s->prim[leaf]; a->h[a->n].leaf = leaf; a->h[a->n].front = front;
  ++a->n;
  return 0;
}
static int hit_cmp(const void* x, const void* y) {
  const hit_t *a = x, *b = y;
  if (a->t != b->t) return a->t < b->t ? -1 : 1;
  return a->prim < b->prim ? -1 : a->prim > b->prim;
}

typedef struct {
  const orc_scene* s;
  const uint32_t* triSlot;   /* global slot per triangle (original index) */
  const float* slots;        /* [nSlots][4] */
  long nSlots;
} optics_t;
static const float* slot_of(const optics_t* op, uint32_t prim) {
  const uint32_t si = op->triSlot ? op->triSlot[prim] : 0;
  return op->slots + 4 * (si < (uint32_t)op->nSlots ? si : (uint32_t)op->nSlots - 1);
}
/* the any-hit traversal on [tmin, inf) with payload 1 (trace_dev.cu:133-146 / :140-141), its any-hit calls in the order of
 * the sorted hits: an opaque hit, or a translucent one that takes the product below 1/510, is reported (index, or -1
 * with the product in *T); products in double */
static long anyhit_pass(const optics_t* op, const hit_t* h, long n, float tmin, const float d[3], double* T) {
  double p = 1.0;
  for (long k = 0; k < n; ++k) {
    if (!(h[k].t >= tmin)) continue;
    const float* sl = slot_of(op, h[k].prim);
    if (!(sl[1] > 0.f)) return k;
    const float* V = op->s->tri + 12 * h[k].leaf;
    p *= pass_factor(sl, h[k].front, V, V + 3, V + 6, d);
    if (!(p >= (double)(1.f / 510.f))) return k;
  }
  *T = p;
  return -1;
}
/* past_own_plane (trace_dev.cu:76-88) with translucent: closest hits from 1e-5 past the sensor's own planes, 4 segments
 * at most; the tmin the rest of the ray goes on from, 0 = blocked */
static float past_own(const optics_t* op, const hit_t* h, long n, const float d[3], const float nrm[3]) {
  float tmin = 1e-5f;
  for (int seg = 0; seg < 4; ++seg) {
    long k = 0;
    while (k < n && !(h[k].t >= tmin)) ++k;
    if (k == n) return tmin;
    const float* V = op->s->tri + 12 * h[k].leaf;
    if (!own_hit(d, nrm, h[k].t, h[k].front, V, V + 3, V + 6)) return slot_of(op, h[k].prim)[1] > 0.f ? tmin : 0.f;
    tmin = f_as(u_as(h[k].t) + 1);
  }
  return 0.f;
}
/* __raygen__bytes for one ray (trace_dev.cu:119-145), already past the hemisphere cull: the transmittance byte */
static uint8_t byte_of(const optics_t* op, const float o[3], const float nrm[3], const float d[3], int cull, all_ctx* a) {
  if (op->s->nt == 0) return 255;
  ray_t r;
  ray_setup(&r, o, d, op->s->pad);
  a->s = op->s; a->r = &r; a->cull = cull; a->n = 0;
  const float tmax = FLT_MAX;
  traverse(op->s, &r, 1e-5f, &tmax, all_leaf, a);
  qsort(a->h, (size_t)a->n, sizeof(hit_t), hit_cmp);
  double T = 0.0;
  const long rep = anyhit_pass(op, a->h, a->n, 1e-5f, d, &T);
  if (rep >= 0) {
    const float* V = op->s->tri + 12 * a->h[rep].leaf;
    if (!own_hit(d, nrm, a->h[rep].t, a->h[rep].front, V, V + 3, V + 6)) return 0;
    const float t0 = past_own(op, a->h, a->n, d, nrm);
    if (!(t0 > 0.f) || anyhit_pass(op, a->h, a->n, t0, d, &T) >= 0) return 0;
  }
  return (uint8_t)rintf((float)T * 255.f);
}

typedef struct {
  job_t j;
  optics_t op;
  uint8_t* bytes;
} bytes_job;
static void* bytes_worker(void* arg) {
  bytes_job* b = arg;
  job_t* j = &b->j;
  all_ctx a = {0};
  for (;;) {
    const long i0 = take(j, 16);
    if (i0 >= j->nc) break;
    const long i1 = i0 + 16 < j->nc ? i0 + 16 : j->nc;
    for (long i = i0; i < i1; ++i) {
      const float* n = j->N + 3 * i;
      float o[3];
      orc_offset_ray(j->P + 3 * i, n, o);
      for (int d = 0; d < j->nDir; ++d) {
        const float* dir = j->D + 3 * d;
        const uint8_t v = n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2] > 0.f ? byte_of(&b->op, o, n, dir, j->cull, &a) : 0;
        b->bytes[(size_t)d * j->nc + i] = v;
      }
    }
  }
  free(a.h);
  return NULL;
}

void orc_trace_bytes(const orc_scene* s, const float* P, const float* N, long nc, const float* D, int nDir, int cull, const uint32_t* triSlot,
                     const float* slots, long nSlots, uint8_t* bytes, int nthreads) {
  bytes_job b;
  memset(&b, 0, sizeof b);
  b.j.s = s; b.j.P = P; b.j.N = N; b.j.nc = nc; b.j.D = D; b.j.nDir = nDir; b.j.cull = cull;
  b.op.s = s; b.op.triSlot = triSlot; b.op.slots = slots; b.op.nSlots = nSlots;
  b.bytes = bytes;
  run_workers(&b.j, bytes_worker, nthreads);
}

/* ---------------------------------------------------------------- reflected light (the bounce pass) */
static int closest(const orc_scene* s, const ray_t* r, float tmin, float tmax, int cull, float* t, long* leaf, int* front);
static uint32_t brev32(uint32_t x) {
  uint32_t r = 0;
  for (int k = 0; k < 32; ++k) r |= ((x >> k) & 1u) << (31 - k);
  return r;
}
/* cosine_dir (trace_dev.cu:235-242): sin^2 theta = u, azimuth 2 pi v, Duff et al.'s basis about unit n */
static void cosine_dir(const float n[3], float u, float v, float out[3]) {
  const float sg = copysignf(1.f, n[2]), a = -1.f / (sg + n[2]), b = n[0] * n[1] * a;
  const float t[3] = {1.f + sg * n[0] * n[0] * a, sg * b, -sg * n[0]}, w[3] = {b, sg + n[1] * n[1] * a, -n[1]};
  float sn, cs;
  sincospi_f(2.f * v, &sn, &cs);
  const float r = sqrtf(u), z = sqrtf(fmaxf(0.f, 1.f - u)), rc = r * cs, rs = r * sn;
  for (int k = 0; k < 3; ++k) out[k] = (rc * t[k] + rs * w[k]) + z * n[k];
}
/* sky_bin (trace_dev.cu:246-258, reinhart.cal rbin): 0 below the horizon, else 1 + the Reinhart patch of the direction
 * turned back to true north; angles in double */
static uint32_t sky_bin(const float d[3], uint32_t mf, float alpha, uint32_t nBins, float cosN, float sinN) {
  static const uint32_t tnaz[7] = {30, 30, 24, 24, 18, 12, 6};
  if (!(d[2] >= 0.f)) return 0;
  const double el = asin((double)fminf(d[2], 1.f));
  const uint32_t r = (uint32_t)(el / (double)alpha);
  if (r >= 7 * mf) return nBins - 1;
  const uint32_t band = r / mf, n = tnaz[band] * mf;
  uint32_t b = 1 + (r % mf) * n;
  for (uint32_t q = 0; q < band; ++q) b += tnaz[q] * mf * mf;
  double az = atan2((double)(d[0] * cosN + d[1] * sinN), (double)(d[1] * cosN - d[0] * sinN));
  if (az < 0.) az += 2. * 3.14159265358979323846;
  return b + (uint32_t)(az * ((double)n / (2. * 3.14159265358979323846)) + .5) % n;
}

typedef struct {
  job_t j;
  optics_t op;
  uint32_t i0, M, bounces, mf, nBins;
  float alpha, cosN, sinN;
  float* dc;
} bounce_job;

/* __raygen__bounce (trace_dev.cu:271-311) for sensor i: its M paths, in sample order, into row i of dc [nBins x nc] */
static void bounce_one(const bounce_job* b, long i) {
  const job_t* j = &b->j;
  const orc_scene* sc = j->s;
  const float* n = j->N + 3 * i;
  const uint32_t gi = b->i0 + (uint32_t)i;
  float o0[3];
  orc_offset_ray(j->P + 3 * i, n, o0);
  uint32_t rot[3];
  orc_pcg3d(gi, ~0u, ~0u, rot);
  const float r0 = u01(rot[0]), r1 = u01(rot[1]), weight = 3.14159265f / (float)b->M;
  for (uint32_t s = 0; s < b->M; ++s) {
    const float u = (float)s / (float)b->M + r0, v = u01(brev32(s)) + r1;
    float o[3], d[3];
    memcpy(o, o0, 12);
    cosine_dir(n, u - floorf(u), v - floorf(v), d);
    float w = 1.f;
    uint32_t diffuse = 0, k = 0;
    int reflected = 0;
    for (uint32_t seg = 0; seg < b->bounces + 16; ++seg) {
      ray_t r;
      ray_setup(&r, o, d, sc->pad);
      float t;
      long leaf;
      int front;
      if (sc->nt == 0 || !closest(sc, &r, 1e-5f, FLT_MAX, j->cull, &t, &leaf, &front)) {
        if (reflected) b->dc[(size_t)sky_bin(d, b->mf, b->alpha, b->nBins, b->cosN, b->sinN) * j->nc + i] += w * weight;
        break;
      }
      const float* V = sc->tri + 12 * leaf;
      const float e1[3] = {V[3] - V[0], V[4] - V[1], V[5] - V[2]}, e2[3] = {V[6] - V[0], V[7] - V[1], V[8] - V[2]};
      const float ng[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      const float l2 = dot3(ng, ng);
      if (!(l2 > 0.f)) break;
      const float scl = front ? 1.f / sqrtf(l2) : -(1.f / sqrtf(l2));
      const float hn[3] = {scl * ng[0], scl * ng[1], scl * ng[2]};
      float bc[2] = {0.f, 0.f}, tt;
      int fr;
      tri_hit_bc(&r, V, V + 3, V + 6, t, t, j->cull, &tt, &fr, bc);
      float p[3], lift[3], thr[3], nh[3];
      const float ext = fmaxf(fmaxf(fmaxf(fabsf(e1[0]), fabsf(e1[1])), fmaxf(fabsf(e1[2]), fabsf(e2[0]))), fmaxf(fabsf(e2[1]), fabsf(e2[2])));
      for (int a = 0; a < 3; ++a) {
        p[a] = (V[a] + bc[0] * e1[a]) + bc[1] * e2[a];
        lift[a] = (ext * 9.5367432e-7f) * hn[a];
        nh[a] = -1.f * hn[a];
      }
      orc_offset_ray(p, nh, thr);
      for (int a = 0; a < 3; ++a) thr[a] = thr[a] - lift[a];   /* through(h): past the triangle */
      if (seg == 0 && own_plane(d, n, t, front, V, V + 3, V + 6)) { memcpy(o, thr, 12); continue; }
      const float* sl = slot_of(&b->op, sc->prim[leaf]);
      uint32_t fl;
      memcpy(&fl, sl + 3, 4);
      if ((fl & (2u | 8u)) == 2u && front == ((fl & 4u) != 0)) { memcpy(o, thr, 12); continue; }   /* a thin solid's exit face */
      const int mirror = (fl & 1u) != 0;
      float T = sl[1], R = sl[0];
      if (mirror) {
        orc_glass_TR((double)fabsf(dot3(d, hn)), (double)sl[2], &T, &R);
        if (!(sl[1] > 0.f)) T = 0.f;
      } else if (diffuse == b->bounces) R = 0.f;
      if (!(T + R > 0.f)) break;
      uint32_t rr[3];
      orc_pcg3d(gi, s, k++, rr);
      w *= T + R;
      if (u01(rr[0]) * (T + R) < T) { memcpy(o, thr, 12); continue; }   /* transmitted */
      reflected = 1;
      if (mirror) {
        const float k2 = 2.f * dot3(d, hn);
        for (int a = 0; a < 3; ++a) d[a] = d[a] - k2 * hn[a];
      } else {
        cosine_dir(hn, u01(rr[1]), u01(rr[2]), d);
        if (++diffuse >= 2) {   /* Russian roulette, survival = w */
          uint32_t q[3];
          orc_pcg3d(gi, s, k++, q);
          if (u01(q[0]) >= w) break;
          w = 1.f;
        }
      }
      orc_offset_ray(p, hn, o);
      for (int a = 0; a < 3; ++a) o[a] = o[a] + lift[a];
    }
  }
}
static void* bounce_worker(void* arg) {
  bounce_job* b = arg;
  for (;;) {
    const long i0 = take(&b->j, 1);
    if (i0 >= b->j.nc) break;
    bounce_one(b, i0);
  }
  return NULL;
}

void orc_bounce(const orc_scene* s, const float* P, const float* N, long nc, unsigned i0, const uint32_t* triSlot, const float* slots, long nSlots,
                unsigned bounces, unsigned M, int mf, double north, int cull, float* dc, int nthreads) {
  bounce_job b;
  memset(&b, 0, sizeof b);
  b.j.s = s; b.j.P = P; b.j.N = N; b.j.nc = nc; b.j.cull = cull;
  b.op.s = s; b.op.triSlot = triSlot; b.op.slots = slots; b.op.nSlots = nSlots;
  b.i0 = i0; b.M = M; b.bounces = bounces; b.mf = (uint32_t)mf; b.nBins = (uint32_t)(144 * mf * mf + 2);
  b.alpha = (float)((3.14159265358979323846 / 2.) / (7 * mf + .5));
  const double rad = north * 3.14159265358979323846 / 180;
  b.cosN = (float)cos(rad); b.sinN = (float)sin(rad);
  b.dc = dc;
  memset(dc, 0, sizeof(float) * (size_t)b.nBins * (size_t)nc);
  run_workers(&b.j, bounce_worker, nthreads);
}

/* ---------------------------------------------------------------- views (closest hits) */
typedef struct {
  const orc_scene* s;
  const ray_t* r;
  float tmin, tmax;
  int cull;
  long best;        /* leaf of the closest hit, -1 none */
  uint32_t bestPrim;
  int front;
} ch_ctx;
static int ch_leaf(void* c, long leaf) {
  ch_ctx* h = c;
  const float* T = h->s->tri + 12 * leaf;
  float t;
  int front;
  if (!tri_hit(h->r, T, T + 3, T + 6, h->tmin, h->tmax, h->cull, &t, &front)) return 0;
  const uint32_t prim = h->s->prim[leaf];
  if (h->best >= 0 && t == h->tmax && prim > h->bestPrim) return 0;   /* ties: the lower triangle index */
  h->tmax = t;
  h->best = leaf;
  h->bestPrim = prim;
  h->front = front;
  return 0;
}
/* closest hit on [tmin, tmax]: 1 with t, leaf, front */
static int closest(const orc_scene* s, const ray_t* r, float tmin, float tmax, int cull, float* t, long* leaf, int* front) {
  ch_ctx h = {s, r, tmin, tmax, cull, -1, 0, 0};
  traverse(s, r, tmin, &h.tmax, ch_leaf, &h);
  if (h.best < 0) return 0;
  *t = h.tmax; *leaf = h.best; *front = h.front;
  return 1;
}

static void view_one(const job_t* j, long i) {
  const orc_view* v = j->v;
  const float* n = j->N + 3 * i;
  const unsigned gi = j->i0 + (unsigned)i, element = gi / v->nS, side = v->nS == 2 ? gi % 2 : (unsigned)v->back;
  float o[3];
  orc_offset_ray(j->P + 3 * i, n, o);
  const float inf = f_as(0x7f800000u);
  const int none = !(dot3(n, n) > 0.f);
  float f[3], r[3], u[3];
  if (v->world) { f[0] = -v->sinN; f[1] = v->cosN; f[2] = 0.f; r[0] = v->cosN; r[1] = v->sinN; r[2] = 0.f; u[0] = 0.f; u[1] = 0.f; u[2] = 1.f; }
  else {
    const float l = sqrtf(n[0] * n[0] + n[1] * n[1]);
    memcpy(f, n, 12);
    if (l > 1e-3f) { r[0] = n[1] / l; r[1] = -n[0] / l; r[2] = 0.f; }
    else {
      r[0] = v->cosN; r[1] = v->sinN; r[2] = 0.f;
      if (l > 0.f) {
        const float k = dot3(r, n);
        for (int a = 0; a < 3; ++a) r[a] = r[a] - k * n[a];
        const float s = 1.f / sqrtf(dot3(r, r));
        for (int a = 0; a < 3; ++a) r[a] = s * r[a];
      }
    }
    u[0] = r[1] * f[2] - r[2] * f[1]; u[1] = r[2] * f[0] - r[0] * f[2]; u[2] = r[0] * f[1] - r[1] * f[0];
  }
  const unsigned K = v->nAz * v->nEl;
  for (unsigned d = 0; d < K; ++d) {
    float l[3];
    const unsigned row = d / v->nAz;
    orc_view_cell(d, d % v->nAz, v->rowSin[2 * row], v->rowSin[2 * row + 1], v->az0, v->dAz, v->jitter, element, side, l);
    float dir[3];
    for (int a = 0; a < 3; ++a) dir[a] = (l[0] * r[a] + l[1] * f[a]) + l[2] * u[a];
    uint8_t cl = 2;
    float t = inf;
    if (none || (v->cull && dot3(n, dir) <= 0.f)) cl = 254;
    else if (v->vground && dir[2] < 0.f && !(o[2] > v->groundZ)) { cl = 1; t = 0.1f; }
    else {
      const int plane = v->vground && dir[2] < 0.f;
      const float tg = plane ? (o[2] - v->groundZ) / -dir[2] : inf;
      const float tmax = plane ? (o[2] - v->groundZ + 0.01f) / -dir[2] : FLT_MAX;
      ray_t ry;
      ray_setup(&ry, o, dir, j->s->pad);
      float tmin = 1e-5f;
      unsigned own = 0, panes = 0;
      for (;;) {
        float th;
        long leaf;
        int front;
        if (!closest(j->s, &ry, tmin, tmax, j->cull, &th, &leaf, &front)) { cl = dir[2] >= 0.f ? 0 : 1; t = tg; break; }
        t = th;
        const float* T = j->s->tri + 12 * leaf;
        if (own_hit(dir, n, t, front, T, T + 3, T + 6)) { if (++own == 4) break; }
        else {
          const uint8_t lb = v->label ? v->label[j->s->prim[leaf]] : 2;
          if (lb != 255) { cl = lb; break; }
          if (++panes > 16) break;
        }
        tmin = f_as(u_as(t) + 1);
      }
      if (panes) cl |= 0x80;
    }
    j->cls[(size_t)d * j->nc + i] = cl;
    j->dist[(size_t)d * j->nc + i] = t;
  }
}

static void* view_worker(void* arg) {
  job_t* j = arg;
  for (;;) {
    const long i0 = take(j, 4);
    if (i0 >= j->nc) break;
    for (long i = i0; i < i0 + 4 && i < j->nc; ++i) view_one(j, i);
  }
  return NULL;
}

void orc_trace_view(const orc_scene* s, const float* P, const float* N, long nc, unsigned i0, const orc_view* v, int cull, uint8_t* cls,
                    float* dist, int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.s = s; j.P = P; j.N = N; j.nc = nc; j.v = v; j.i0 = i0; j.cls = cls; j.dist = dist; j.cull = cull;
  run_workers(&j, view_worker, nthreads);
}

/* Brute force over every triangle (no BVH): the hit sets the traversal must reproduce. */
int orc_visible_brute(const float* V, const uint32_t* I, long nt, const float p[3], const float n[3], const float d[3], int cull) {
  float o[3];
  orc_offset_ray(p, n, o);
  if (n[0] * d[0] + n[1] * d[1] + n[2] * d[2] <= 0.f) return 0;
  ray_t r;
  ray_setup(&r, o, d, 0.f);
  float ts[3];
  int nts = 0;
  for (long k = 0; k < nt; ++k) {
    const float *a = V + 3 * I[3 * k], *b = V + 3 * I[3 * k + 1], *c = V + 3 * I[3 * k + 2];
    float t;
    int front;
    if (!tri_hit(&r, a, b, c, 1e-5f, FLT_MAX, cull, &t, &front)) continue;
    if (!own_hit(d, n, t, front, a, b, c)) return 0;
    int seen = 0;
    for (int q = 0; q < nts; ++q) seen |= ts[q] == t;
    if (seen) continue;
    if (nts == 3) return 0;
    ts[nts++] = t;
  }
  return 1;
}
