/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * Oracle (test infrastructure only): the accumulation kernels of reference src/gpu.cu restated on the CPU.
 *  - k_reduce<Bits> / k_reduce<Bytes> (gpu.cu:76-114): double sums in ascending direction order of float products
 *    (dotf without fma; byte mode: x byte / 255.f).
 *  - k_series (gpu.cu:119-149): in FP32 as the reference's kernel accumulates (fma per direction, in direction order),
 *    and the FP64 contraction every GPU form is measured against.
 *  - k_view / k_iso (gpu.cu:171-274): per sensor the class weights, log distances, minima, own surface, reach, glazing
 *    and the isovist. The reference sums a warp of 32 lanes through a shuffle tree; the library's wave64 kernels sum
 *    64 lanes (lane l takes directions l, l + 64, ...) through the tree s = 32, 16, ..., 1, and so does this oracle.
 *    Distances' logs are double log here (the reference's float logf: ~1e-7 relative in a log mean).
 */
#include "orc.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define LANES 64

static float dotf(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
/* the visibility of (sensor i, direction d) as the fraction that arrives (gpu.cu:76-92): Bits word-major, 1 for a set
 * bit; Bytes direction-major (byte mode), byte / 255.f; 0 = blocked or culled */
typedef struct { const uint32_t* bits; const uint8_t* bytes; } vis_t;
static float vis_at(const vis_t* v, long nc, long i, unsigned d) {
  if (v->bytes) { const uint8_t b = v->bytes[(size_t)d * nc + i]; return b ? (float)b / 255.f : 0.f; }
  return ((v->bits[(size_t)(d >> 5) * nc + i] >> (d & 31)) & 1u) ? 1.f : 0.f;
}

static void reduce(const vis_t* v, long nc, int nDir, const float* N, const float* D, const float* W, int useCos, double* out) {
  for (long i = 0; i < nc; ++i) {
    const float* n = N + 3 * i;
    double acc = 0;
    if (v->bits) {   /* the set bits, ascending */
      for (int w = 0; w < (nDir + 31) / 32; ++w)
        for (uint32_t m = v->bits[(size_t)w * nc + i]; m; m &= m - 1) {
          const int d = w * 32 + __builtin_ctz(m);
          const float c = useCos ? fmaxf(dotf(n, D + 3 * d), 0.f) : 1.f;
          acc += (double)(c * 1.f) * (double)W[d];
        }
    } else {
      for (int d = 0; d < nDir; ++d) {
        const float T = vis_at(v, nc, i, (unsigned)d);
        if (T == 0.f) continue;
        const float c = useCos ? fmaxf(dotf(n, D + 3 * d), 0.f) : 1.f;
        acc += (double)(c * T) * (double)W[d];
      }
    }
    out[i] = acc;
  }
}
void orc_reduce(const uint32_t* bits, long nc, int nDir, const float* N, const float* D, const float* W, int useCos, double* out) {
  const vis_t v = {bits, NULL};
  reduce(&v, nc, nDir, N, D, W, useCos, out);
}
void orc_reduce_bytes(const uint8_t* bytes, long nc, int nDir, const float* N, const float* D, const float* W, int useCos, double* out) {
  const vis_t v = {NULL, bytes};
  reduce(&v, nc, nDir, N, D, W, useCos, out);
}

typedef struct {
  vis_t vis; long nc; const float *N, *D; unsigned nSky; const int* row; const float* wsky; const float* S; int T;
  const float *g, *G; unsigned sun0, nSun; const int* sunStep; const float *sunVec, *sunW; double* E;
  long next; pthread_mutex_t mu;
} s64_job;

static void* s64_worker(void* arg) {
  s64_job* j = arg;
  const int T = j->T;
  double* acc = malloc(sizeof(double) * (size_t)T);
  for (;;) {
    pthread_mutex_lock(&j->mu);
    const long i = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (i >= j->nc) break;
    const float* n = j->N + 3 * i;
    for (int t = 0; t < T; ++t) acc[t] = 0;
    for (unsigned d = 0; d < j->nSky; ++d) {
      const float Tv = vis_at(&j->vis, j->nc, i, d);
      if (Tv == 0.f) continue;
      const double c = (double)((fmaxf(dotf(n, j->D + 3 * d), 0.f) * Tv) * j->wsky[d]);
      const float* s = j->S + (size_t)j->row[d] * T;
      for (int t = 0; t < T; ++t) acc[t] += c * (double)s[t];
    }
    double* Ei = j->E + (size_t)i * T;
    for (int t = 0; t < T; ++t) Ei[t] = acc[t] + (double)j->G[i] * (double)j->g[t];
    for (unsigned q = 0; q < j->nSun; ++q) {
      const float Tv = vis_at(&j->vis, j->nc, i, j->sun0 + q);
      if (Tv != 0.f) Ei[j->sunStep[q]] += (double)(fmaxf(dotf(n, j->sunVec + 3 * q), 0.f) * Tv) * (double)j->sunW[q];
    }
  }
  free(acc);
  return NULL;
}

static void series64(const vis_t* v, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                     const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                     const float* sunVec, const float* sunW, double* E, int nthreads) {
  s64_job j = {*v, nc, N, D, nSky, row, wsky, S, T, g, G, sun0, nSun, sunStep, sunVec, sunW, E, 0, PTHREAD_MUTEX_INITIALIZER};
  pthread_mutex_init(&j.mu, NULL);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, s64_worker, &j);
  s64_worker(&j);
  for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&j.mu);
}

void orc_series64(const uint32_t* bits, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                  const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                  const float* sunVec, const float* sunW, double* E, int nthreads) {
  const vis_t v = {bits, NULL};
  series64(&v, nc, N, D, nSky, row, wsky, S, T, g, G, sun0, nSun, sunStep, sunVec, sunW, E, nthreads);
}
void orc_series64_bytes(const uint8_t* bytes, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                        const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                        const float* sunVec, const float* sunW, double* E, int nthreads) {
  const vis_t v = {NULL, bytes};
  series64(&v, nc, N, D, nSky, row, wsky, S, T, g, G, sun0, nSun, sunStep, sunVec, sunW, E, nthreads);
}

static void series32(const vis_t* v, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                     const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                     const float* sunVec, const float* sunW, float* E) {
  for (long i = 0; i < nc; ++i) {
    const float* n = N + 3 * i;
    float* Ei = E + (size_t)i * T;
    for (int t = 0; t < T; ++t) Ei[t] = 0.f;
    for (unsigned d = 0; d < nSky; ++d) {
      const float Tv = vis_at(v, nc, i, d);
      if (Tv == 0.f) continue;
      const float c = (fmaxf(dotf(n, D + 3 * d), 0.f) * Tv) * wsky[d];
      const float* s = S + (size_t)row[d] * T;
      for (int t = 0; t < T; ++t) Ei[t] = fmaf(c, s[t], Ei[t]);
    }
    for (int t = 0; t < T; ++t) Ei[t] = fmaf(G[i], g[t], Ei[t]);
    for (unsigned q = 0; q < nSun; ++q) {
      const float Tv = vis_at(v, nc, i, sun0 + q);
      if (Tv != 0.f) Ei[sunStep[q]] = fmaf(fmaxf(dotf(n, sunVec + 3 * q), 0.f) * Tv, sunW[q], Ei[sunStep[q]]);
    }
  }
}
void orc_series32(const uint32_t* bits, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                  const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                  const float* sunVec, const float* sunW, float* E) {
  const vis_t v = {bits, NULL};
  series32(&v, nc, N, D, nSky, row, wsky, S, T, g, G, sun0, nSun, sunStep, sunVec, sunW, E);
}
void orc_series32_bytes(const uint8_t* bytes, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                        const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                        const float* sunVec, const float* sunW, float* E) {
  const vis_t v = {NULL, bytes};
  series32(&v, nc, N, D, nSky, row, wsky, S, T, g, G, sun0, nSun, sunStep, sunVec, sunW, E);
}

/* k_series_dc (gpu.cu:153-169): reflected light added to a series, E[i T + t] += sum_b dc[b nc + i] Sdc[b T + t]: E32 in
 * FP32 as the reference's kernel (an fma chain over the bins, then added), E64 in double (either may be NULL) */
void orc_series_dc(const float* dc, long nc, unsigned nBins, const float* Sdc, int T, float* E32, double* E64) {
  for (long i = 0; i < nc; ++i)
    for (int t = 0; t < T; ++t) {
      float e = 0.f;
      double e64 = 0.0;
      for (unsigned b = 0; b < nBins; ++b) {
        const float c = dc[(size_t)b * nc + i], s = Sdc[(size_t)b * T + t];
        e = fmaf(c, s, e);
        e64 += (double)c * (double)s;
      }
      if (E32) E32[(size_t)i * T + t] += e;
      if (E64) E64[(size_t)i * T + t] += e64;
    }
}

/* lane values summed as the wave64 tree: v[l] += v[l + s] for s = 32 .. 1, result in lane 0 */
static double tree64(double* v) {
  for (int s = LANES / 2; s; s >>= 1)
    for (int l = 0; l < s; ++l) v[l] += v[l + s];
  return v[0];
}

static float det3(const float a[3], const float b[3], const float c[3]) {
  const float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  return a[0] * (u[1] * v[2] - u[2] * v[1]) + a[1] * (u[2] * v[0] - u[0] * v[2]) + a[2] * (u[0] * v[1] - u[1] * v[0]);
}

/* iso_lane (gpu.cu:204-229) for wave64 lanes */
static double iso_sensor(const uint8_t* cls, const float* dist, long nc, long i, const orc_view* v, unsigned i0, unsigned dim, float farF) {
  const unsigned gi = i0 + (unsigned)i, element = gi / v->nS, side = v->nS == 2 ? gi % 2 : (unsigned)v->back;
  const unsigned nAz = v->nAz, nEl = v->nEl;
#define LEN(k) (cls[(size_t)(k) * nc + i] == 254 ? 0.f : fminf(dist[(size_t)(k) * nc + i], farF))
  double lane[LANES];
  for (unsigned ln = 0; ln < LANES; ++ln) {
    double sum = 0;
    for (unsigned base = 0; base < nAz; base += LANES) {
      const unsigned col = base + ln < nAz ? base + ln : nAz - 1, c1 = col + 1 == nAz ? 0 : col + 1;
      const int on = base + ln < nAz && (v->world || col + 1 < nAz);
      float a[3], b[3];
      orc_view_cell(col, col, v->rowSin[0], v->rowSin[1], v->az0, v->dAz, v->jitter, element, side, a);
      orc_view_cell(c1, c1, v->rowSin[0], v->rowSin[1], v->az0, v->dAz, v->jitter, element, side, b);
      float l0x = LEN(col), l0y = LEN(c1);
      if (dim == 2) {
        if (on) sum += (double)(0.5f * l0x * l0y * (a[1] * b[0] - a[0] * b[1]));
        continue;
      }
      for (unsigned r = 1; r < nEl; ++r) {
        const unsigned d = r * nAz + col;
        const float l1x = LEN(r * nAz + col), l1y = LEN(r * nAz + c1);
        float c[3], e[3];
        orc_view_cell(d, col, v->rowSin[2 * r], v->rowSin[2 * r + 1], v->az0, v->dAz, v->jitter, element, side, c);
        orc_view_cell(d - col + c1, c1, v->rowSin[2 * r], v->rowSin[2 * r + 1], v->az0, v->dAz, v->jitter, element, side, e);
        if (on) sum += (double)(l0x * l0y * l1x * det3(a, b, c) + l0y * l1y * l1x * det3(b, e, c));
        memcpy(a, c, 12); memcpy(b, e, 12);
        l0x = l1x; l0y = l1y;
      }
    }
    lane[ln] = sum;
  }
#undef LEN
  const double s = tree64(lane);
  return dim == 2 ? s : s / 6;
}

void orc_reduce_view(const uint8_t* cls, const float* dist, long nc, unsigned K, unsigned nAz, const double* wRow, unsigned nClass,
                     double farDist, const float* reach, unsigned nReach, int glazing, unsigned isoDim, const orc_view* v, unsigned i0,
                     double* out) {
  const unsigned NA = 3 * nClass + 1 + nReach + (glazing ? 1 : 0) + (isoDim ? 1 : 0);
  const float farF = (float)farDist;
  const double lnFar = log(farDist);
  for (long i = 0; i < nc; ++i) {
    double w[19][LANES], wl[19][LANES], rc[8][LANES], own[LANES], gz[LANES];
    float mn[19];
    memset(w, 0, sizeof w); memset(wl, 0, sizeof wl); memset(rc, 0, sizeof rc); memset(own, 0, sizeof own); memset(gz, 0, sizeof gz);
    for (unsigned c = 0; c < 19; ++c) mn[c] = INFINITY;
    for (unsigned ln = 0; ln < LANES; ++ln)
      for (unsigned d = ln; d < K; d += LANES) {
        unsigned c = cls[(size_t)d * nc + i];
        const float t = dist[(size_t)d * nc + i];
        const double wd = wRow[d / nAz];
        if (c == 254) { own[ln] += wd; continue; }
        if (c & 0x80) { gz[ln] += wd; c &= 0x7f; }
        const float dd = fminf(fmaxf(t, 0.1f), farF);
        w[c][ln] += wd;
        wl[c][ln] += wd * (dd == farF ? lnFar : dd == 0.1f ? -2.302585092994045684 : log((double)dd));
        mn[c] = fminf(mn[c], dd);
        for (unsigned r = 0; r < nReach; ++r) if (t >= reach[r]) rc[r][ln] += wd;
      }
    double* o = out + (size_t)i * NA;
    for (unsigned c = 0; c < nClass; ++c) { o[3 * c] = tree64(w[c]); o[3 * c + 1] = tree64(wl[c]); o[3 * c + 2] = mn[c]; }
    o[3 * nClass] = tree64(own);
    for (unsigned r = 0; r < nReach; ++r) o[3 * nClass + 1 + r] = tree64(rc[r]);
    if (glazing) o[3 * nClass + 1 + nReach] = tree64(gz);
    if (isoDim) o[NA - 1] = iso_sensor(cls, dist, nc, i, v, i0, isoDim, farF);
  }
}
