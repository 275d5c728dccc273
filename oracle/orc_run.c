/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * Oracle (test infrastructure only): run_daylight of reference src/run.cpp:216-738 restated for the physical mode
 * and the ladybug mode with opaque occluders (no transmission, no reflected light, no --no-occlude): steps, suns,
 * the direction table, the Perez sky per step, the per-quantity weights, the trace, the reduces and the optional FP64
 * series. Pinned end to end by docs/daylight.md:133-150 (config 1) and :163-180.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
static const double kDegRad = ORC_PI / 180.0;
static const double kRf = 0.265074126, kGf = 0.670114631, kBf = 0.064811243;   /* Radiance CIE_rf/gf/bf */
static double bright(const float* c) { return kRf * c[0] + kGf * c[1] + kBf * c[2]; }

static int fail(char* err, int n, const char* msg) { if (err && n > 0) snprintf(err, (size_t)n, "%s", msg); return -1; }

typedef struct {
  double lat, lon, tz, shift;
  int leap;
} clock_t_;

static int run_clock(const orc_daylight_opt* o, int haveW, double wlat, double wlon, double wtz, int leap, clock_t_* c) {
  c->lat = haveW ? wlat : o->lat; c->lon = haveW ? wlon : o->lon; c->tz = haveW ? wtz : o->tz;
  if (!isnan(o->lat)) c->lat = o->lat;
  if (!isnan(o->lon)) c->lon = o->lon;
  if (haveW) c->tz = wtz + (c->lon - wlon) / 15;
  c->shift = haveW && !isnan(o->tz) ? o->tz - c->tz : 0;
  c->leap = haveW ? leap : 0;
  return isnan(c->lat) || isnan(c->lon) || isnan(c->tz) ? -1 : 0;
}

/* run.cpp:362-470, ladybug branch: per step the sun of sun.c at the .wea-rounded site, gendaymtx -d and -s at integer
 * DNI / DHI, float sums over all rows, the averages printed %.3g, broadband x omega x hours / 1000. */
void orc_ladybug_sky(int mf, long T, const int* jd, const double* hour, const float* dni, const float* dhi, double lat, double lon,
                     double fileTz, double grefl, double dur, const int* colOf, int Tup, float* S, float* gS, double* dir, double* dif,
                     double* rad, double* gnd, float* dump) {
  const int P = orc_patch_count(mf);
  float* om = malloc(sizeof(float) * P); float* pa = malloc(sizeof(float) * P); float* pz = malloc(sizeof(float) * P);
  int* pr = malloc(sizeof(int) * P); int* pn = malloc(sizeof(int) * P);
  orc_patches(mf, pa, pz, om, pr, pn, NULL, NULL);
  const double latW = orc_round_fmt(lat, "%.2f"), lonW = orc_round_fmt(-lon, "%.2f"), merW = (double)(int)(-fileTz * 15.0);
  float* dr = malloc(sizeof(float) * 3 * (P + 1));
  float* df = malloc(sizeof(float) * 3 * (P + 1));
  float* accDir = calloc((size_t)P, sizeof(float));
  float* accDif = calloc((size_t)3 * P, sizeof(float));
  float gDir[3] = {0, 0, 0}, gDif[3] = {0, 0, 0};
  for (long t = 0; t < T; ++t) {
    double a, z;
    orc_sun_radiance(latW, lonW, merW, jd[t], hour[t], &a, &z);
    const double dn = (double)(int)dni[t], dh = (double)(int)dhi[t];
    orc_gendaymtx_step(mf, jd[t], a, z, dn, dh, grefl, 1, 0, 0, dr);
    orc_gendaymtx_step(mf, jd[t], a, z, dn, dh, grefl, 0, 1, 0, df);
    for (int ch = 0; ch < 3; ++ch) { gDir[ch] += dr[ch]; gDif[ch] += df[ch]; }
    const int col = colOf ? colOf[t] : -1;
    double mean = 0;
    for (int p = 0; p < P; ++p) {
      accDir[p] += dr[3 * (p + 1)];
      for (int ch = 0; ch < 3; ++ch) accDif[3 * p + ch] += df[3 * (p + 1) + ch];
      if (S && col >= 0) {
        const double L = bright(df + 3 * (p + 1)) + dr[3 * (p + 1)];
        S[(size_t)p * Tup + col] = (float)L;
        mean += L * om[p];
      }
    }
    if (S && col >= 0) gS[col] = (float)(grefl * mean / P);
  }
  const float inv = (float)(1.0 / (double)T);
  const size_t row = 3 * (size_t)(P + 1);
  for (int ch = 0; ch < 3; ++ch) {
    dump[ch] = (float)orc_round_fmt(T > 1 ? gDif[ch] * inv : gDif[ch], "%.3g");
    dump[row + ch] = (float)orc_round_fmt(T > 1 ? gDir[ch] * inv : gDir[ch], "%.3g");
  }
  double mean = 0;
  for (int p = 0; p < P; ++p) {
    const float ad = T > 1 ? accDir[p] * inv : accDir[p];
    float c3[3];
    for (int ch = 0; ch < 3; ++ch) c3[ch] = T > 1 ? accDif[3 * p + ch] * inv : accDif[3 * p + ch];
    const double rd = orc_round_fmt(ad, "%.3g"), rr = orc_round_fmt(c3[0], "%.3g"), rg = orc_round_fmt(c3[1], "%.3g"),
                 rb = orc_round_fmt(c3[2], "%.3g");
    dump[3 * (p + 1)] = (float)rr; dump[3 * (p + 1) + 1] = (float)rg; dump[3 * (p + 1) + 2] = (float)rb;
    dump[row + 3 * (p + 1)] = dump[row + 3 * (p + 1) + 1] = dump[row + 3 * (p + 1) + 2] = (float)rd;
    dir[p] = rd * om[p] * dur / 1000.0;
    dif[p] = (kRf * rr + kGf * rg + kBf * rb) * om[p] * dur / 1000.0;
    rad[p] = dir[p] + dif[p];
    mean += rad[p];
  }
  *gnd = grefl * mean / (double)P;
  free(om); free(pa); free(pz); free(pr); free(pn); free(dr); free(df); free(accDir); free(accDif);
}

int orc_daylight_size(const orc_mesh* m, int nMesh, const float* dni, const float* dhi, long nRows, double wlat, double wlon, double wtz,
                      int leap, const orc_daylight_opt* o, long* nElem, long* nCols, int* Tup, long* nDir) {
  const int nS = o->sides == 2 ? 2 : 1;
  long ne = 0;
  for (int k = 0; k < nMesh; ++k) if (m[k].sensor) ne += o->onVertices ? m[k].nPoints : m[k].nFaces;
  *nElem = ne;
  *nCols = (long)o->nQ * nS;
  clock_t_ c;
  if (run_clock(o, dni != NULL, wlat, wlon, wtz, leap, &c)) return -1;
  orc_step* st = malloc(sizeof(orc_step) * 600000);
  const long T = orc_make_steps(&o->time, dni, dhi, nRows, c.leap, c.shift, st, 600000);
  if (T <= 0) { free(st); return -1; }
  int needW = 0;
  for (int q = 0; q < o->nQ; ++q) needW |= o->quantities[q] != ORC_SUNHOURS;
  int tup = 0, nSun = 0;
  for (long t = 0; t < T; ++t) {
    if (st[t].dni + st[t].dhi > 1e-4) ++tup;
    double alt, azi;
    int up;
    float v[3];
    orc_sun_noaa(c.lat, c.lon, c.tz, st[t].month, st[t].day, st[t].hour, c.leap, &alt, &azi, &up, v);
    nSun += up;
  }
  *Tup = tup;
  const int K = o->ladybug ? 1 : o->patchSamples > 1 ? o->patchSamples : 1;
  *nDir = (needW ? 2L * orc_patch_count(o->mf) * K : 0) + nSun;
  free(st);
  return 0;
}

int orc_daylight(const orc_mesh* m, int nMesh, const float* dni, const float* dhi, long nRows, double wlat, double wlon, double wtz, int leap,
                 const orc_daylight_opt* o, orc_daylight_out* out, char* err, int errLen) {
  int needW = 0, wantVis = 0, wantRad = 0;
  for (int q = 0; q < o->nQ; ++q)
    if (o->quantities[q] != ORC_SUNHOURS) { needW = 1; if (o->quantities[q] == ORC_ILLUMINANCE) wantVis = 1; else wantRad = 1; }
  if (needW && !dni) return fail(err, errLen, "weather required");
  if (o->ladybug && wantVis) return fail(err, errLen, "illuminance is not available in ladybug mode");
  if (o->ladybug && o->mf != 1 && o->mf != 2) return fail(err, errLen, "ladybug mode supports --sky-density 1 or 2 only");
  if (o->mf != 1 && o->mf != 2 && o->mf != 4) return fail(err, errLen, "--sky-density must be 1, 2 or 4");
  const int K = o->ladybug ? 1 : o->patchSamples > 1 ? o->patchSamples : 1;
  { const int s = (int)lround(sqrt((double)K)); if (s * s != K) return fail(err, errLen, "--patch-samples must be a square number"); }
  const int series = o->wantSeries;
  const int seriesVis = series && o->quantities[0] == ORC_ILLUMINANCE;
  clock_t_ ck;
  if (run_clock(o, dni != NULL, wlat, wlon, wtz, leap, &ck)) return fail(err, errLen, "location unknown");

  orc_step* st = malloc(sizeof(orc_step) * 600000);
  const long T = orc_make_steps(&o->time, dni, dhi, nRows, ck.leap, ck.shift, st, 600000);
  if (T <= 0) { free(st); return fail(err, errLen, "no timesteps selected"); }
  double dur = 0;
  for (long t = 0; t < T; ++t) dur += st[t].dt;
  out->durationHours = dur;

  const int nS = o->sides == 2 ? 2 : 1;
  long nElem = 0;
  for (int k = 0; k < nMesh; ++k) if (m[k].sensor) nElem += o->onVertices ? m[k].nPoints : m[k].nFaces;
  const long nSens = nElem * nS;
  if (nSens == 0) { free(st); return fail(err, errLen, "no sensor elements"); }

  /* suns, turned by north */
  double* sAlt = malloc(sizeof(double) * T);
  double* sAzi = malloc(sizeof(double) * T);
  int* sUp = malloc(sizeof(int) * T);
  float* sVec = malloc(sizeof(float) * 3 * T);
  for (long t = 0; t < T; ++t) {
    float v[3];
    orc_sun_noaa(ck.lat, ck.lon, ck.tz, st[t].month, st[t].day, st[t].hour, ck.leap, &sAlt[t], &sAzi[t], &sUp[t], v);
    orc_rotate_z(v, o->north, sVec + 3 * t);
  }

  /* direction table [sky P K | ground = -sky | sun up-steps] */
  const int P = orc_patch_count(o->mf);
  float* pAlt = malloc(sizeof(float) * P); float* pAzi = malloc(sizeof(float) * P); float* omega = malloc(sizeof(float) * P);
  int* prow = malloc(sizeof(int) * P); int* nin = malloc(sizeof(int) * P);
  float* lbv = malloc(sizeof(float) * 3 * P);
  orc_patches(o->mf, pAlt, pAzi, omega, prow, nin, NULL, lbv);
  long nSun = 0;
  for (long t = 0; t < T; ++t) nSun += sUp[t];
  const long nSky = needW ? (long)P * K : 0;
  const long nDir = 2 * nSky + nSun, sun0 = 2 * nSky;
  float* D = malloc(sizeof(float) * 3 * (size_t)(nDir ? nDir : 1));
  int* rowOf = malloc(sizeof(int) * (size_t)(nSky ? nSky : 1));
  for (int p = 0; p < P && needW; ++p)
    for (int k = 0; k < K; ++k) {
      float v[3];
      if (o->ladybug) memcpy(v, lbv + 3 * p, 12);
      else orc_patch_sample(o->mf, p, k, K, v);
      orc_rotate_z(v, o->north, D + 3 * (p * K + k));
      rowOf[p * K + k] = p;
    }
  for (long d = 0; d < nSky; ++d) for (int a = 0; a < 3; ++a) D[3 * (nSky + d) + a] = -D[3 * d + a];
  int* sunStep = malloc(sizeof(int) * (size_t)(nSun ? nSun : 1));
  long* sunOf = malloc(sizeof(long) * T);
  for (long t = 0, j = 0; t < T; ++t) {
    sunOf[t] = -1;
    if (!sUp[t]) continue;
    memcpy(D + 3 * (sun0 + j), sVec + 3 * t, 12);
    sunStep[j] = (int)t;
    sunOf[t] = j++;
  }
  out->nSky = nSky; out->nSun = nSun; out->nDir = nDir;

  /* sky model per step */
  int* colOf = malloc(sizeof(int) * T);
  int Tup = 0;
  for (long t = 0; t < T; ++t) colOf[t] = st[t].dni + st[t].dhi > 1e-4 ? Tup++ : -1;
  out->Tup = Tup;
  float* S = series ? calloc((size_t)P * (Tup ? Tup : 1), sizeof(float)) : NULL;
  float* gS = series ? calloc((size_t)(Tup ? Tup : 1), sizeof(float)) : NULL;
  double* skyRad = calloc((size_t)P, sizeof(double));
  double* skyVis = calloc((size_t)P, sizeof(double));
  double gndRad = 0, gndVis = 0;
  double* sunRad = calloc((size_t)(nSun ? nSun : 1), sizeof(double));
  double* sunVis = calloc((size_t)(nSun ? nSun : 1), sizeof(double));
  float* buf = malloc(sizeof(float) * 3 * (P + 1));
  const double grefl = o->virtualGround ? o->ground : 0;
  double *lbDir = NULL, *lbDif = NULL, lbGnd = 0;
  if (o->ladybug && needW) {
    int* jdv = malloc(sizeof(int) * T); double* hv = malloc(sizeof(double) * T);
    float* dn = malloc(sizeof(float) * T); float* dh = malloc(sizeof(float) * T);
    for (long t = 0; t < T; ++t) { jdv[t] = st[t].jd; hv[t] = st[t].hour; dn[t] = st[t].dni; dh[t] = st[t].dhi; }
    lbDir = malloc(sizeof(double) * P); lbDif = malloc(sizeof(double) * P);
    float* dump = malloc(sizeof(float) * 6 * (P + 1));
    orc_ladybug_sky(o->mf, T, jdv, hv, dn, dh, ck.lat, ck.lon, dni ? wtz : ck.tz, grefl, dur, colOf, Tup, series ? S : NULL, series ? gS : NULL,
                    lbDir, lbDif, skyRad, &lbGnd, dump);
    free(jdv); free(hv); free(dn); free(dh); free(dump);
  }
  for (long t = 0; t < T && needW && !o->ladybug; ++t) {
    const orc_step* s = &st[t];
    const double dt = s->dt;
    const int col = series ? colOf[t] : -1;
    const double alt = sAlt[t] * kDegRad, azi = sAzi[t] * kDegRad;
    if (wantRad || (series && !seriesVis)) {
      orc_gendaymtx_step(o->mf, s->jd, alt, azi, s->dni, s->dhi, grefl, 0, 1, 0, buf);
      const double g = bright(buf);
      gndRad += g * dt / 1000.0;
      for (int p = 0; p < P; ++p) {
        const double L = bright(buf + 3 * (p + 1));
        skyRad[p] += L * dt / 1000.0;
        if (col >= 0 && !seriesVis) S[(size_t)p * Tup + col] = (float)L;
      }
      if (col >= 0 && !seriesVis) gS[col] = (float)g;
    }
    if (wantVis || seriesVis) {
      orc_gendaymtx_step(o->mf, s->jd, alt, azi, s->dni, s->dhi, grefl, 0, 1, 1, buf);
      const double g = 179.0 * bright(buf);
      gndVis += g * dt;
      for (int p = 0; p < P; ++p) {
        const double L = 179.0 * bright(buf + 3 * (p + 1));
        skyVis[p] += L * dt;
        if (col >= 0 && seriesVis) S[(size_t)p * Tup + col] = (float)L;
      }
      if (col >= 0 && seriesVis) gS[col] = (float)g;
      orc_gendaymtx_step(o->mf, s->jd, alt, azi, s->dni, s->dhi, grefl, 1, 0, 1, buf);
      double dirIll = 0;
      for (int p = 0; p < P; ++p) dirIll += buf[3 * (p + 1) + 1] * omega[p];
      if (sunOf[t] >= 0) sunVis[sunOf[t]] = 179.0 * dirIll * dt;
    }
    if (sunOf[t] >= 0) sunRad[sunOf[t]] = s->dni * dt / 1000.0;
  }

  /* per-quantity weights (run.cpp:475-505) */
  const int nQ = o->nQ;
  float* W = calloc((size_t)nQ * (nDir ? nDir : 1), sizeof(float));
  float* Wg = calloc((size_t)(nDir ? nDir : 1), sizeof(float));
  for (long d = nSky; d < 2 * nSky; ++d) Wg[d] = o->ladybug ? 1.f : omega[rowOf[d - nSky]] / (float)K;
  for (int qi = 0; qi < nQ; ++qi) {
    const int q = o->quantities[qi];
    float* w = W + (size_t)qi * nDir;
    if (q == ORC_SUNHOURS) { for (long j = 0; j < nSun; ++j) w[sun0 + j] = (float)st[sunStep[j]].dt; continue; }
    if (o->ladybug) {   /* the sun inside the patches: sky and ground only (run.cpp:489-502) */
      for (long d = 0; d < nSky; ++d) {
        const int p = rowOf[d];
        w[d] = (float)(q == ORC_DIRECT ? lbDir[p] : q == ORC_DIFFUSE ? lbDif[p] : skyRad[p]);
      }
      if (q != ORC_DIRECT) for (long d = nSky; d < 2 * nSky; ++d) w[d] = (float)lbGnd;
      continue;
    }
    const int vis = q == ORC_ILLUMINANCE, wantSky = q != ORC_DIRECT, wantSun = q != ORC_DIFFUSE;
    for (long d = 0; d < nSky; ++d) {
      const int p = rowOf[d];
      w[d] = (float)(wantSky ? (vis ? skyVis[p] : skyRad[p]) * omega[p] / K : 0.0);
    }
    if (q != ORC_DIRECT)
      for (long d = nSky; d < 2 * nSky; ++d) w[d] = (float)((vis ? gndVis : gndRad) * omega[rowOf[d - nSky]] / K);
    if (wantSun) for (long j = 0; j < nSun; ++j) w[sun0 + j] = (float)(vis ? sunVis[j] : sunRad[j]);
  }

  /* scene and sensors: recentred by the bbox centre of every mesh; sensor meshes block too (selfOcclude) */
  double c[3];
  orc_bbox_centre(m, nMesh, c);
  long nv = 0, ntCap = 0;
  for (int k = 0; k < nMesh; ++k) {
    const int blocks = !m[k].sensor || o->selfOcclude;
    if (!blocks) continue;
    nv += m[k].nPoints;
    for (long f = 0; f < m[k].nFaces; ++f) ntCap += m[k].counts[f] >= 3 ? m[k].counts[f] - 2 : 0;
  }
  float* V = malloc(sizeof(float) * 3 * (size_t)(nv ? nv : 1));
  uint32_t* I = malloc(sizeof(uint32_t) * 3 * (size_t)(ntCap ? ntCap : 1));
  long vb = 0, nt = 0;
  int skipped = 0;
  for (int k = 0; k < nMesh; ++k) {
    if (!(!m[k].sensor || o->selfOcclude)) continue;
    nt += orc_triangulate(&m[k], c, V + 3 * vb, I + 3 * nt, (uint32_t)vb, &skipped);
    vb += m[k].nPoints;
  }
  float* Ps = malloc(sizeof(float) * 3 * (size_t)nSens);
  float* Ns = malloc(sizeof(float) * 3 * (size_t)nSens);
  long at = 0;
  for (int k = 0; k < nMesh; ++k)
    if (m[k].sensor) {
      const long made = orc_sensors_mode(&m[k], o->onVertices, o->sides, o->offset, o->ladybug, c, Ps + 3 * at, Ns + 3 * at);
      if (made < 0) { free(st); return fail(err, errLen, "error: ladybug mode supports tri/quad faces only"); }
      at += nS * made;
    }

  out->nElem = nElem; out->nSens = nSens; out->nCols = (long)nQ * nS;
  const long C = out->nCols;
  for (long e = 0; e < nElem * C; ++e) out->values[e] = 0;
  if (nDir > 0) {
    orc_scene* sc = orc_scene_new(V, nv, I, nt);
    const int nWords = (int)((nDir + 31) / 32);
    uint32_t* bits = malloc(sizeof(uint32_t) * (size_t)nWords * nSens);
    orc_trace_bits(sc, Ps, Ns, nSens, D, (int)nDir, o->cullBackfaces, bits, o->nthreads);
    double* tmp = malloc(sizeof(double) * nSens);
    for (int qi = 0; qi < nQ; ++qi) {
      const int q = o->quantities[qi];
      orc_reduce(bits, nSens, (int)nDir, Ns, D, W + (size_t)qi * nDir, q != ORC_SUNHOURS, tmp);
      for (long si = 0; si < nSens; ++si) {
        double v = tmp[si];
        if (q == ORC_IRRADIANCE) v = dur > 0 ? v * 1000.0 / dur : 0.0;
        else if (q == ORC_ILLUMINANCE) v = dur > 0 ? v / dur : 0.0;
        out->values[(si / nS) * C + qi * nS + si % nS] = v;
      }
    }
    if (series && out->series && Tup > 0) {
      orc_reduce(bits, nSens, (int)nDir, Ns, D, Wg, 1, tmp);
      float* G = malloc(sizeof(float) * nSens);
      for (long i = 0; i < nSens; ++i) G[i] = (float)tmp[i];
      float* wsky = malloc(sizeof(float) * (size_t)(nSky ? nSky : 1));
      for (long d = 0; d < nSky; ++d) wsky[d] = o->ladybug ? omega[rowOf[d]] : omega[rowOf[d]] / (float)K;
      float* sunW = calloc((size_t)(nSun ? nSun : 1), sizeof(float));
      int* sunCol = calloc((size_t)(nSun ? nSun : 1), sizeof(int));
      for (long j = 0; j < nSun; ++j) {
        const int cidx = colOf[sunStep[j]];
        sunCol[j] = cidx < 0 ? 0 : cidx;
        const double dtj = st[sunStep[j]].dt;
        sunW[j] = (o->ladybug || cidx < 0 || dtj <= 0) ? 0.f : (float)((seriesVis ? sunVis[j] : sunRad[j] * 1000.0) / dtj);
      }
      orc_series64(bits, nSens, Ns, D, (unsigned)nSky, rowOf, wsky, S, Tup, gS, G, (unsigned)sun0, (unsigned)nSun, sunCol, D + 3 * sun0, sunW,
                   out->series, o->nthreads);
      free(G); free(wsky); free(sunW); free(sunCol);
    }
    free(tmp); free(bits);
    orc_scene_free(sc);
  }
  free(st); free(sAlt); free(sAzi); free(sUp); free(sVec); free(pAlt); free(pAzi); free(omega); free(prow); free(nin);
  free(D); free(rowOf); free(sunStep); free(sunOf); free(colOf); free(S); free(gS); free(skyRad); free(skyVis); free(sunRad);
  free(sunVis); free(buf); free(W); free(Wg); free(V); free(I); free(Ps); free(Ns); free(lbv); free(lbDir); free(lbDif);
  return 0;
}
