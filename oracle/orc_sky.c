/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * Oracle (test infrastructure only): the sky model of reference src/sky.cpp restated in C. That file ports Radiance
 * rad5R4 gendaymtx.c / sun.c / tcos.c (Radiance Software License 2.0, LBNL); the Perez coefficient tables below are the
 * published numbers of Perez, Seals & Michalsky 1993 (Solar Energy 50(3), Table 1) and Perez et al. 1990 (Solar Energy
 * 44(5), Table 4), which any implementation of the model has to carry as they are.
 * Float / double rounding points follow the reference so that rows agree after float rounding.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* Perez 1993 Table 1: a, b, c, d, e (each as x0 + x1 z + delta (x2 + x3 z)), per clearness category 0..7. */
static const double kPerez[8][5][4] = {
    {{1.3525, -0.2576, -0.2690, -1.4366}, {-0.7670, 0.0007, 1.2734, -0.1233}, {2.8000, 0.6004, 1.2375, 1.0000},
     {1.8734, 0.6297, 0.9738, 0.2809}, {0.0356, -0.1246, -0.5718, 0.9938}},
    {{-1.2219, -0.7730, 1.4148, 1.1016}, {-0.2054, 0.0367, -3.9128, 0.9156}, {6.9750, 0.1774, 6.4477, -0.1239},
     {-1.5798, -0.5081, -1.7812, 0.1080}, {0.2624, 0.0672, -0.2190, -0.4285}},
    {{-1.1000, -0.2515, 0.8952, 0.0156}, {0.2782, -0.1812, -4.5000, 1.1766}, {24.7219, -13.0812, -37.7000, 34.8438},
     {-5.0000, 1.5218, 3.9229, -2.6204}, {-0.0156, 0.1597, 0.4199, -0.5562}},
    {{-0.5484, -0.6654, -0.2672, 0.7117}, {0.7234, -0.6219, -5.6812, 2.6297}, {33.3389, -18.3000, -62.2500, 52.0781},
     {-3.5000, 0.0016, 1.1477, 0.1062}, {0.4659, -0.3296, -0.0876, -0.0329}},
    {{-0.6000, -0.3566, -2.5000, 2.3250}, {0.2937, 0.0496, -5.6812, 1.8415}, {21.0000, -4.7656, -21.5906, 7.2492},
     {-3.5000, -0.1554, 1.4062, 0.3988}, {0.0032, 0.0766, -0.0656, -0.1294}},
    {{-1.0156, -0.3670, 1.0078, 1.4051}, {0.2875, -0.5328, -3.8500, 3.3750}, {14.0000, -0.9999, -7.1406, 7.5469},
     {-3.4000, -0.1078, -1.0750, 1.5702}, {-0.0672, 0.4016, 0.3017, -0.4844}},
    {{-1.0000, 0.0211, 0.5025, -0.5119}, {-0.3000, 0.1922, 0.7023, -1.6317}, {19.0000, -5.0000, 1.2438, -1.9094},
     {-4.0000, 0.0250, 0.3844, 0.2656}, {1.0468, -0.3788, -2.4517, 1.4656}},
    {{-1.0500, 0.0289, 0.4260, 0.3590}, {-0.3250, 0.1156, 0.7781, 0.0025}, {31.0625, -14.5000, -46.1148, 55.3750},
     {-7.2312, 0.4050, 13.3500, 0.6234}, {1.5000, -0.6426, 1.8564, 0.5636}}};
/* clearness category bounds [lower, upper) */
static const double kClearLo[8] = {1.000, 1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200};
static const double kClearHi[8] = {1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200, 12.01};
/* Perez 1990 Table 4: luminous efficacy a, b, c, d (diffuse: Eqn 7, direct: Eqn 8) */
static const double kEffDiff[8][4] = {{97.24, -0.46, 12.00, -8.91},   {107.22, 1.15, 0.59, -3.95},  {104.97, 2.96, -5.53, -8.77},
                                      {102.39, 5.59, -13.95, -13.90}, {100.71, 5.94, -22.75, -23.74}, {106.42, 3.83, -36.15, -28.83},
                                      {141.88, 1.90, -53.24, -14.03}, {152.23, 0.35, -45.27, -7.98}};
static const double kEffDir[8][4] = {{57.20, -4.55, -2.98, 117.12}, {98.99, -3.46, -1.21, 12.38},  {109.83, -4.90, -1.71, -8.81},
                                     {110.34, -5.84, -1.99, -4.56}, {106.36, -3.97, -1.75, -6.16}, {107.19, -1.25, -1.51, -26.73},
                                     {105.75, 0.77, -1.26, -34.44}, {101.18, 1.58, -1.10, -8.29}};
static const int kRowAz[7] = {30, 30, 24, 24, 18, 12, 6};   /* Tregenza patches per row */

static double deg2rad(double d) { return d * (ORC_PI / 180.); }   /* sky.cpp DegToRad */
static double rad2deg(double r) { return r * (180. / ORC_PI); }

/* tcos.c: 1024-entry cosine table over [0, pi/2], nearest entry, no interpolation (sky.cpp:202-227) */
static double g_cos[1025];
static int g_cosReady = 0;
double orc_tcos(double x) {
  if (!g_cosReady) {
    for (int k = 0; k <= 1024; ++k) g_cos[k] = cos((ORC_PI / 2. / 1024) * k);
    g_cosReady = 1;
  }
  if (x < 0.) x = -x;
  int k = (int)((1024 * 2. / ORC_PI) * x + 0.5);
  while (k >= 4096) k -= 4096;
  switch (k >> 10) {
    case 0: return g_cos[k];
    case 1: return -g_cos[2048 - k];
    case 2: return -g_cos[k - 2048];
    default: return g_cos[4096 - k];
  }
}
double orc_tsin(double x) { return orc_tcos(x - ORC_PI / 2.); }

/* (cos alt sin azi, cos alt cos azi, sin alt) in double, the gendaymtx vector() macro's order of operations */
static void altaz_vec(double alt, double azi, double v[3]) {
  const double ca = cos(alt);
  v[0] = ca * sin(azi);
  v[1] = ca * cos(azi);
  v[2] = sin(alt);
}
void orc_dir_from_altaz(double altRad, double aziRad, float out[3]) {
  double v[3];
  altaz_vec(altRad, aziRad, v);
  out[0] = (float)v[0]; out[1] = (float)v[1]; out[2] = (float)v[2];
}
void orc_rotate_z(const float v[3], double deg, float out[3]) {
  const double t = deg2rad(deg), c = cos(t), s = sin(t);
  const float x = v[0], y = v[1];
  out[0] = (float)(x * c - y * s);
  out[1] = (float)(x * s + y * c);
  out[2] = v[2];
}

int orc_patch_count(int mf) {
  int n = 1;
  for (int r = 0; r < 7 * mf; ++r) n += kRowAz[r / mf] * mf;
  return n;
}

/* sky.cpp:240-270: Reinhart rows of height alpha, the zenith cap last; angles and solid angles rounded to float. */
float orc_patches(int mf, float* alt, float* azi, float* omega, int* row, int* ninrow, float* centre, float* lb) {
  const double alpha = (ORC_PI / 2.) / (7 * mf + .5);
  int p = 0;
  for (int r = 0; r < 7 * mf; ++r) {
    const int n = kRowAz[r / mf] * mf;
    const float a = (float)(alpha * (r + .5));
    const float w = (float)(2. * ORC_PI * (sin(alpha * (r + 1)) - sin(alpha * r)) / (double)n);
    for (int j = 0; j < n; ++j, ++p) {
      alt[p] = a;
      azi[p] = (float)(2. * ORC_PI * j / (double)n);
      omega[p] = w;
      row[p] = r;
      ninrow[p] = n;
    }
  }
  alt[p] = (float)(ORC_PI / 2.);
  azi[p] = 0.f;
  omega[p] = (float)(2. * ORC_PI * (1. - cos(alpha * .5)));
  row[p] = 7 * mf;
  ninrow[p] = 1;
  const int P = p + 1;
  for (int q = 0; q < P; ++q) {
    if (centre) orc_dir_from_altaz(alt[q], azi[q], centre + 3 * q);
    if (lb) {
      if (q == P - 1) { lb[3 * q] = 0.f; lb[3 * q + 1] = 0.f; lb[3 * q + 2] = 1.f; }
      else orc_dir_from_altaz(atan(cos(ORC_PI / ninrow[q]) * tan((double)alt[q])), azi[q], lb + 3 * q);
    }
  }
  return (float)alpha;
}

void orc_patch_sample(int mf, int p, int k, int K, float out[3]) {
  const int P = orc_patch_count(mf);
  float* alt = malloc(sizeof(float) * P); float* azi = malloc(sizeof(float) * P); float* om = malloc(sizeof(float) * P);
  int* row = malloc(sizeof(int) * P); int* nin = malloc(sizeof(int) * P);
  const float alphaF = orc_patches(mf, alt, azi, om, row, nin, NULL, NULL);
  if (K <= 1) orc_dir_from_altaz(alt[p], azi[p], out);
  else {
    const int m = (int)lround(sqrt((double)K));
    const double u = ((k % m) + 0.5) / m, v = ((k / m) + 0.5) / m, a = alphaF;   /* the float row height */
    if (p == P - 1) {   /* zenith cap: cos theta uniform in [cos(alpha / 2), 1] */
      const double ct = 1 - v * (1 - cos(a / 2)), st = sqrt(1 - ct * ct), az = 2 * ORC_PI * u;
      out[0] = (float)(st * sin(az)); out[1] = (float)(st * cos(az)); out[2] = (float)ct;
    } else {
      const int r = row[p];
      const double s0 = sin(a * r), s1 = sin(a * (r + 1));
      orc_dir_from_altaz(asin(s0 + v * (s1 - s0)), azi[p] + (u - 0.5) * 2 * ORC_PI / nin[p], out);
    }
  }
  free(alt); free(azi); free(om); free(row); free(nin);
}

double orc_round_fmt(double x, const char* fmt) {
  char buf[64];
  snprintf(buf, sizeof buf, fmt, x);
  return strtod(buf, NULL);
}

/* A cached patch table per sky density (the step function is called T times per run). */
typedef struct { int mf, P; float *alt, *azi, *omega; int *row, *nin; } patch_tab;
static patch_tab g_tab[5];
static const patch_tab* patch_table(int mf) {
  patch_tab* t = &g_tab[mf];
  if (t->P == 0) {
    const int P = orc_patch_count(mf);
    t->alt = malloc(sizeof(float) * P); t->azi = malloc(sizeof(float) * P); t->omega = malloc(sizeof(float) * P);
    t->row = malloc(sizeof(int) * P); t->nin = malloc(sizeof(int) * P);
    orc_patches(mf, t->alt, t->azi, t->omega, t->row, t->nin, NULL, NULL);
    t->mf = mf;
    t->P = P;
  }
  return t;
}

/* gendaymtx ComputeSky + AddDirect for one step (sky.cpp:297-474). */
void orc_gendaymtx_step(int mf, int jd, double alt, double azi, double dni, double dhi, double grefl_in, int sun, int sky,
                        int visible, float* out) {
  const patch_tab* pt = patch_table(mf);
  const int P = pt->P;
  if (dni + dhi <= 1e-4) { memset(out, 0, sizeof(float) * 3 * (size_t)(P + 1)); return; }
  const float colSky[3] = {sky ? .96f : 0.f, sky ? 1.004f : 0.f, sky ? 1.118f : 0.f};
  const float colSun = sun ? 1.f : 0.f;
  const float grefl = sky ? (float)grefl_in : 0.f;
  const double pwc = exp(0.07 * 11.0 - 0.075);   /* precipitable water, dew point 11 C */

  double zen;   /* sun zenith: horizon clamp, and at least 3 degrees */
  if (alt <= 0.0) zen = deg2rad(90.0);
  else if (alt >= deg2rad(87.0)) zen = deg2rad(3.0);
  else zen = deg2rad(90.0) - alt;

  const double da = (jd - 1.0) * (2.0 * ORC_PI / 365.0);
  const double ecc = 1.00011 + 0.034221 * cos(da) + 0.00128 * sin(da) + 0.000719 * cos(2.0 * da) + 0.000077 * sin(2.0 * da);
  const double am = 1.0 / (cos(zen) + 0.15 * pow(93.885 - rad2deg(zen), -1.253));   /* Kasten 1966 */
  double bright = dhi * am / (1367.0 * ecc);
  const double z3 = zen * zen * zen;
  double clear = ((dhi + dni) / dhi + 1.041 * z3) / (1.0 + 1.041 * z3);
  if (clear > 11.9) clear = 11.9;
  if (bright < 0.01) bright = 0.01;
  int cat = 0;
  while (cat < 8 && !(clear >= kClearLo[cat] && clear < kClearHi[cat])) ++cat;
  if (cat > 7) cat = 7;

  double difIll, dirIll;
  if (visible) {   /* -O0: Perez 1990 luminous efficacy */
    const double* e = kEffDiff[cat];
    difIll = dhi * (e[0] + e[1] * pwc + e[2] * cos(zen) + e[3] * log(bright));
    e = kEffDir[cat];
    const double r = e[0] + e[1] * pwc + e[2] * exp(5.73 * zen - 5.0) + e[3] * bright;
    dirIll = dni * (r > 0.0 ? r : 0.0);
  } else {         /* -O1: radiometric (179 lm/W in, 179 out) */
    difIll = dhi * 179.;
    dirIll = dni * 179.;
  }

  /* ground row: float accumulation as the C code's float array */
  float g = (float)difIll;
  if (alt > 0) g = (float)(g + dirIll * sin(alt));
  g = (float)(g * (1. / ORC_PI / 179.));
  out[0] = out[1] = out[2] = g * grefl;

  if (!sky) memset(out + 3, 0, sizeof(float) * 3 * (size_t)P);
  else {
    double delta = bright, c[5];
    if (clear > 1.065 && clear < 2.8 && delta < 0.2) delta = 0.2;
    const double(*x)[4] = kPerez[cat];
    for (int i = 0; i < 5; ++i) c[i] = x[i][0] + x[i][1] * zen + delta * (x[i][2] + x[i][3] * zen);
    if (cat == 0) {   /* overcast category: c and d have their own forms (Perez 1993 Eqns 7 and 8) */
      c[2] = exp(pow(delta * (x[2][0] + x[2][1] * zen), x[2][2])) - x[2][3];
      c[3] = -exp(delta * (x[3][0] + x[3][1] * zen)) + x[3][2] + delta * x[3][3];
    }
    const double cz = cos(zen), sz = sin(zen);
    for (int p = 0; p < P; ++p) {
      const double dAz = fabs(pt->azi[p] - azi);
      const double zp = ORC_PI * 0.5 - pt->alt[p];
      const double gam = acos(cz * cos(zp) + sz * sin(zp) * cos(dAz));
      float L = (float)((1.0 + c[0] * exp(c[1] / cos(zp))) * (1.0 + c[2] * exp(c[3] * gam) + c[4] * cos(gam) * cos(gam)));
      if (L < 0) L = 0;
      out[3 * (p + 1)] = out[3 * (p + 1) + 1] = out[3 * (p + 1) + 2] = L;
    }
    double norm = 0.0;   /* horizontal illuminance of the relative sky, table sine */
    for (int p = 0; p < P; ++p) norm += out[3 * (p + 1) + 1] * orc_tsin(pt->alt[p]) * pt->omega[p];
    if (norm <= 1e-6) {
      for (int p = 0; p < P; ++p) out[3 * (p + 1)] = out[3 * (p + 1) + 1] = out[3 * (p + 1) + 2] = 1.f;
      norm = ORC_PI;
    }
    norm = difIll / norm;
    const double f = norm * (1. / 179.);
    for (int p = 0; p < P; ++p)
      for (int ch = 0; ch < 3; ++ch) out[3 * (p + 1) + ch] = (float)(out[3 * (p + 1) + ch] * f) * colSky[ch];
  }

  /* AddDirect: the sun into its 4 nearest patch centres, weighted 1 / (1.002 - cos) */
  if (dirIll <= 1e-4 || !sun) return;
  double best[4] = {-1., -1., -1., -1.};
  int bestP[4] = {0, 0, 0, 0};
  double sv[3];
  altaz_vec(alt, azi, sv);
  for (int p = 1; p <= P; ++p) {
    double pv[3];
    altaz_vec(pt->alt[p - 1], pt->azi[p - 1], pv);
    const double dp = pv[0] * sv[0] + pv[1] * sv[1] + pv[2] * sv[2];
    for (int i = 0; i < 4; ++i)
      if (dp > best[i]) {
        for (int j = 3; j > i; --j) { best[j] = best[j - 1]; bestP[j] = bestP[j - 1]; }
        best[i] = dp;
        bestP[i] = p;
        break;
      }
  }
  double w[4], wsum = 0;
  for (int i = 3; i >= 0; --i) wsum += w[i] = 1. / (1.002 - best[i]);
  for (int i = 3; i >= 0; --i) {
    float add = (float)(w[i] * dirIll / (179. * wsum));
    add /= pt->omega[bestP[i] - 1];
    for (int ch = 0; ch < 3; ++ch) out[3 * bestP[i] + ch] += add * colSun;
  }
}

/* sun.c (rad5R4): sdec, stadj, salt, sazi (sky.cpp:477-486) */
void orc_sun_radiance(double latDeg, double lonWestDeg, double meridianWestDeg, int jd, double hour, double* alt, double* azi) {
  const double la = deg2rad(latDeg), lo = deg2rad(lonWestDeg), me = deg2rad(meridianWestDeg);
  const double dec = 0.4093 * sin((2. * ORC_PI / 365.) * (jd - 81));
  const double adj = 0.170 * sin((4. * ORC_PI / 373.) * (jd - 80)) - 0.129 * sin((2. * ORC_PI / 355.) * (jd - 8)) + (12. / ORC_PI) * (me - lo);
  const double st = hour + adj;
  *alt = asin(sin(la) * sin(dec) - cos(la) * cos(dec) * cos(st * (ORC_PI / 12.)));
  *azi = -atan2(cos(dec) * sin(st * (ORC_PI / 12.)), -cos(la) * sin(dec) - sin(la) * cos(dec) * cos(st * (ORC_PI / 12.))) + ORC_PI;
}
