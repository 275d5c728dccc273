/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * Oracle (test infrastructure only): reference src/solar.cpp restated in C: the NOAA/Meeus solar position with the
 * Ladybug Sunpath conventions (fixed year, minute resolution, 0.01-day rounding, NOAA refraction), Spencer's
 * extraterrestrial irradiance, the Ineichen-Perez clear sky (pvlib formulation) and the evaluation steps.
 */
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* solar.cpp's degree helpers: (d pi) / 180 and (r 180) / pi, in that order */
static double rad(double d) { return d * ORC_PI / 180; }
static double deg(double r) { return r * 180 / ORC_PI; }
static double pmod(double a, double m) { const double r = fmod(a, m); return r < 0 ? r + m : r; }
static int month_days(int m, int leap) {
  static const int n[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  return n[m - 1] + (leap && m == 2 ? 1 : 0);
}
/* Excel-style serial of Jan 1 of `year` minus one day, counted from 1900 (which is not leap): 2017 -> 42734 */
static int serial_1900(int year) {
  const int y = year - 1, y0 = 1899;
  return (year - 1900) * 365 + (y / 4 - y / 100 + y / 400) - (y0 / 4 - y0 / 100 + y0 / 400);
}

void orc_sun_noaa(double lat, double lon, double tz, int month, int day, double hourFrac, int leap, double* altOut, double* aziOut,
                  int* up, float vec[3]) {
  int hour = (int)hourFrac, minute = (int)lround((hourFrac - hour) * 60);
  if (minute == 60) { ++hour; minute = 0; }
  int doy = day;
  for (int m = 1; m < month; ++m) doy += month_days(m, leap);
  const int base = serial_1900(leap ? 2016 : 2017) + doy + 1;
  const double jd = base + 2415018.5 + orc_round_fmt((minute + hour * 60) / 1440.0, "%.2f") - tz / 24;
  const double jc = (jd - 2451545) / 36525;
  const double L0 = pmod(280.46646 + jc * (36000.76983 + jc * 0.0003032), 360);
  const double M = 357.52911 + jc * (35999.05029 - 0.0001537 * jc);
  const double ecc = 0.016708634 - jc * (0.000042037 + 0.0000001267 * jc);
  const double C = sin(rad(M)) * (1.914602 - jc * (0.004817 + 0.000014 * jc)) + sin(rad(2 * M)) * (0.019993 - 0.000101 * jc) +
                   sin(rad(3 * M)) * 0.000289;
  const double om = rad(125.04 - 1934.136 * jc);
  const double lam = rad(L0 + C - 0.00569 - 0.00478 * sin(om));
  const double eps0 = 23 + (26 + (21.448 - jc * (46.815 + jc * (0.00059 - jc * 0.001813))) / 60) / 60;
  const double eps = rad(eps0 + 0.00256 * cos(om));
  const double dec = asin(sin(eps) * sin(lam));
  double y = tan(eps / 2);
  y *= y;
  const double eot = 4 * deg(y * sin(2 * rad(L0)) - 2 * ecc * sin(rad(M)) + 4 * ecc * y * sin(rad(M)) * cos(2 * rad(L0)) -
                             0.5 * y * y * sin(4 * rad(L0)) - 1.25 * ecc * ecc * sin(2 * rad(M)));
  const double tst = pmod(hourFrac * 60 + eot + 4 * lon - 60 * tz, 1440);
  const double H = tst < 0 ? tst / 4 + 180 : tst / 4 - 180;
  const double la = rad(lat);
  const double zen = acos(sin(la) * sin(dec) + cos(la) * cos(dec) * cos(rad(H)));
  const double alt = 90 - deg(zen), t = tan(rad(alt));
  double r;   /* NOAA refraction, arcseconds */
  if (alt > 85) r = 0;
  else if (alt > 5) r = 58.1 / t - 0.07 / (t * t * t) + 0.000086 / (t * t * t * t * t);
  else if (alt > -0.575) r = 1735 + alt * (-518.2 + alt * (103.4 + alt * (-12.79 + alt * 0.711)));
  else r = -20.772 / t;
  const double altR = alt + r / 3600;
  const double a = (sin(la) * cos(zen) - sin(dec)) / (cos(la) * sin(zen));
  double azi;
  if (a >= -1 && a <= 1) azi = pmod(H > 0 ? deg(acos(a)) + 180 : 540 - deg(acos(a)), 360);
  else azi = a > 0 ? 180 : 0;
  *altOut = altR;
  *aziOut = azi;
  *up = altR >= 0;
  orc_dir_from_altaz(rad(altR), rad(azi), vec);
}

double orc_extraterrestrial(int doy) {
  const double B = 2 * ORC_PI * (doy - 1) / 365;
  return 1366.1 * (1.00011 + 0.034221 * cos(B) + 0.00128 * sin(B) + 0.000719 * cos(2 * B) + 0.000077 * sin(2 * B));
}

void orc_ineichen(double z, double TL, double altitude, double dniExtra, double* ghi, double* dni, double* dhi) {
  z = fmin(fmax(z, 0.0), 90.0);
  const double cz = cos(rad(z));
  const double am = 1 / (cz + 0.50572 * pow(6.07995 + (90 - z), -1.6364));   /* Kasten-Young 1989 */
  const double pres = 100 * pow((44331.514 - altitude) / 11880.516, 1 / 0.1902632);
  const double ama = am * pres / 101325;
  const double fh1 = exp(-altitude / 8000), fh2 = exp(-altitude / 1250);
  const double cg1 = 5.09e-5 * altitude + 0.868, cg2 = 3.92e-5 * altitude + 0.0387;
  const double g = fmax(0.0, cg1 * dniExtra * cz * exp(-cg2 * ama * (fh1 + fh2 * (TL - 1))));
  const double b = 0.664 + 0.163 / fh1;
  const double bnci = dniExtra * fmax(0.0, b * exp(-0.09 * ama * (TL - 1)));
  const double bnci2 = g * fmin(fmax((1 - (0.1 - 0.2 * exp(-TL)) / (0.1 + 0.882 / fh1)) / cz, 0.0), 1e20);
  *ghi = g;
  *dni = fmin(bnci, bnci2);
  *dhi = g - *dni * cz;
}

void orc_hoy_to_date(double hoy, int leap, int* month, int* day, double* hour) {
  int d = (int)floor(hoy / 24);
  *hour = hoy - d * 24.0;
  int m = 1;
  while (m < 12 && d >= month_days(m, leap)) d -= month_days(m++, leap);
  *month = m;
  *day = d + 1;
}

int orc_clear_sky(double lat, double lon, double tz, double TL, double altitude, float* dni, float* dhi) {
  if (TL < 1 || fabs(lat) > 90) return -1;
  for (int r = 0; r < 8760; ++r) {
    dni[r] = dhi[r] = 0.f;
    int mo, dy;
    double hr, alt, azi;
    int up;
    float v[3];
    orc_hoy_to_date(r + 0.5, 0, &mo, &dy, &hr);
    orc_sun_noaa(lat, lon, tz, mo, dy, hr, 0, &alt, &azi, &up, v);
    if (!up) continue;
    double g, b, d;
    orc_ineichen(90 - alt, TL, altitude, orc_extraterrestrial(r / 24 + 1), &g, &b, &d);
    dni[r] = (float)b;
    dhi[r] = (float)d;
  }
  return 0;
}

int orc_parse_hoy(const char* s, int leap, double* hoy) {
  int y, m, d, h, mi;
  if (sscanf(s, "%d-%d-%dT%d:%d", &y, &m, &d, &h, &mi) != 5 && sscanf(s, "%d-%dT%d:%d", &m, &d, &h, &mi) != 4) return -1;
  if (m < 1 || m > 12 || d < 1 || d > month_days(m, leap) || h < 0 || h > 23 || mi < 0 || mi > 59) return -1;
  double v = 24.0 * (d - 1) + h + mi / 60.0;
  for (int k = 1; k < m; ++k) v += 24 * month_days(k, leap);
  *hoy = v;
  return 0;
}

/* weather between the centres of hour-ending rows, cyclic (solar.cpp:29-34) */
static float lerp_rows(const float* v, long n, double t) {
  const double x = t - 0.5, f = x - floor(x);
  const long a = (((long)floor(x) % n) + n) % n, b = (a + 1) % n;
  return (float)(v[a] * (1 - f) + v[b] * f);
}

long orc_make_steps(const orc_time* tm, const float* dni, const float* dhi, long nRows, int leap, double shift, orc_step* out, long cap) {
  const double nh = leap ? 8784 : 8760;
#define WRAP(t) ((t) - nh * floor((t) / nh))
  long n = 0;
  if (isnan(tm->t0)) {
    if (!dni) return -1;
    for (long r = 0; r < nRows; ++r) {
      if (n >= cap) return -1;
      orc_step s;
      s.hoy = WRAP(r + 0.5 + shift);
      s.jd = (int)(r / 24) + 1;
      s.dt = 1;
      s.dni = dni[r];
      s.dhi = dhi[r];
      orc_hoy_to_date(r + 0.5, leap, &s.month, &s.day, &s.hour);
      out[n++] = s;
    }
    return n;
  }
  if (tm->stepMin <= 0) return -1;
  const int instant = isnan(tm->t1), window = tm->dayFrom > 0 || tm->dayTo < 1440;
  const double step = tm->stepMin / 60.0;
  double t1 = instant ? tm->t0 + step : tm->t1;
  if (t1 > nh) t1 = nh;
  if (tm->t0 < 0 || tm->t0 >= nh || t1 <= tm->t0) return -1;
  for (int k = 0;; ++k) {
    const double hoy = tm->t0 + k * step;
    if (hoy >= t1 - 1e-9) break;
    double hod = (hoy - 24 * floor(hoy / 24)) * 60;
    if (hod >= 1440 - 1e-6) hod -= 1440;
    if (hod < tm->dayFrom - 1e-6 || hod >= tm->dayTo - 1e-6) continue;
    double dt = step;
    if (!instant) dt = fmin(dt, t1 - hoy);
    if (!instant && window) dt = fmin(dt, (tm->dayTo - hod) / 60);
    if (n >= cap) return -1;
    orc_step s;
    s.hoy = instant ? hoy : hoy + dt / 2;
    const double at = WRAP(s.hoy - shift);
    s.jd = (int)(at / 24) + 1;
    s.dt = dt;
    orc_hoy_to_date(at, leap, &s.month, &s.day, &s.hour);
    s.dni = s.dhi = 0.f;
    if (dni && nRows > 0) { s.dni = lerp_rows(dni, nRows, at); s.dhi = lerp_rows(dhi, nRows, at); }
    out[n++] = s;
  }
#undef WRAP
  return n ? n : -1;
}
