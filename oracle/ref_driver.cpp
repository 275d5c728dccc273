/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
// Oracle build helper (test infrastructure only): a C entry point per reference host function of the sky / sun model,
// compiled together with the reference's own src/sky.cpp and src/solar.cpp (read in place, never copied) into
// _ref/libdlref.so. Tests call these beside the C restatement (liborc.so) to pin it bit for bit.
#include "internal.h"
#include <cmath>
#include <cstring>

using namespace daylight;

extern "C" {

int ref_patches(int mf, float* alt, float* azi, float* omega, int* row, int* ninrow, float* centre, float* lb, float* alpha) {
  const Patches p = make_patches(mf);
  for (int k = 0; k < p.P; ++k) {
    alt[k] = p.alt[k]; azi[k] = p.azi[k]; omega[k] = p.omega[k]; row[k] = p.row[k]; ninrow[k] = p.ninrow[k];
    std::memcpy(centre + 3 * k, &p.centre[k], 12);
    std::memcpy(lb + 3 * k, &p.lb[k], 12);
  }
  *alpha = p.alpha;
  return p.P;
}

void ref_patch_sample(int mf, int p, int k, int K, float* out) {
  const float3 v = patch_sample(make_patches(mf), p, k, K);
  out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

void ref_gendaymtx_step(int mf, int jd, double alt, double azi, double dni, double dhi, double grefl, int sun, int sky, int vis, float* out) {
  static Patches cache[5];
  if (cache[mf].P == 0) cache[mf] = make_patches(mf);
  gendaymtx_step(cache[mf], jd, alt, azi, dni, dhi, grefl, sun != 0, sky != 0, vis != 0, out);
}

void ref_sun_radiance(double lat, double lonW, double merW, int jd, double hour, double* alt, double* azi) {
  sun_radiance(lat, lonW, merW, jd, hour, *alt, *azi);
}

void ref_sun_noaa(double lat, double lon, double tz, int month, int day, double hour, int leap, double* alt, double* azi, int* up, float* vec) {
  const Sun s = sun_noaa(lat, lon, tz, month, day, hour, leap != 0);
  *alt = s.alt; *azi = s.azi; *up = s.up ? 1 : 0;
  vec[0] = s.vec.x; vec[1] = s.vec.y; vec[2] = s.vec.z;
}

void ref_ineichen(double z, double TL, double altitude, double dniExtra, double* ghi, double* dni, double* dhi) {
  ineichen(z, TL, altitude, dniExtra, *ghi, *dni, *dhi);
}

double ref_extraterrestrial(int doy) { return extraterrestrial_irradiance(doy); }

int ref_clear_sky(double lat, double lon, double tz, double TL, double altitude, float* dni, float* dhi) {
  Weather w;
  std::string err;
  if (!clear_sky_weather(lat, lon, tz, TL, altitude, w, err)) return -1;
  std::memcpy(dni, w.dni.data(), 4 * w.dni.size());
  std::memcpy(dhi, w.dhi.data(), 4 * w.dhi.size());
  return 0;
}

double ref_round_fmt(double x, const char* fmt) { return round_fmt(x, fmt); }

void ref_rotate_z(const float* v, double deg, float* out) {
  const float3 r = rotate_z(float3{v[0], v[1], v[2]}, deg);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// make_steps over a weather of nRows (dni / dhi may be null: none); out as (month, day, jd, hour, hoy, dt, dni, dhi) rows.
long ref_make_steps(double t0, double t1, int stepMin, double dayFrom, double dayTo, const float* dni, const float* dhi, long nRows,
                    int leap, double shift, double* out, long cap) {
  Options o;
  o.t0 = t0; o.t1 = t1; o.stepMin = stepMin; o.dayFrom = dayFrom; o.dayTo = dayTo;
  Weather w;
  w.leap = leap != 0;
  if (dni) { w.dni.assign(dni, dni + nRows); w.dhi.assign(dhi, dhi + nRows); }
  std::vector<Step> st;
  std::string err;
  if (!make_steps(o, dni ? &w : nullptr, st, err, shift)) return -1;
  if ((long)st.size() > cap) return -1;
  for (size_t k = 0; k < st.size(); ++k) {
    double* r = out + 8 * k;
    r[0] = st[k].month; r[1] = st[k].day; r[2] = st[k].jd; r[3] = st[k].hour; r[4] = st[k].hoy; r[5] = st[k].dt; r[6] = st[k].dni;
    r[7] = st[k].dhi;
  }
  return (long)st.size();
}

}  // extern "C"
