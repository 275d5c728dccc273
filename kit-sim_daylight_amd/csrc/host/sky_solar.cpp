// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Radiance notice: the Perez tables, tcos and glass formulas here follow Radiance 5.4 (gendaymtx.c, tcos.c, m_glass.c),
// Copyright (c) 1990 to 2022 The Regents of the University of California, through LBNL; Radiance Software License 2.0,
// reproduced with its conditions and disclaimer in NOTICES.md.
// Host inputs of the sky and the sun: the Reinhart patch table and its samples (reference sky.cpp:240-294), Radiance's
// tcos table, evaluation steps, timestamps, the EPW reader with its PVGIS realignment, Spencer / Ineichen clear sky
// (reference solar.cpp). The per-step sun and Perez sky of a run are computed on the GPU (sky.hip); clear_sky_weather,
// which makes an input file's worth of weather, calls the same __host__ __device__ sun as the kernel.
#include <algorithm>
#include <array>
#include <charconv>
#include <cmath>
#include <cstring>
#include <memory>
#if defined(__SSE2__)
#include <emmintrin.h>
#endif
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "core.h"
#include "dl_math.h"

namespace daylight {
namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr int kRowAz[7] = {30, 30, 24, 24, 18, 12, 6};
double rad(double d) { return d * kPi / 180; }
double pmod(double a, double m) { const double r = std::fmod(a, m); return r < 0 ? r + m : r; }
int month_days(int m, bool leap) { return dlm::month_days(m, leap); }

float3 altaz(double alt, double azi) { return dlm::altaz_vec(alt, azi); }

// The weather at hour t of a cyclic year of hour-ending rows (row r stands for r + 0.5): linear between the two rows
// whose centres bracket t.
float lerp_rows(const std::vector<float>& v, double t) {
  const long n = long(v.size());
  const double x = t - 0.5, below = std::floor(x), f = x - below;
  const long lo = ((long(below) % n) + n) % n, hi = (lo + 1) % n;
  return float(v[size_t(lo)] * (1 - f) + v[size_t(hi)] * f);
}

// PVGIS labels satellite irradiance ahead of its hours: the mean offset of each day's lit midpoint (first and last rows
// with GHI > 0, days lit at midnight left out) from solar noon is its lead; over 30 days or more and beyond 0.3 h, DNI
// and DHI move later by it (reference solar.cpp:36-63). A file that declares its -0.5 h shift is left as it is.
void pvgis_align(Weather& w, const std::vector<float>& ghi, double declared) {
  if (std::fabs(declared + 0.5) < 0.05) return;
  const double noon = 12 - (w.lon / 15 - w.tz);
  std::vector<double> leads;
  for (size_t d0 = 0; d0 + 24 <= ghi.size(); d0 += 24) {
    const auto lit = [&](int h) { return ghi[d0 + size_t(h)] > 0; };
    int first = 0, last = 23;
    while (first < 24 && !lit(first)) ++first;
    while (last >= 0 && !lit(last)) --last;
    if (first >= 24 || first == 0 || last == 23) continue;   // dark, or lit at either end of the day
    leads.push_back(pmod((first + last) / 2.0 + 0.5 - noon + 12, 24) - 12);
  }
  if (leads.size() < 30) return;
  double sum = 0;
  for (const double l : leads) sum += l;
  const double lead = sum / int(leads.size());
  if (std::fabs(lead) <= 0.3) return;
  const std::vector<float> dni = w.dni, dhi = w.dhi;
  for (size_t r = 0; r < dni.size(); ++r) {
    w.dni[r] = lerp_rows(dni, double(r) + 0.5 + lead);
    w.dhi[r] = lerp_rows(dhi, double(r) + 0.5 + lead);
  }
  w.shiftH = -lead;
}

// the day of the year each month starts on (0-based), and the year's length in days after them
std::array<int, 13> month_starts(bool leap) {
  std::array<int, 13> s{};
  for (int m = 1; m <= 12; ++m) s[size_t(m)] = s[size_t(m - 1)] + month_days(m, leap);
  return s;
}
}  // namespace

Patches make_patches(int mf) {
  Patches pa;
  const double alpha = (kPi / 2.) / (7 * mf + .5);
  pa.alpha = float(alpha);
  for (int r = 0; r < 7 * mf; ++r) {
    const int n = kRowAz[r / mf] * mf;
    const float a = float(alpha * (r + .5));
    const float w = float(2. * kPi * (std::sin(alpha * (r + 1)) - std::sin(alpha * r)) / double(n));
    for (int j = 0; j < n; ++j) {
      pa.alt.push_back(a);
      pa.azi.push_back(float(2. * kPi * j / double(n)));
      pa.omega.push_back(w);
      pa.row.push_back(r);
      pa.ninrow.push_back(n);
    }
  }
  pa.alt.push_back(float(kPi / 2.));
  pa.azi.push_back(0.f);
  pa.omega.push_back(float(2. * kPi * (1. - std::cos(alpha * .5))));
  pa.row.push_back(7 * mf);
  pa.ninrow.push_back(1);
  pa.P = int(pa.alt.size());
  for (int p = 0; p < pa.P; ++p) {
    pa.centre.push_back(altaz(pa.alt[p], pa.azi[p]));
    pa.lb.push_back(p == pa.P - 1 ? make_float3(0.f, 0.f, 1.f)
                                  : altaz(std::atan(std::cos(kPi / pa.ninrow[p]) * std::tan(double(pa.alt[p]))), pa.azi[p]));
  }
  return pa;
}

float3 patch_sample(const Patches& pa, int p, int k, int K) {
  if (K <= 1) return pa.centre[size_t(p)];
  const int m = int(std::lround(std::sqrt(double(K))));
  const double u = ((k % m) + 0.5) / m, v = ((k / m) + 0.5) / m, a = pa.alpha;
  if (p == pa.P - 1) {   // zenith cap: cos theta uniform in [cos(alpha / 2), 1]
    const double ct = 1 - v * (1 - std::cos(a / 2)), st = std::sqrt(1 - ct * ct), az = 2 * kPi * u;
    return make_float3(float(st * std::sin(az)), float(st * std::cos(az)), float(ct));
  }
  const int r = pa.row[size_t(p)];
  const double s0 = std::sin(a * r), s1 = std::sin(a * (r + 1));
  return altaz(std::asin(s0 + v * (s1 - s0)), pa.azi[size_t(p)] + (u - 0.5) * 2 * kPi / pa.ninrow[size_t(p)]);
}

const double* tcos_table() {
  static const std::vector<double> t = [] {
    std::vector<double> v(1025);
    for (int k = 0; k <= 1024; ++k) v[size_t(k)] = std::cos((kPi / 2. / 1024) * k);
    return v;
  }();
  return t.data();
}

void hoy_to_date(double hoy, bool leap, int& month, int& day, double& hour) {
  const int doy = int(std::floor(hoy / 24));
  hour = hoy - doy * 24.0;
  const std::array<int, 13> start = month_starts(leap);
  month = 1;   // the last month that starts on or before doy (December takes the days past the year)
  while (month < 12 && doy >= start[size_t(month)]) ++month;
  day = doy - start[size_t(month - 1)] + 1;
}

bool parse_hoy(const std::string& s, bool leap, double& hoy, std::string& err) {
  int year, m, d, h, mi;
  const bool dated = std::sscanf(s.c_str(), "%d-%d-%dT%d:%d", &year, &m, &d, &h, &mi) == 5;
  if (!dated && std::sscanf(s.c_str(), "%d-%dT%d:%d", &m, &d, &h, &mi) != 4) {
    err = "bad timestamp '" + s + "', expected [YYYY-]MM-DDTHH:MM";
    return false;
  }
  const bool inRange = m >= 1 && m <= 12 && d >= 1 && d <= month_days(m, leap) && h >= 0 && h <= 23 && mi >= 0 && mi <= 59;
  if (!inRange) {
    err = "timestamp out of range '" + s + "'";
    return false;
  }
  hoy = 24.0 * (d - 1) + h + mi / 60.0;
  for (int k = 1; k < m; ++k) hoy += 24 * month_days(k, leap);   // whole months before, added one by one
  return true;
}

namespace {

// One step per weather row, evaluated at the row's middle (moved by `shift` into the weather's clock).
void annual_steps(const Weather& w, double shift, double yearH, std::vector<Step>& out) {
  out.resize(w.dni.size());
  for (size_t r = 0; r < out.size(); ++r) {
    Step& s = out[r];
    const double mid = double(r) + 0.5, at = mid + shift;
    s.hoy = at - yearH * std::floor(at / yearH);
    s.jd = int(r / 24) + 1;
    s.dt = 1;
    s.dni = w.dni[r];
    s.dhi = w.dhi[r];
    hoy_to_date(mid, w.leap, s.month, s.day, s.hour);
  }
}

// Steps of stepMin minutes from t0 up to t1 (one step at t0 when t1 is NaN), inside the [dayFrom, dayTo) minutes of
// every day: a step lasts to the next one, the range's end or the day window's end, whichever is first, and is
// evaluated at its instant (--time) or its middle; the weather is interpolated at that time in the weather's clock.
bool ranged_steps(const Options& opt, const Weather* w, double shift, double yearH, std::vector<Step>& out, std::string& err) {
  const bool leap = w && w->leap;
  if (opt.stepMin <= 0) { err = "step must be a positive number of minutes"; return false; }
  const bool instant = std::isnan(opt.t1), window = opt.dayFrom > 0 || opt.dayTo < 1440;
  const double step = opt.stepMin / 60.0, end = std::min(instant ? opt.t0 + step : opt.t1, yearH);
  if (opt.t0 < 0 || opt.t0 >= yearH || end <= opt.t0) { err = "time range must satisfy 0 <= from < to <= end of year"; return false; }
  for (int k = 0;; ++k) {
    const double start = opt.t0 + k * step;
    if (start >= end - 1e-9) break;
    double minute = (start - 24 * std::floor(start / 24)) * 60;   // into the day (the --hours window)
    if (minute >= 1440 - 1e-6) minute -= 1440;                     // a step at 24 n - eps is midnight's
    if (minute < opt.dayFrom - 1e-6 || minute >= opt.dayTo - 1e-6) continue;
    double dt = step;
    if (!instant) {
      dt = std::min(dt, end - start);
      if (window) dt = std::min(dt, (opt.dayTo - minute) / 60);
    }
    Step s;
    s.hoy = instant ? start : start + dt / 2;
    const double at = (s.hoy - shift) - yearH * std::floor((s.hoy - shift) / yearH);
    s.jd = int(at / 24) + 1;
    s.dt = dt;
    hoy_to_date(at, leap, s.month, s.day, s.hour);
    if (w && !w->dni.empty()) {
      s.dni = lerp_rows(w->dni, at);
      s.dhi = lerp_rows(w->dhi, at);
    }
    out.push_back(s);
  }
  if (out.empty()) { err = "no step falls inside the --hours window"; return false; }
  return true;
}

}  // namespace

bool make_steps(const Options& opt, const Weather* w, std::vector<Step>& out, std::string& err, double shift) {
  out.clear();
  const double yearH = w && w->leap ? 8784 : 8760;
  if (!std::isnan(opt.t0)) return ranged_steps(opt, w, shift, yearH, out, err);
  if (!w) { err = "no weather and no time range: nothing to evaluate"; return false; }
  annual_steps(*w, shift, yearH, out);
  return true;
}

double extraterrestrial_irradiance(int doy) {
  const double B = 2 * kPi * (doy - 1) / 365;
  return 1366.1 * (1.00011 + 0.034221 * std::cos(B) + 0.00128 * std::sin(B) + 0.000719 * std::cos(2 * B) + 0.000077 * std::sin(2 * B));
}

void ineichen(double z, double TL, double altitude, double dniExtra, double& ghi, double& dni, double& dhi) {
  z = std::min(std::max(z, 0.0), 90.0);
  const double cz = std::cos(rad(z));
  const double am = 1 / (cz + 0.50572 * std::pow(6.07995 + (90 - z), -1.6364));
  const double pres = 100 * std::pow((44331.514 - altitude) / 11880.516, 1 / 0.1902632);
  const double ama = am * pres / 101325;
  const double fh1 = std::exp(-altitude / 8000), fh2 = std::exp(-altitude / 1250);
  const double cg1 = 5.09e-5 * altitude + 0.868, cg2 = 3.92e-5 * altitude + 0.0387;
  ghi = std::max(0.0, cg1 * dniExtra * cz * std::exp(-cg2 * ama * (fh1 + fh2 * (TL - 1))));
  const double b = 0.664 + 0.163 / fh1;
  const double bnci = dniExtra * std::max(0.0, b * std::exp(-0.09 * ama * (TL - 1)));
  const double bnci2 = ghi * std::min(std::max((1 - (0.1 - 0.2 * std::exp(-TL)) / (0.1 + 0.882 / fh1)) / cz, 0.0), 1e20);
  dni = std::min(bnci, bnci2);
  dhi = ghi - dni * cz;
}

bool clear_sky_weather(double lat, double lon, double tz, double turbidity, double altitude, Weather& out, std::string& err) {
  if (turbidity < 1) { err = "turbidity must be >= 1"; return false; }
  if (std::fabs(lat) > 90) { err = "latitude out of range"; return false; }
  out.lat = lat; out.lon = lon; out.tz = tz; out.leap = false; out.missing = 0; out.shiftH = 0;
  out.dni.assign(8760, 0.f);
  out.dhi.assign(8760, 0.f);
  for (int r = 0; r < 8760; ++r) {
    int month, day;
    double hour;
    hoy_to_date(r + 0.5, false, month, day, hour);
    const dlm::SunPos s = dlm::sun_noaa(lat, lon, tz, month, day, hour, false);
    if (!s.up) continue;
    double ghi, dni, dhi;
    ineichen(90 - s.alt, turbidity, altitude, extraterrestrial_irradiance(r / 24 + 1), ghi, dni, dhi);
    out.dni[size_t(r)] = float(dni);
    out.dhi[size_t(r)] = float(dhi);
  }
  return true;
}

namespace {

// The first k commas of [p, eol) (their addresses into out): 16 bytes at a time where the line has them, so that no
// character costs a branch. Returns how many were found (<= k).
size_t first_commas(const char* p, const char* eol, const char** out, size_t k) {
  size_t n = 0;
  const char* c = p;
#if defined(__SSE2__)
  const __m128i comma = _mm_set1_epi8(',');
  for (; c + 16 <= eol && n < k; c += 16) {
    unsigned m = unsigned(_mm_movemask_epi8(_mm_cmpeq_epi8(_mm_loadu_si128(reinterpret_cast<const __m128i*>(c)), comma)));
    for (; m && n < k; m &= m - 1) out[n++] = c + __builtin_ctz(m);
  }
#endif
  for (; c < eol && n < k; ++c)
    if (*c == ',') out[n++] = c;
  return n;
}

// The EPW text in memory: the 8 header lines, then the data rows, scanned in place (columns 13-15: GHI, DNI, DHI, read
// as atof reads a column).
bool parse_epw(const char* data, size_t size, const std::string& label, Weather& w, std::string& err) {
  const char* p = data;
  const char* const end = data + size;
  std::string line;
  std::vector<std::string> cols;
  const auto split = [&](const std::string& l) {
    cols.clear();
    std::stringstream ss(l);
    for (std::string c; std::getline(ss, c, ',');) cols.push_back(c);
  };
  bool hdrLeap = false, pvgis = false;
  double declared = NAN;
  for (int i = 0; i < 8; ++i) {   // getline's lines: up to '\n' (not kept); the last one may end at the buffer's end
    if (p >= end) { err = "EPW header truncated: " + label; return false; }
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', size_t(end - p)));
    line.assign(p, eol ? eol : end);
    p = eol ? eol + 1 : end;
    split(line);
    if (i == 0) {
      if (cols.size() < 10 || cols[0] != "LOCATION") { err = "EPW: bad LOCATION line: " + label; return false; }
      w.lat = std::atof(cols[6].c_str()); w.lon = std::atof(cols[7].c_str()); w.tz = std::atof(cols[8].c_str());
    }
    if (i == 4 && cols.size() > 1 && cols[1] == "Yes") hdrLeap = true;
    if (const size_t q = line.find("Irradiance Time Offset"); q != std::string::npos) {
      pvgis = true;
      if (const size_t c = line.find(':', q); c != std::string::npos) declared = std::atof(line.c_str() + c + 1);
    }
  }
  w.dni.clear(); w.dhi.clear(); w.missing = 0; w.shiftH = 0;
  w.dni.reserve(8784);
  w.dhi.reserve(8784);
  std::vector<float> ghi;
  std::string field;
  // atof of a field. A plain decimal ([-]digits[.digits], at most 15 digits, then the field's end) is exactly
  // mantissa / 10^k with both parts exact in double, so one correctly rounded division gives what strtod gives (Clinger
  // 1990's fast path); anything else goes to from_chars (correctly rounded, as glibc's strtod) where it takes the whole
  // number, else to atof itself.
  static const double kPow10[16] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11, 1e12, 1e13, 1e14, 1e15};
  const auto number = [&](const char* a, const char* b) {
    const char* c = a;
    while (c < b && (*c == ' ' || *c == '\t')) ++c;
    if (c < b && *c == '+' && c + 1 < b && *(c + 1) != '-' && *(c + 1) != '+') ++c;
    {
      const char* d = c;
      const bool neg = d < b && *d == '-';
      if (neg) ++d;
      uint64_t m = 0;
      int digits = 0, frac = 0;
      bool dot = false, ok = d < b;
      for (; d < b && *d != '\r'; ++d) {
        if (*d >= '0' && *d <= '9') { m = m * 10 + uint64_t(*d - '0'); ++digits; frac += dot; }
        else if (*d == '.' && !dot) dot = true;
        else { ok = false; break; }
      }
      if (ok && digits > 0 && digits <= 15 && (d == b || *d == '\r')) {
        const double x = double(m) / kPow10[frac];
        return neg ? -x : x;
      }
    }
    double x = 0;
    const auto r = std::from_chars(c, b, x);
    if (r.ec == std::errc() && (r.ptr == b || *r.ptr == '\r' || *r.ptr == ' ' || *r.ptr == '\t')) return x;
    field.assign(a, b);
    return std::atof(field.c_str());
  };
  while (p < end) {
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', size_t(end - p)));
    if (!eol) eol = end;
    const char* q = p;
    while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q == eol) { p = eol + 1; continue; }
    // field starts up to the 17th: the field count (a getline split's: commas + 1, less a trailing empty field) only
    // has to reach 16, which 16 commas settle; short rows are counted out in full
    const char* cm[16];
    const char* at[17];
    size_t n = 1 + first_commas(p, eol, cm, 16);
    at[0] = p;
    for (size_t k = 1; k < n; ++k) at[k] = cm[k - 1] + 1;
    if (n < 17 && eol > p && eol[-1] == ',') --n;
    if (n < 16) { err = "EPW: short data row " + std::to_string(w.dni.size() + 1) + ": " + label; return false; }
    const auto colEnd = [&](size_t k) { return k + 1 < n ? at[k + 1] - 1 : eol; };
    double dni = number(at[14], colEnd(14)), dhi = number(at[15], colEnd(15));
    if (dni >= 9999 || dhi >= 9999) { ++w.missing; if (dni >= 9999) dni = 0; if (dhi >= 9999) dhi = 0; }
    w.dni.push_back(float(dni));
    w.dhi.push_back(float(dhi));
    if (pvgis) { const double g = number(at[13], colEnd(13)); ghi.push_back(g < 9999 ? float(g) : 0.f); }
    p = eol + 1;
  }
  const size_t n = w.dni.size();
  w.leap = hdrLeap || n == 8784;
  if (n != (w.leap ? 8784u : 8760u)) {
    err = "EPW: " + std::to_string(n) + " data rows, expected " + (w.leap ? "8784" : "8760") + ": " + label;
    return false;
  }
  if (pvgis) pvgis_align(w, ghi, declared);
  return true;
}

}  // namespace

bool read_epw(const std::string& path, Weather& w, std::string& err) {
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path.c_str(), "rb"), std::fclose);
  if (!f) { err = "cannot open EPW " + path; return false; }
  std::string data;
  if (std::fseek(f.get(), 0, SEEK_END) == 0) {   // one read of the whole file
    const long size = std::ftell(f.get());
    if (size > 0 && std::fseek(f.get(), 0, SEEK_SET) == 0) {
      data.resize(size_t(size));
      data.resize(std::fread(data.data(), 1, data.size(), f.get()));
    }
  }
  if (data.empty()) {   // not seekable: read it in pieces
    std::rewind(f.get());
    char buf[1 << 16];
    for (size_t k; (k = std::fread(buf, 1, sizeof buf, f.get())) > 0;) data.append(buf, k);
  }
  return parse_epw(data.data(), data.size(), path, w, err);
}

bool read_epw(std::istream& f, const std::string& label, Weather& w, std::string& err) {
  std::string data;
  char buf[1 << 16];
  while (f.read(buf, sizeof buf) || f.gcount() > 0) data.append(buf, size_t(f.gcount()));
  return parse_epw(data.data(), data.size(), label, w, err);
}

}  // namespace daylight
