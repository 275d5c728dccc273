// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Radiance notice: the Perez tables, tcos and glass formulas here follow Radiance 5.4 (gendaymtx.c, tcos.c, m_glass.c),
// Copyright (c) 1990 to 2022 The Regents of the University of California, through LBNL; Radiance Software License 2.0,
// reproduced with its conditions and disclaimer in NOTICES.md.
// Ladybug mode's sky (reference run.cpp:362-470): gendaymtx -A as Ladybug's RadiationStudy runs it, computed on the host
// in double with the host's libm, since the emulation is exact only where its operations are the reference's: per step
// the sun of Radiance's sun.c at the site as Ladybug writes it into the .wea (latitude and west longitude %.2f, the
// meridian of the file's whole-hour zone, DNI and DHI truncated to integers), the -d and -s rows of gendaymtx, float
// sums over all rows in step order, the averages printed %.3g, then broadband x omega x hours / 1000 (kWh/m2).
// The steps' rows are made in parallel blocks on the host threads and summed in order after each block.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "core.h"
#include "dl_math.h"

namespace daylight {

namespace {
constexpr double kPi = 3.14159265358979323846;
constexpr double kRf = 0.265074126, kGf = 0.670114631, kBf = 0.064811243;   // Radiance color.h CIE_rf/gf/bf
double bright(const float* c) { return kRf * c[0] + kGf * c[1] + kBf * c[2]; }
}  // namespace

// Radiance sun.c (rad5R4 sdec, stadj, salt, sazi), reference sky.cpp:476-486: latitude north, longitude and meridian
// west-positive, degrees in; altitude and azimuth (from north through east, + pi as gendaymtx turns it) in radians.
void sun_radiance(double latDeg, double lonWestDeg, double meridianWestDeg, int jd, double hour, double& alt, double& azi) {
  const double la = latDeg * (kPi / 180.), lo = lonWestDeg * (kPi / 180.), me = meridianWestDeg * (kPi / 180.);
  const double sd = 0.4093 * std::sin((2. * kPi / 365.) * (jd - 81));
  const double sta = 0.170 * std::sin((4. * kPi / 373.) * (jd - 80)) - 0.129 * std::sin((2. * kPi / 355.) * (jd - 8)) + (12. / kPi) * (me - lo);
  const double st = hour + sta;
  alt = std::asin(std::sin(la) * std::sin(sd) - std::cos(la) * std::cos(sd) * std::cos(st * (kPi / 12.)));
  azi = -std::atan2(std::cos(sd) * std::sin(st * (kPi / 12.)), -std::cos(la) * std::sin(sd) - std::sin(la) * std::cos(sd) * std::cos(st * (kPi / 12.))) + kPi;
}

// a value as printf would print it with fmt and a reader would read it back (sky.cpp:488-493)
double round_fmt(double x, const char* fmt) {
  char buf[64];
  std::snprintf(buf, sizeof buf, fmt, x);
  return std::strtod(buf, nullptr);
}

// gendaymtx for one step, radiometric (-O1; sky.cpp:297-474): out[3 (P + 1)], row 0 the ground (the sky's reflectance
// only with the sky), rows 1 + p the patches: the Perez sky scaled to the diffuse horizontal (table sine) and coloured,
// and/or the sun split over the 4 patch centres nearest it (AddDirect, weights 1 / (1.002 - cos)).
void gendaymtx_rgb(const Patches& pa, const double* tsinP, int jd, double alt, double azi, double dni, double dhi, double grefl, bool sun, bool sky,
                   float* out) {
  const size_t P = size_t(pa.P);
  std::memset(out, 0, sizeof(float) * 3 * (P + 1));
  if (dni + dhi <= 1e-4) return;
  const dlm::PerezStep q = dlm::perez_step(jd, alt, dni, dhi);
  const float grf = sky ? float(grefl) : 0.f;
  out[0] = out[1] = out[2] = q.gRad * grf;
  if (sky) {
    const float col[3] = {.96f, 1.004f, 1.118f};
    double norm = 0.0;
    for (size_t p = 0; p < P; ++p) {
      const float L = dlm::perez_patch(q, pa.alt[p], pa.azi[p], azi);
      out[3 * (p + 1)] = out[3 * (p + 1) + 1] = out[3 * (p + 1) + 2] = L;
      norm += L * tsinP[p] * pa.omega[p];
    }
    if (norm <= 1e-6) {
      for (size_t p = 0; p < P; ++p) out[3 * (p + 1)] = out[3 * (p + 1) + 1] = out[3 * (p + 1) + 2] = 1.f;
      norm = kPi;
    }
    const double f = (q.difRad / norm) * (1. / 179.);
    for (size_t p = 0; p < P; ++p)
      for (int ch = 0; ch < 3; ++ch) out[3 * (p + 1) + size_t(ch)] = float(out[3 * (p + 1) + size_t(ch)] * f) * col[ch];
  }
  if (q.dirRad <= 1e-4 || !sun) return;
  double best[4] = {-1., -1., -1., -1.};
  size_t bestP[4] = {0, 0, 0, 0};
  const double ca = std::cos(alt), sv0 = ca * std::sin(azi), sv1 = ca * std::cos(azi), sv2 = std::sin(alt);
  for (size_t p = 1; p <= P; ++p) {
    const double pc = std::cos(double(pa.alt[p - 1]));
    const double dp = (pc * std::sin(double(pa.azi[p - 1]))) * sv0 + (pc * std::cos(double(pa.azi[p - 1]))) * sv1 + std::sin(double(pa.alt[p - 1])) * sv2;
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
    float add = float(w[i] * q.dirRad / (179. * wsum));
    add /= pa.omega[bestP[i] - 1];
    for (int ch = 0; ch < 3; ++ch) out[3 * bestP[i] + size_t(ch)] += add * 1.f;
  }
}

// Radiance tcos.c's table sine of each patch altitude (nearest of 1024 entries per quadrant)
std::vector<double> patch_table_sine(const Patches& pa) {
  const double* tc = tcos_table();
  std::vector<double> s(size_t(pa.P));
  for (size_t p = 0; p < s.size(); ++p) {
    double x = double(pa.alt[p]) - kPi / 2.;
    if (x < 0.) x = -x;
    int k = int((1024 * 2. / kPi) * x + 0.5);
    while (k >= 4096) k -= 4096;
    const int r = k >> 10;
    s[p] = r == 0 ? tc[k] : r == 1 ? -tc[2048 - k] : r == 2 ? -tc[k - 2048] : tc[4096 - k];
  }
  return s;
}

void ladybug_sky(const Patches& pa, const std::vector<Step>& steps, double lat, double lon, double fileTz, double grefl, double durationHours,
                 const std::vector<int>& colOf, float* S, float* gS, int Tup, LadybugSky& out) {
  const size_t P = size_t(pa.P), T = steps.size(), row = 3 * (P + 1);
  const std::vector<double> tsinP = patch_table_sine(pa);
  const double latW = round_fmt(lat, "%.2f"), lonWestW = round_fmt(-lon, "%.2f"), meridW = double(int(-fileTz * 15.0));
  std::vector<float> accDir(P, 0.f), accDif(3 * P, 0.f), gDif(3, 0.f), gDir(3, 0.f);
  constexpr size_t kBlock = 512;
  std::vector<float> rows(std::min(T, kBlock) * 2 * row);
  for (size_t t0 = 0; t0 < T; t0 += kBlock) {
    const size_t n = std::min(kBlock, T - t0);
    par_for(n, [&](size_t k) {
      const Step& s = steps[t0 + k];
      double altS, aziS;
      sun_radiance(latW, lonWestW, meridW, s.jd, s.hour, altS, aziS);
      const double dniW = double(int(s.dni)), dhiW = double(int(s.dhi));
      float* dr = rows.data() + 2 * row * k;
      gendaymtx_rgb(pa, tsinP.data(), s.jd, altS, aziS, dniW, dhiW, grefl, true, false, dr);          // -d
      gendaymtx_rgb(pa, tsinP.data(), s.jd, altS, aziS, dniW, dhiW, grefl, false, true, dr + row);    // -s
    });
    for (size_t k = 0; k < n; ++k) {   // the float sums, in step order
      const float* dr = rows.data() + 2 * row * k;
      const float* df = dr + row;
      for (int ch = 0; ch < 3; ++ch) { gDir[size_t(ch)] += dr[ch]; gDif[size_t(ch)] += df[ch]; }
      const int col = colOf.empty() ? -1 : colOf[t0 + k];
      double meanPatch = 0;
      for (size_t p = 0; p < P; ++p) {
        const float* d = dr + 3 * (p + 1);
        const float* f = df + 3 * (p + 1);
        accDir[p] += d[0];
        for (int ch = 0; ch < 3; ++ch) accDif[3 * p + size_t(ch)] += f[ch];
        if (S && col >= 0) {
          const double Lp = bright(f) + d[0];
          S[p * size_t(Tup) + size_t(col)] = float(Lp);
          meanPatch += Lp * pa.omega[p];
        }
      }
      if (S && col >= 0) gS[col] = float(grefl * meanPatch / double(P));
    }
  }
  const float inv = float(1.0 / double(T));
  const auto avg = [&](float a) { return T > 1 ? a * inv : a; };
  out.dump.assign(2 * row, 0.f);
  for (size_t ch = 0; ch < 3; ++ch) {
    out.dump[ch] = float(round_fmt(avg(gDif[ch]), "%.3g"));
    out.dump[row + ch] = float(round_fmt(avg(gDir[ch]), "%.3g"));
  }
  out.dir.assign(P, 0.0);
  out.dif.assign(P, 0.0);
  out.rad.assign(P, 0.0);
  for (size_t p = 0; p < P; ++p) {
    const double rd = round_fmt(avg(accDir[p]), "%.3g"), rr = round_fmt(avg(accDif[3 * p]), "%.3g"), rg = round_fmt(avg(accDif[3 * p + 1]), "%.3g"),
                 rb = round_fmt(avg(accDif[3 * p + 2]), "%.3g");
    out.dump[3 * (p + 1)] = float(rr);
    out.dump[3 * (p + 1) + 1] = float(rg);
    out.dump[3 * (p + 1) + 2] = float(rb);
    out.dump[row + 3 * (p + 1)] = out.dump[row + 3 * (p + 1) + 1] = out.dump[row + 3 * (p + 1) + 2] = float(rd);
    out.dir[p] = rd * pa.omega[p] * durationHours / 1000.0;
    out.dif[p] = (kRf * rr + kGf * rg + kBf * rb) * pa.omega[p] * durationHours / 1000.0;
    out.rad[p] = out.dir[p] + out.dif[p];
  }
  double mean = 0;
  for (size_t p = 0; p < P; ++p) mean += out.rad[p];
  out.gnd = grefl * mean / double(P);
}

}  // namespace daylight
