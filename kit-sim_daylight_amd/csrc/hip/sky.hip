// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Radiance notice: the Perez tables, tcos and glass formulas here follow Radiance 5.4 (gendaymtx.c, tcos.c, m_glass.c),
// Copyright (c) 1990 to 2022 The Regents of the University of California, through LBNL; Radiance Software License 2.0,
// reproduced with its conditions and disclaimer in NOTICES.md.
// The sky of a daylight run on the GPU in FP64: replaces the host loops of reference src/run.cpp:311-446 (sun_noaa per
// step, gendaymtx_step per step, the cumulative per-patch weights) with kernels over steps and (step, patch).
//   k_steps   a thread per step: sun position (NOAA, Ladybug conventions), Perez parameters, ground row, the sun's
//             direct illuminance as AddDirect spreads it (for illuminance), the sun's weights
//   k_lum     a thread per (step, patch): Perez relative luminance (float, as gendaymtx's array)
//   k_norm    a thread per step: the horizontal-illuminance normalisation in patch order (table sine), scale factors
//   k_apply   a thread per (step, patch): the patch radiance bright() per mode (double) and the series matrix S
//   k_accum   a thread per patch: sum over steps in step order (the reference's double sums, same order)
// Float / double rounding points follow gendaymtx (reference src/sky.cpp:297-474), so S and the weights agree with the
// host model after float rounding; libm differences (device vs host double transcendental functions) stay below the
// float rounding of the outputs except in rare last-bit cases.
#include <vector>

#include <algorithm>

#include "dl_math.h"
#include "gpu_internal.h"

namespace {

using namespace dlm;

struct StepOut {
  PerezStep q;
  double altRad, aziRad;
  double sfRad, sfVis;
  int uniform;                 // zero sky: every patch 1
  double gRadB, gVisB;         // ground bright (radiometric W/m2/sr, visible x179)
  double sunVis, sunRad;       // the sun's weights of this step
  // the sun as AddDirect spreads it (gendaymtx -d) over its 4 nearest patches: the patches and each one's value, visible
  // and radiometric (0: not spread, e.g. at night)
  int binP[4];
  float binVis[4], binRad[4];
};

// the patch centres' unit vectors as AddDirect's search dots them with the sun (cos alt sin azi, cos alt cos azi,
// sin alt, each from the float angles in FP64): made once per sky instead of once per step and patch
__global__ void k_patch_dirs(int P, const float* __restrict__ palt, const float* __restrict__ pazi, double* __restrict__ pd) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double pa = cos((double)palt[p]);
  pd[3 * p] = pa * sin((double)pazi[p]);
  pd[3 * p + 1] = pa * cos((double)pazi[p]);
  pd[3 * p + 2] = sin((double)palt[p]);
}

__global__ void k_steps(int T, const int* __restrict__ month, const int* __restrict__ day, const int* __restrict__ jd,
                        const double* __restrict__ hour, const double* __restrict__ dt, const float* __restrict__ dni,
                        const float* __restrict__ dhi, double lat, double lon, double tz, double north, int leap, float grefl, int wantVis,
                        int reflect, int P, const double* __restrict__ pd, const float* __restrict__ omega,
                        StepOut* __restrict__ so, double* __restrict__ alt, double* __restrict__ azi, int* __restrict__ up,
                        float* __restrict__ vec) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const SunPos s = sun_noaa(lat, lon, tz, month[t], day[t], hour[t], leap != 0);
  alt[t] = s.alt;
  azi[t] = s.azi;
  up[t] = s.up ? 1 : 0;
  const float3 v = rotate_z(s.vec, north);
  vec[3 * t] = v.x; vec[3 * t + 1] = v.y; vec[3 * t + 2] = v.z;
  StepOut o{};
  o.altRad = s.alt * (kPi / 180.0);
  o.aziRad = s.azi * (kPi / 180.0);
  o.q = perez_step(jd[t], o.altRad, dni[t], dhi[t]);
  o.gRadB = o.q.night ? 0.0 : ground_bright(o.q.gRad, grefl);
  o.gVisB = o.q.night ? 0.0 : 179.0 * ground_bright(o.q.gVis, grefl);
  o.sunRad = dni[t] * dt[t] / 1000.0;
  const bool vis = (wantVis || reflect) && o.q.dirVis > 1e-4, rad = reflect && o.q.dirRad > 1e-4;
  if (!o.q.night && (vis || rad)) {   // AddDirect (-d): the 4 patch centres nearest the sun (sky.cpp:435-473)
    double best[4] = {-1., -1., -1., -1.};
    int bestP[4] = {0, 0, 0, 0};
    const double ca = cos(o.altRad), sv0 = ca * sin(o.aziRad), sv1 = ca * cos(o.aziRad), sv2 = sin(o.altRad);
    constexpr int kRun = 8;   // the patch vectors' loads of a run issued ahead of its ordered insertions
    for (int p0 = 0; p0 < P; p0 += kRun) {
      double dps[kRun];
#pragma unroll
      for (int k = 0; k < kRun; ++k) {
        const int p = p0 + k < P ? p0 + k : P - 1;
        dps[k] = pd[3 * p] * sv0 + pd[3 * p + 1] * sv1 + pd[3 * p + 2] * sv2;
      }
      for (int k = 0; k < kRun && p0 + k < P; ++k) {
        const double dp = dps[k];
        for (int i = 0; i < 4; ++i)
          if (dp > best[i]) {
            for (int j = 3; j > i; --j) { best[j] = best[j - 1]; bestP[j] = bestP[j - 1]; }
            best[i] = dp;
            bestP[i] = p0 + k;
            break;
          }
      }
    }
    double w[4], wsum = 0;
    for (int i = 3; i >= 0; --i) wsum += w[i] = 1. / (1.002 - best[i]);
    // each patch's value after AddDirect (the row zeroed first, -d): the added float, per mode
    for (int i = 0; i < 4; ++i) {
      o.binP[i] = bestP[i];
      float addV = (float)(w[i] * o.q.dirVis / (179. * wsum)), addR = (float)(w[i] * o.q.dirRad / (179. * wsum));
      addV /= omega[bestP[i]];
      addR /= omega[bestP[i]];
      o.binVis[i] = vis ? 0.f + addV * 1.f : 0.f;
      o.binRad[i] = rad ? 0.f + addR * 1.f : 0.f;
    }
    if (vis && wantVis) {   // the direct illuminance: the patches' green channel x omega summed in patch order
      int ord[4] = {0, 1, 2, 3};
      for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b)
          if (bestP[ord[b]] < bestP[ord[a]]) { const int x = ord[a]; ord[a] = ord[b]; ord[b] = x; }
      double dirIll = 0;
      for (int k = 0; k < 4; ++k) {
        const int i = ord[k];
        dirIll += (double)(o.binVis[i] * omega[bestP[i]]);   // float x float, as run.cpp:462 multiplies the float row by omega
      }
      o.sunVis = 179.0 * dirIll * dt[t];
    }
  }
  so[t] = o;
}

__global__ void k_lum(int T, int P, const StepOut* __restrict__ so, const float* __restrict__ palt, const float* __restrict__ pazi,
                      float* __restrict__ L) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
  if (p >= P) return;
  const StepOut& o = so[t];
  L[size_t(t) * P + p] = o.q.night ? 0.f : perez_patch(o.q, palt[p], pazi[p], o.aziRad);
}

// A step's normalisation, the sum over its patches in patch order of L tsin omega (sky.cpp's order): a lane per step,
// its row read 16 patches at a time with the loads issued ahead of the 16 dependent adds.
constexpr int kNormRun = 16;
__global__ void k_norm(int T, int P, StepOut* __restrict__ so, const float* __restrict__ L, const double* __restrict__ tsinP,
                       const float* __restrict__ omega) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  StepOut& o = so[t];
  if (o.q.night) return;
  double norm = 0.0;
  const float* Lt = L + size_t(t) * P;
  for (int p0 = 0; p0 < P; p0 += kNormRun) {
    float l[kNormRun], om[kNormRun];
    double ts[kNormRun];
#pragma unroll
    for (int k = 0; k < kNormRun; ++k) {
      const int p = p0 + k < P ? p0 + k : P - 1;
      l[k] = Lt[p]; ts[k] = tsinP[p]; om[k] = omega[p];
    }
#pragma unroll
    for (int k = 0; k < kNormRun; ++k)
      if (p0 + k < P) norm += l[k] * ts[k] * om[k];
  }
  o.uniform = norm <= 1e-6;
  if (o.uniform) norm = kPi;
  o.sfRad = (o.q.difRad / norm) * (1. / 179.);
  o.sfVis = (o.q.difVis / norm) * (1. / 179.);
}

__global__ void k_apply(int T, int P, const StepOut* __restrict__ so, const float* __restrict__ L, const double* __restrict__ dt, int wantRad, int wantVis,
                        const int* __restrict__ colOf, int Tup, int seriesMode, double* __restrict__ LpRad, double* __restrict__ LpVis,
                        float* __restrict__ S) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
  if (p >= P) return;
  const StepOut& o = so[t];
  const size_t at = size_t(t) * P + p;
  double r = 0.0, v = 0.0;
  if (!o.q.night) {
    const float l = o.uniform ? 1.f : L[at];
    if (wantRad || seriesMode == 1) r = sky_bright(l, o.sfRad);
    if (wantVis || seriesMode == 2) v = 179.0 * sky_bright(l, o.sfVis);
  }
  // the step's contribution to the cumulative weights (run.cpp:436-446), 0 at night
  if (wantRad) LpRad[at] = r * dt[t] / 1000.0;
  if (wantVis) LpVis[at] = v * dt[t];
  const int col = colOf[t];
  if (seriesMode && col >= 0) S[size_t(p) * Tup + col] = (float)(seriesMode == 1 ? r : v);
}

// skyRad[p] = sum over the steps, in step order, of the contributions c[t P + p] (a step at night adds 0.0, which
// leaves the sum's bits unchanged): 4 patches a block (145 blocks at P = 577), tiles of 256 steps staged through LDS so
// that the sequential sum of each patch reads LDS, not HBM, while the next tile's loads are in flight.
constexpr int kAccP = 4, kAccT = 256;
__global__ __launch_bounds__(256) void k_accum(int T, int P, const double* __restrict__ c, double* __restrict__ sum) {
  __shared__ double tile[kAccT][kAccP];
  // thread (k, lp): patch p0 + lp at steps t0 + k + 64 j, j < 4; the next tile is loaded into registers while the
  // lanes lp < kAccP add this one's kAccT values in step order
  const int lp = threadIdx.x % kAccP, k = threadIdx.x / kAccP, p = blockIdx.x * kAccP + lp;
  double nxt[4];
  const auto load = [&](int t0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int t = t0 + k + 64 * j;
      nxt[j] = (t < T && p < P) ? c[size_t(t) * P + p] : 0.0;
    }
  };
  load(0);
  double r = 0.0;
  for (int t0 = 0; t0 < T; t0 += kAccT) {
#pragma unroll
    for (int j = 0; j < 4; ++j) tile[k + 64 * j][lp] = nxt[j];
    __syncthreads();
    if (t0 + kAccT < T) load(t0 + kAccT);
    if (threadIdx.x < kAccP) {
      const int n = T - t0 < kAccT ? T - t0 : kAccT;
      for (int q = 0; q < n; ++q) r += tile[q][lp];
    }
    __syncthreads();
  }
  if (threadIdx.x < kAccP && p < P) sum[p] = r;
}

// per step: the sun's weights, the ground's series value and the ground's contributions to the cumulative weights
__global__ void k_ground_steps(int T, const StepOut* __restrict__ so, const double* __restrict__ dt, const int* __restrict__ colOf,
                               int seriesMode, double* __restrict__ cR, double* __restrict__ cV, float* __restrict__ gS,
                               double* __restrict__ sunRad, double* __restrict__ sunVis) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const StepOut& o = so[t];
  sunRad[t] = o.sunRad;
  sunVis[t] = o.sunVis;
  cR[t] = o.q.night ? 0.0 : o.gRadB * dt[t] / 1000.0;
  cV[t] = o.q.night ? 0.0 : o.gVisB * dt[t];
  if (seriesMode && !o.q.night && colOf[t] >= 0) gS[colOf[t]] = (float)(seriesMode == 1 ? o.gRadB : o.gVisB);
}

// gnd[0] / gnd[1]: the ground contributions summed in step order (one block; tiles of 256 steps through LDS)
__global__ __launch_bounds__(256) void k_ground_sum(int T, const double* __restrict__ cR, const double* __restrict__ cV, double* __restrict__ gnd) {
  __shared__ double a[256], b[256];
  double r = 0.0, v = 0.0;
  for (int t0 = 0; t0 < T; t0 += 256) {
    const int t = t0 + int(threadIdx.x);
    a[threadIdx.x] = t < T ? cR[t] : 0.0;
    b[threadIdx.x] = t < T ? cV[t] : 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int n = T - t0 < 256 ? T - t0 : 256;
      for (int k = 0; k < n; ++k) { r += a[k]; v += b[k]; }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { gnd[0] = r; gnd[1] = v; }
}

// reflected light's bin radiance per step (run.cpp:412-442): row 0 the ground g, row 1 + p the patch's sky S plus the sun
// binned by AddDirect (float S + float(bright of the sun's value)); a thread per (bin, column)
__global__ void k_sdc_sky(int P, int Tup, const float* __restrict__ S, const float* __restrict__ gS, float* __restrict__ Sdc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (c >= Tup) return;
  Sdc[size_t(b) * Tup + c] = b == 0 ? gS[c] : S[size_t(b - 1) * Tup + c];
}
__global__ void k_sdc_sun(int T, int Tup, const StepOut* __restrict__ so, const int* __restrict__ colOf, int seriesVis, const float* __restrict__ S,
                          float* __restrict__ Sdc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T || colOf[t] < 0) return;
  const StepOut& o = so[t];
  for (int i = 0; i < 4; ++i) {
    const float v = seriesVis ? o.binVis[i] : o.binRad[i];
    if (v == 0.f) continue;   // not spread: S + 0 = S
    const double Ls = (seriesVis ? 179.0 : 1.0) * sun_bright(v);
    const int p = o.binP[i];
    Sdc[size_t(p + 1) * Tup + colOf[t]] = S[size_t(p) * Tup + colOf[t]] + (float)Ls;
  }
}

inline unsigned blocks(int64_t n, int b) { return unsigned((n + b - 1) / b); }

}  // namespace

extern "C" int32_t dlg_sky(dlg_gpu* g, const dlg_sky_in* in, dlg_sky_out* out, char* err, int32_t errLen) {
  const int T = in->T, mf = in->mf;
  static const int kRowAz[7] = {30, 30, 24, 24, 18, 12, 6};
  // the patch table (reference sky.cpp:240-270), as the oracle and gendaymtx compute it: float angles and solid angles
  std::vector<float> palt, pazi, om;
  const double alpha = (kPi / 2.) / (7 * mf + .5);
  for (int r = 0; r < 7 * mf; ++r) {
    const int n = kRowAz[r / mf] * mf;
    const float a = (float)(alpha * (r + .5));
    const float w = (float)(2. * kPi * (std::sin(alpha * (r + 1)) - std::sin(alpha * r)) / (double)n);
    for (int j = 0; j < n; ++j) { palt.push_back(a); pazi.push_back((float)(2. * kPi * j / (double)n)); om.push_back(w); }
  }
  palt.push_back((float)(kPi / 2.)); pazi.push_back(0.f); om.push_back((float)(2. * kPi * (1. - std::cos(alpha * .5))));
  const int P = int(palt.size());
  std::vector<double> tsinP(static_cast<size_t>(P));
  for (int p = 0; p < P; ++p) {   // Radiance tcos.c table sine of each patch altitude (1024 entries, nearest, no interpolation)
    double x = (double)palt[p] - kPi / 2.;
    if (x < 0.) x = -x;
    int k = (int)((1024 * 2. / kPi) * x + 0.5);
    while (k >= 4096) k -= 4096;
    const int q = k >> 10;
    tsinP[p] = q == 0 ? in->tcos[k] : q == 1 ? -in->tcos[2048 - k] : q == 2 ? -in->tcos[k - 2048] : in->tcos[4096 - k];
  }
  std::vector<int> colOf(static_cast<size_t>(T));
  int Tup = 0;
  for (int t = 0; t < T; ++t) colOf[t] = in->dni[t] + in->dhi[t] > 1e-4 ? Tup++ : -1;
  *out->Tup = Tup;
  const int seriesMode = in->series ? (in->seriesVis ? 2 : 1) : 0;
  // device work area: inputs, step records, L, Lp (per mode), outputs
  size_t off = 0;
  const auto take = [&](size_t bytes) { const size_t at = off; off += (bytes + 255) & ~size_t(255); return at; };
  const size_t oMonth = take(4 * T), oDay = take(4 * T), oJd = take(4 * T), oHour = take(8 * T), oDt = take(8 * T), oDni = take(4 * T),
               oDhi = take(4 * T), oPalt = take(4 * P), oPazi = take(4 * P), oOm = take(4 * P), oTs = take(8 * P), oCol = take(4 * T),
               oSo = take(sizeof(StepOut) * T), oL = take(size_t(4) * T * P), oLr = take(in->wantRad ? size_t(8) * T * P : 8),
               oLv = take(in->wantVis ? size_t(8) * T * P : 8), oAlt = take(8 * T), oAzi = take(8 * T), oUp = take(4 * T),
               oVec = take(12 * T), oSkyR = take(8 * P), oSkyV = take(8 * P), oGnd = take(16), oSunR = take(8 * T), oSunV = take(8 * T),
               oGr = take(8 * T), oGv = take(8 * T), oPd = take(24 * size_t(P));
  DLG_CHECK(g->skyWork.grow(off));
  char* b = g->skyWork.as<char>();
  const auto up = [&](size_t at, const void* src, size_t bytes) { return hipMemcpyAsync(b + at, src, bytes, hipMemcpyHostToDevice, g->stream); };
  DLG_CHECK(up(oMonth, in->month, 4 * T)); DLG_CHECK(up(oDay, in->day, 4 * T)); DLG_CHECK(up(oJd, in->jd, 4 * T));
  DLG_CHECK(up(oHour, in->hour, 8 * T)); DLG_CHECK(up(oDt, in->dt, 8 * T)); DLG_CHECK(up(oDni, in->dni, 4 * T));
  DLG_CHECK(up(oDhi, in->dhi, 4 * T)); DLG_CHECK(up(oPalt, palt.data(), 4 * P)); DLG_CHECK(up(oPazi, pazi.data(), 4 * P));
  DLG_CHECK(up(oOm, om.data(), 4 * P)); DLG_CHECK(up(oTs, tsinP.data(), 8 * P)); DLG_CHECK(up(oCol, colOf.data(), 4 * T));
  if (seriesMode) {
    DLG_CHECK(g->skyS.grow(size_t(P) * size_t(Tup > 0 ? Tup : 1) * 4));
    DLG_CHECK(g->skyG.grow(size_t(Tup > 0 ? Tup : 1) * 4));
    g->skyP = P;
    g->skyT = Tup;
    g->skyReflect = in->reflect ? 1 : 0;
    if (in->reflect) DLG_CHECK(g->skySdc.grow(size_t(P + 1) * size_t(Tup > 0 ? Tup : 1) * 4));
  }
#define PTR(T_, o) reinterpret_cast<T_*>(b + (o))
  dlg_tbegin(g);
  k_patch_dirs<<<blocks(P, 256), 256, 0, g->stream>>>(P, PTR(float, oPalt), PTR(float, oPazi), PTR(double, oPd));
  k_steps<<<blocks(T, 64), 64, 0, g->stream>>>(T, PTR(int, oMonth), PTR(int, oDay), PTR(int, oJd), PTR(double, oHour), PTR(double, oDt),
                                               PTR(float, oDni), PTR(float, oDhi), in->lat, in->lon, in->tz, in->north, in->leap,
                                               (float)in->grefl, in->wantVis, in->reflect, P, PTR(double, oPd), PTR(float, oOm),
                                               PTR(StepOut, oSo), PTR(double, oAlt), PTR(double, oAzi), PTR(int, oUp), PTR(float, oVec));
  if (T > 0) {
    k_lum<<<dim3(blocks(P, 256), unsigned(T)), 256, 0, g->stream>>>(T, P, PTR(StepOut, oSo), PTR(float, oPalt), PTR(float, oPazi), PTR(float, oL));
    k_norm<<<blocks(T, 64), 64, 0, g->stream>>>(T, P, PTR(StepOut, oSo), PTR(float, oL), PTR(double, oTs), PTR(float, oOm));
    k_apply<<<dim3(blocks(P, 256), unsigned(T)), 256, 0, g->stream>>>(T, P, PTR(StepOut, oSo), PTR(float, oL), PTR(double, oDt), in->wantRad,
                                                                      in->wantVis, PTR(int, oCol), Tup, seriesMode, PTR(double, oLr),
                                                                      PTR(double, oLv), seriesMode ? g->skyS.as<float>() : nullptr);
    if (in->wantRad) k_accum<<<blocks(P, kAccP), 256, 0, g->stream>>>(T, P, PTR(double, oLr), PTR(double, oSkyR));
    if (in->wantVis) k_accum<<<blocks(P, kAccP), 256, 0, g->stream>>>(T, P, PTR(double, oLv), PTR(double, oSkyV));
    k_ground_steps<<<blocks(T, 256), 256, 0, g->stream>>>(T, PTR(StepOut, oSo), PTR(double, oDt), PTR(int, oCol), seriesMode, PTR(double, oGr),
                                                          PTR(double, oGv), seriesMode ? g->skyG.as<float>() : nullptr, PTR(double, oSunR),
                                                          PTR(double, oSunV));
    k_ground_sum<<<1, 256, 0, g->stream>>>(T, PTR(double, oGr), PTR(double, oGv), PTR(double, oGnd));
    if (in->reflect && seriesMode && Tup > 0) {
      k_sdc_sky<<<dim3(blocks(Tup, 256), unsigned(P + 1)), 256, 0, g->stream>>>(P, Tup, g->skyS.as<float>(), g->skyG.as<float>(), g->skySdc.as<float>());
      k_sdc_sun<<<blocks(T, 256), 256, 0, g->stream>>>(T, Tup, PTR(StepOut, oSo), PTR(int, oCol), seriesMode == 2, g->skyS.as<float>(),
                                                        g->skySdc.as<float>());
    }
  }
  DLG_CHECK(hipGetLastError());
  const float ms = dlg_tend(g);
  g->stats.skyMs += ms;
  g->stats.lastSkyMs = ms;
  const auto down = [&](void* dst, size_t at, size_t bytes) { return hipMemcpyAsync(dst, b + at, bytes, hipMemcpyDeviceToHost, g->stream); };
  DLG_CHECK(down(out->alt, oAlt, 8 * T)); DLG_CHECK(down(out->azi, oAzi, 8 * T)); DLG_CHECK(down(out->up, oUp, 4 * T));
  DLG_CHECK(down(out->vec, oVec, 12 * T));
  if (out->skyRad) {
    if (in->wantRad && T > 0) DLG_CHECK(down(out->skyRad, oSkyR, 8 * P));
    else std::fill_n(out->skyRad, P, 0.0);
  }
  if (out->skyVis) {
    if (in->wantVis && T > 0) DLG_CHECK(down(out->skyVis, oSkyV, 8 * P));
    else std::fill_n(out->skyVis, P, 0.0);
  }
  double gnd[2] = {0, 0};
  DLG_CHECK(down(gnd, oGnd, 16));
  if (out->sunRad) DLG_CHECK(down(out->sunRad, oSunR, 8 * T));
  if (out->sunVis) DLG_CHECK(down(out->sunVis, oSunV, 8 * T));
  std::vector<StepOut> steps(in->reflect ? size_t(T) : 0);
  if (!steps.empty()) DLG_CHECK(down(steps.data(), oSo, sizeof(StepOut) * T));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  if (out->gndRad) *out->gndRad = gnd[0];
  if (out->gndVis) *out->gndVis = gnd[1];
  // reflected light: the sun binned (AddDirect) summed per patch in step order, as the sky's weights (run.cpp:414-441)
  if (out->sunBinRad) std::fill_n(out->sunBinRad, P, 0.0);
  if (out->sunBinVis) std::fill_n(out->sunBinVis, P, 0.0);
  for (int t = 0; t < int(steps.size()); ++t) {
    const StepOut& o = steps[size_t(t)];
    if (o.q.night) continue;
    for (int i = 0; i < 4; ++i) {
      if (out->sunBinRad && in->wantRad && o.binRad[i] != 0.f) out->sunBinRad[o.binP[i]] += sun_bright(o.binRad[i]) * in->dt[t] / 1000.0;
      if (out->sunBinVis && in->wantVis && o.binVis[i] != 0.f) out->sunBinVis[o.binP[i]] += 179.0 * sun_bright(o.binVis[i]) * in->dt[t];
    }
  }
#undef PTR
  return 0;
}

extern "C" int32_t dlg_set_sky_series(dlg_gpu* g, const float* S, const float* gS, int32_t P, int32_t Tup, char* err, int32_t errLen) {
  if (P <= 0 || Tup <= 0) { dlg_fail(err, errLen, "dlg_set_sky_series: empty series"); return 1; }
  DLG_CHECK(g->skyS.grow(size_t(P) * size_t(Tup) * 4));
  DLG_CHECK(g->skyG.grow(size_t(Tup) * 4));
  DLG_CHECK(hipMemcpyAsync(g->skyS.p, S, size_t(P) * size_t(Tup) * 4, hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipMemcpyAsync(g->skyG.p, gS, size_t(Tup) * 4, hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  g->skyP = P;
  g->skyT = Tup;
  g->skyReflect = 0;
  return 0;
}

extern "C" int32_t dlg_sky_series(dlg_gpu* g, float* S, float* gS, float* Sdc, char* err, int32_t errLen) {
  if (!g->skyT) { dlg_fail(err, errLen, "dlg_sky_series: the last dlg_sky kept no series"); return 1; }
  if (Sdc && !g->skyReflect) { dlg_fail(err, errLen, "dlg_sky_series: the last dlg_sky ran without reflect"); return 1; }
  DLG_CHECK(hipMemcpyAsync(S, g->skyS.p, size_t(g->skyP) * size_t(g->skyT) * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipMemcpyAsync(gS, g->skyG.p, size_t(g->skyT) * 4, hipMemcpyDeviceToHost, g->stream));
  if (Sdc) DLG_CHECK(hipMemcpyAsync(Sdc, g->skySdc.p, size_t(g->skyP + 1) * size_t(g->skyT) * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}
