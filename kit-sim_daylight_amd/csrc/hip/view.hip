// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Views on CDNA4: replaces __raygen__view (reference src/trace_dev.cu:177-232), gpu_view (gpu.cu:572-610), k_view /
// k_iso / gpu_reduce_view (gpu.cu:171-274, 671-697).
//   k_view_rays  persistent waves, each taking (sensor, group of 4 tiles of 8 x 8 cells) items in turn: frame, jittered
//                cell, virtual ground, closest hit through the 4-wide records of the tree (exact ties by the lower
//                triangle index), the sensor's own planes passed (at most 3; the 4th an obstruction there), the hit
//                triangle's label. Sensor-major cls / dist [nc x K].
//   k_view_sums  a wave per sensor, lane l sums directions l, l + 64, ... (class weights, weight x ln d by a
//                table-driven FP64 log, minima, own surface, reach, glazing), then a fixed shuffle tree (s = 32 .. 1):
//                the same bits for any chunking.
//   k_view_iso   a wave per sensor: the isovist's area (one row) or volume (tetrahedra between adjacent rows and
//                columns), each piece in float, summed in double in the same fixed order; a lane makes its own column's
//                cells and takes the next column's from its neighbour.
#include <cmath>
#include <type_traits>
#include <vector>

#include "dl_math.h"
#include "gpu_internal.h"
#include "closest.h"
#include "lanestack.h"
#include "wide.h"

namespace {

using namespace dlm;
using dlc::Closest;
using dlc::closest_hit;
constexpr int kWave = 64;
constexpr unsigned kClasses = 3 + 16;

struct ViewArgs {
  const float4* wide; const float4* leaves; int64_t nTris; float pad;
  const float3* P; const float3* N;
  const uint8_t* label;
  const float* rowSin;       // [2 nEl]
  uint32_t nAz, nEl, nc, i0, nS, back, world, cull, jitter, vground, cullBack;
  float az0, dAz, groundZ, cosN, sinN;
  uint8_t* cls; float* dist;
  uint32_t nTG, nItems;   // tile groups per sensor, sensors x tile groups
  uint32_t sBlock;        // sensors per block of the item order
  uint32_t* ovf;          // the stacks' global part, [gridDim.x][STACK - LDS part][64]
};

// Work items: a sensor and a group of kTiles cell tiles; lane l takes cell (row, col) of an 8 x 8 tile (1 x 64 on a flat
// field), so a wave's rays leave one origin in neighbouring directions and meet the same part of the tree. The grid is
// persistent (as many one-wave blocks as the CUs hold): block b takes items b, b + gridDim.x, ..., numbered in blocks of
// kSensorBlock neighbouring sensors, tile-group major inside a block (item = (sb nTG + tg) kSensorBlock + j), so the
// waves in flight trace about ten groups of directions from a few hundred neighbouring sensors and share in L2 the
// records those rays meet; each block keeps the top of its lanes' stacks in LDS and the rest in its own area of a
// global buffer (lanestack.h), so that a wave's LDS stays small and the waves per SIMD are what the registers allow.
// Measured over sensor blocks of 1 (sensor-major) to the whole chunk (tile-group major), the 1M-sensor / 10M-triangle
// views bench at its default field (3.66e9 rays) took 851 / 810 / 577 / 463 / 452 ms at 1 / 32 / 128 / 512 / 2048, its
// 360-degree field 763-811 ms, the interior (1.08e9 rays) 190.3-197.1 ms; 512 is within 2 % of the best on each.
constexpr uint32_t kTiles = 4;
constexpr uint32_t kSensorBlock = 512;
// the rays of sensor i's tile group tg
template <int STACK, bool OFF32>
__device__ __forceinline__ void view_item(const ViewArgs& a, uint32_t i, uint32_t tg, uint32_t lane, const dls::LaneStack<STACK>& stk) {
  const uint32_t gi = a.i0 + i, element = gi / a.nS, side = a.nS == 2 ? gi % 2 : a.back;
  const float3 n = a.N[i];
  const float3 o = offset_ray(a.P[i], n);
  const float inf = __int_as_float(0x7f800000);
  const bool none = !(dot3(n, n) > 0.f);
  float3 f, r, u;
  if (a.world) { f = make_float3(-a.sinN, a.cosN, 0.f); r = make_float3(a.cosN, a.sinN, 0.f); u = make_float3(0.f, 0.f, 1.f); }
  else {
    const float l = sqrtf(n.x * n.x + n.y * n.y);
    f = n;
    if (l > 1e-3f) r = make_float3(n.y / l, -n.x / l, 0.f);
    else {
      r = make_float3(a.cosN, a.sinN, 0.f);
      if (l > 0.f) {
        const float k = dot3(r, n);
        r = make_float3(r.x - k * n.x, r.y - k * n.y, r.z - k * n.z);
        const float s = 1.f / sqrtf(dot3(r, r));
        r = make_float3(s * r.x, s * r.y, s * r.z);
      }
    }
    u = make_float3(r.y * f.z - r.z * f.y, r.z * f.x - r.x * f.z, r.x * f.y - r.y * f.x);
  }
  const uint32_t K = a.nAz * a.nEl, th = a.nEl > 1 ? 8u : 1u, tw = kWave / th, tilesAz = (a.nAz + tw - 1) / tw;
  for (uint32_t tk = 0; tk < kTiles; ++tk) {
    const uint32_t tile = tg * kTiles + tk, row = (tile / tilesAz) * th + lane / tw, col = (tile % tilesAz) * tw + lane % tw;
    if (row >= a.nEl || col >= a.nAz) continue;
    const uint32_t d = row * a.nAz + col;
    const float3 lc = view_cell(d, col, a.rowSin[2 * row], a.rowSin[2 * row + 1], a.az0, a.dAz, a.jitter != 0, element, side);
    const float3 dir = make_float3((lc.x * r.x + lc.y * f.x) + lc.z * u.x, (lc.x * r.y + lc.y * f.y) + lc.z * u.y, (lc.x * r.z + lc.y * f.z) + lc.z * u.z);
    uint8_t cl = 2;
    float t = inf;
    if (none || (a.cull && dot3(n, dir) <= 0.f)) cl = 254;
    else if (a.vground && dir.z < 0.f && !(o.z > a.groundZ)) { cl = 1; t = 0.1f; }
    else {
      const bool plane = a.vground && dir.z < 0.f;
      const float tg = plane ? (o.z - a.groundZ) / -dir.z : inf;
      const float tmax = plane ? (o.z - a.groundZ + 0.01f) / -dir.z : 3.40282347e38f;
      Ray ray;
      ray_setup(ray, o, dir, a.pad);
      float tmin = 1e-5f;
      uint32_t own = 0, panes = 0;
      for (;;) {
        Closest c{};
        if (a.nTris > 0) c = closest_hit<OFF32>(a.wide, a.leaves, ray, tmin, tmax, a.cullBack != 0, stk, int(lane));
        if (!c.hit) { cl = dir.z >= 0.f ? 0 : 1; t = tg; break; }
        t = c.t;
        const float4 A = a.leaves[3 * c.leaf], B = a.leaves[3 * c.leaf + 1], E = a.leaves[3 * c.leaf + 2];
        if (own_hit(dir, n, t, c.front, make_float3(A.x, A.y, A.z), make_float3(A.w, B.x, B.y), make_float3(B.z, B.w, E.x))) {
          if (++own == 4) break;
        } else {
          const uint8_t lb = a.label ? a.label[c.prim] : 2;
          if (lb != 255) { cl = lb; break; }
          if (++panes > 16) break;
        }
        tmin = __uint_as_float(__float_as_uint(t) + 1);
      }
      if (panes) cl |= 0x80;
    }
    const size_t at = size_t(i) * K + d;   // sensor-major
    a.cls[at] = cl;
    a.dist[at] = t;
  }
}

// OFF32: fewer than 2^25 records (wide.h visit); 5 waves per SIMD under a register bound (96 registers, a few spilled)
// trace 7 % faster than the 4 that 111 registers allow
template <int STACK, bool OFF32>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(5))) void k_view_rays(ViewArgs a) {
  __shared__ uint32_t lds[dls::lds_stack<STACK>() * kWave];
  const dls::LaneStack<STACK> stk{(dls::lds_u32*)lds, a.ovf + size_t(blockIdx.x) * dls::ovf_stack<STACK>() * kWave};
  const uint32_t lane = threadIdx.x;
  const uint32_t per = a.sBlock * a.nTG;
  for (uint32_t item = blockIdx.x; item < a.nItems; item += gridDim.x) {
    const uint32_t sb = item / per, r = item - sb * per, s0 = sb * a.sBlock, sn = a.nc - s0 < a.sBlock ? a.nc - s0 : a.sBlock;
    view_item<STACK, OFF32>(a, s0 + r % sn, r / sn, lane, stk);
  }
}

__device__ __forceinline__ double tree_sum(double v) {
  for (int s = kWave / 2; s; s >>= 1) v += __shfl_down(v, s, kWave);
  return v;
}
__device__ __forceinline__ float tree_min(float v) {
  for (int s = kWave / 2; s; s >>= 1) v = fminf(v, __shfl_down(v, s, kWave));
  return v;
}

struct SumArgs {
  const uint8_t* cls; const float* dist; const double* wRow;
  uint32_t nc, K, nAz, nClass, nReach, glazing, NA;
  float farF; double lnFar;
  float reach[8];
  double* out;
};

// ln x in double for a positive normal float x, within about 2 ulp of the correctly rounded value: x = 2^e m with
// m in [1, 2) taken to the nearest c = 1 + j / 128 (j = 0 .. 128), ln x = e ln 2 + ln c + log1p(r), r = m / c - 1 with
// |r| <= 2^-8 and log1p(r) by its series through r^7 (the first term left out is below 2^-66). ln 2 in Cody and
// Waite's two parts, so that e ln2_hi is exact; tab[j] = (ln c, 1 / c), made once per block (ln_table).
constexpr int kLnTab = 129;
__device__ __forceinline__ double ln_float(float x, const double2* tab) {
  const uint32_t u = __float_as_uint(x), mb = u & 0x7fffffu;
  const int e = int(u >> 23) - 127;
  const double2 t = tab[(mb + (1u << 15)) >> 16];
  const double m = 1.0 + double(mb) * 0x1p-23;   // exact
  const double r = fma(m, t.y, -1.0);
  double q = fma(r, 1.0 / 7.0, -1.0 / 6.0);
  q = fma(r, q, 0.2);
  q = fma(r, q, -0.25);
  q = fma(r, q, 1.0 / 3.0);
  q = fma(r, q, -0.5);
  const double p = fma(r * r, q, r);
  const double ln2Hi = 6.93147180369123816490e-01, ln2Lo = 1.90821492927058770002e-10;   // ln2Hi: 32 bits, e ln2Hi exact
  return (double(e) * ln2Hi + t.x) + fma(double(e), ln2Lo, p);
}
__device__ __forceinline__ void ln_table(double2* tab) {
  for (int j = threadIdx.x; j < kLnTab; j += blockDim.x) {
    const double c = 1.0 + double(j) / 128.0;
    tab[j] = make_double2(log(c), 1.0 / c);
  }
  __syncthreads();
}

// NC >= the run's classes: the per-class sums live in registers, so the common runs (sky, ground, obstruction and a
// target group or two) get a small NC and the occupancy its few registers allow. NR >= the reach distances (0, 3 or 8):
// a run without reach (or with EN 17037's three) adds no work for the slots it does not use.
template <uint32_t NC, uint32_t NR>
__global__ __launch_bounds__(256) void k_view_sums(SumArgs a) {
  __shared__ double2 lnTab[kLnTab];
  ln_table(lnTab);
  const uint32_t lane = threadIdx.x & (kWave - 1), i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (i >= a.nc) return;   // whole waves
  double w[NC], wl[NC], rc[NR > 0 ? NR : 1], own = 0, gz = 0;
  float mn[NC];
  for (uint32_t c = 0; c < NC; ++c) { w[c] = 0; wl[c] = 0; mn[c] = __int_as_float(0x7f800000); }
  for (uint32_t r = 0; r < NR; ++r) rc[r] = 0;
  const auto take = [&](uint32_t c, float t, double wd) {
    if (c == 254) { own += wd; return; }
    if (c & 0x80) { gz += wd; c &= 0x7f; }
    const float dd = fminf(fmaxf(t, 0.1f), a.farF);
    const double ld = dd == a.farF ? a.lnFar : dd == 0.1f ? -2.302585092994045684 : ln_float(dd, lnTab);
    // registers indexed by class: a small switch keeps them out of scratch
#pragma unroll
    for (uint32_t q = 0; q < NC; ++q)
      if (q == c) { w[q] += wd; wl[q] += wd * ld; mn[q] = fminf(mn[q], dd); }
#pragma unroll
    for (uint32_t r = 0; r < NR; ++r) if (r < a.nReach && t >= a.reach[r]) rc[r] += wd;
  };
  // lane l takes directions l, l + 64, ... in ascending order (the oracle's order), four at a time so that their loads
  // are in flight together; the row of each (its weight) is followed without a division per ray
  const uint8_t* cls = a.cls + size_t(i) * a.K;
  const float* dist = a.dist + size_t(i) * a.K;
  uint32_t row = a.nAz ? lane / a.nAz : 0, col = a.nAz ? lane % a.nAz : 0;
  const auto next = [&]() {
    col += kWave;
    while (col >= a.nAz) { col -= a.nAz; ++row; }
  };
  uint32_t d = lane;
  for (; d + 3 * kWave < a.K; d += 4 * kWave) {
    const uint32_t c0 = cls[d], c1 = cls[d + kWave], c2 = cls[d + 2 * kWave], c3 = cls[d + 3 * kWave];
    const float t0 = dist[d], t1 = dist[d + kWave], t2 = dist[d + 2 * kWave], t3 = dist[d + 3 * kWave];
    const double w0 = a.wRow[row];
    next();
    const double w1 = a.wRow[row];
    next();
    const double w2 = a.wRow[row];
    next();
    const double w3 = a.wRow[row];
    next();
    take(c0, t0, w0);
    take(c1, t1, w1);
    take(c2, t2, w2);
    take(c3, t3, w3);
  }
  for (; d < a.K; d += kWave) {
    take(cls[d], dist[d], a.wRow[row]);
    next();
  }
  double* o = a.out + size_t(i) * a.NA;
#pragma unroll
  for (uint32_t c = 0; c < NC; ++c) {
    if (c >= a.nClass) break;
    const double s0 = tree_sum(w[c]), s1 = tree_sum(wl[c]);
    const float m = tree_min(mn[c]);
    if (!lane) { o[3 * c] = s0; o[3 * c + 1] = s1; o[3 * c + 2] = m; }
  }
  const double ow = tree_sum(own);
  if (!lane) o[3 * a.nClass] = ow;
#pragma unroll
  for (uint32_t r = 0; r < NR; ++r) {
    if (r >= a.nReach) break;
    const double s = tree_sum(rc[r]);
    if (!lane) o[3 * a.nClass + 1 + r] = s;
  }
  if (a.glazing) {
    const double s = tree_sum(gz);
    if (!lane) o[3 * a.nClass + 1 + a.nReach] = s;
  }
}

struct IsoArgs {
  const uint8_t* cls; const float* dist; const float* rowSin;
  uint32_t nc, nAz, nEl, i0, nS, back, world, jitter, dim, NA;
  float az0, dAz, farF;
  double* out;
};

__device__ __forceinline__ float det3(float3 a, float3 b, float3 c) {
  const float3 u = make_float3(b.x - a.x, b.y - a.y, b.z - a.z), v = make_float3(c.x - a.x, c.y - a.y, c.z - a.z);
  return a.x * (u.y * v.z - u.z * v.y) + a.y * (u.z * v.x - u.x * v.z) + a.z * (u.x * v.y - u.y * v.x);
}

// A wave per sensor; lane l takes columns l, l + 64, ... and walks the rows: the pieces between its column and the
// next one (c1) in each pair of adjacent rows. The next column's cell and length are lane l + 1's own (a shuffle); the
// lanes whose next column is not lane l + 1's (lane 63, and the wrap to column 0) take the pass's one extra column cx,
// whose cells and lengths the lanes make beforehand, one row each, 64 rows at a time. Every value is the one the
// direct evaluation gives (the same function of the same cell), so the pieces and their order are unchanged.
__global__ __launch_bounds__(256) void k_view_iso(IsoArgs a) {
  const uint32_t lane = threadIdx.x & (kWave - 1), i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (i >= a.nc) return;   // whole waves
  const uint32_t gi = a.i0 + i, element = gi / a.nS, side = a.nS == 2 ? gi % 2 : a.back;
  const size_t K = size_t(a.nAz) * a.nEl;
  const auto len = [&](uint32_t k) { const size_t at = size_t(i) * K + k; return a.cls[at] == 254 ? 0.f : fminf(a.dist[at], a.farF); };
  const auto cell = [&](uint32_t r, uint32_t col) {
    return view_cell(r * a.nAz + col, col, a.rowSin[2 * r], a.rowSin[2 * r + 1], a.az0, a.dAz, a.jitter != 0, element, side);
  };
  const auto shfl3 = [](float3 v, int src) { return make_float3(__shfl(v.x, src), __shfl(v.y, src), __shfl(v.z, src)); };
  double sum = 0;
  for (uint32_t base = 0; base < a.nAz; base += kWave) {
    const uint32_t col = base + lane < a.nAz ? base + lane : a.nAz - 1, c1 = col + 1 == a.nAz ? 0 : col + 1;
    const bool on = base + lane < a.nAz && (a.world || col + 1 < a.nAz);
    // the pass's extra column: lane 63's next (base + 64, or 0 past the last column), which is also the wrap's
    const uint32_t last = base + kWave - 1 < a.nAz ? base + kWave - 1 : a.nAz - 1, cx = last + 1 == a.nAz ? 0 : last + 1;
    const bool extra = lane == kWave - 1 || c1 == 0;
    float3 exP = make_float3(0.f, 0.f, 0.f);
    float exL = 0.f;
    const auto extras = [&](uint32_t r0) {   // rows r0 + lane
      const uint32_t r = r0 + lane < a.nEl ? r0 + lane : a.nEl - 1;
      exP = cell(r, cx);
      exL = len(r * a.nAz + cx);
    };
    const auto nextOf = [&](uint32_t r, float3 p, float l, float3& pn, float& ln) {   // row r's cell and length at c1
      pn = shfl3(p, int(lane) + 1 < kWave ? int(lane) + 1 : int(lane));
      ln = __shfl(l, int(lane) + 1 < kWave ? int(lane) + 1 : int(lane));
      const int src = int(r % kWave);
      const float3 xp = shfl3(exP, src);
      const float xl = __shfl(exL, src);
      if (extra) { pn = xp; ln = xl; }
    };
    extras(0);
    float3 pa = cell(0, col), pb;
    float l0x = len(col), l0y;
    nextOf(0, pa, l0x, pb, l0y);
    if (a.dim == 2) {
      if (on) sum += (double)(0.5f * l0x * l0y * (pa.y * pb.x - pa.x * pb.y));
      continue;
    }
    for (uint32_t r = 1; r < a.nEl; ++r) {
      if (r % kWave == 0) extras(r);
      const float3 pc = cell(r, col);
      const float l1x = len(r * a.nAz + col);
      float3 pe;
      float l1y;
      nextOf(r, pc, l1x, pe, l1y);
      if (on) sum += (double)(l0x * l0y * l1x * det3(pa, pb, pc) + l0y * l1y * l1x * det3(pb, pe, pc));
      pa = pc; pb = pe; l0x = l1x; l0y = l1y;
    }
  }
  const double s = tree_sum(sum);
  if (!lane) a.out[size_t(i) * a.NA + a.NA - 1] = a.dim == 2 ? s : s / 6;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

extern "C" int32_t dlg_set_labels(dlg_gpu* g, const uint8_t* label, int64_t n, char* err, int32_t errLen) {
  g->labels.release();
  if (n <= 0) return 0;
  DLG_CHECK(g->labels.grow(size_t(n)));
  DLG_CHECK(hipMemcpyAsync(g->labels.p, label, size_t(n), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int32_t dlg_view(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, const dlg_view_in* v, char* err, int32_t errLen) {
  if (i0 + nc > g->nSens) { dlg_fail(err, errLen, "dlg_view: sensor range exceeds the scene"); return 1; }
  const size_t K = size_t(v->nAz) * v->nEl;
  const size_t oDist = align256(size_t(nc) * K), oSin = oDist + align256(size_t(nc) * K * 4);
  DLG_CHECK(g->view.grow(oSin + align256(size_t(v->nEl) * 8)));
  char* b = g->view.as<char>();
  DLG_CHECK(hipMemcpyAsync(b + oSin, v->rowSin, size_t(v->nEl) * 8, hipMemcpyHostToDevice, g->stream));
  g->viewN = nc; g->viewK = int64_t(K); g->viewI0 = i0; g->viewIn = *v; g->viewIn.rowSin = nullptr;
  g->i0 = i0;
  if (nc == 0 || K == 0) { DLG_CHECK(hipStreamSynchronize(g->stream)); return 0; }
  const dlg_scene& s = g->scene;
  ViewArgs a;
  a.wide = s.wide.as<float4>(); a.leaves = s.leaves.as<float4>(); a.nTris = s.nTris; a.pad = s.pad;
  a.P = g->P.as<float3>() + i0; a.N = g->N.as<float3>() + i0;
  a.label = g->labels.as<uint8_t>();
  a.rowSin = reinterpret_cast<const float*>(b + oSin);
  a.nAz = v->nAz; a.nEl = v->nEl; a.nc = uint32_t(nc); a.i0 = uint32_t(i0); a.nS = v->nS; a.back = v->back; a.world = v->world; a.cull = v->cull;
  a.jitter = v->jitter; a.vground = v->vground; a.cullBack = (rayFlags & 16u) ? 1u : 0u;
  a.az0 = v->az0; a.dAz = v->dAz; a.groundZ = v->groundZ;
  const double rad = v->north * 3.14159265358979323846 / 180;
  a.cosN = float(std::cos(rad)); a.sinN = float(std::sin(rad));
  a.cls = reinterpret_cast<uint8_t*>(b); a.dist = reinterpret_cast<float*>(b + oDist);
  const uint32_t th = v->nEl > 1 ? 8u : 1u, tw = uint32_t(kWave) / th;
  const size_t tiles = size_t((v->nEl + th - 1) / th) * ((v->nAz + tw - 1) / tw);
  const size_t nTG = (tiles + kTiles - 1) / kTiles, nItems = size_t(nc) * nTG;
  if (nItems > size_t(0xffffffffu)) { dlg_fail(err, errLen, "dlg_view: too many sensors x cell tiles in one chunk"); return 1; }
  a.nTG = uint32_t(nTG);
  a.nItems = uint32_t(nItems);
  a.sBlock = uint32_t(nc) < kSensorBlock ? uint32_t(nc) : kSensorBlock;
  // the smallest stack that covers the tree, with the entries above the top that the unconditional pushes write; the
  // persistent grid (whole rounds of the 8 XCDs) and each block's global stack area
  const int need = dlw::wide_stack(s.height) + 3;
  const auto launch = [&](auto S) -> int32_t {
    constexpr int ST = decltype(S)::value;
    const bool off32 = s.nNodes < (int64_t(1) << 25);
    const void* fn = off32 ? reinterpret_cast<const void*>(&k_view_rays<ST, true>) : reinterpret_cast<const void*>(&k_view_rays<ST, false>);
    uint64_t grid = uint64_t(dlg_blocks_per_cu(fn)) * uint64_t(g->cus > 0 ? g->cus : 1);
    grid = grid < 8 ? 8 : grid / 8 * 8;
    if (grid > nItems) grid = nItems;
    const size_t ovf = size_t(dls::ovf_stack<ST>()) * kWave * 4;
    if (ovf) DLG_CHECK(g->viewOvf.grow(ovf * grid));
    a.ovf = g->viewOvf.as<uint32_t>();
    dlg_tbegin(g);
    if (off32) k_view_rays<ST, true><<<unsigned(grid), kWave, 0, g->stream>>>(a);
    else k_view_rays<ST, false><<<unsigned(grid), kWave, 0, g->stream>>>(a);
    return 0;
  };
  int32_t rc = 0;
  if (need <= 24) rc = launch(std::integral_constant<int, 24>{});
  else if (need <= 32) rc = launch(std::integral_constant<int, 32>{});
  else if (need <= 48) rc = launch(std::integral_constant<int, 48>{});
  else if (need <= 64) rc = launch(std::integral_constant<int, 64>{});
  else if (need <= 96) rc = launch(std::integral_constant<int, 96>{});
  else if (need <= 160) rc = launch(std::integral_constant<int, 160>{});
  else { dlg_fail(err, errLen, "dlg_view: the scene's tree is too deep for the traversal stack"); return 1; }
  if (rc) return rc;
  DLG_CHECK(hipGetLastError());
  const float ms = dlg_tend(g);
  g->stats.viewMs += ms;
  g->stats.traceMs += ms;
  g->stats.lastViewMs = ms;
  g->stats.rays += int64_t(nc) * int64_t(K);
  g->stats.launches += 1;
  return 0;
}

extern "C" int32_t dlg_reduce_view(dlg_gpu* g, int64_t nc, const double* wRow, uint32_t nClass, double farDist, const float* reach, uint32_t nReach,
                                   int32_t glazing, uint32_t isoDim, double* out, char* err, int32_t errLen) {
  if (nc == 0) return 0;
  if (nc != g->viewN) { dlg_fail(err, errLen, "dlg_reduce_view: no dlg_view for this chunk"); return 1; }
  if (nClass > kClasses || nReach > 8) { dlg_fail(err, errLen, "dlg_reduce_view: at most 16 groups and 8 reach distances"); return 1; }
  const dlg_view_in& v = g->viewIn;
  const uint32_t NA = 3 * nClass + 1 + nReach + (glazing ? 1 : 0) + (isoDim ? 1 : 0);
  DLG_CHECK(g->scratch.grow(size_t(v.nEl) * 8));
  DLG_CHECK(g->out.grow(size_t(nc) * NA * 8));
  DLG_CHECK(hipMemcpyAsync(g->scratch.p, wRow, size_t(v.nEl) * 8, hipMemcpyHostToDevice, g->stream));
  const size_t K = size_t(g->viewK), oDist = align256(size_t(nc) * K), oSin = oDist + align256(size_t(nc) * K * 4);
  char* b = g->view.as<char>();
  SumArgs s;
  s.cls = reinterpret_cast<const uint8_t*>(b); s.dist = reinterpret_cast<const float*>(b + oDist); s.wRow = g->scratch.as<double>();
  s.nc = uint32_t(nc); s.K = uint32_t(K); s.nAz = v.nAz; s.nClass = nClass; s.nReach = nReach; s.glazing = glazing ? 1u : 0u; s.NA = NA;
  s.farF = float(farDist); s.lnFar = std::log(farDist);
  for (uint32_t r = 0; r < 8; ++r) s.reach[r] = r < nReach ? reach[r] : 0.f;
  s.out = g->out.as<double>();
  dlg_tbegin(g);
  const unsigned grid = unsigned((nc + 3) / 4);
  const auto sums = [&](auto nc) {
    constexpr uint32_t NC = decltype(nc)::value;
    if (nReach == 0) k_view_sums<NC, 0><<<grid, 256, 0, g->stream>>>(s);
    else if (nReach <= 3) k_view_sums<NC, 3><<<grid, 256, 0, g->stream>>>(s);
    else k_view_sums<NC, 8><<<grid, 256, 0, g->stream>>>(s);
  };
  if (nClass <= 4) sums(std::integral_constant<uint32_t, 4>{});
  else if (nClass <= 8) sums(std::integral_constant<uint32_t, 8>{});
  else sums(std::integral_constant<uint32_t, kClasses>{});
  if (isoDim) {
    IsoArgs a;
    a.cls = s.cls; a.dist = s.dist; a.rowSin = reinterpret_cast<const float*>(b + oSin);
    a.nc = uint32_t(nc); a.nAz = v.nAz; a.nEl = v.nEl; a.i0 = uint32_t(g->viewI0); a.nS = v.nS; a.back = v.back; a.world = v.world;
    a.jitter = v.jitter; a.dim = isoDim; a.NA = NA; a.az0 = v.az0; a.dAz = v.dAz; a.farF = float(farDist); a.out = s.out;
    k_view_iso<<<unsigned((nc + 3) / 4), 256, 0, g->stream>>>(a);
  }
  DLG_CHECK(hipGetLastError());
  g->stats.lastViewReduceMs = dlg_tend(g);
  g->stats.reduceMs += g->stats.lastViewReduceMs;
  DLG_CHECK(hipMemcpyAsync(out, g->out.p, size_t(nc) * NA * 8, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

extern "C" int32_t dlg_download_view(dlg_gpu* g, int64_t nc, uint8_t* cls, float* dist, char* err, int32_t errLen) {
  if (nc != g->viewN) { dlg_fail(err, errLen, "dlg_download_view: no dlg_view for this chunk"); return 1; }
  const size_t K = size_t(g->viewK), n = size_t(nc) * K;
  if (!n) return 0;
  const char* b = g->view.as<char>();   // sensor-major on the device, as the download is
  DLG_CHECK(hipMemcpyAsync(cls, b, n, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipMemcpyAsync(dist, b + align256(n), n * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}
