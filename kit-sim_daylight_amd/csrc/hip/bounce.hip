// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Reflected light on CDNA4: replaces __raygen__bounce (reference src/trace_dev.cu:260-311), gpu_bounce (gpu.cu:533-568),
// gpu_reduce_dc (gpu.cu:627-637) and gpu_download_dc (gpu.cu:700-705).
//   k_bounce     a wave per sensor of the chunk walks its M samples, 64 at a time, in order. Sample s leaves the sensor
//                cosine-weighted (Hammersley point s, rotated by a hash of the global sensor index) and goes from
//                closest hit to closest hit (closest.h): the sensor's own plane (first segment) and a thin solid's exit
//                face pass it unchanged; elsewhere it is transmitted with probability T / (T + R) (porous tau, Radiance
//                glass T(cos)), its weight times T + R, else reflected: mirrored on glass, diffusely by rho elsewhere
//                (these count toward the bounces; Russian roulette from the second on, survival = the weight). A path
//                that reflected and escapes adds weight x pi / M to its bin (dl_math.h sky_bin) of the sensor's row.
//                Random numbers pcg3d(global sensor, sample, event): rows do not depend on chunks or launches; a row is
//                summed in sample order (no atomics).
//   k_reduce_dc  a lane per sensor: out[i] = sum over bins b (ascending) of double(dc[b nc + i]) x double(W[b]).
// Rows dc[b nc + i] float, bin 0 the ground, 1 + p the Reinhart patches of the sky density (gendaymtx order).
#include <algorithm>
#include <string>

#include "closest.h"
#include "dl_math.h"
#include "gpu_internal.h"

namespace {

using namespace dlm;
using dlc::Closest;
using dlc::closest_hit;
constexpr int kWave = 64;

struct BounceArgs {
  const float4* wide;
  const float4* leaves;
  int64_t nTris;
  float pad;
  const float3* P;
  const float3* N;
  const float4* slots;   // rho, tau, tn, flags (bits)
  uint32_t slotLast;
  float* dc;
  uint32_t nc, i0, M, bounces, mf, nBins, cull;
  float alpha, cosN, sinN;
  const uint32_t* order;   // the sensor at each position of the chunk (the trace's orientation + Morton order), or null
};

DL_HD float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
DL_HD float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
DL_HD float3 mul3(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }

// One path: sample s of global sensor gi (normal n, origin o0), the rotation (r0, r1) of its sample pattern. True when
// it reflected and escaped, with its bin and weight.
template <int STACK, bool OFF32>
__device__ __forceinline__ bool trace_path(const BounceArgs& a, float3 n, float3 o0, uint32_t gi, uint32_t s, float r0, float r1,
                                           const dls::LaneStack<STACK, STACK>& stk, int lane, uint32_t& bin, float& val) {
  const bool cull = a.cull != 0;
  const float u = float(s) / float(a.M) + r0, v = u01(brev32(s)) + r1;
  float3 o = o0, d = cosine_dir(n, u - floorf(u), v - floorf(v));
  float w = 1.f;
  uint32_t diffuse = 0, k = 0;   // diffuse reflections, random draws (events)
  bool reflected = false;
  for (uint32_t seg = 0; seg < a.bounces + 16; ++seg) {
    Ray ray;
    ray_setup(ray, o, d, a.pad);
    Closest c{};
    if (a.nTris > 0) c = closest_hit<OFF32>(a.wide, a.leaves, ray, 1e-5f, 3.40282347e38f, cull, stk, lane);
    if (!c.hit) {
      if (!reflected) return false;
      bin = sky_bin(d, a.mf, a.alpha, a.nBins, a.cosN, a.sinN);
      val = w * (3.14159265f / float(a.M));
      return true;
    }
    // the hit triangle (trace_dev.cu:43-56): its point by barycentrics, unit normal on the side the ray came from,
    // extent, and the lift new origins leave it by
    const float4 A = a.leaves[3 * c.leaf], B = a.leaves[3 * c.leaf + 1], E = a.leaves[3 * c.leaf + 2];
    const float3 v0 = make_float3(A.x, A.y, A.z), v1 = make_float3(A.w, B.x, B.y), v2 = make_float3(B.z, B.w, E.x);
    const float3 e1 = sub3(v1, v0), e2 = sub3(v2, v0);
    const float3 ng = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    const float l2 = dot3(ng, ng);
    if (!(l2 > 0.f)) return false;
    const float sc = c.front ? 1.f / sqrtf(l2) : -(1.f / sqrtf(l2));
    const float3 hn = mul3(sc, ng);
    float bc[2] = {0.f, 0.f}, tt;
    bool fr;
    tri_hit(ray, v0, v1, v2, c.t, c.t, cull, tt, fr, bc);   // the closest hit again, for its barycentrics
    const float3 p = add3(add3(v0, mul3(bc[0], e1)), mul3(bc[1], e2));
    const float ext = fmaxf(fmaxf(fmaxf(fabsf(e1.x), fabsf(e1.y)), fmaxf(fabsf(e1.z), fabsf(e2.x))), fmaxf(fabsf(e2.y), fabsf(e2.z)));
    const float3 lift = mul3(ext * 9.5367432e-7f, hn);
    const float3 through = sub3(offset_ray(p, mul3(-1.f, hn)), lift);   // the origin past the triangle
    if (seg == 0 && own_plane(d, n, c.t, c.front, v0, v1, v2)) { o = through; continue; }
    const uint32_t si = __float_as_uint(E.z);
    const float4 sl = a.slots[si < a.slotLast ? si : a.slotLast];
    const uint32_t fl = __float_as_uint(sl.w);
    if ((fl & (kSlotThin | kSlotTwin)) == kSlotThin && c.front == ((fl & kSlotInward) != 0)) { o = through; continue; }   // exit face
    const bool mirror = (fl & kSlotGlass) != 0;
    float T = sl.y, R = sl.x;   // opaque: tau 0 (every slot with --transmission off)
    if (mirror) {
      glass_TR(double(fabsf(dot3(d, hn))), double(sl.z), T, R);
      if (!(sl.y > 0.f)) T = 0.f;   // glass with --transmission off mirrors R and absorbs T
    } else if (diffuse == a.bounces) R = 0.f;   // no diffuse reflection left: transmitted or absorbed
    if (!(T + R > 0.f)) return false;
    const uint3 rr = pcg3d(gi, s, k++);
    w *= T + R;
    if (u01(rr.x) * (T + R) < T) { o = through; continue; }   // transmitted
    reflected = true;
    if (mirror) d = sub3(d, mul3(2.f * dot3(d, hn), hn));
    else {
      d = cosine_dir(hn, u01(rr.y), u01(rr.z));
      if (++diffuse >= 2) {   // Russian roulette, survival = w (<= 1)
        if (u01(pcg3d(gi, s, k++).x) >= w) return false;
        w = 1.f;
      }
    }
    o = add3(offset_ray(p, hn), lift);
  }
  return false;
}

// A wave per position of [c0, c0 + gridDim.x) of the chunk's sensor order (the trace's orientation + Morton order when
// the chunk has one: the waves in flight trace a compact patch of sensors; seeds and rows belong to the sensor, so the
// order changes no result): lane l traces samples l, l + 64, ... (rounds of 64 from one origin, so the top of the tree
// is shared); each round's escaping paths add to the sensor's row in LDS in lane order, i.e. in sample
// order, which keeps every bin's float sum that of one thread walking the samples in turn; the row is written at the end.
// OFF32: fewer than 2^25 records (wide.h visit)
template <int STACK, bool OFF32>
// 3 waves per SIMD (168 VGPRs, a few spilled) trace 17 % faster than the 202 VGPRs of 2 waves
__global__ __launch_bounds__(kWave, 3) void k_bounce(BounceArgs a, uint32_t c0) {
  __shared__ uint32_t stkLds[STACK * kWave];
  const dls::LaneStack<STACK, STACK> stk{(dls::lds_u32*)stkLds, nullptr};   // the whole stack in LDS
  extern __shared__ float sRow[];   // [nBins]
  const uint32_t lane = threadIdx.x, p = c0 + blockIdx.x;
  if (p >= a.nc) return;   // whole blocks
  const uint32_t i = a.order ? a.order[p] : p;
  for (uint32_t b = lane; b < a.nBins; b += kWave) sRow[b] = 0.f;
  __syncthreads();
  const uint32_t gi = a.i0 + i;
  const float3 n = a.N[i], o0 = offset_ray(a.P[i], n);
  const uint3 rot = pcg3d(gi, ~0u, ~0u);
  const float r0 = u01(rot.x), r1 = u01(rot.y);
  for (uint32_t s0 = 0; s0 < a.M; s0 += kWave) {
    const uint32_t s = s0 + lane;
    uint32_t bin = 0;
    float val = 0.f;
    const bool has = s < a.M && trace_path<STACK, OFF32>(a, n, o0, gi, s, r0, r1, stk, int(lane), bin, val);
    // in lane order: one lane at a time adds its value (a bin's sum is the sequential one)
    for (uint64_t m = __ballot(has); m; m &= m - 1) {
      const uint32_t l = uint32_t(__ffsll((unsigned long long)m) - 1);
      if (lane == l) sRow[bin] += val;
      __syncthreads();   // one wave per block: orders the LDS updates
    }
  }
  __syncthreads();
  float* row = a.dc + i;
  for (uint32_t b = lane; b < a.nBins; b += kWave) row[size_t(b) * a.nc] = sRow[b];
}

template <int STACK>
void launch_bounce(dlg_gpu* g, const BounceArgs& a, uint32_t c0, uint32_t n) {
  if (g->scene.nNodes < (int64_t(1) << 25)) k_bounce<STACK, true><<<n, kWave, a.nBins * 4, g->stream>>>(a, c0);
  else k_bounce<STACK, false><<<n, kWave, a.nBins * 4, g->stream>>>(a, c0);
}

__global__ __launch_bounds__(256) void k_reduce_dc(const float* __restrict__ dc, uint32_t nc, uint32_t nBins, const float* __restrict__ W,
                                                   double* __restrict__ out) {
  extern __shared__ float sW[];
  for (uint32_t b = threadIdx.x; b < nBins; b += blockDim.x) sW[b] = W[b];
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  double acc = 0.0;
  for (uint32_t b = 0; b < nBins; ++b) acc += (double)(1.f * dc[size_t(b) * nc + i]) * (double)sW[b];
  out[i] = acc;
}

}  // namespace

extern "C" {

uint32_t dlg_bounce_flags(int32_t cullBackfaces) { return 1u | (cullBackfaces ? 16u : 0u); }

int32_t dlg_bounce(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, const dlg_bounce_in* b, char* err, int32_t errLen) {
  if (i0 + nc > g->nSens) { dlg_fail(err, errLen, "dlg_bounce: sensor range exceeds the scene"); return 1; }
  if (b->mf != 1 && b->mf != 2 && b->mf != 4) { dlg_fail(err, errLen, "dlg_bounce: sky density 1, 2 or 4"); return 1; }
  const dlg_scene& s = g->scene;
  if (s.nTris > 0 && (g->nSlots == 0 || int64_t(s.slotMax) >= g->nSlots)) {
    dlg_fail(err, errLen, "dlg_bounce: the scene names slot " + std::to_string(s.slotMax) + " and " + std::to_string(g->nSlots) +
                              " slots are set (dlg_set_slots)");
    return 1;
  }
  const uint32_t nBins = uint32_t(144 * b->mf * b->mf + 2);   // 1 + the patches of the sky density
  g->dcBins = nBins;
  g->dcN = nc;
  g->dcI0 = i0;
  DLG_CHECK(g->dc.grow(size_t(nBins) * size_t(nc) * 4));
  if (nc == 0) return 0;
  DLG_CHECK(hipMemsetAsync(g->dc.p, 0, size_t(nBins) * size_t(nc) * 4, g->stream));
  const int need = dlw::wide_stack(s.height) + 3;
  if (need > 160) { dlg_fail(err, errLen, "dlg_bounce: the scene's tree is too deep for the traversal stack"); return 1; }
  const double rad = b->north * 3.14159265358979323846 / 180;
  BounceArgs a{};
  a.wide = s.wide.as<float4>(); a.leaves = s.leaves.as<float4>(); a.nTris = s.nTris; a.pad = s.pad;
  a.P = g->P.as<float3>() + i0; a.N = g->N.as<float3>() + i0;
  a.slots = g->slots.as<float4>(); a.slotLast = uint32_t(g->nSlots > 0 ? g->nSlots - 1 : 0);
  a.dc = g->dc.as<float>(); a.nc = uint32_t(nc); a.i0 = uint32_t(i0); a.M = b->M; a.bounces = b->bounces; a.mf = uint32_t(b->mf);
  a.nBins = nBins; a.cull = (rayFlags & 16u) ? 1u : 0u;
  a.alpha = float((3.14159265358979323846 / 2.) / (7 * b->mf + .5));   // the patch rows' height (the sky's Patches::alpha)
  a.cosN = float(std::cos(rad)); a.sinN = float(std::sin(rad));
  a.order = g->orderGen == g->sensGen && g->orderI0 == i0 && g->orderN == nc ? g->tileOrder.as<uint32_t>() : nullptr;
  // sensors in launches of about 2^24 paths (a launch's runtime bounded; each sensor's row is one block's)
  const uint32_t per = std::max<uint32_t>(1, (1u << 24) / std::max<uint32_t>(1, b->M));
  for (uint32_t c0 = 0; c0 < uint32_t(nc); c0 += per) {
    const uint32_t m = std::min<uint32_t>(per, uint32_t(nc) - c0);
    dlg_tbegin(g);
    if (need <= 24) launch_bounce<24>(g, a, c0, m);
    else if (need <= 32) launch_bounce<32>(g, a, c0, m);
    else if (need <= 48) launch_bounce<48>(g, a, c0, m);
    else if (need <= 64) launch_bounce<64>(g, a, c0, m);
    else if (need <= 96) launch_bounce<96>(g, a, c0, m);
    else launch_bounce<160>(g, a, c0, m);
    DLG_CHECK(hipGetLastError());
    const float ms = dlg_tend(g);
    g->stats.bounceMs += ms;
    g->stats.paths += int64_t(m) * b->M;
    g->stats.bounceLaunches += 1;
  }
  return 0;
}

int32_t dlg_reduce_dc(dlg_gpu* g, int64_t nc, const float* W, double* out, char* err, int32_t errLen) {
  if (nc == 0) return 0;
  if (nc != g->dcN) { dlg_fail(err, errLen, "dlg_reduce_dc: no dlg_bounce rows for this chunk"); return 1; }
  const uint32_t nBins = g->dcBins;
  DLG_CHECK(g->scratch.grow(size_t(nBins) * 4));
  DLG_CHECK(g->out.grow(size_t(nc) * 8));
  DLG_CHECK(hipMemcpyAsync(g->scratch.p, W, size_t(nBins) * 4, hipMemcpyHostToDevice, g->stream));
  k_reduce_dc<<<unsigned((nc + 255) / 256), 256, nBins * 4, g->stream>>>(g->dc.as<float>(), uint32_t(nc), nBins, g->scratch.as<float>(),
                                                                       g->out.as<double>());
  DLG_CHECK(hipGetLastError());
  DLG_CHECK(hipMemcpyAsync(out, g->out.p, size_t(nc) * 8, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int32_t dlg_download_dc(dlg_gpu* g, int64_t nc, float* out, char* err, int32_t errLen) {
  if (nc != g->dcN) { dlg_fail(err, errLen, "dlg_download_dc: no dlg_bounce rows for this chunk"); return 1; }
  const size_t n = size_t(g->dcBins) * size_t(nc);
  if (n) DLG_CHECK(hipMemcpyAsync(out, g->dc.p, n * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"
