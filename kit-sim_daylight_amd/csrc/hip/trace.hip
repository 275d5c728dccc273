// SPDX-License-Identifier: Apache-2.0
// Visibility bits on CDNA4: replaces OptiX __raygen__bits (reference src/trace_dev.cu:91-108) and gpu_trace's launch loop
// (src/gpu.cu:499-531). No RT cores: a software traversal of the LBVH (lbvh.hip) per ray.
//
// Launch: a wave64 per 64 consecutive sensors and one 32-direction word (grid y). All lanes of a wave trace the same
// direction from neighbouring origins, so their node and triangle fetches mostly coincide (one 64-byte node line per
// visit, served by L1/L2 for the whole wave). Each lane keeps its traversal stack in LDS (stack[depth][lane], no bank
// conflicts), sized by the tree's height.
// A ray is open when every hit on [1e-5, inf) is an own-plane hit (dl_math.h own_hit) and they lie at 3 distinct
// distances at most (the reference's closest-hit walk past the sensor's own planes, trace_dev.cu:76-88); the first other
// hit ends the traversal (any-hit, as TERMINATE_ON_FIRST_HIT).
#include <type_traits>

#include "dl_math.h"
#include "gpu_internal.h"

namespace {

using namespace dlm;

constexpr int kWave = 64;

struct BoxHit { bool hit; float t; };
__device__ __forceinline__ BoxHit box_hit(const Ray& r, float lx, float ly, float lz, float hx, float hy, float hz, float tmin, float tmax) {
  const float nx = r.inv.x >= 0.f ? lx : hx, fx = r.inv.x >= 0.f ? hx : lx;
  const float ny = r.inv.y >= 0.f ? ly : hy, fy = r.inv.y >= 0.f ? hy : ly;
  const float nz = r.inv.z >= 0.f ? lz : hz, fz = r.inv.z >= 0.f ? hz : lz;
  const float t0 = fmaxf(fmaxf(fmaxf(tmin, (nx - r.oNear.x) * r.inv.x), (ny - r.oNear.y) * r.inv.y), (nz - r.oNear.z) * r.inv.z);
  const float t1 = fminf(fminf(fminf(tmax, (fx - r.oFar.x) * r.inv.x), (fy - r.oFar.y) * r.inv.y), (fz - r.oFar.z) * r.inv.z);
  return {t0 <= t1, t0};
}

// the own-plane bookkeeping of one visibility ray: distinct distances of own-plane hits so far
struct OwnSet {
  int n = 0;
  float t0 = 0.f, t1 = 0.f, t2 = 0.f;
  // false: the ray is blocked (a hit that is not an own plane, or a fourth own plane)
  __device__ __forceinline__ bool add(const Ray& r, float3 nrm, const float4* __restrict__ leaves, uint32_t leaf, bool cull) {
    const float4 a = leaves[3 * leaf], b = leaves[3 * leaf + 1], c = leaves[3 * leaf + 2];
    const float3 v0 = make_float3(a.x, a.y, a.z), v1 = make_float3(a.w, b.x, b.y), v2 = make_float3(b.z, b.w, c.x);
    float t;
    bool front;
    if (!tri_hit(r, v0, v1, v2, 1e-5f, 3.40282347e38f, cull, t, front)) return true;
    if (!own_hit(r.d, nrm, t, front, v0, v1, v2)) return false;
    if ((n > 0 && t == t0) || (n > 1 && t == t1) || (n > 2 && t == t2)) return true;
    if (n == 3) return false;
    if (n == 0) t0 = t; else if (n == 1) t1 = t; else t2 = t;
    ++n;
    return true;
  }
};

// fetch counts of one lane (the counting variant of the kernel): nodes visited, triangles tested
struct Fetches { uint32_t nodes = 0, tris = 0; };

template <int STACK, bool COUNT>
__device__ bool ray_open(const float4* __restrict__ nodes, const float4* __restrict__ leaves, float pad, float3 o, float3 nrm, float3 d,
                         bool cull, uint32_t* __restrict__ stk, int lane, Fetches& f) {
  Ray r;
  ray_setup(r, o, d, pad);
  OwnSet own;
  uint32_t cur = 0;
  int sp = 0;
  const float tmax = 3.40282347e38f;
  for (;;) {
    if (COUNT) ++f.nodes;
    const float4 A = nodes[4 * cur], B = nodes[4 * cur + 1], C = nodes[4 * cur + 2], E = nodes[4 * cur + 3];
    const BoxHit hl = box_hit(r, A.x, A.y, A.z, A.w, B.x, B.y, 1e-5f, tmax);
    const BoxHit hr = box_hit(r, B.z, B.w, C.x, C.y, C.z, C.w, 1e-5f, tmax);
    const uint32_t cl = __float_as_uint(E.x), cr = __float_as_uint(E.y);
    if (COUNT) f.tris += (hl.hit && (cl & 0x80000000u)) + (hr.hit && (cr & 0x80000000u));
    if (hl.hit && (cl & 0x80000000u) && !own.add(r, nrm, leaves, cl & 0x7fffffffu, cull)) return false;
    if (hr.hit && (cr & 0x80000000u) && !own.add(r, nrm, leaves, cr & 0x7fffffffu, cull)) return false;
    const bool il = hl.hit && !(cl & 0x80000000u), ir = hr.hit && !(cr & 0x80000000u);
    if (il && ir) {
      const bool leftFirst = hl.t <= hr.t;
      stk[sp * kWave + lane] = leftFirst ? cr : cl;
      ++sp;
      cur = leftFirst ? cl : cr;
    } else if (il) cur = cl;
    else if (ir) cur = cr;
    else {
      if (sp == 0) return true;
      --sp;
      cur = stk[sp * kWave + lane];
    }
  }
}

// COUNT: also adds the lane's node visits and triangle tests to counts[0] / counts[1] (the algorithmic fetches)
template <int STACK, bool COUNT>
__global__ __launch_bounds__(kWave) void k_bits(const float4* __restrict__ nodes, const float4* __restrict__ leaves, int64_t nTris, float pad,
                                               const float3* __restrict__ P, const float3* __restrict__ N, const float3* __restrict__ D,
                                               uint32_t nDir, uint32_t nc, uint32_t cull, uint32_t* __restrict__ bits,
                                               unsigned long long* __restrict__ counts) {
  __shared__ uint32_t stk[STACK * kWave];
  const uint32_t lane = threadIdx.x, i = blockIdx.x * kWave + lane, w = blockIdx.y;
  if (i >= nc) return;
  const float3 n = N[i];
  const float3 o = offset_ray(P[i], n);
  uint32_t word = 0;
  Fetches f;
  for (uint32_t k = 0; k < 32; ++k) {
    const uint32_t d = w * 32 + k;
    if (d >= nDir) break;
    const float3 dir = D[d];
    if (n.x * dir.x + n.y * dir.y + n.z * dir.z <= 0.f) continue;   // hemisphere cull (trace_dev.cu:101)
    if (nTris == 0 || ray_open<STACK, COUNT>(nodes, leaves, pad, o, n, dir, cull != 0, stk, int(lane), f)) word |= 1u << k;
  }
  bits[size_t(w) * nc + i] = word;
  if (COUNT) {
    atomicAdd(&counts[0], (unsigned long long)f.nodes);
    atomicAdd(&counts[1], (unsigned long long)f.tris);
  }
}

template <int STACK, bool COUNT>
void launch_bits(dlg_gpu* g, dim3 grid, const float3* P, const float3* N, uint32_t nc, uint32_t cull, unsigned long long* counts) {
  const dlg_scene& s = g->scene;
  k_bits<STACK, COUNT><<<grid, kWave, 0, g->stream>>>(s.nodes.as<float4>(), s.leaves.as<float4>(), s.nTris, s.pad, P, N, g->D.as<float3>(),
                                                      uint32_t(g->nDir), nc, cull, g->bits.as<uint32_t>(), counts);
}

}  // namespace

extern "C" int32_t dlg_trace(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, char* err, int32_t errLen) {
  if (i0 + nc > g->nSens) { dlg_fail(err, errLen, "dlg_trace: sensor range exceeds the scene"); return 1; }
  if (!(rayFlags & 1u)) { dlg_fail(err, errLen, "dlg_trace: transmitting slots (any-hit) are not part of this build"); return 1; }
  DLG_CHECK(g->bits.grow(size_t(g->nWords) * size_t(nc) * 4));
  g->i0 = i0;
  g->nc = nc;
  if (nc == 0 || g->nWords == 0) return 0;
  const dlg_scene& s = g->scene;
  const uint32_t cull = (rayFlags & 16u) ? 1u : 0u;
  const dim3 grid(unsigned((nc + kWave - 1) / kWave), unsigned(g->nWords));
  const float3* P = g->P.as<float3>() + i0;
  const float3* N = g->N.as<float3>() + i0;
  if (s.height > 160) { dlg_fail(err, errLen, "dlg_trace: the scene's tree is deeper than 160 levels"); return 1; }
  unsigned long long* counts = nullptr;
  if (g->countFetches) {
    DLG_CHECK(g->counts.grow(2 * sizeof(unsigned long long)));
    counts = g->counts.as<unsigned long long>();
    DLG_CHECK(hipMemsetAsync(counts, 0, 2 * sizeof(unsigned long long), g->stream));
  }
  // the stack holds at most one entry per level below the root: the smallest variant that covers the tree's height
  // (a smaller LDS stack lets more waves share a CU)
  const auto launch = [&](auto count) {
    constexpr bool C_ = decltype(count)::value;
    if (s.height <= 24) launch_bits<24, C_>(g, grid, P, N, uint32_t(nc), cull, counts);
    else if (s.height <= 32) launch_bits<32, C_>(g, grid, P, N, uint32_t(nc), cull, counts);
    else if (s.height <= 48) launch_bits<48, C_>(g, grid, P, N, uint32_t(nc), cull, counts);
    else if (s.height <= 64) launch_bits<64, C_>(g, grid, P, N, uint32_t(nc), cull, counts);
    else launch_bits<160, C_>(g, grid, P, N, uint32_t(nc), cull, counts);
  };
  dlg_tbegin(g);
  if (counts) launch(std::true_type{});
  else launch(std::false_type{});
  DLG_CHECK(hipGetLastError());
  const float ms = dlg_tend(g);
  if (counts) {
    unsigned long long c[2] = {0, 0};
    DLG_CHECK(hipMemcpy(c, counts, sizeof c, hipMemcpyDeviceToHost));
    g->stats.nodeFetches += int64_t(c[0]);
    g->stats.triFetches += int64_t(c[1]);
    g->stats.countedRays += nc * g->nDir;
  }
  g->stats.traceMs += ms;
  g->stats.lastTraceMs = ms;
  g->stats.rays += nc * g->nDir;
  g->stats.launches += 1;
  return 0;
}
