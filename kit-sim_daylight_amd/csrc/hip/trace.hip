// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Visibility bits on CDNA4: replaces OptiX __raygen__bits (reference src/trace_dev.cu:91-108), byte mode's
// __raygen__bytes + __anyhit__tr (:115-165) and gpu_trace's launch loop (src/gpu.cu:499-531). No RT cores: a software traversal per ray of the 4-wide records built from the LBVH (lbvh.hip,
// wide.h), each visit two levels of the binary tree.
//
// Launch: persistent waves, as many as the CUs hold at once; each takes tiles of 64 consecutive sensors x one 32-direction
// word in turn. All lanes of a wave trace the same direction from neighbouring origins, so their record and triangle
// fetches mostly coincide (one 128-byte record per visit, served by L1/L2 for the whole wave). Each lane keeps the top
// of its traversal stack in LDS (stack[depth][lane], no bank conflicts) and the deeper entries, which only deep trees
// reach, in its wave's area of a global buffer: the LDS a wave needs stays small, so the waves per SIMD are what the
// registers allow (a 3M-triangle city: 6 instead of 3; the traversal waits on dependent fetches, and more waves hide
// them).
// A ray is open when every hit on [1e-5, inf) is an own-plane hit (dl_math.h own_hit) and they lie at 3 distinct
// distances at most (the reference's closest-hit walk past the sensor's own planes, trace_dev.cu:76-88); the first other
// hit ends the traversal (any-hit, as TERMINATE_ON_FIRST_HIT).
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "dl_math.h"
#include "gpu_internal.h"
#include "lanestack.h"
#include "wide.h"

namespace {

using namespace dlm;

constexpr int kWave = 64;

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

// fetch counts of one lane (the counting variant of the kernel): records visited, triangles tested, the traversal
// iterations its wave executed and the triangle-test iterations its wave executed (the leaf loop's SIMD efficiency is
// triangle tests / (64 x leaf iterations))
struct Fetches { uint32_t nodes = 0, tris = 0, waveIters = 0, leafIters = 0; };

// the lanes' traversal stacks: the top entries in LDS, the deeper ones in the wave's area of a global buffer
using dls::LaneStack;
using dls::lds_stack;
using dls::lds_u32;
using dls::ovf_stack;

// A ray's record fetches through a buffer descriptor (fewer than 2^25 records: byte offsets below 2^32). Each of a
// visit's six plane float4 is one buffer load at the record's byte offset n * 128 (the only per-lane address) plus
// the plane's offset for the ray's direction signs in a scalar register (soffset: axis a's near planes at float4 a or
// a + 3, wide.h Slab), the references at the instruction's immediate 96. The offsets are the same numbers the global
// addresses of wide.h visit() name; one descriptor and six 32-bit offsets instead of seven 64-bit bases.
struct PlaneRsrc {
  __amdgpu_buffer_rsrc_t r;
  uint32_t so[6];
};
__device__ __forceinline__ PlaneRsrc plane_rsrc(const float4* __restrict__ wide, const dlw::Slab& sl) {
  PlaneRsrc b;
  // raw buffer, 32-bit data format (descriptor word 3), no range limit below 4 GiB
  b.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(wide), short(0), int(0xffffffffu), 0x00020000);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b.so[a] = uint32_t(sl.nOff[a]) * 16u;
    b.so[3 + a] = uint32_t(2 * a + 3 - sl.nOff[a]) * 16u;
  }
  return b;
}
__device__ __forceinline__ void visit_buf(const PlaneRsrc& b, uint32_t n, const dlw::Slab& sl, float tmin, float tmax, dlw::Visit& v) {
  const uint32_t off = n * 128u;
  const auto f4 = [&](int k) { return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(b.r, int(off), int(b.so[k]), 0)); };
  const uint4 rf = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(b.r, int(off + 96u), 0, 0));
  dlw::decide<false>(f4(0), f4(1), f4(2), f4(3), f4(4), f4(5), rf, sl, tmin, tmax, v);
}

// every hit through the 4-wide records (wide.h) until hit(leaf) returns false (then false: the ray is blocked); true when
// the walk ends. The stack holds record indices, nearest child taken first.
// OFF32: fewer than 2^25 records, fetched through the buffer descriptor bs (visit_buf)
template <int STACK, bool COUNT, bool OFF32, class Hit>
__device__ __forceinline__ bool walk(const float4* __restrict__ wide, const dlw::Slab& sl, const PlaneRsrc& bs, const LaneStack<STACK>& stk,
                                     int lane, Fetches& f, Hit&& hit) {
  uint32_t cur = 0;
  int sp = 0;
  for (;;) {
    if (COUNT) {
      ++f.nodes;
      if (__builtin_amdgcn_readfirstlane(lane) == lane) ++f.waveIters;   // one per iteration the wave executes
    }
    dlw::Visit v;
    if (OFF32) visit_buf(bs, cur, sl, 1e-5f, 3.40282347e38f, v);
    else dlw::visit<false>(wide, cur, sl, 1e-5f, 3.40282347e38f, v);
    // the lane's hit leaves, one triangle test per iteration (lanes with fewer wait less than with a block per slot); a
    // pair's second triangle is the next iteration's (one test site: the rule's code is not duplicated)
    uint32_t second = dlw::kEmpty;
    for (uint32_t m = v.leafMask; m || second != dlw::kEmpty;) {
      uint32_t leaf = second;
      if (second != dlw::kEmpty) second = dlw::kEmpty;
      else {
        const int c = __ffs(m) - 1;
        m &= m - 1;
        const uint32_t ref = c == 0 ? v.leaf[0] : c == 1 ? v.leaf[1] : c == 2 ? v.leaf[2] : v.leaf[3];
        leaf = dlw::leaf_first(ref);
        if (ref & dlw::kPair) second = leaf + 1;
      }
      if (COUNT) {
        ++f.tris;
        if (__builtin_amdgcn_readfirstlane(lane) == lane) ++f.leafIters;
      }
      if (!hit(leaf)) return false;
    }
    if (v.nInner == 0) {
      if (sp == 0) return true;
      --sp;
      cur = stk.at(sp, lane);
      continue;
    }
    // slot 0 holds the nearest hit inner child; the other hit ones (finite entry distance) are pushed by unconditional
    // writes at positions that skip the misses, and the top moves by the hits (a miss's entry lies above the top, dead;
    // the stack holds two entries beyond the deepest top for them)
    const float inf = __int_as_float(0x7f800000);
    const int h1 = v.tInner[1] < inf, h2 = v.tInner[2] < inf, h3 = v.tInner[3] < inf;
    stk.push3(sp, h1, h2, lane, v.inner[1], v.inner[2], v.inner[3]);
    sp += h1 + h2 + h3;
    cur = v.inner[0];
  }
}

// any hit of a visibility ray: open when every hit is an own-plane hit at 3 distinct distances at most
template <int STACK, bool COUNT, bool OFF32>
__device__ bool ray_open(const float4* __restrict__ wide, const float4* __restrict__ leaves, float pad, float3 o, float3 nrm, float3 d,
                         float4 dirA, float4 dirB, bool cull, const LaneStack<STACK>& stk, int lane, Fetches& f) {
  Ray r;
  dir_unpack(r, d, dirA, dirB);   // the direction's constants, made once per direction (k_dir_table)
  ray_origin(r, o, pad);
  const dlw::Slab sl = dlw::slab(r);
  const PlaneRsrc bs = plane_rsrc(wide, sl);
  OwnSet own;
  return walk<STACK, COUNT, OFF32>(wide, sl, bs, stk, lane, f, [&](uint32_t leaf) { return own.add(r, nrm, leaves, leaf, cull); });
}

// The tiles of a persistent wave: the chunk's sensor blocks (64 positions of the trace's order each) in groups of kGroup
// neighbouring blocks, a group's tiles word-major (tile t of a group: block t % n of word t / n, n its blocks), so the
// waves in flight trace a few words' directions over a few groups of neighbouring sensors. Measured over groups of 16
// blocks to the whole chunk (word-major throughout): the headline city's trace 255.3-259.8 ms, `daylight --bench
// 1000000 12000000` 549-623 ms, 64 blocks within 0.3 % of the best of each. Tiles are dealt to the 8 XCDs in runs
// of kRun (XCD x owns the runs r with r % 8 == x), so the waves of an XCD take neighbouring sensors of one word, which
// share records in that XCD's L2, while every XCD draws from the whole chunk. A wave takes its next tile from its XCD's
// counter (workgroup b runs on XCD b % 8) and, once that XCD's tiles are gone, from the other XCDs' counters in turn:
// tiles vary in cost, and taking them as waves come free keeps every wave busy to the end. Each counter only grows, so
// every wave leaves once all 8 hand out tiles past the last. The counters (kTileCounters uint32, one per 64-byte line)
// start at zero each launch.
constexpr uint32_t kRun = 16, kGroup = 64;
constexpr uint32_t kTileCounters = 8 * 16;
static_assert(kTileCounters * 4 == dlg_tile_counter_bytes, "the counters' bytes");
struct Tiles {
  uint32_t nSB, nTiles;
  uint32_t nW, grp;        // words; sensor blocks per group of the tile order (word-major within a group)
  uint32_t* next;          // [8 x 16]: XCD x's count at 16 x
  const uint32_t* order;   // the sensor at each position of the chunk (dlg_gpu::tileOrder), or null: position = sensor
  __device__ __forceinline__ uint64_t tile(uint32_t j, uint32_t xcd) const { return (uint64_t(j / kRun) * 8 + xcd) * kRun + j % kRun; }
  __device__ __forceinline__ uint32_t sensor(uint32_t p) const { return order ? order[p] : p; }
};
template <class F>
__device__ __forceinline__ void for_tiles(const Tiles& tl, F&& body) {
  const uint32_t home = blockIdx.x % 8;
  for (uint32_t v = 0; v < 8; ++v) {
    const uint32_t x = (home + v) % 8;
    for (;;) {
      uint32_t j = 0;
      if (threadIdx.x == 0) j = atomicAdd(&tl.next[16 * x], 1u);
      j = __builtin_amdgcn_readfirstlane(j);
      const uint64_t t = tl.tile(j, x);
      if (t >= tl.nTiles) break;
      const uint32_t per = tl.grp * tl.nW, gi = uint32_t(t) / per, r = uint32_t(t) - gi * per, s0 = gi * tl.grp;
      const uint32_t sn = tl.nSB - s0 < tl.grp ? tl.nSB - s0 : tl.grp;
      body(s0 + r % sn, r / sn);
    }
  }
}

// The trace's sensor order (dlg_gpu::tileOrder): each sensor's key is its normal's components rounded to -1 / 0 / 1
// (27 orientations, bits 57-61) above the Morton code of its position's 1 m cell (19 bits an axis around the recentred
// origin, clamped), and the chunk-local indices are sorted by it; the sort is stable, so equal keys keep their order.
// Which wave traces a sensor changes nothing in its bits: the order only makes a wave's 64 rays start close together.
__device__ __forceinline__ uint64_t spread3(uint32_t v) {   // bit b of v to bit 3 b
  uint64_t x = v & 0x7ffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}
__global__ void k_order_keys(const float3* __restrict__ P, const float3* __restrict__ N, uint32_t nc, uint64_t* __restrict__ key,
                             uint32_t* __restrict__ idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const float3 p = P[i], n = N[i];
  const auto side = [](float c) { return c > 0.5f ? 2u : c < -0.5f ? 0u : 1u; };
  const auto cell = [](float c) { return uint32_t(fminf(fmaxf(floorf(c) + 262144.f, 0.f), 524287.f)); };
  const uint64_t orient = side(n.x) * 9u + side(n.y) * 3u + side(n.z);
  key[i] = (orient << 57) | spread3(cell(p.x)) | (spread3(cell(p.y)) << 1) | (spread3(cell(p.z)) << 2);
  idx[i] = i;
}
constexpr int kOrderKeyBits = 62;

__global__ void k_zero_counters(uint32_t* c) {
  if (threadIdx.x < kTileCounters) c[threadIdx.x] = 0u;
}

// the rays a wave traced (past the hemisphere cull) added to *traced: a full-wave sum, one atomic per wave
__device__ __forceinline__ void add_traced(uint32_t n, uint32_t lane, unsigned long long* __restrict__ traced) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) n += uint32_t(__shfl_xor(int(n), m));
  if (lane == 0 && n) atomicAdd(traced, (unsigned long long)n);
}

// COUNT: also adds the lane's node visits and triangle tests to counts[0] / counts[1] (the algorithmic fetches) and the
// traversal iterations the wave executed to counts[2] (SIMD efficiency = node visits / (64 x wave iterations))
template <int STACK, bool COUNT, bool OFF32>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(6))) void k_bits(const float4* __restrict__ wide, const float4* __restrict__ leaves, int64_t nTris, float pad,
                                               const float3* __restrict__ P, const float3* __restrict__ N, const float3* __restrict__ D,
                                               const float4* __restrict__ dirTab, uint32_t nDir, uint32_t nc, uint32_t cull, uint32_t* __restrict__ bits,
                                               unsigned long long* __restrict__ counts, unsigned long long* __restrict__ traced,
                                               uint32_t* __restrict__ ovf, Tiles tl) {
  __shared__ uint32_t lds[lds_stack<STACK>() * kWave];
  const LaneStack<STACK> stk{(lds_u32*)lds, ovf + size_t(blockIdx.x) * ovf_stack<STACK>() * kWave};
  const uint32_t lane = threadIdx.x;
  uint32_t nTraced = 0;
  Fetches f;
  for_tiles(tl, [&](uint32_t sb, uint32_t w) {
    const uint32_t p = sb * kWave + lane;   // a tile is 64 consecutive positions of the trace's sensor order
    const bool live = p < nc;   // a lane past the chunk keeps a zero normal: every direction culled, nothing stored
    const uint32_t i = live ? tl.sensor(p) : p;
    const float3 n = live ? N[i] : make_float3(0.f, 0.f, 0.f);
    const float3 o = live ? offset_ray(P[i], n) : n;
    uint32_t word = 0;
    for (uint32_t k = 0; k < 32; ++k) {
      const uint32_t d = w * 32 + k;
      if (d >= nDir) break;
      const float3 dir = D[d];
      if (n.x * dir.x + n.y * dir.y + n.z * dir.z <= 0.f) continue;   // hemisphere cull (trace_dev.cu:101)
      ++nTraced;
      if (nTris == 0 || ray_open<STACK, COUNT, OFF32>(wide, leaves, pad, o, n, dir, dirTab[2 * d], dirTab[2 * d + 1], cull != 0, stk, int(lane), f))
        word |= 1u << k;
    }
    if (live) bits[size_t(w) * nc + i] = word;
  });
  add_traced(nTraced, lane, traced);
  if (COUNT) {
    atomicAdd(&counts[0], (unsigned long long)f.nodes);
    atomicAdd(&counts[1], (unsigned long long)f.tris);
    if (f.waveIters) atomicAdd(&counts[2], (unsigned long long)f.waveIters);
    if (f.leafIters) atomicAdd(&counts[3], (unsigned long long)f.leafIters);
  }
}

// Byte mode (reference __raygen__bytes + __anyhit__tr, trace_dev.cu:115-165): what a ray passes, one walk over every hit
// on [1e-5, inf):
//  - a hit that is not an own plane: opaque blocks the ray; translucent multiplies its factor (dl_math.h pass_factor)
//    into pNon, which only shrinks (below 1/510 the ray is blocked: its byte would be 0), and m is the nearest such hit;
//  - own-plane hits (dl_math.h own_hit) are kept by distance (4 distinct at most; a fifth blocks): whether an opaque one
//    lies there and the product of the translucent ones' factors.
// Then without an opaque own plane the ray passes the product of every factor, the reference's any-hit product (its
// terminate below 1/510 blocks it). With one, the reference reports that hit and walks past the sensor's own planes by
// closest hits (past_own_plane, :76-88): the own planes nearer than m are passed (3 distances at most, else blocked)
// and the ray from there passes the product of what lies beyond them, any opaque hit there blocking it.
// Factors multiply in double, so any hit order rounds to the same float (the reference's float product in OptiX's hit
// order agrees to an ulp). Where the reference's result depends on its hit order (a translucent own plane in a ray
// whose product falls below 1/510), the rule above blocks the ray.
struct ByteRay {
  double pNon = 1.0, f0 = 1.0, f1 = 1.0, f2 = 1.0, f3 = 1.0;   // f_k: the translucent factors of own planes at t_k
  float m = 3.40282347e38f, t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
  int n = 0;
  uint32_t opq = 0;   // bit k: an opaque own plane at t_k
  static constexpr double kMin = double(1.f / 510.f);

  __device__ __forceinline__ bool own(float tt, bool opaque, double fac) {
    int k = (n > 0 && t0 == tt) ? 0 : (n > 1 && t1 == tt) ? 1 : (n > 2 && t2 == tt) ? 2 : (n > 3 && t3 == tt) ? 3 : -1;
    if (k < 0) {
      if (n == 4) return false;
      k = n++;
      if (k == 0) t0 = tt; else if (k == 1) t1 = tt; else if (k == 2) t2 = tt; else t3 = tt;
    }
    if (opaque) opq |= 1u << k;
    else if (k == 0) f0 *= fac; else if (k == 1) f1 *= fac; else if (k == 2) f2 *= fac; else f3 *= fac;
    return true;
  }
  __device__ __forceinline__ bool add(const Ray& r, float3 nrm, const float4* __restrict__ leaves, const float4* __restrict__ slots,
                                      uint32_t slotLast, uint32_t leaf, bool cull) {
    const float4 a = leaves[3 * leaf], b = leaves[3 * leaf + 1], c = leaves[3 * leaf + 2];
    const float3 v0 = make_float3(a.x, a.y, a.z), v1 = make_float3(a.w, b.x, b.y), v2 = make_float3(b.z, b.w, c.x);
    float tt;
    bool front;
    if (!tri_hit(r, v0, v1, v2, 1e-5f, 3.40282347e38f, cull, tt, front)) return true;
    const uint32_t si = __float_as_uint(c.z);
    const float4 s = slots[si < slotLast ? si : slotLast];   // rho, tau, tn, flags
    const bool isOwn = own_hit(r.d, nrm, tt, front, v0, v1, v2);
    if (!(s.y > 0.f)) return isOwn && own(tt, true, 1.0);
    const double fac = pass_factor(s.y, s.z, __float_as_uint(s.w), front, v0, v1, v2, r.d);
    if (isOwn) return own(tt, false, fac);
    pNon *= fac;
    m = fminf(m, tt);
    return pNon >= kMin;
  }
  // the own plane k at t with factor f into the product p (false: blocked)
  __device__ __forceinline__ bool take(int k, float t, double f, double& p, int& below) const {
    if (k >= n) return true;
    if (opq && t < m) { ++below; return true; }   // walked past (an opaque own plane was reported)
    if ((opq >> k) & 1u) return false;
    p *= f;
    return true;
  }
  __device__ __forceinline__ float value() const {
    double p = pNon;
    int below = 0;
    if (!take(0, t0, f0, p, below) || !take(1, t1, f1, p, below) || !take(2, t2, f2, p, below) || !take(3, t3, f3, p, below)) return 0.f;
    if (below >= 4 || !(p >= kMin)) return 0.f;
    return float(p);
  }
};

// bytes in quads: word (32 w + 4 q) / 4 of sensor i at vis[(8 w + q) nc + i], direction 32 w + 4 q + j in byte j; every
// quad of the word is written (0 beyond the last direction)
template <int STACK, bool OFF32>
// 4 waves per SIMD where the stack's LDS allows (128 VGPRs, 8 B spilled): 9 % faster than 137 VGPRs at 3
__global__ __launch_bounds__(kWave, 4) void k_bytes(const float4* __restrict__ wide, const float4* __restrict__ leaves, int64_t nTris, float pad,
                                                const float3* __restrict__ P, const float3* __restrict__ N, const float3* __restrict__ D,
                                                const float4* __restrict__ dirTab, uint32_t nDir, uint32_t nc, uint32_t cull,
                                                const float4* __restrict__ slots, uint32_t slotLast, uint32_t* __restrict__ vis,
                                                unsigned long long* __restrict__ traced, uint32_t* __restrict__ ovf, Tiles tl) {
  __shared__ uint32_t lds[lds_stack<STACK>() * kWave];
  const LaneStack<STACK> stk{(lds_u32*)lds, ovf + size_t(blockIdx.x) * ovf_stack<STACK>() * kWave};
  const uint32_t lane = threadIdx.x;
  Fetches f;
  uint32_t nTraced = 0;
  for_tiles(tl, [&](uint32_t sb, uint32_t w) {
    const uint32_t p = sb * kWave + lane;
    const bool live = p < nc;   // as k_bits: a zero normal past the chunk, nothing stored
    const uint32_t i = live ? tl.sensor(p) : p;
    const float3 n = live ? N[i] : make_float3(0.f, 0.f, 0.f);
    const float3 o = live ? offset_ray(P[i], n) : n;
    uint32_t quad = 0;
    for (uint32_t k = 0; k < 32; ++k) {
      const uint32_t d = w * 32 + k;
      uint32_t b = 0;
      if (d < nDir) {
        const float3 dir = D[d];
        if (n.x * dir.x + n.y * dir.y + n.z * dir.z > 0.f) {   // hemisphere cull (trace_dev.cu:131)
          ++nTraced;
          if (nTris == 0) b = 255;
          else {
            Ray r;
            dir_unpack(r, dir, dirTab[2 * d], dirTab[2 * d + 1]);
            ray_origin(r, o, pad);
            const dlw::Slab sl = dlw::slab(r);
            const PlaneRsrc bs = plane_rsrc(wide, sl);
            ByteRay br;
            if (walk<STACK, false, OFF32>(wide, sl, bs, stk, int(lane), f,
                                          [&](uint32_t leaf) { return br.add(r, n, leaves, slots, slotLast, leaf, cull != 0); }))
              b = T_byte(br.value());
          }
        }
      }
      quad |= b << (8 * (k & 3));
      if ((k & 3) == 3) {
        if (live) vis[size_t(w * 8 + k / 4) * nc + i] = quad;
        quad = 0;
      }
    }
  });
  add_traced(nTraced, lane, traced);
}

// the launch's buffers: sensors, the visibility out, the traced-ray count (added to), the counting variant's counters,
// the stacks' global part and the persistent grid
struct TraceBufs {
  const float3* P;
  const float3* N;
  const uint32_t* order;   // null: the chunk's own order
  uint32_t nc;
  uint32_t* vis;
  unsigned long long* traced;
  unsigned long long* counts;
  uint32_t* ovf;
  uint32_t grid;
  uint32_t* next;   // the tile counters
};

// f(std::integral_constant<int, STACK>) for the instantiated stack sizes
template <class F>
void with_stack(int stack, F&& f) {
  switch (stack) {
    case 24: f(std::integral_constant<int, 24>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    case 48: f(std::integral_constant<int, 48>{}); break;
    case 64: f(std::integral_constant<int, 64>{}); break;
    case 96: f(std::integral_constant<int, 96>{}); break;
    default: f(std::integral_constant<int, 160>{}); break;
  }
}

template <int STACK>
const void* trace_kernel(bool bytes, bool count, bool off32) {
  if (bytes) return off32 ? reinterpret_cast<const void*>(&k_bytes<STACK, true>) : reinterpret_cast<const void*>(&k_bytes<STACK, false>);
  if (count) return off32 ? reinterpret_cast<const void*>(&k_bits<STACK, true, true>) : reinterpret_cast<const void*>(&k_bits<STACK, true, false>);
  return off32 ? reinterpret_cast<const void*>(&k_bits<STACK, false, true>) : reinterpret_cast<const void*>(&k_bits<STACK, false, false>);
}

template <int STACK>
void launch_bytes(dlg_gpu* g, const Tiles& tl, const TraceBufs& b, uint32_t cull) {
  const dlg_scene& s = g->scene;
  const auto k = s.nNodes < (int64_t(1) << 25) ? k_bytes<STACK, true> : k_bytes<STACK, false>;
  k<<<b.grid, kWave, 0, g->stream>>>(s.wide.as<float4>(), s.leaves.as<float4>(), s.nTris, s.pad, b.P, b.N, g->D.as<float3>(), g->dirTab.as<float4>(),
                                     uint32_t(g->nDir), b.nc, cull, g->slots.as<float4>(), uint32_t(g->nSlots - 1), b.vis, b.traced, b.ovf, tl);
}

template <int STACK, bool COUNT>
void launch_bits(dlg_gpu* g, const Tiles& tl, const TraceBufs& b, uint32_t cull) {
  const dlg_scene& s = g->scene;
  const auto k = s.nNodes < (int64_t(1) << 25) ? k_bits<STACK, COUNT, true> : k_bits<STACK, COUNT, false>;
  k<<<b.grid, kWave, 0, g->stream>>>(s.wide.as<float4>(), s.leaves.as<float4>(), s.nTris, s.pad, b.P, b.N, g->D.as<float3>(), g->dirTab.as<float4>(),
                                     uint32_t(g->nDir), b.nc, cull, b.vis, b.counts, b.traced, b.ovf, tl);
}

// the trace kernel of the stack size and the mode, over the buffers b
void launch_trace(dlg_gpu* g, const TraceBufs& b, uint32_t rayFlags, int stack, bool bytes) {
  if (b.nc == 0 || g->nWords == 0) return;
  const uint32_t cull = (rayFlags & 16u) ? 1u : 0u;
  const uint32_t nSB = (b.nc + kWave - 1) / kWave;
  const Tiles tl{nSB, nSB * uint32_t(g->nWords), uint32_t(g->nWords), nSB < kGroup ? nSB : kGroup, b.next, b.order};
  k_zero_counters<<<1, kTileCounters, 0, g->stream>>>(b.next);
  with_stack(stack, [&](auto S) {
    constexpr int ST = decltype(S)::value;
    if (bytes) launch_bytes<ST>(g, tl, b, cull);
    else if (b.counts) launch_bits<ST, true>(g, tl, b, cull);
    else launch_bits<ST, false>(g, tl, b, cull);
  });
}

}  // namespace

uint32_t dlg_blocks_per_cu(const void* fn) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, int>> known;
  std::lock_guard<std::mutex> lk(mu);
  for (const auto& [f, n] : known)
    if (f == fn) return uint32_t(n);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kWave, 0) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    n = 4;
  }
  known.push_back({fn, n});
  return uint32_t(n);
}

// The checks and buffers of a trace of sensors [i0, i0 + nc): 0, or 1 with err. stack: the LDS stack entries the
// scene's tree needs (the smallest instantiation that covers it, with the two entries above the top that the
// unconditional pushes write).
int32_t dlg_trace_check(const dlg_gpu* g, uint32_t rayFlags, int& stack, char* err, int32_t errLen) {
  const bool bytes = !(rayFlags & 1u);   // any-hit on: a slot transmits
  const dlg_scene& s = g->scene;
  if (bytes && s.nTris > 0 && (g->nSlots == 0 || int64_t(s.slotMax) >= g->nSlots)) {
    dlg_fail(err, errLen, "dlg_trace: the scene names slot " + std::to_string(s.slotMax) + " and " + std::to_string(g->nSlots) +
                              " slots are set (dlg_set_slots)");
    return 1;
  }
  const int need = dlw::wide_stack(s.height) + 2;
  if (need > 160) { dlg_fail(err, errLen, "dlg_trace: the scene's tree is too deep for the traversal stack"); return 1; }
  stack = need <= 24 ? 24 : need <= 32 ? 32 : need <= 48 ? 48 : need <= 64 ? 64 : need <= 96 ? 96 : 160;
  return 0;
}

dlg_trace_shape dlg_trace_grid(const dlg_gpu* g, int64_t nc, int stack, bool bytes, bool count) {
  dlg_trace_shape t{};
  const uint64_t nTiles = uint64_t((nc + kWave - 1) / kWave) * uint64_t(g->nWords);
  with_stack(stack, [&](auto S) {
    constexpr int ST = decltype(S)::value;
    uint64_t grid = uint64_t(dlg_blocks_per_cu(trace_kernel<ST>(bytes, count, g->scene.nNodes < (int64_t(1) << 25)))) * uint64_t(g->cus > 0 ? g->cus : 1);
    grid = grid < 8 ? 8 : grid / 8 * 8;   // whole rounds of the 8 XCDs
    const uint64_t need = (nTiles + 7) / 8 * 8;
    t.grid = uint32_t(need < grid ? (need < 8 ? 8 : need) : grid);
    t.ovfPerBlock = size_t(ovf_stack<ST>()) * kWave * 4;
  });
  return t;
}

int32_t dlg_trace_prepare(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, int& stack, char* err, int32_t errLen) {
  if (i0 + nc > g->nSens) { dlg_fail(err, errLen, "dlg_trace: sensor range exceeds the scene"); return 1; }
  if (nc > int64_t(0x7fffffff)) { dlg_fail(err, errLen, "dlg_trace: at most 2^31 - 1 sensors a chunk"); return 1; }
  if (dlg_trace_check(g, rayFlags, stack, err, errLen)) return 1;
  const bool bytes = !(rayFlags & 1u);
  DLG_CHECK(g->bits.grow(size_t(g->nWords) * size_t(nc) * (bytes ? 32 : 4)));
  const dlg_trace_shape sh = dlg_trace_grid(g, nc, stack, bytes, g->countFetches != 0);
  g->traceGrid = sh.grid;
  if (sh.ovfPerBlock) DLG_CHECK(g->stackOvf.grow(sh.ovfPerBlock * sh.grid));
  DLG_CHECK(g->tileNext.grow(kTileCounters * 4));
  if (!g->traceCount.p) {   // the running count of traced rays, read after each launch
    DLG_CHECK(g->traceCount.grow(sizeof(unsigned long long)));
    DLG_CHECK(hipMemsetAsync(g->traceCount.p, 0, sizeof(unsigned long long), g->stream));
  }
  if (g->orderGen != g->sensGen || g->orderI0 != i0 || g->orderN != nc) {   // the chunk's sensor order
    DLG_CHECK(g->tileOrder.grow(size_t(nc > 0 ? nc : 1) * 4));
    if (nc > 0) {
      size_t tmp = 0;
      DLG_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, static_cast<const uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                                                   static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), int(nc), 0, kOrderKeyBits,
                                                   g->stream));
      const size_t oKey2 = size_t(nc) * 8, oIdx = 2 * oKey2, oTmp = (oIdx + size_t(nc) * 4 + 255) & ~size_t(255);
      DLG_CHECK(g->orderWork.grow(oTmp + tmp));
      char* w = g->orderWork.as<char>();
      uint64_t* key = reinterpret_cast<uint64_t*>(w);
      uint32_t* idx = reinterpret_cast<uint32_t*>(w + oIdx);
      k_order_keys<<<unsigned((nc + 255) / 256), 256, 0, g->stream>>>(g->P.as<float3>() + i0, g->N.as<float3>() + i0, uint32_t(nc), key, idx);
      DLG_CHECK(hipcub::DeviceRadixSort::SortPairs(w + oTmp, tmp, key, reinterpret_cast<uint64_t*>(w + oKey2), idx, g->tileOrder.as<uint32_t>(),
                                                   int(nc), 0, kOrderKeyBits, g->stream));
    }
    g->orderGen = g->sensGen;
    g->orderI0 = i0;
    g->orderN = nc;
  }
  g->i0 = i0;
  g->nc = nc;
  g->bytes = bytes ? 1 : 0;
  return 0;
}

// The trace's kernel enqueued on the library's stream (no wait: a graph may capture it); counts non-null runs the
// counting variant.
void dlg_trace_enqueue(dlg_gpu* g, uint32_t rayFlags, int stack, unsigned long long* counts) {
  const TraceBufs b{g->P.as<float3>() + g->i0, g->N.as<float3>() + g->i0, g->tileOrder.as<uint32_t>(), uint32_t(g->nc), g->bits.as<uint32_t>(),
                    g->traceCount.as<unsigned long long>(), counts, g->stackOvf.as<uint32_t>(), g->traceGrid, g->tileNext.as<uint32_t>()};
  launch_trace(g, b, rayFlags, stack, g->bytes != 0);
}

void dlg_trace_enqueue_into(dlg_gpu* g, const float* P, const float* N, uint32_t nc, uint32_t rayFlags, int stack, uint32_t* vis,
                            unsigned long long* traced, uint32_t* ovf, uint32_t grid, uint32_t* next) {
  const TraceBufs b{reinterpret_cast<const float3*>(P), reinterpret_cast<const float3*>(N), nullptr, nc, vis, traced, nullptr, ovf, grid, next};
  launch_trace(g, b, rayFlags, stack, !(rayFlags & 1u));
}

// the statistics of a finished trace launch of ms milliseconds
int32_t dlg_trace_account(dlg_gpu* g, float ms, char* err, int32_t errLen) {
  unsigned long long traced = 0;
  DLG_CHECK(hipMemcpy(&traced, g->traceCount.p, sizeof traced, hipMemcpyDeviceToHost));
  g->stats.traced = int64_t(traced);
  g->stats.traceMs += ms;
  g->stats.lastTraceMs = ms;
  g->stats.rays += g->nc * g->nDir;
  g->stats.launches += 1;
  return 0;
}

extern "C" int32_t dlg_trace(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, char* err, int32_t errLen) {
  int stack = 0;
  if (dlg_trace_prepare(g, i0, nc, rayFlags, stack, err, errLen)) return 1;
  if (nc == 0 || g->nWords == 0) return 0;
  unsigned long long* counts = nullptr;
  if (g->countFetches && !g->bytes) {
    DLG_CHECK(g->counts.grow(4 * sizeof(unsigned long long)));
    counts = g->counts.as<unsigned long long>();
    DLG_CHECK(hipMemsetAsync(counts, 0, 4 * sizeof(unsigned long long), g->stream));
  }
  dlg_tbegin(g);
  dlg_trace_enqueue(g, rayFlags, stack, counts);
  DLG_CHECK(hipGetLastError());
  const float ms = dlg_tend(g);
  if (counts) {
    unsigned long long c[4] = {0, 0, 0, 0};
    DLG_CHECK(hipMemcpy(c, counts, sizeof c, hipMemcpyDeviceToHost));
    g->stats.nodeFetches += int64_t(c[0]);
    g->stats.triFetches += int64_t(c[1]);
    g->stats.waveIters += int64_t(c[2]);
    g->stats.leafIters += int64_t(c[3]);
    g->stats.countedRays += nc * g->nDir;
  }
  return dlg_trace_account(g, ms, err, errLen);
}
