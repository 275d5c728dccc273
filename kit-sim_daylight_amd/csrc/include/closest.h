// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The closest hit of a ray through the 4-wide records (wide.h), for the kernels that walk closest hits: views' rays
// (view.hip) and the paths of reflected light (bounce.hip). Exact ties of t go to the lower triangle index.
#pragma once
#include "dl_math.h"
#include "lanestack.h"
#include "wide.h"

namespace dlc {

using dlm::Ray;
constexpr int kWave = 64;

struct Closest { float t; uint32_t leaf, prim; bool front, hit; };

using dls::lds_u32;

__device__ __forceinline__ void leaf_test(const Ray& r, const float4* __restrict__ leaves, uint32_t leaf, float tmin, bool cull, Closest& c) {
  const float4 a = leaves[3 * leaf], b = leaves[3 * leaf + 1], e = leaves[3 * leaf + 2];
  float t;
  bool front;
  if (!dlm::tri_hit(r, make_float3(a.x, a.y, a.z), make_float3(a.w, b.x, b.y), make_float3(b.z, b.w, e.x), tmin, c.t, cull, t, front)) return;
  const uint32_t prim = __float_as_uint(e.y);
  if (c.hit && t == c.t && prim > c.prim) return;
  c.t = t; c.leaf = leaf; c.prim = prim; c.front = front; c.hit = true;
}

// closest hit through the 4-wide records (wide.h): children nearest first, those starting beyond the best hit dropped;
// stk: the lane's stack (lanestack.h), STACK entries; OFF32: fewer than 2^25 records (wide.h visit)
template <bool OFF32 = false, int STACK, int KL>
__device__ Closest closest_hit(const float4* __restrict__ wide, const float4* __restrict__ leaves, const Ray& r, float tmin, float tmax, bool cull,
                               const dls::LaneStack<STACK, KL>& stk, int lane) {
  Closest c{tmax, 0u, 0u, false, false};
  const dlw::Slab sl = dlw::slab(r);
  uint32_t cur = 0;
  int sp = 0;
  for (;;) {
    dlw::Visit v;
    dlw::visit<true, OFF32>(wide, cur, sl, tmin, c.t, v);
    uint32_t second = dlw::kEmpty;   // a pair's second triangle, the next iteration's
    for (uint32_t m = v.leafMask; m || second != dlw::kEmpty;) {
      uint32_t leaf = second;
      if (second != dlw::kEmpty) second = dlw::kEmpty;
      else {
        const int k = __ffs(m) - 1;
        m &= m - 1;
        const uint32_t ref = k == 0 ? v.leaf[0] : k == 1 ? v.leaf[1] : k == 2 ? v.leaf[2] : v.leaf[3];
        leaf = dlw::leaf_first(ref);
        if (ref & dlw::kPair) second = leaf + 1;
      }
      leaf_test(r, leaves, leaf, tmin, cull, c);
    }
    // slot 0 holds the nearest hit inner child; a child is taken where it starts within the best hit so far
    if (!(v.tInner[0] <= c.t)) {
      if (sp == 0) return c;
      --sp;
      cur = stk.at(sp, lane);
      continue;
    }
    // the sorted slots 1..3 that start within the best hit are a prefix (n of them): pushed farthest first, so that
    // the nearest is popped next, by unconditional writes; the ones not taken land above the new top, dead (at most
    // two entries above the deepest top: the stack holds three beyond it)
    const int n = (v.tInner[1] <= c.t) + (v.tInner[2] <= c.t) + (v.tInner[3] <= c.t);
    stk.put3(sp, n >= 1 ? n - 1 : 0, n >= 2 ? n - 2 : 1, n >= 3 ? 0 : 2, lane, v.inner[1], v.inner[2], v.inner[3]);
    sp += n;
    cur = v.inner[0];
  }
}

}  // namespace dlc
