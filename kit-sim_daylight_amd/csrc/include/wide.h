// SPDX-License-Identifier: Apache-2.0
// The 4-wide records the traversal kernels walk (built from the LBVH by k_wide, lbvh.hip).
//
// Record n, one per LBVH inner node: the boxes and references of n's children, an inner child replaced by its own two
// children, so that one visit covers two levels of the binary tree (half the dependent fetches per ray; the traversal
// is bound by their latency, not by bytes). 8 float4 = 128 bytes:
//   q[0..5]  lo x, lo y, lo z, hi x, hi y, hi z of the 4 slots (SoA: one 16-byte load per coordinate)
//   q[6]     the 4 references as bits: bit 31 set = leaf (index in the low 31 bits), else the record of an inner node;
//            0xffffffff = empty slot (its box is empty: lo +inf, hi -inf, never hit)
//   q[7]     unused
// Visiting a record: the slot boxes are tested against the ray, hit leaves go to the caller's triangle test, hit inner
// slots come back sorted near to far (a 4-input sorting network). Which order the children are visited in does not
// change any result: the visibility rule depends on the set of hits and the closest hit breaks exact ties by the lower
// triangle index.
#pragma once
#include "dl_math.h"

namespace dlw {

using dlm::Ray;

constexpr uint32_t kEmpty = 0xffffffffu;

struct BoxHit { bool hit; float t; };

// slab test of one box on [tmin, tmax], robust to the box padding through the ray's near / far origins
DL_HD BoxHit box_hit(const Ray& r, float lx, float ly, float lz, float hx, float hy, float hz, float tmin, float tmax) {
  const float nx = r.inv.x >= 0.f ? lx : hx, fx = r.inv.x >= 0.f ? hx : lx;
  const float ny = r.inv.y >= 0.f ? ly : hy, fy = r.inv.y >= 0.f ? hy : ly;
  const float nz = r.inv.z >= 0.f ? lz : hz, fz = r.inv.z >= 0.f ? hz : lz;
  const float t0 = fmaxf(fmaxf(fmaxf(tmin, (nx - r.oNear.x) * r.inv.x), (ny - r.oNear.y) * r.inv.y), (nz - r.oNear.z) * r.inv.z);
  const float t1 = fminf(fminf(fminf(tmax, (fx - r.oFar.x) * r.inv.x), (fy - r.oFar.y) * r.inv.y), (fz - r.oFar.z) * r.inv.z);
  return {t0 <= t1, t0};
}

// The ray as the slab test of a record uses it: per axis the inverse direction and the two origins' products with it,
// the one meeting a slot's lo plane and the one meeting its hi plane (oNear for the near plane, oFar for the far one,
// by the sign of the direction). A plane's distance is then one fma, (plane inv - o inv); its rounding (2^-24 |o| along
// the ray) stays far inside the box padding (2^-18 (|coord| + 1)), so the test remains conservative. Per axis the
// near distance is min(lo's, hi's) and the far one max: with lo <= hi that is the plane the sign selects.
struct Slab { float3 inv, oLo, oHi; };
DL_HD Slab slab(const Ray& r) {
  const float3 lo = make_float3(r.inv.x >= 0.f ? r.oNear.x : r.oFar.x, r.inv.y >= 0.f ? r.oNear.y : r.oFar.y, r.inv.z >= 0.f ? r.oNear.z : r.oFar.z);
  const float3 hi = make_float3(r.inv.x >= 0.f ? r.oFar.x : r.oNear.x, r.inv.y >= 0.f ? r.oFar.y : r.oNear.y, r.inv.z >= 0.f ? r.oFar.z : r.oNear.z);
  return {r.inv, make_float3(lo.x * r.inv.x, lo.y * r.inv.y, lo.z * r.inv.z), make_float3(hi.x * r.inv.x, hi.y * r.inv.y, hi.z * r.inv.z)};
}
DL_HD BoxHit slab_hit(const Slab& s, float lx, float ly, float lz, float hx, float hy, float hz, float tmin, float tmax) {
  const float ax = fmaf(lx, s.inv.x, -s.oLo.x), bx = fmaf(hx, s.inv.x, -s.oHi.x);
  const float ay = fmaf(ly, s.inv.y, -s.oLo.y), by = fmaf(hy, s.inv.y, -s.oHi.y);
  const float az = fmaf(lz, s.inv.z, -s.oLo.z), bz = fmaf(hz, s.inv.z, -s.oHi.z);
  const float t0 = fmaxf(fmaxf(tmin, fminf(ax, bx)), fmaxf(fminf(ay, by), fminf(az, bz)));
  const float t1 = fminf(fminf(tmax, fmaxf(ax, bx)), fminf(fmaxf(ay, by), fmaxf(az, bz)));
  return {t0 <= t1, t0};
}

// One record visit: leaf slots hit (bit c of leafMask, reference leaf[c]) and the inner slots with their entry distance
// (+inf where not hit), slot 0 the nearest hit one.
struct Visit {
  uint32_t leafMask;
  uint32_t leaf[4];
  int nInner;
  uint32_t inner[4];
  float tInner[4];
};

DL_HD void cx(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
  const bool s = kb < ka;
  const float k0 = s ? kb : ka, k1 = s ? ka : kb;
  const uint32_t r0 = s ? rb : ra, r1 = s ? ra : rb;
  ka = k0; kb = k1; ra = r0; rb = r1;
}

// FULL: all four inner slots sorted (closest hit: the pops come nearest first, which prunes); else only the nearest is
// moved to slot 0 (any hit: the order of the rest does not matter, and three fewer compare-exchanges per visit)
template <bool FULL>
DL_HD void visit(const float4* __restrict__ wide, uint32_t n, const Slab& sl, float tmin, float tmax, Visit& v) {
  const float4* q = wide + 8 * size_t(n);
  const float4 lx = q[0], ly = q[1], lz = q[2], hx = q[3], hy = q[4], hz = q[5];
  const uint4 rf = *reinterpret_cast<const uint4*>(q + 6);
  BoxHit h0 = slab_hit(sl, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, tmin, tmax);
  BoxHit h1 = slab_hit(sl, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, tmin, tmax);
  BoxHit h2 = slab_hit(sl, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, tmin, tmax);
  BoxHit h3 = slab_hit(sl, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, tmin, tmax);
  // an empty slot (lo +inf, hi -inf) is never followed: excluded by its reference
  h0.hit = h0.hit && rf.x != kEmpty; h1.hit = h1.hit && rf.y != kEmpty; h2.hit = h2.hit && rf.z != kEmpty; h3.hit = h3.hit && rf.w != kEmpty;
  const bool l0 = (rf.x & 0x80000000u) != 0, l1 = (rf.y & 0x80000000u) != 0, l2 = (rf.z & 0x80000000u) != 0, l3 = (rf.w & 0x80000000u) != 0;
  v.leafMask = (h0.hit && l0 ? 1u : 0u) | (h1.hit && l1 ? 2u : 0u) | (h2.hit && l2 ? 4u : 0u) | (h3.hit && l3 ? 8u : 0u);
  v.leaf[0] = rf.x & 0x7fffffffu; v.leaf[1] = rf.y & 0x7fffffffu; v.leaf[2] = rf.z & 0x7fffffffu; v.leaf[3] = rf.w & 0x7fffffffu;
  const float inf = dlm::as_float(0x7f800000u);
  // inner slots hit keep their entry distance (finite: t0 <= t1 <= tmax), the others +inf; then sort
  float k0 = h0.hit && !l0 ? h0.t : inf, k1 = h1.hit && !l1 ? h1.t : inf, k2 = h2.hit && !l2 ? h2.t : inf, k3 = h3.hit && !l3 ? h3.t : inf;
  uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
  cx(k0, r0, k1, r1);
  cx(k2, r2, k3, r3);
  cx(k0, r0, k2, r2);
  if (FULL) {
    cx(k1, r1, k3, r3);
    cx(k1, r1, k2, r2);
  }
  v.nInner = (k0 < inf) + (k1 < inf) + (k2 < inf) + (k3 < inf);
  v.inner[0] = r0; v.inner[1] = r1; v.inner[2] = r2; v.inner[3] = r3;
  v.tInner[0] = k0; v.tInner[1] = k1; v.tInner[2] = k2; v.tInner[3] = k3;
}

// Stack entries a traversal of a tree of the given LBVH height can hold at once: records sit at even depths (at most
// ceil(height / 2) on a path) and a visit pushes at most 3.
inline int wide_stack(int height) { return 3 * ((height + 1) / 2) + 1; }

}  // namespace dlw
