// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The 4-wide records the traversal kernels walk (built from the LBVH by k_wide, lbvh.hip).
//
// Record n, one per LBVH inner node: the boxes and references of n's children, an inner child replaced by its own two
// children, so that one visit covers two levels of the binary tree (half the dependent fetches per ray; the traversal
// is bound by their latency, not by bytes). 8 float4 = 128 bytes:
//   q[0..5]  lo x, lo y, lo z, hi x, hi y, hi z of the 4 slots (SoA: one 16-byte load per coordinate)
//   q[6]     the 4 references as bits: bit 31 set = leaf, else the record of an inner node; a leaf's index in the low 30
//            bits, and bit 30 set = a pair: that leaf and the next (an inner node of two adjacent leaves, the two
//            triangles of a quad in a mesh of quads, tested from this record instead of a record of its own);
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
constexpr uint32_t kPair = 0x40000000u;   // in a leaf reference (Visit::leaf): the leaf and the next one
DL_HD uint32_t leaf_first(uint32_t ref) { return ref & 0x3fffffffu; }

struct BoxHit { bool hit; float t; };

// The ray as the slab test of a record uses it. Per axis the near plane of every box is its lo plane when the direction
// is positive and its hi plane otherwise (and the far plane the other), so the record's six SoA float4 are read as
// near x/y/z and far x/y/z by the direction's signs (nOff: the float4 index of axis a's near planes, a or a + 3), and
// a plane's distance is one fma, plane inv - c: c = oNear inv for near planes, oFar inv for far ones (the ray's padded
// origins, dl_math.h ray_origin). Its rounding (2^-24 |o| along the ray) stays far inside the box padding (2^-18
// (|coord| + 1)), so the test remains conservative. Entry and exit are the same numbers as the min / max of the two
// planes' distances would give (for a box, lo <= hi, the near plane's distance is the smaller one, fma being monotonic).
struct Slab { float3 inv, cN, cF; int nOff[3]; };
DL_HD Slab slab(const Ray& r) {
  Slab s;
  s.inv = r.inv;
  s.cN = make_float3(r.oNear.x * r.inv.x, r.oNear.y * r.inv.y, r.oNear.z * r.inv.z);
  s.cF = make_float3(r.oFar.x * r.inv.x, r.oFar.y * r.inv.y, r.oFar.z * r.inv.z);
  s.nOff[0] = r.inv.x >= 0.f ? 0 : 3;
  s.nOff[1] = r.inv.y >= 0.f ? 1 : 4;
  s.nOff[2] = r.inv.z >= 0.f ? 2 : 5;
  return s;
}
// TMAX false: the exit is not clamped to tmax (any hit, whose tmax is FLT_MAX: the clamp only matters where the exit
// is +inf, and a box kept there is one more visit, never a lost hit)
template <bool TMAX>
DL_HD BoxHit slab_hit(const Slab& s, float nx, float ny, float nz, float fx, float fy, float fz, float tmin, float tmax) {
  const float t0 = fmaxf(fmaxf(fmaf(nx, s.inv.x, -s.cN.x), fmaf(ny, s.inv.y, -s.cN.y)), fmaxf(fmaf(nz, s.inv.z, -s.cN.z), tmin));
  const float f = fminf(fminf(fmaf(fx, s.inv.x, -s.cF.x), fmaf(fy, s.inv.y, -s.cF.y)), fmaf(fz, s.inv.z, -s.cF.z));
  const float t1 = TMAX ? fminf(f, tmax) : f;
  return {t0 <= t1, t0};
}

// One record visit: leaf slots hit (bit c of leafMask, reference leaf[c]: leaf_first and kPair) and the inner slots with
// their entry distance (+inf where not hit), slot 0 the nearest hit one.
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

// FULL: all four inner slots sorted (closest hit: the pops come nearest first, which prunes) and the boxes cut at tmax;
// else only the nearest is moved to slot 0 (any hit: the order of the rest does not matter, and three fewer
// compare-exchanges per visit) and tmax, FLT_MAX there, is not applied.
// nx..fz: the record's near and far plane float4s (slab), rf its references.
template <bool FULL>
DL_HD void decide(float4 nx, float4 ny, float4 nz, float4 fx, float4 fy, float4 fz, uint4 rf, const Slab& sl, float tmin, float tmax,
                  Visit& v) {
  // an empty slot (lo +inf, hi -inf) is never hit: its entry is +inf and its exit -inf whatever the direction
  const BoxHit h0 = slab_hit<FULL>(sl, nx.x, ny.x, nz.x, fx.x, fy.x, fz.x, tmin, tmax);
  const BoxHit h1 = slab_hit<FULL>(sl, nx.y, ny.y, nz.y, fx.y, fy.y, fz.y, tmin, tmax);
  const BoxHit h2 = slab_hit<FULL>(sl, nx.z, ny.z, nz.z, fx.z, fy.z, fz.z, tmin, tmax);
  const BoxHit h3 = slab_hit<FULL>(sl, nx.w, ny.w, nz.w, fx.w, fy.w, fz.w, tmin, tmax);
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

// record n, its planes read by the ray's own signs (any lane its own direction). OFF32: fewer than 2^25 records, so a
// plane's byte offset n * 128 + 16 k fits 32 bits: each load is the records' base plus one 32-bit offset (the same
// addresses; no 64-bit address arithmetic per plane)
template <bool FULL, bool OFF32 = false>
DL_HD void visit(const float4* __restrict__ wide, uint32_t n, const Slab& sl, float tmin, float tmax, Visit& v) {
  if (OFF32) {
    const char* b = reinterpret_cast<const char*>(wide);
    const uint32_t o = n * 128u;
    const auto at = [&](int k) { return *reinterpret_cast<const float4*>(b + (o + uint32_t(k) * 16u)); };
    decide<FULL>(at(sl.nOff[0]), at(sl.nOff[1]), at(sl.nOff[2]), at(3 - sl.nOff[0]), at(5 - sl.nOff[1]), at(7 - sl.nOff[2]),
                 *reinterpret_cast<const uint4*>(b + (o + 96u)), sl, tmin, tmax, v);
    return;
  }
  const float4* q = wide + 8 * size_t(n);
  decide<FULL>(q[sl.nOff[0]], q[sl.nOff[1]], q[sl.nOff[2]], q[3 - sl.nOff[0]], q[5 - sl.nOff[1]], q[7 - sl.nOff[2]],
               *reinterpret_cast<const uint4*>(q + 6), sl, tmin, tmax, v);   // axis a's far planes: 2a + 3 - nOff[a]
}

// Stack entries a traversal of a tree of the given LBVH height can hold at once: records sit at even depths (at most
// ceil(height / 2) on a path) and a visit pushes at most 3.
inline int wide_stack(int height) { return 3 * ((height + 1) / 2) + 1; }

}  // namespace dlw
