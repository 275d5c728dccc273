// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The traversal stack of a lane in the one-wave walks (trace.hip, view.hip, bounce.hip), STACK entries: [0, KL) in LDS
// (lds[depth][lane], no bank conflicts), [KL, STACK) in the wave's area of a global buffer (ovf[depth - KL][lane]),
// which only deep trees reach. A lane reads back only what it wrote itself, so no wave waits on another. Keeping the
// LDS part short keeps a wave's LDS small, so the waves per SIMD are what the registers allow (the walks wait on
// dependent fetches, and more waves hide them).
#pragma once
#include <stdint.h>

namespace dls {

constexpr int kWave = 64;
constexpr int kLdsStack = 24;
// LDS-typed pointers: the stack's LDS accesses stay ds_read / ds_write (through a generic pointer the compiler may
// merge an LDS and a global access of the two paths into one flat access, which takes the vector-memory path)
typedef __attribute__((address_space(3))) uint32_t lds_u32;

template <int STACK, int KL = (STACK < kLdsStack ? STACK : kLdsStack)>
struct LaneStack {
  static constexpr int kL = KL;
  static constexpr bool kDeep = STACK > kL;
  lds_u32* lds;    // [kL][64]
  uint32_t* ovf;   // [STACK - kL][64] of this wave (unused unless kDeep)
  __device__ __forceinline__ void put(int k, int lane, uint32_t v) const {
    if (!kDeep || k < kL) lds[k * kWave + lane] = v;
    else ovf[(k - kL) * kWave + lane] = v;
  }
  // three pushes at sp + o1, sp + o2, sp + o3 (each offset 0..2); LDS alone while every lane is low
  __device__ __forceinline__ void put3(int sp, int o1, int o2, int o3, int lane, uint32_t a, uint32_t b, uint32_t c) const {
    if (!kDeep || __ballot(sp + 2 >= kL) == 0) {
      lds[(sp + o1) * kWave + lane] = a;
      lds[(sp + o2) * kWave + lane] = b;
      lds[(sp + o3) * kWave + lane] = c;
    } else {
      put(sp + o1, lane, a);
      put(sp + o2, lane, b);
      put(sp + o3, lane, c);
    }
  }
  // the three pushes of an any-hit visit, at sp, sp + h1, sp + h1 + h2
  __device__ __forceinline__ void push3(int sp, int h1, int h2, int lane, uint32_t a, uint32_t b, uint32_t c) const {
    put3(sp, 0, h1, h1 + h2, lane, a, b, c);
  }
  __device__ __forceinline__ uint32_t at(int k, int lane) const {
    if (!kDeep || __ballot(k >= kL) == 0) return lds[k * kWave + lane];
    uint32_t v;
    if (k < kL) v = lds[k * kWave + lane];
    else v = ovf[(k - kL) * kWave + lane];
    return v;
  }
};
// the LDS and global entries a lane's stack of STACK entries takes (the default split)
template <int STACK>
constexpr int lds_stack() { return LaneStack<STACK>::kL; }
template <int STACK>
constexpr int ovf_stack() { return STACK - LaneStack<STACK>::kL; }

}  // namespace dls
