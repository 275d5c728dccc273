// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The occluder acceleration structure on the GPU: a linear BVH (Karras 2012, "Maximizing parallelism in the construction
// of BVHs, octrees and k-d trees"). Replaces the OptiX GAS build of reference src/gpu.cu:386-457 (RT cores have no
// CDNA4 counterpart; the traversal kernels read this layout).
//   1. centroid ((v0 + v1) + v2) / 3 of every triangle and the centroid box (exact min / max by ordered-int atomics)
//   2. 63-bit Morton code (21 bits per axis of the position in the centroid box, x highest)
//   3. stable radix sort of (code, triangle index) pairs (hipCUB): leaves in (code, index) order
//   4. per internal node its children by the longest common prefix of neighbouring codes (ties of equal codes by index)
//   5. bottom-up boxes: a thread per leaf climbs, the second thread to reach a node writes it (exact, so any order)
// Layout (HBM): node n = 16 floats: left child box lo xyz hi xyz, right child box, left, right (child >= 2^31: leaf
// child & 0x7fffffff), 2 unused: one 64-byte line per node visit. Leaf k = 12 floats: v0, v1, v2 of the k-th triangle
// in sorted order, its original index (as bits), its global slot (as bits, 0 until dlg_set_tri_slots), 1 unused: three
// 16-byte loads.
// The same arrays come out of the host build in the oracle, bit for bit (tests compare them).
//   6. the 4-wide records the traversals walk (wide.h): one per inner node, from the node and its inner children
// By default steps 4-5 are replaced by the binned-SAH tree of sah.hip over the sorted leaves (same layout); the LBVH
// stays for dlg_set_tree(DLG_TREE_LBVH) and for SAH trees deeper than the traversal stacks take.
#include <hipcub/hipcub.hpp>

#include "gpu_internal.h"
#include "wide.h"

namespace {

__device__ __forceinline__ uint32_t ordered(float f) {   // float -> uint with the same order
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

__global__ void k_centroids(const float* __restrict__ V, const uint32_t* __restrict__ I, int64_t nt, float* __restrict__ cen,
                            uint32_t* __restrict__ box) {
  __shared__ uint32_t sb[6];
  if (threadIdx.x < 6) sb[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < nt) {
    const uint32_t a = I[3 * t], b = I[3 * t + 1], c = I[3 * t + 2];
    for (int k = 0; k < 3; ++k) {
      const float x = ((V[3 * a + k] + V[3 * b + k]) + V[3 * c + k]) * (1.f / 3.f);
      cen[3 * t + k] = x;
      atomicMin(&sb[k], ordered(x));
      atomicMax(&sb[3 + k], ordered(x));
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicMin(&box[threadIdx.x], sb[threadIdx.x]);
  else if (threadIdx.x < 6) atomicMax(&box[threadIdx.x], sb[threadIdx.x]);
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

__global__ void k_morton(const float* __restrict__ cen, int64_t nt, const uint32_t* __restrict__ box, uint64_t* __restrict__ keys,
                         uint32_t* __restrict__ idx) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float lo = unordered(box[k]), hi = unordered(box[3 + k]);
    const float e = hi - lo, sc = e > 0.f ? 1.f / e : 0.f;
    const float u = (cen[3 * t + k] - lo) * sc;
    q[k] = (uint32_t)fminf(u * 2097152.f, 2097151.f);
  }
  keys[t] = spread21(q[0]) << 2 | spread21(q[1]) << 1 | spread21(q[2]);
  idx[t] = uint32_t(t);
}

__global__ void k_gather(const float* __restrict__ V, const uint32_t* __restrict__ I, const uint32_t* __restrict__ order, int64_t nt,
                         float* __restrict__ leaves) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  const uint32_t t = order[k];
  float* o = leaves + 12 * k;
  for (int c = 0; c < 3; ++c) {
    const uint32_t v = I[3 * t + c];
    o[3 * c] = V[3 * v]; o[3 * c + 1] = V[3 * v + 1]; o[3 * c + 2] = V[3 * v + 2];
  }
  o[9] = __uint_as_float(t);
  o[10] = 0.f;
  o[11] = 0.f;
}

__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz(uint32_t(i) ^ uint32_t(j));
  return __clzll(a ^ b);
}

__global__ void k_topology(const uint64_t* __restrict__ k, int64_t n, uint32_t* __restrict__ L, uint32_t* __restrict__ R,
                           uint32_t* __restrict__ parentLeaf, uint32_t* __restrict__ parentInner) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = delta(k, n, i, i + 1) - delta(k, n, i, i - 1) > 0 ? 1 : -1;
  const int dmin = delta(k, n, i, i - d);
  int64_t lmax = 2;
  while (delta(k, n, i, i + lmax * d) > dmin) lmax *= 2;
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2)
    if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta(k, n, i, j);
  int64_t s = 0;
  for (int64_t div = 2;; div *= 2) {
    const int64_t t = (l + div - 1) / div;
    if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int64_t g = i + s * d + (d < 0 ? -1 : 0);
  const int64_t mn = i < j ? i : j, mx = i < j ? j : i;
  const bool leafL = mn == g, leafR = mx == g + 1;
  L[i] = leafL ? 0x80000000u | uint32_t(g) : uint32_t(g);
  R[i] = leafR ? 0x80000000u | uint32_t(g + 1) : uint32_t(g + 1);
  if (leafL) parentLeaf[g] = uint32_t(i); else parentInner[g] = uint32_t(i);
  if (leafR) parentLeaf[g + 1] = uint32_t(i); else parentInner[g + 1] = uint32_t(i);
}

__device__ __forceinline__ void child_box(uint32_t c, const float* __restrict__ leaves, const float* innerBox, float b[6]) {
  if (c & 0x80000000u) {
    const float* t = leaves + 12 * (c & 0x7fffffffu);
    for (int a = 0; a < 3; ++a) {
      b[a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
      b[3 + a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
    }
  } else {
    const volatile float* s = innerBox + 6 * c;
    for (int a = 0; a < 6; ++a) b[a] = s[a];
  }
}

// a thread per leaf climbs; the second arrival at a node writes it. height: 1 + the higher child's (leaves 0).
__global__ void k_refit(int64_t n, const uint32_t* __restrict__ L, const uint32_t* __restrict__ R, const uint32_t* __restrict__ parentLeaf,
                        const uint32_t* __restrict__ parentInner, const float* __restrict__ leaves, float* innerBox, int* height,
                        uint32_t* flags, float* __restrict__ nodes) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t p = n == 1 ? 0u : parentLeaf[k];
  for (;;) {
    __threadfence();
    if (n > 1 && atomicAdd(&flags[p], 1u) == 0) return;   // the sibling's subtree is not done yet
    __threadfence();
    float bl[6], br[6];
    child_box(L[p], leaves, innerBox, bl);
    child_box(R[p], leaves, innerBox, br);
    float* nd = nodes + 16 * size_t(p);
    for (int a = 0; a < 6; ++a) { nd[a] = bl[a]; nd[6 + a] = br[a]; }
    nd[12] = __uint_as_float(L[p]);
    nd[13] = __uint_as_float(R[p]);
    nd[14] = 0.f;
    nd[15] = 0.f;
    volatile float* ib = innerBox + 6 * size_t(p);
    for (int a = 0; a < 3; ++a) { ib[a] = fminf(bl[a], br[a]); ib[3 + a] = fmaxf(bl[3 + a], br[3 + a]); }
    const int hl = (L[p] & 0x80000000u) ? 0 : ((volatile int*)height)[L[p]];
    const int hr = (R[p] & 0x80000000u) ? 0 : ((volatile int*)height)[R[p]];
    ((volatile int*)height)[p] = 1 + (hl > hr ? hl : hr);
    if (p == 0) return;
    p = parentInner[p];
  }
}

// record n of the 4-wide tree (wide.h): n's children, an inner child replaced by its two children; a grandchild whose
// children are two adjacent leaves becomes a pair slot (both triangles tested from here, no record visit of its own)
__global__ void k_wide(int64_t nn, const float* __restrict__ nodes, float4* __restrict__ wide) {
  const int64_t n = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const float inf = __int_as_float(0x7f800000);
  float lo[3][4], hi[3][4];
  uint32_t rf[4];
  for (int k = 0; k < 4; ++k) {
    for (int a = 0; a < 3; ++a) { lo[a][k] = inf; hi[a][k] = -inf; }
    rf[k] = dlw::kEmpty;
  }
  const float* nd = nodes + 16 * n;
  int k = 0;
  for (int side = 0; side < 2; ++side) {
    const uint32_t c = __float_as_uint(nd[12 + side]);
    if (c & 0x80000000u) {
      for (int a = 0; a < 3; ++a) { lo[a][k] = nd[6 * side + a]; hi[a][k] = nd[6 * side + 3 + a]; }
      rf[k++] = c;
    } else {
      const float* cn = nodes + 16 * size_t(c);
      for (int s2 = 0; s2 < 2; ++s2) {
        for (int a = 0; a < 3; ++a) { lo[a][k] = cn[6 * s2 + a]; hi[a][k] = cn[6 * s2 + 3 + a]; }
        uint32_t g = __float_as_uint(cn[12 + s2]);
        if (!(g & 0x80000000u)) {   // an inner grandchild: a pair slot when its children are leaves l, l + 1 (l < 2^30 - 1)
          const float* gn = nodes + 16 * size_t(g);
          const uint32_t a0 = __float_as_uint(gn[12]), a1 = __float_as_uint(gn[13]);
          if ((a0 & a1 & 0x80000000u) && (a0 & 0x7fffffffu) < dlw::kPair - 1 && (a1 & 0x7fffffffu) == (a0 & 0x7fffffffu) + 1)
            g = 0x80000000u | dlw::kPair | (a0 & 0x7fffffffu);
        }
        rf[k++] = g;
      }
    }
  }
  float4* q = wide + 8 * n;
  for (int a = 0; a < 3; ++a) {
    q[a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
    q[3 + a] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
  }
  q[6] = make_float4(__uint_as_float(rf[0]), __uint_as_float(rf[1]), __uint_as_float(rf[2]), __uint_as_float(rf[3]));
  q[7] = make_float4(0.f, 0.f, 0.f, 0.f);
}

inline unsigned blocks(int64_t n, int b) { return unsigned((n + b - 1) / b); }

}  // namespace

extern "C" int64_t dlg_build_bytes(int64_t nv, int64_t nt) {
  if (nt <= 0) return 0;
  const int64_t in = nv * 12 + nt * 12, keys = nt * (8 + 4) * 2, cen = nt * 12, tree = nt * (4 * 4 + 24 + 4 + 4);
  const int64_t out = (nt > 1 ? nt - 1 : 1) * (64 + 128) + nt * 48;
  // + the sort's temp storage (a few MB) + the SAH build's temporaries (the larger tree build counted for both)
  return in + keys + cen + tree + out + dlg_sah_bytes(nt) + (int64_t(64) << 20);
}

int32_t dlg_build_device(dlg_gpu* g, const float* dV, int64_t nv, const uint32_t* dI, int64_t nt, float mx, char* err, int32_t errLen) {
  (void)nv;
  dlg_scene& s = g->scene;
  s.release();
  s.nNodes = s.nTris = 0;
  s.height = 0;
  s.slotMax = 0;
  s.pad = (mx + 1.f) * 3.81469727e-6f;   // 2^-18 (|coordinate| + 1): box slack beyond the triangle test's rounding
  if (nt <= 0) return 0;
  dlg_tbegin(g);
  DevBuf cen, box, keys, keys2, idx, idx2, tmp, L, R, pl, pi, ibox, height, flags;
  const int64_t nn = nt > 1 ? nt - 1 : 1;
  const auto cleanup = [&] { for (DevBuf* b : {&cen, &box, &keys, &keys2, &idx, &idx2, &tmp, &L, &R, &pl, &pi, &ibox, &height, &flags}) b->release(); };
  const auto fail = [&](hipError_t e, const char* what) {
    dlg_fail(err, errLen, std::string("dlg_build: ") + what + ": " + hipGetErrorString(e));
    cleanup();
    return 1;
  };
  hipError_t e;
#define TRY(x, what) if ((e = (x)) != hipSuccess) return fail(e, what)
  TRY(cen.grow(size_t(nt) * 12), "centroids");
  TRY(box.grow(24), "box");
  const uint32_t init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  TRY(hipMemcpyAsync(box.p, init, 24, hipMemcpyHostToDevice, g->stream), "box init");
  k_centroids<<<blocks(nt, 256), 256, 0, g->stream>>>(dV, dI, nt, cen.as<float>(), box.as<uint32_t>());
  TRY(keys.grow(size_t(nt) * 8), "keys"); TRY(keys2.grow(size_t(nt) * 8), "keys");
  TRY(idx.grow(size_t(nt) * 4), "index"); TRY(idx2.grow(size_t(nt) * 4), "index");
  k_morton<<<blocks(nt, 256), 256, 0, g->stream>>>(cen.as<float>(), nt, box.as<uint32_t>(), keys.as<uint64_t>(), idx.as<uint32_t>());
  size_t tb = 0;
  TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(), idx.as<uint32_t>(), idx2.as<uint32_t>(),
                                         int(nt), 0, 64, g->stream), "sort size");
  TRY(tmp.grow(tb), "sort temp");
  TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.as<uint64_t>(), keys2.as<uint64_t>(), idx.as<uint32_t>(), idx2.as<uint32_t>(),
                                         int(nt), 0, 64, g->stream), "sort");
  TRY(s.leaves.grow(size_t(nt) * 48), "leaves");
  k_gather<<<blocks(nt, 256), 256, 0, g->stream>>>(dV, dI, idx2.as<uint32_t>(), nt, s.leaves.as<float>());
  if (g->tree == DLG_TREE_SAH && nt >= 2) {
    // the binned-SAH tree over the Morton order (sah.hip); a tree deeper than the traversal stacks take (status 2,
    // the leaves untouched) falls through to the LBVH
    TRY(s.nodes.grow(size_t(nn) * 64), "nodes");
    int32_t h = 0;
    const int32_t rc = dlg_sah_tree(g, nt, s.leaves.as<float>(), s.nodes.as<float>(), h, err, errLen);
    if (rc == 1) { cleanup(); return 1; }
    if (rc == 0) {
      TRY(s.wide.grow(size_t(nn) * 128), "wide records");
      k_wide<<<blocks(nn, 256), 256, 0, g->stream>>>(nn, s.nodes.as<float>(), s.wide.as<float4>());
      TRY(hipGetLastError(), "kernels");
      TRY(hipStreamSynchronize(g->stream), "build");
      s.nNodes = nn;
      s.nTris = nt;
      s.height = h;
      g->stats.buildMs += dlg_tend(g);
      cleanup();
      return 0;
    }
  }
  TRY(L.grow(size_t(nn) * 4), "tree"); TRY(R.grow(size_t(nn) * 4), "tree");
  TRY(pl.grow(size_t(nt) * 4), "tree"); TRY(pi.grow(size_t(nn) * 4), "tree");
  if (nt == 1) {
    const uint32_t leaf = 0x80000000u;
    TRY(hipMemcpyAsync(L.p, &leaf, 4, hipMemcpyHostToDevice, g->stream), "tree");
    TRY(hipMemcpyAsync(R.p, &leaf, 4, hipMemcpyHostToDevice, g->stream), "tree");
    TRY(hipStreamSynchronize(g->stream), "tree");
  } else {
    k_topology<<<blocks(nt - 1, 256), 256, 0, g->stream>>>(keys2.as<uint64_t>(), nt, L.as<uint32_t>(), R.as<uint32_t>(), pl.as<uint32_t>(),
                                                            pi.as<uint32_t>());
  }
  TRY(ibox.grow(size_t(nn) * 24), "boxes");
  TRY(height.grow(size_t(nn) * 4), "heights");
  TRY(flags.grow(size_t(nn) * 4), "flags");
  TRY(hipMemsetAsync(flags.p, 0, size_t(nn) * 4, g->stream), "flags");
  TRY(s.nodes.grow(size_t(nn) * 64), "nodes");
  k_refit<<<blocks(nt, 256), 256, 0, g->stream>>>(nt, L.as<uint32_t>(), R.as<uint32_t>(), pl.as<uint32_t>(), pi.as<uint32_t>(),
                                                  s.leaves.as<float>(), ibox.as<float>(), height.as<int>(), flags.as<uint32_t>(),
                                                  s.nodes.as<float>());
  TRY(s.wide.grow(size_t(nn) * 128), "wide records");
  k_wide<<<blocks(nn, 256), 256, 0, g->stream>>>(nn, s.nodes.as<float>(), s.wide.as<float4>());
  TRY(hipGetLastError(), "kernels");
  int h = 0;
  TRY(hipMemcpyAsync(&h, height.p, 4, hipMemcpyDeviceToHost, g->stream), "height");
  TRY(hipStreamSynchronize(g->stream), "build");
#undef TRY
  s.nNodes = nn;
  s.nTris = nt;
  s.height = h;
  g->stats.buildMs += dlg_tend(g);
  cleanup();
  return 0;
}

extern "C" int32_t dlg_build(dlg_gpu* g, const float* V, int64_t nv, const uint32_t* I, int64_t nt, char* err, int32_t errLen) {
  float mx = 0.f;
  for (int64_t i = 0; i < 3 * nv; ++i) mx = fmaxf(mx, fabsf(V[i]));
  if (nt <= 0) return dlg_build_device(g, nullptr, nv, nullptr, 0, mx, err, errLen);
  g->scene.release();   // the old scene's room first: the upload and the build share it
  DevBuf dV, dI;
  hipError_t e = dV.grow(size_t(nv) * 12);
  if (e == hipSuccess) e = dI.grow(size_t(nt) * 12);
  if (e == hipSuccess) e = hipMemcpyAsync(dV.p, V, size_t(nv) * 12, hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dI.p, I, size_t(nt) * 12, hipMemcpyHostToDevice, g->stream);
  if (e != hipSuccess) {
    dlg_fail(err, errLen, std::string("dlg_build: upload: ") + hipGetErrorString(e));
    g->scene.release();
    g->scene.nNodes = g->scene.nTris = 0;
    return 1;
  }
  const int32_t rc = dlg_build_device(g, dV.as<float>(), nv, dI.as<uint32_t>(), nt, mx, err, errLen);
  dV.release();
  dI.release();
  return rc;
}
