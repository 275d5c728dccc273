// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The default acceleration structure: a binned-SAH tree (16 centroid bins per axis, surface-area heuristic down to
// single-triangle leaves), built top-down on the GPU over the Morton-ordered leaves of lbvh.hip and stored in the LBVH's
// node layout, so that the 4-wide records (wide.h) and every traversal are shared. Replaces, as lbvh.hip does, the OptiX
// GAS build of reference src/gpu.cu:386-457.
//
// Why: the LBVH splits by Morton bits, blind to triangle sizes. In scenes of large overlapping triangles (extruded boxes
// over a fine sensor grid) its node boxes overlap deeply: a ray visits ~2-3x the records it visits in an SAH tree, and
// neighbouring rays end after very different walks, which costs SIMD efficiency on top.
//
// The rules (the oracle's host build restates them and must give the same arrays bit for bit):
//   - a triangle's box: ordered min / max of its vertices; its centroid c = (lo + hi) * 0.5 per axis
//   - a range's bins, per axis a with centroid bounds [l, h]: ext = h - l, scale = 16 / ext; the axis is used when
//     ext > 0 and scale is finite; bin = min(15, int((c - l) * scale)); a bin holds its triangles' box and count
//   - candidate (a, s), s = 0..14: bins 0..s left, s+1..15 right, valid when both sides hold triangles; cost
//     A(L) n(L) + A(R) n(R) with A the half area (dx dy + dy dz) + dz dx; the least cost wins, ties to the lower a * 15 + s
//   - no valid candidate: the range is halved (left = the first n / 2)
//   - the partition is stable (a range keeps its order, so an unsplittable range halves in Morton order)
//   - node ids: an inner node over leaf positions [b, e) split at m has id m - 1 (the n - 1 split positions of a binary
//     tree over n ordered leaves are distinct: no allocation, no dependence on the order blocks run in); the root's id
//     and id 0 swap places, since the traversals start at node 0
//   - leaf k of the output is the k-th triangle of the final order
// Min / max (ordered-int keys) and counts commute, and every cost is evaluated by one lane in a fixed order, so the tree
// does not depend on scheduling.
//
// The build runs level by level over three sizes of task (a range still to split):
//   big (> 4096 triangles): chunks of 4096, a block each - bin (LDS, merged into the task's bins by global atomics),
//       choose (a wave per task), count (per chunk), offsets (a wave per task), scatter (per chunk: stable partition
//       into the other buffer, child boxes and centroid bounds), emit (a thread per task)
//   medium (65..4096): a block per task does all of that in one kernel
//   small (2..64): after the levels, a 64-lane block per task builds the whole subtree in LDS
// emit writes the node, its parent's reference to it and the child tasks (leaves are written as references at once).
#include <algorithm>
#include <vector>

#include "gpu_internal.h"

namespace {

constexpr int kBins = 16, kPlanes = kBins - 1, kCand = 3 * kPlanes;
constexpr int kBinWords = 3 * kBins * 7;   // per axis and bin: lo x y z, hi x y z (ordered keys), count
constexpr uint32_t kChunk = 4096, kSmall = 64, kRootSlot = 0xffffffffu;
constexpr int kBlock = 256;
constexpr int kMaxHeight = 100;   // the traversal stacks hold wide_stack(100) + 3 = 154 <= 160 entries

struct Task {
  uint32_t b, e;        // leaf positions [b, e)
  uint32_t slot;        // the parent's id << 1 | side, or kRootSlot
  uint32_t depth;       // root 1
  uint32_t cb[6];       // centroid bounds: lo x y z, hi x y z (ordered keys)
  uint32_t chunk0, nChunk;
  uint32_t buf;         // the reference buffer holding the range (0 / 1)
  uint32_t pad;
};

struct Split { int32_t axis, plane; uint32_t nL, pad; };   // axis 3: halve

// the build's outputs and task lists for the kernels of one level
struct Lists {
  float* nodes;
  uint32_t* finalRef;
  uint32_t* rootId;
  int* height;
  uint32_t* ref[2];
  Task* big; uint32_t* nBig;           // next level
  uint32_t* chunkTask; uint32_t* nChunk;
  Task* med; uint32_t* nMed;
  Task* small; uint32_t* nSmall;
};

__device__ __forceinline__ uint32_t okey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ofl(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }
__device__ __forceinline__ float omin(float a, float b) { return okey(b) < okey(a) ? b : a; }
__device__ __forceinline__ float omax(float a, float b) { return okey(b) > okey(a) ? b : a; }

__device__ __forceinline__ uint32_t map_id(uint32_t id, uint32_t root) { return id == root ? 0u : id == 0u ? root : id; }

// per axis: lower bound and scale, or !use
struct Binning {
  float lo[3], scale[3];
  bool use[3];
  __device__ void set(const uint32_t cb[6]) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = ofl(cb[a]);
      const float ext = ofl(cb[3 + a]) - lo[a];
      scale[a] = 16.f / ext;
      use[a] = ext > 0.f && scale[a] < __int_as_float(0x7f800000);
    }
  }
  __device__ int bin(float c, int a) const {
    const int q = int((c - lo[a]) * scale[a]);
    return q < kBins - 1 ? q : kBins - 1;
  }
};

__device__ __forceinline__ float half_area(const uint32_t k[6]) {
  const float dx = ofl(k[3]) - ofl(k[0]), dy = ofl(k[4]) - ofl(k[1]), dz = ofl(k[5]) - ofl(k[2]);
  return (dx * dy + dy * dz) + dz * dx;
}

__device__ __forceinline__ void init_bins(uint32_t* sb, int tid, int nthreads) {
  for (int w = tid; w < kBinWords; w += nthreads) sb[w] = (w % 7) < 3 ? 0xffffffffu : 0u;
}

// one triangle into the LDS bins of each used axis
__device__ __forceinline__ void bin_one(uint32_t* sb, const Binning& bn, float4 l, float4 h) {
  const float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
  const uint32_t kl[3] = {okey(l.x), okey(l.y), okey(l.z)}, kh[3] = {okey(h.x), okey(h.y), okey(h.z)};
  for (int a = 0; a < 3; ++a) {
    if (!bn.use[a]) continue;
    uint32_t* B = sb + (a * kBins + bn.bin(c[a], a)) * 7;
    atomicMin(B + 0, kl[0]); atomicMin(B + 1, kl[1]); atomicMin(B + 2, kl[2]);
    atomicMax(B + 3, kh[0]); atomicMax(B + 4, kh[1]); atomicMax(B + 5, kh[2]);
    atomicAdd(B + 6, 1u);
  }
}

// The best split of a range of n triangles from its bins, by one wave (lane < 64): every lane returns the same Split.
__device__ Split choose(const uint32_t* bins, const Binning& bn, uint32_t n, int lane) {
  unsigned long long key = ~0ull;
  if (lane < kCand) {
    const int a = lane / kPlanes, s = lane % kPlanes;
    if (bn.use[a]) {
      uint32_t L[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u}, R[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
      uint32_t nl = 0, nr = 0;
      for (int q = 0; q < kBins; ++q) {
        const uint32_t* B = bins + (a * kBins + q) * 7;
        uint32_t* X = q <= s ? L : R;
        for (int k = 0; k < 3; ++k) { X[k] = min(X[k], B[k]); X[3 + k] = max(X[3 + k], B[3 + k]); }
        (q <= s ? nl : nr) += B[6];
      }
      if (nl > 0 && nr > 0) {
        const float cost = half_area(L) * float(nl) + half_area(R) * float(nr);
        key = (static_cast<unsigned long long>(__float_as_uint(cost)) << 32) | static_cast<unsigned long long>(lane);
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_xor(key, o, 64);
    key = other < key ? other : key;
  }
  Split sp{3, 0, n / 2, 0};
  if (key != ~0ull) {
    const int idx = int(key & 0xffffffffull), a = idx / kPlanes, s = idx % kPlanes;
    uint32_t nl = 0;
    for (int q = 0; q <= s; ++q) nl += bins[(a * kBins + q) * 7 + 6];
    sp = Split{a, s, nl, 0};
  }
  return sp;
}

__device__ __forceinline__ bool goes_left(const Split& sp, const Binning& bn, float4 l, float4 h, uint32_t posInTask) {
  if (sp.axis == 3) return posInTask < sp.nL;
  const float lo = sp.axis == 0 ? l.x : sp.axis == 1 ? l.y : l.z, hi = sp.axis == 0 ? h.x : sp.axis == 1 ? h.y : h.z;
  return bn.bin((lo + hi) * 0.5f, sp.axis) <= sp.plane;
}

// per side: box keys [6] and centroid-bound keys [6] accumulated by one thread
struct SideAcc {
  uint32_t k[2][12];
  __device__ void init() { for (int s = 0; s < 2; ++s) for (int j = 0; j < 12; ++j) k[s][j] = (j % 6) < 3 ? 0xffffffffu : 0u; }
  __device__ void add(int s, float4 l, float4 h) {
    const float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
    const float lo[3] = {l.x, l.y, l.z}, hi[3] = {h.x, h.y, h.z};
    for (int a = 0; a < 3; ++a) {
      k[s][a] = min(k[s][a], okey(lo[a])); k[s][3 + a] = max(k[s][3 + a], okey(hi[a]));
      k[s][6 + a] = min(k[s][6 + a], okey(c[a])); k[s][9 + a] = max(k[s][9 + a], okey(c[a]));
    }
  }
  // into 24 shared / global words (min for lo halves, max for hi halves)
  __device__ void merge(uint32_t* w) const {
    for (int s = 0; s < 2; ++s)
      for (int j = 0; j < 12; ++j) {
        const uint32_t v = k[s][j];
        if ((j % 6) < 3) { if (v != 0xffffffffu) atomicMin(w + 12 * s + j, v); }
        else if (v != 0u) atomicMax(w + 12 * s + j, v);
      }
  }
};
__device__ __forceinline__ void init_acc(uint32_t* w, int tid, int nthreads) {
  for (int j = tid; j < 24; j += nthreads) w[j] = (j % 6) < 3 ? 0xffffffffu : 0u;
}

// The node of task T (split nL), its parent's reference, its children: leaves at once, the rest as tasks. acc: 24 words
// (per side: box lo/hi, centroid lo/hi). dst holds the task's partitioned references (buffer dstBuf).
__device__ void emit(const Task& T, uint32_t nL, const uint32_t* acc, const uint32_t* dst, uint32_t dstBuf, const Lists& o) {
  const uint32_t id = T.b + nL - 1;
  uint32_t root;
  if (T.slot == kRootSlot) { root = id; *o.rootId = id; }
  else root = *o.rootId;
  const uint32_t slot = map_id(id, root);
  float* nd = o.nodes + 16 * size_t(slot);
  for (int s = 0; s < 2; ++s)
    for (int j = 0; j < 6; ++j) nd[6 * s + j] = ofl(acc[12 * s + j]);
  nd[14] = 0.f;
  nd[15] = 0.f;
  if (T.slot != kRootSlot) o.nodes[16 * size_t(map_id(T.slot >> 1, root)) + 12 + (T.slot & 1u)] = __uint_as_float(slot);
  atomicMax(o.height, int(T.depth));
  for (int s = 0; s < 2; ++s) {
    const uint32_t b = s ? T.b + nL : T.b, e = s ? T.e : T.b + nL, n = e - b;
    if (n == 1) {
      nd[12 + s] = __uint_as_float(0x80000000u | b);
      o.finalRef[b] = dst[b];
      continue;
    }
    Task C;
    C.b = b; C.e = e; C.slot = id << 1 | uint32_t(s); C.depth = T.depth + 1;
    for (int j = 0; j < 6; ++j) C.cb[j] = acc[12 * s + 6 + j];
    C.chunk0 = 0; C.nChunk = 1; C.buf = dstBuf; C.pad = 0;
    nd[12 + s] = 0.f;   // the child's emit writes its id here
    if (n <= kSmall) o.small[atomicAdd(o.nSmall, 1u)] = C;
    else if (n <= kChunk) o.med[atomicAdd(o.nMed, 1u)] = C;
    else {
      C.nChunk = (n + kChunk - 1) / kChunk;
      C.chunk0 = atomicAdd(o.nChunk, C.nChunk);
      const uint32_t t = atomicAdd(o.nBig, 1u);
      o.big[t] = C;
      for (uint32_t j = 0; j < C.nChunk; ++j) o.chunkTask[C.chunk0 + j] = t;
    }
  }
}

// ---------------------------------------------------------------------------------------------------- set-up

// per triangle (in Morton order): its box; the root's centroid bounds
__global__ void k_boxes(const float4* __restrict__ leaves, uint32_t nt, float4* __restrict__ pbox, uint32_t* __restrict__ rootCb) {
  __shared__ uint32_t s[6];
  if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? 0xffffffffu : 0u;
  __syncthreads();
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nt) {
    const float4 a = leaves[3 * k], b = leaves[3 * k + 1], c = leaves[3 * k + 2];
    const float4 l = make_float4(omin(omin(a.x, a.w), b.z), omin(omin(a.y, b.x), b.w), omin(omin(a.z, b.y), c.x), 0.f);
    const float4 h = make_float4(omax(omax(a.x, a.w), b.z), omax(omax(a.y, b.x), b.w), omax(omax(a.z, b.y), c.x), 0.f);
    pbox[2 * k] = l;
    pbox[2 * k + 1] = h;
    const float cc[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
    for (int a2 = 0; a2 < 3; ++a2) { atomicMin(&s[a2], okey(cc[a2])); atomicMax(&s[3 + a2], okey(cc[a2])); }
  }
  __syncthreads();
  if (threadIdx.x < 3) atomicMin(&rootCb[threadIdx.x], s[threadIdx.x]);
  else if (threadIdx.x < 6) atomicMax(&rootCb[threadIdx.x], s[threadIdx.x]);
}

__global__ void k_iota(uint32_t* __restrict__ r, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) r[k] = k;
}

// the root task into the list its size belongs to
__global__ void k_root(uint32_t nt, const uint32_t* __restrict__ rootCb, Lists o) {
  Task T;
  T.b = 0; T.e = nt; T.slot = kRootSlot; T.depth = 1;
  for (int j = 0; j < 6; ++j) T.cb[j] = rootCb[j];
  T.chunk0 = 0; T.nChunk = 1; T.buf = 0; T.pad = 0;
  if (nt <= kSmall) o.small[atomicAdd(o.nSmall, 1u)] = T;
  else if (nt <= kChunk) o.med[atomicAdd(o.nMed, 1u)] = T;
  else {
    T.nChunk = (nt + kChunk - 1) / kChunk;
    T.chunk0 = atomicAdd(o.nChunk, T.nChunk);
    const uint32_t t = atomicAdd(o.nBig, 1u);
    o.big[t] = T;
    for (uint32_t j = 0; j < T.nChunk; ++j) o.chunkTask[T.chunk0 + j] = t;
  }
}

// ---------------------------------------------------------------------------------------------------- big tasks

__global__ void k_big_init(uint32_t nTask, uint32_t* __restrict__ bins) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nTask * kBinWords) bins[w] = ((w % kBinWords) % 7) < 3 ? 0xffffffffu : 0u;
}

__device__ __forceinline__ void chunk_range(const Task& T, uint32_t c, uint32_t& k0, uint32_t& k1) {
  k0 = T.b + (c - T.chunk0) * kChunk;
  k1 = min(T.e, k0 + kChunk);
}

__global__ __launch_bounds__(kBlock) void k_big_bin(const Task* __restrict__ tasks, const uint32_t* __restrict__ chunkTask,
                                                    const uint32_t* __restrict__ ref0, const uint32_t* __restrict__ ref1,
                                                    const float4* __restrict__ pbox, uint32_t* __restrict__ bins) {
  __shared__ uint32_t sb[kBinWords];
  const uint32_t c = blockIdx.x, t = chunkTask[c];
  const Task T = tasks[t];
  const uint32_t* src = T.buf ? ref1 : ref0;
  Binning bn;
  bn.set(T.cb);
  init_bins(sb, threadIdx.x, kBlock);
  __syncthreads();
  uint32_t k0, k1;
  chunk_range(T, c, k0, k1);
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += kBlock) {
    const uint32_t p = src[k];
    bin_one(sb, bn, pbox[2 * p], pbox[2 * p + 1]);
  }
  __syncthreads();
  uint32_t* G = bins + size_t(t) * kBinWords;
  for (int w = threadIdx.x; w < kBinWords; w += kBlock) {
    const uint32_t v = sb[w];
    const int f = w % 7;
    if (f < 3) { if (v != 0xffffffffu) atomicMin(G + w, v); }
    else if (f < 6) { if (v != 0u) atomicMax(G + w, v); }
    else if (v) atomicAdd(G + w, v);
  }
}

// a wave per task: the split; the task's child accumulators cleared
__global__ __launch_bounds__(64) void k_big_choose(const Task* __restrict__ tasks, const uint32_t* __restrict__ bins, Split* __restrict__ split,
                                                   uint32_t* __restrict__ acc) {
  const uint32_t t = blockIdx.x;
  const Task T = tasks[t];
  Binning bn;
  bn.set(T.cb);
  const Split sp = choose(bins + size_t(t) * kBinWords, bn, T.e - T.b, int(threadIdx.x));
  if (threadIdx.x == 0) split[t] = sp;
  init_acc(acc + 24 * size_t(t), int(threadIdx.x), 64);
}

// per chunk: the triangles going left
__global__ __launch_bounds__(kBlock) void k_big_count(const Task* __restrict__ tasks, const uint32_t* __restrict__ chunkTask,
                                                      const uint32_t* __restrict__ ref0, const uint32_t* __restrict__ ref1,
                                                      const float4* __restrict__ pbox, const Split* __restrict__ split,
                                                      uint32_t* __restrict__ chunkLeft) {
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const uint32_t c = blockIdx.x, t = chunkTask[c];
  const Task T = tasks[t];
  const uint32_t* src = T.buf ? ref1 : ref0;
  const Split sp = split[t];
  Binning bn;
  bn.set(T.cb);
  uint32_t k0, k1;
  chunk_range(T, c, k0, k1);
  uint32_t mine = 0;
  for (uint32_t k = k0 + threadIdx.x; k < k1; k += kBlock) {
    const uint32_t p = src[k];
    mine += goes_left(sp, bn, pbox[2 * p], pbox[2 * p + 1], k - T.b) ? 1u : 0u;
  }
  for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) chunkLeft[c] = cnt;
}

// a wave per task: exclusive scan of its chunks' left counts (in place)
__global__ __launch_bounds__(64) void k_big_offsets(const Task* __restrict__ tasks, uint32_t* __restrict__ chunkLeft) {
  const Task T = tasks[blockIdx.x];
  uint32_t carry = 0;
  const uint32_t lane = threadIdx.x;
  for (uint32_t j0 = 0; j0 < T.nChunk; j0 += 64) {
    const uint32_t j = j0 + lane;
    const uint32_t v = j < T.nChunk ? chunkLeft[T.chunk0 + j] : 0u;
    uint32_t x = v;   // inclusive scan
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= uint32_t(o)) x += y;
    }
    if (j < T.nChunk) chunkLeft[T.chunk0 + j] = carry + x - v;
    carry += __shfl(x, 63, 64);
  }
}

// Stable partition of [k0, k1) of task T (positions relative to the task's b) by block-wide scans in rounds of kBlock:
// left ones to b + leftBase + (left ones before), right ones to b + nL + rightBase + (right ones before). Accumulates
// the children's boxes and centroid bounds in acc (the thread's own) as it goes.
__device__ void partition_range(const Task& T, const Split& sp, const Binning& bn, uint32_t k0, uint32_t k1, uint32_t leftBase,
                                uint32_t rightBase, const uint32_t* src, uint32_t* dst, const float4* pbox, SideAcc& acc,
                                uint32_t* sWave /* [kBlock / 64] */) {
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (uint32_t base = k0; base < k1; base += kBlock) {
    const uint32_t k = base + threadIdx.x;
    const bool valid = k < k1;
    uint32_t p = 0;
    bool left = false;
    float4 l = make_float4(0.f, 0.f, 0.f, 0.f), h = l;
    if (valid) {
      p = src[k];
      l = pbox[2 * p];
      h = pbox[2 * p + 1];
      left = goes_left(sp, bn, l, h, k - T.b);
      acc.add(left ? 0 : 1, l, h);
    }
    const unsigned long long bal = __ballot(left);
    const unsigned long long below = lane ? (bal & ((1ull << lane) - 1ull)) : 0ull;
    const uint32_t before = uint32_t(__popcll(below));
    __syncthreads();
    if (lane == 0) sWave[wv] = uint32_t(__popcll(bal));
    __syncthreads();
    uint32_t wavesBefore = 0, roundLeft = 0;
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
      if (w < wv) wavesBefore += sWave[w];
      roundLeft += sWave[w];
    }
    if (valid) {
      const uint32_t leftBefore = wavesBefore + before;
      const uint32_t idx = k - base;   // position within the round
      const uint32_t pos = left ? T.b + leftBase + leftBefore : T.b + sp.nL + rightBase + (idx - leftBefore);
      if (pos < T.e) dst[pos] = p;   // always (the counts agree with the bins); kept inside the task's range regardless
    }
    leftBase += roundLeft;
    rightBase += min(k1 - base, uint32_t(kBlock)) - roundLeft;
  }
}

__global__ __launch_bounds__(kBlock) void k_big_scatter(const Task* __restrict__ tasks, const uint32_t* __restrict__ chunkTask,
                                                        uint32_t* __restrict__ ref0, uint32_t* __restrict__ ref1,
                                                        const float4* __restrict__ pbox, const Split* __restrict__ split,
                                                        const uint32_t* __restrict__ chunkLeft, uint32_t* __restrict__ acc) {
  __shared__ uint32_t sWave[kBlock / 64];
  __shared__ uint32_t sAcc[24];
  const uint32_t c = blockIdx.x, t = chunkTask[c];
  const Task T = tasks[t];
  const uint32_t* src = T.buf ? ref1 : ref0;
  uint32_t* dst = T.buf ? ref0 : ref1;
  const Split sp = split[t];
  Binning bn;
  bn.set(T.cb);
  init_acc(sAcc, threadIdx.x, kBlock);
  uint32_t k0, k1;
  chunk_range(T, c, k0, k1);
  const uint32_t leftBase = chunkLeft[c], rightBase = (k0 - T.b) - leftBase;
  SideAcc a;
  a.init();
  partition_range(T, sp, bn, k0, k1, leftBase, rightBase, src, dst, pbox, a, sWave);
  a.merge(sAcc);
  __syncthreads();
  if (threadIdx.x < 24) {
    const uint32_t v = sAcc[threadIdx.x];
    if ((threadIdx.x % 6) < 3) atomicMin(acc + 24 * size_t(t) + threadIdx.x, v);
    else atomicMax(acc + 24 * size_t(t) + threadIdx.x, v);
  }
}

__global__ void k_big_emit(const Task* __restrict__ tasks, uint32_t nTask, const Split* __restrict__ split, const uint32_t* __restrict__ acc,
                           Lists o) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nTask) return;
  const Task T = tasks[t];
  emit(T, split[t].nL, acc + 24 * size_t(t), T.buf ? o.ref[0] : o.ref[1], T.buf ? 0u : 1u, o);
}

// ---------------------------------------------------------------------------------------------------- medium tasks

__global__ __launch_bounds__(kBlock) void k_medium(const Task* __restrict__ tasks, const float4* __restrict__ pbox, Lists o) {
  __shared__ uint32_t sb[kBinWords];
  __shared__ Split sSplit;
  __shared__ uint32_t sWave[kBlock / 64];
  __shared__ uint32_t sAcc[24];
  const Task T = tasks[blockIdx.x];
  const uint32_t* src = T.buf ? o.ref[1] : o.ref[0];
  uint32_t* dst = T.buf ? o.ref[0] : o.ref[1];
  Binning bn;
  bn.set(T.cb);
  init_bins(sb, threadIdx.x, kBlock);
  init_acc(sAcc, threadIdx.x, kBlock);
  __syncthreads();
  for (uint32_t k = T.b + threadIdx.x; k < T.e; k += kBlock) {
    const uint32_t p = src[k];
    bin_one(sb, bn, pbox[2 * p], pbox[2 * p + 1]);
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const Split sp = choose(sb, bn, T.e - T.b, int(threadIdx.x));
    if (threadIdx.x == 0) sSplit = sp;
  }
  __syncthreads();
  const Split sp = sSplit;
  SideAcc a;
  a.init();
  partition_range(T, sp, bn, T.b, T.e, 0, 0, src, dst, pbox, a, sWave);
  a.merge(sAcc);
  __syncthreads();
  if (threadIdx.x == 0) emit(T, sp.nL, sAcc, dst, T.buf ? 0u : 1u, o);
}

// ---------------------------------------------------------------------------------------------------- small tasks

// min / max of ordered keys over the wave
__device__ __forceinline__ uint32_t wmin(uint32_t v) { for (int o = 32; o > 0; o >>= 1) v = min(v, uint32_t(__shfl_xor(v, o, 64))); return v; }
__device__ __forceinline__ uint32_t wmax(uint32_t v) { for (int o = 32; o > 0; o >>= 1) v = max(v, uint32_t(__shfl_xor(v, o, 64))); return v; }

// a 64-lane block per task: the subtree of the range, depth first, one range at a time
__global__ __launch_bounds__(64) void k_small(const Task* __restrict__ tasks, const float4* __restrict__ pbox, Lists o) {
  __shared__ uint32_t sb[kBinWords];
  __shared__ uint32_t sRef[kSmall];
  __shared__ uint32_t sStack[kSmall][4];   // rb, re, slot, depth
  __shared__ int sSp;
  const uint32_t lane = threadIdx.x;
  const Task T = tasks[blockIdx.x];
  const uint32_t n = T.e - T.b;
  const uint32_t* src = T.buf ? o.ref[1] : o.ref[0];
  if (lane < n) sRef[lane] = src[T.b + lane];
  if (lane == 0) {
    sStack[0][0] = 0; sStack[0][1] = n; sStack[0][2] = T.slot; sStack[0][3] = T.depth;
    sSp = 1;
  }
  __syncthreads();
  uint32_t root = T.slot == kRootSlot ? 0u : *o.rootId;
  int maxDepth = 0;   // lane 0's deepest node: one atomic per task, not one per node on the one height word
  while (sSp > 0) {
    const int sp0 = sSp - 1;
    const uint32_t rb = sStack[sp0][0], re = sStack[sp0][1], slotP = sStack[sp0][2], depth = sStack[sp0][3];
    const uint32_t m = re - rb;
    const bool act = lane < m;
    const uint32_t p = act ? sRef[rb + lane] : 0u;
    const float4 l = act ? pbox[2 * p] : make_float4(0.f, 0.f, 0.f, 0.f), h = act ? pbox[2 * p + 1] : l;
    const float c[3] = {(l.x + h.x) * 0.5f, (l.y + h.y) * 0.5f, (l.z + h.z) * 0.5f};
    uint32_t cb[6];
    for (int a = 0; a < 3; ++a) {
      cb[a] = wmin(act ? okey(c[a]) : 0xffffffffu);
      cb[3 + a] = wmax(act ? okey(c[a]) : 0u);
    }
    Binning bn;
    bn.set(cb);
    init_bins(sb, int(lane), 64);
    __syncthreads();
    if (act) bin_one(sb, bn, l, h);
    __syncthreads();
    const Split sp = choose(sb, bn, m, int(lane));
    const bool left = act && goes_left(sp, bn, l, h, lane);
    const unsigned long long bal = __ballot(left), actMask = m >= 64 ? ~0ull : ((1ull << m) - 1ull);
    const unsigned long long below = lane ? ((1ull << lane) - 1ull) : 0ull;
    const uint32_t nL = uint32_t(__popcll(bal));
    const uint32_t pos = left ? uint32_t(__popcll(bal & below)) : nL + uint32_t(__popcll(~bal & actMask & below));
    // children's boxes and centroid bounds
    uint32_t acc[24];
    for (int s = 0; s < 2; ++s) {
      const bool mine = act && (left == (s == 0));
      const float lo[3] = {l.x, l.y, l.z}, hi[3] = {h.x, h.y, h.z};
      for (int a = 0; a < 3; ++a) {
        acc[12 * s + a] = wmin(mine ? okey(lo[a]) : 0xffffffffu);
        acc[12 * s + 3 + a] = wmax(mine ? okey(hi[a]) : 0u);
        acc[12 * s + 6 + a] = wmin(mine ? okey(c[a]) : 0xffffffffu);
        acc[12 * s + 9 + a] = wmax(mine ? okey(c[a]) : 0u);
      }
    }
    __syncthreads();   // every lane has read sRef and the stack entry
    if (act) sRef[rb + pos] = p;
    const uint32_t id = T.b + rb + nL - 1;
    if (slotP == kRootSlot) root = id;
    if (lane == 0) {
      if (slotP == kRootSlot) *o.rootId = id;
      const uint32_t slot = map_id(id, root);
      float* nd = o.nodes + 16 * size_t(slot);
      for (int s = 0; s < 2; ++s)
        for (int j = 0; j < 6; ++j) nd[6 * s + j] = ofl(acc[12 * s + j]);
      nd[14] = 0.f;
      nd[15] = 0.f;
      if (slotP != kRootSlot) o.nodes[16 * size_t(map_id(slotP >> 1, root)) + 12 + (slotP & 1u)] = __uint_as_float(slot);
      maxDepth = max(maxDepth, int(depth));
      int spn = sp0;
      for (int s = 1; s >= 0; --s) {   // right pushed first: the left range is taken next
        const uint32_t b = s ? rb + nL : rb, e = s ? re : rb + nL;
        if (e - b == 1) { nd[12 + s] = __uint_as_float(0x80000000u | (T.b + b)); continue; }
        nd[12 + s] = 0.f;
        sStack[spn][0] = b; sStack[spn][1] = e; sStack[spn][2] = id << 1 | uint32_t(s); sStack[spn][3] = depth + 1;
        ++spn;
      }
      sSp = spn;
    }
    __syncthreads();
  }
  if (lane < n) o.finalRef[T.b + lane] = sRef[lane];
  if (lane == 0) atomicMax(o.height, maxDepth);
}

// ---------------------------------------------------------------------------------------------------- output

__global__ void k_reorder(const float4* __restrict__ in, const uint32_t* __restrict__ finalRef, uint32_t nt, float4* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  const uint32_t p = finalRef[k];
  out[3 * k] = in[3 * p];
  out[3 * k + 1] = in[3 * p + 1];
  out[3 * k + 2] = in[3 * p + 2];
}

inline unsigned nblocks(uint64_t n, unsigned b) { return unsigned((n + b - 1) / b); }

}  // namespace

// Device bytes of the SAH build's temporaries beyond the LBVH build's (dlg_build_bytes adds them).
int64_t dlg_sah_bytes(int64_t nt) {
  if (nt < 2) return 0;
  const int64_t bigMax = nt / int64_t(kChunk) + 2, chunkMax = 2 * bigMax + nt / int64_t(kChunk);
  return nt * (32 + 3 * 4 + 48) + 2 * bigMax * int64_t(sizeof(Task)) + 2 * chunkMax * 4 + bigMax * (kBinWords * 4 + 16 + 96) +
         chunkMax * 4 + 2 * (nt / int64_t(kSmall + 1) + 2) * int64_t(sizeof(Task)) + (nt / 2 + 1) * int64_t(sizeof(Task));
}

// The SAH tree over nt >= 2 leaves in Morton order (leaves [nt][12] floats): nodes [nt - 1][16] written, the leaves
// reordered in place (through a temporary), the height returned. 2: the tree would be deeper than the traversal
// stacks take (height > kMaxHeight); the leaves are then left as they were, for the LBVH.
int32_t dlg_sah_tree(dlg_gpu* g, int64_t nt64, float* leaves, float* nodes, int32_t& height, char* err, int32_t errLen) {
  const uint32_t nt = uint32_t(nt64);
  const int64_t bigMax = nt64 / int64_t(kChunk) + 2, chunkMax = 2 * bigMax + nt64 / int64_t(kChunk);
  const int64_t medMax = nt64 / int64_t(kSmall + 1) + 2, smallMax = nt64 / 2 + 1;
  DevBuf pbox, ref0, ref1, fin, leavesTmp, big[2], chunkTask[2], med[2], small, bins, split, acc, chunkLeft, ctr;
  const auto cleanup = [&] {
    for (DevBuf* b : {&pbox, &ref0, &ref1, &fin, &leavesTmp, &big[0], &big[1], &chunkTask[0], &chunkTask[1], &med[0], &med[1], &small, &bins,
                      &split, &acc, &chunkLeft, &ctr})
      b->release();
  };
  hipError_t e;
#define TRY(x, what)                                                                                 \
  if ((e = (x)) != hipSuccess) {                                                                     \
    dlg_fail(err, errLen, std::string("dlg_build (SAH): ") + what + ": " + hipGetErrorString(e)); \
    cleanup();                                                                                       \
    return 1;                                                                                        \
  }
  TRY(pbox.grow(size_t(nt) * 32), "boxes");
  TRY(ref0.grow(size_t(nt) * 4), "references"); TRY(ref1.grow(size_t(nt) * 4), "references"); TRY(fin.grow(size_t(nt) * 4), "references");
  for (int k = 0; k < 2; ++k) {
    TRY(big[k].grow(size_t(bigMax) * sizeof(Task)), "tasks");
    TRY(chunkTask[k].grow(size_t(chunkMax) * 4), "chunks");
    TRY(med[k].grow(size_t(medMax) * sizeof(Task)), "tasks");
  }
  TRY(small.grow(size_t(smallMax) * sizeof(Task)), "tasks");
  TRY(bins.grow(size_t(bigMax) * kBinWords * 4), "bins");
  TRY(split.grow(size_t(bigMax) * sizeof(Split)), "splits");
  TRY(acc.grow(size_t(bigMax) * 96), "child bounds");
  TRY(chunkLeft.grow(size_t(chunkMax) * 4), "chunk counts");
  // counters: [0..2] list A (big, chunks, medium), [3..5] list B, [6] small, [7] root id, [8..13] root centroid bounds,
  // [14] height
  TRY(ctr.grow(16 * 4), "counters");
  uint32_t init[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0, 0, 0, 0, 0};
  TRY(hipMemcpyAsync(ctr.p, init, sizeof init, hipMemcpyHostToDevice, g->stream), "counters");
  uint32_t* C = ctr.as<uint32_t>();
  k_boxes<<<nblocks(nt, 256), 256, 0, g->stream>>>(reinterpret_cast<const float4*>(leaves), nt, pbox.as<float4>(), C + 8);
  k_iota<<<nblocks(nt, 256), 256, 0, g->stream>>>(ref0.as<uint32_t>(), nt);
  const auto lists = [&](int next) {
    Lists o;
    o.nodes = nodes; o.finalRef = fin.as<uint32_t>(); o.rootId = C + 7; o.height = reinterpret_cast<int*>(C + 14);
    o.ref[0] = ref0.as<uint32_t>(); o.ref[1] = ref1.as<uint32_t>();
    o.big = big[next].as<Task>(); o.nBig = C + 3 * next;
    o.chunkTask = chunkTask[next].as<uint32_t>(); o.nChunk = C + 3 * next + 1;
    o.med = med[next].as<Task>(); o.nMed = C + 3 * next + 2;
    o.small = small.as<Task>(); o.nSmall = C + 6;
    return o;
  };
  k_root<<<1, 1, 0, g->stream>>>(nt, C + 8, lists(0));
  TRY(hipGetLastError(), "kernels");
  int cur = 0;
  uint32_t h[16];
  for (int level = 0;; ++level) {
    TRY(hipMemcpyAsync(h, C, sizeof h, hipMemcpyDeviceToHost, g->stream), "counters");
    TRY(hipStreamSynchronize(g->stream), "levels");
    const uint32_t nBig = h[3 * cur], nChunk = h[3 * cur + 1], nMed = h[3 * cur + 2];
    if (nBig + nMed == 0) break;
    if (level >= kMaxHeight) { cleanup(); return 2; }   // the tasks of this level sit at depth level + 1
    const int next = 1 - cur;
    TRY(hipMemsetAsync(C + 3 * next, 0, 12, g->stream), "counters");
    const Lists o = lists(next);
    const Task* T = big[cur].as<Task>();
    const uint32_t* CT = chunkTask[cur].as<uint32_t>();
    if (nBig) {
      k_big_init<<<nblocks(uint64_t(nBig) * kBinWords, 256), 256, 0, g->stream>>>(nBig, bins.as<uint32_t>());
      k_big_bin<<<nChunk, kBlock, 0, g->stream>>>(T, CT, ref0.as<uint32_t>(), ref1.as<uint32_t>(), pbox.as<float4>(), bins.as<uint32_t>());
      k_big_choose<<<nBig, 64, 0, g->stream>>>(T, bins.as<uint32_t>(), split.as<Split>(), acc.as<uint32_t>());
      k_big_count<<<nChunk, kBlock, 0, g->stream>>>(T, CT, ref0.as<uint32_t>(), ref1.as<uint32_t>(), pbox.as<float4>(), split.as<Split>(),
                                                    chunkLeft.as<uint32_t>());
      k_big_offsets<<<nBig, 64, 0, g->stream>>>(T, chunkLeft.as<uint32_t>());
      k_big_scatter<<<nChunk, kBlock, 0, g->stream>>>(T, CT, ref0.as<uint32_t>(), ref1.as<uint32_t>(), pbox.as<float4>(), split.as<Split>(),
                                                      chunkLeft.as<uint32_t>(), acc.as<uint32_t>());
      k_big_emit<<<nblocks(nBig, 64), 64, 0, g->stream>>>(T, nBig, split.as<Split>(), acc.as<uint32_t>(), o);
    }
    if (nMed) k_medium<<<nMed, kBlock, 0, g->stream>>>(med[cur].as<Task>(), pbox.as<float4>(), o);
    TRY(hipGetLastError(), "kernels");
    cur = next;
  }
  const uint32_t nSmall = h[6];
  if (nSmall) k_small<<<nSmall, 64, 0, g->stream>>>(small.as<Task>(), pbox.as<float4>(), lists(cur));
  TRY(hipGetLastError(), "kernels");
  int hh = 0;
  TRY(hipMemcpyAsync(&hh, C + 14, 4, hipMemcpyDeviceToHost, g->stream), "height");
  TRY(hipStreamSynchronize(g->stream), "build");
  if (hh > kMaxHeight) { cleanup(); return 2; }
  TRY(leavesTmp.grow(size_t(nt) * 48), "leaves");
  TRY(hipMemcpyAsync(leavesTmp.p, leaves, size_t(nt) * 48, hipMemcpyDeviceToDevice, g->stream), "leaves");
  k_reorder<<<nblocks(nt, 256), 256, 0, g->stream>>>(leavesTmp.as<float4>(), fin.as<uint32_t>(), nt, reinterpret_cast<float4*>(leaves));
  TRY(hipGetLastError(), "kernels");
  TRY(hipStreamSynchronize(g->stream), "build");
#undef TRY
  height = hh;
  cleanup();
  return 0;
}
