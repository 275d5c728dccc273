// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The time-series contraction E = A S + G g + sun of a daylight chunk: replaces k_series / gpu_series (reference
// src/gpu.cu:119-149, 639-669). A[i, d] = max(0, dotf(n_i, D_d)) wsky_d on the visible sky directions (the daylight
// coefficients, exact 0 where a direction is blocked; byte mode: x the transmittance byte / 255.f, gpu.cu:83-92), S[row_d,
// t] the patch radiance of the non-night steps (dlg_sky).
//   method 0, FP32: each E[i, t] an fma chain over d in ascending order, then fma(G_i, g_t, .), then the sun's fma at its
//             step: the reference kernel's arithmetic (nvcc contracts it so), tiled 64 x 64 through LDS.
//   method 1, BF16x3: A and S split into bf16 hi + lo; hi.hi + hi.lo + lo.hi on MFMA (v_mfma_f32_32x32x16 bf16),
//             FP32 accumulation in blocks of 256 directions summed into a running total; epilogue in FP64.
//   method 2, BF16x6: hi + mid + lo; the six products down to 2^-16 relative, as method 1 otherwise: FP32-level.
//   method 3, FP16x3: A and S scaled by powers of two into fp16's range and split into fp16 hi + lo (11 significant bits
//             each: 22 bits kept, against 16 for two bf16 pieces); hi.hi + hi.lo + lo.hi on v_mfma_f32_32x32x16_f16,
//             accumulated as method 1: method 1's cost at about method 2's accuracy (2^-21 per product at worst).
// Operands: fragment-major bf16 planes (Kp = nSky rounded up to 32, zero padded), described at k_split_a. Reflected light
// (dlg_bounce rows dc and the sky's bin radiance Sdc) extends K: E += dc Sdc in the same products (methods 1 / 2) or a
// second FP32 pass after the reference's (method 0).
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "dl_math.h"
#include "gpu_internal.h"

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));   // 16 bytes of operand (a native vector: kept in registers)

__device__ __forceinline__ uint16_t bf16_rn(float f) {   // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(f);
  return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// the visibility of (sensor i, direction d) as a transmittance byte: a set bit is 255; byte mode's quads hold the byte
__device__ __forceinline__ uint32_t vis_byte(const uint32_t* __restrict__ vis, uint32_t nc, uint32_t i, uint32_t d, uint32_t bytes) {
  if (bytes) return (vis[size_t(d >> 2) * nc + i] >> (8 * (d & 3))) & 255u;
  return ((vis[size_t(d >> 5) * nc + i] >> (d & 31)) & 1u) * 255u;
}
// the fraction a byte stands for (a bit's 255: exactly 1)
__device__ __forceinline__ float vis_T(uint32_t v) { return v == 255u ? 1.f : dlm::byte_T(v); }

// the daylight coefficient of (sensor i, sky direction d), as k_series forms it (gpu.cu:137): cos x T x wsky in float
__device__ __forceinline__ float coeff(const uint32_t* __restrict__ vis, uint32_t bytes, uint32_t nc, uint32_t i, uint32_t d, float3 n,
                                       const float3* __restrict__ D, const float* __restrict__ wsky) {
  const uint32_t v = vis_byte(vis, nc, i, d, bytes);
  if (!v) return 0.f;
  return (fmaxf(dlm::dot3(n, D[d]), 0.f) * vis_T(v)) * wsky[d];
}

// ---------------------------------------------------------------------------------------------- method 0: FP32
// Each E[i, t]: the fma chain over the sky directions in ascending order, then fma(G_i, g_t, .), then the sun of step t
// (its one direction, when visible): per element the reference kernel's order of operations (gpu.cu:128-148).
constexpr int kTile = 64, kKt = 16;
__global__ __launch_bounds__(256) void k_series_f32(const uint32_t* __restrict__ bits, uint32_t bytes, uint32_t nc, const float3* __restrict__ N,
                                                    const float3* __restrict__ D, uint32_t nSky, const int* __restrict__ row,
                                                    const float* __restrict__ wsky, const float* __restrict__ S, int T,
                                                    const float* __restrict__ G, const float* __restrict__ g, uint32_t sun0,
                                                    const int* __restrict__ sunOfCol, const float* __restrict__ sunW, float* __restrict__ E) {
  __shared__ float As[kKt][kTile];
  __shared__ float Bs[kKt][kTile + 4];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const uint32_t i0 = blockIdx.y * kTile, t0 = blockIdx.x * kTile;
  float e[4][4] = {};
  for (uint32_t d0 = 0; d0 < nSky; d0 += kKt) {
    for (int k = threadIdx.x; k < kKt * kTile; k += 256) {   // A tile: 16 directions x 64 sensors, from the bits
      const int dd = k / kTile, ii = k % kTile;
      const uint32_t d = d0 + dd, i = i0 + ii;
      As[dd][ii] = (d < nSky && i < nc) ? coeff(bits, bytes, nc, i, d, N[i], D, wsky) : 0.f;
      const int t = int(t0) + ii;
      Bs[dd][ii] = (d < nSky && t < T) ? S[size_t(row[d]) * T + t] : 0.f;
    }
    __syncthreads();
    const int kk = int(nSky - d0) < kKt ? int(nSky - d0) : kKt;
    for (int k = 0; k < kk; ++k) {
      float a[4], b[4];
      for (int r = 0; r < 4; ++r) { a[r] = As[k][ty * 4 + r]; b[r] = Bs[k][tx * 4 + r]; }
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) e[r][c] = fmaf(a[r], b[c], e[r][c]);
    }
    __syncthreads();
  }
  for (int r = 0; r < 4; ++r) {
    const uint32_t i = i0 + ty * 4 + r;
    if (i >= nc) continue;
    const float3 n = N[i];
    for (int c = 0; c < 4; ++c) {
      const int t = int(t0) + tx * 4 + c;
      if (t >= T) continue;
      float v = fmaf(G[i], g[t], e[r][c]);
      const int j = sunOfCol[t];
      const uint32_t b = j >= 0 ? vis_byte(bits, nc, i, sun0 + uint32_t(j), bytes) : 0u;
      if (b) v = fmaf(fmaxf(dlm::dot3(n, D[sun0 + uint32_t(j)]), 0.f) * vis_T(b), sunW[j], v);
      E[size_t(i) * T + t] = v;
    }
  }
}

// reflected light in FP32 (k_series_dc, gpu.cu:153-169): E[i, t] += (an fma chain over the bins in order of dc(b, i)
// Sdc[b, t]), the reference kernel's arithmetic, tiled 64 x 64 through LDS as k_series_f32
__global__ __launch_bounds__(256) void k_series_dc_f32(const float* __restrict__ dc, uint32_t nc, uint32_t nBins, const float* __restrict__ Sdc,
                                                       int T, float* __restrict__ E) {
  __shared__ float As[kKt][kTile];
  __shared__ float Bs[kKt][kTile + 4];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const uint32_t i0 = blockIdx.y * kTile, t0 = blockIdx.x * kTile;
  float e[4][4] = {};
  for (uint32_t b0 = 0; b0 < nBins; b0 += kKt) {
    for (int k = threadIdx.x; k < kKt * kTile; k += 256) {
      const int bb = k / kTile, ii = k % kTile;
      const uint32_t b = b0 + bb, i = i0 + ii;
      As[bb][ii] = (b < nBins && i < nc) ? dc[size_t(b) * nc + i] : 0.f;
      const int t = int(t0) + ii;
      Bs[bb][ii] = (b < nBins && t < T) ? Sdc[size_t(b) * T + t] : 0.f;
    }
    __syncthreads();
    const int kk = int(nBins - b0) < kKt ? int(nBins - b0) : kKt;
    for (int k = 0; k < kk; ++k) {
      float a[4], bv[4];
      for (int r = 0; r < 4; ++r) { a[r] = As[k][ty * 4 + r]; bv[r] = Bs[k][tx * 4 + r]; }
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) e[r][c] = fmaf(a[r], bv[c], e[r][c]);
    }
    __syncthreads();
  }
  for (int r = 0; r < 4; ++r) {
    const uint32_t i = i0 + ty * 4 + r;
    if (i >= nc) continue;
    for (int c = 0; c < 4; ++c) {
      const int t = int(t0) + tx * 4 + c;
      if (t < T) E[size_t(i) * T + t] += e[r][c];
    }
  }
}

// ---------------------------------------------------------------------------------------------- methods 1 / 2: MFMA
// split x into NS pieces, each the rounding of what the previous ones leave (r - piece is exact in float): bf16, or
// (H) fp16 of x 2^e, e the operand's scale exponent (scale_exp)
template <int NS, bool H>
__device__ __forceinline__ void split(float x, int e, uint16_t* out) {
  if (H) {
    float r = __builtin_ldexpf(x, e);
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const _Float16 h = (_Float16)r;
      out[s] = __builtin_bit_cast(uint16_t, h);
      r = r - (float)h;
    }
  } else {
    float r = x;
#pragma unroll
    for (int s = 0; s < NS; ++s) { out[s] = bf16_rn(r); r = r - bf16_f(out[s]); }
  }
}

// The fp16 operands' scales (method 3). fp16 is normal from 2^-14 to 65504: each operand is multiplied by 2^e so that
// its largest magnitude m lies below 2^14 (e = 14 - k for m < 2^k), which keeps every value within 2^28 of m at full
// precision; what lies below that adds at most 2^-25 absolute per piece (2^-39 of m). Powers of two scale exactly, and
// the epilogue takes 2^-(eA + eS) back off. mx[0]: the largest |A| bound (the sky weights, as A = cos T wsky with cos
// and T at most 1, and the bounce rows), mx[1]: the largest |S| (sky rows and bins), both as float bits (for
// non-negative floats the integer order is the float order), cleared by k_scale_zero and raised by k_scale_max.
__device__ __forceinline__ int scale_exp(uint32_t bits) {
  if (bits == 0u) return 0;
  const int k = int(bits >> 23) - 126;   // m < 2^k
  const int e = 14 - k;
  return e < -120 ? -120 : e > 120 ? 120 : e;
}
__global__ void k_scale_zero(uint32_t* mx) {
  if (threadIdx.x < 2) mx[threadIdx.x] = 0u;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    const uint32_t o = uint32_t(__shfl_xor(int(v), m));
    v = o > v ? o : v;
  }
  return v;
}
// grid-stride over the A bound's inputs (wsky [nW], dc [nDc]) and S's (S [nS], Sdc [nSdc]); one atomic per wave and operand
__global__ __launch_bounds__(256) void k_scale_max(const float* __restrict__ wsky, uint32_t nW, const float* __restrict__ dc, size_t nDc,
                                                   const float* __restrict__ S, size_t nS, const float* __restrict__ Sdc, size_t nSdc,
                                                   uint32_t* __restrict__ mx) {
  const size_t stride = size_t(gridDim.x) * blockDim.x, t0 = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t a = 0u, b = 0u;
  const auto mag = [](float f) { return __float_as_uint(f) & 0x7fffffffu; };
  const auto up = [](uint32_t& m, uint32_t v) { m = v > m ? v : m; };
  for (size_t k = t0; k < nW; k += stride) up(a, mag(wsky[k]));
  for (size_t k = t0; k < nDc; k += stride) up(a, mag(dc[k]));
  for (size_t k = t0; k < nS; k += stride) up(b, mag(S[k]));
  for (size_t k = t0; k < nSdc; k += stride) up(b, mag(Sdc[k]));
  a = wave_max(a);
  b = wave_max(b);
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicMax(&mx[0], a);
    if (b) atomicMax(&mx[1], b);
  }
}

// Operand layouts: fragment-major, so that an MFMA operand is one 16-byte load per lane and a wave's load one
// contiguous kilobyte. Planes A [NS][nRG][nKS][64][8] (nRG row groups of 32 sensors, zero padded to the tile) and
// B = S^T [NS][nCG][nKS][64][8] (nCG column groups of 32 steps); fragment (group g, k-step ks) lane l holds the 8 values
// of row (or column) 32 g + l % 32 at k = 16 ks + 8 (l / 32) + 0..7, the v_mfma_f32_32x32x16_bf16 operand layout.
__device__ __forceinline__ size_t frag_index(uint32_t g, uint32_t r, uint32_t k, uint32_t nKS) {
  return ((size_t(g) * nKS + k / 16) * 64 + r + 32 * ((k / 8) & 1)) * 8 + (k & 7);
}

// A's columns: the sky directions' coefficients, then (reflected light) the chunk's bounce rows dc[b nc + i]. A wave
// writes one fragment per plane: lane l the 8 values of row 32 g + l % 32 at k = 16 ks + 8 (l / 32) + 0..7, as one
// 16-byte store (a wave's store one contiguous kilobyte); its visibility word is one load for the 8 directions. A block's
// four waves take kSplitKs k-steps of one row group in turn (few blocks, each with a run of stores).
constexpr uint32_t kSplitKs = 32;
// nz[ks nRG + g]: 1 where fragment (g, ks) holds a coefficient other than 0 (a wave's ballot), else 0; the contraction
// steps over the slices whose fragments are all 0 for a block's rows (their products are exact zeros).
template <int NS, bool H>
__global__ __launch_bounds__(256) void k_split_a(const uint32_t* __restrict__ bits, uint32_t bytes, uint32_t nc, const float3* __restrict__ N,
                                                 const float3* __restrict__ D, uint32_t nSky, uint32_t nKS, uint32_t nRG, const float* __restrict__ wsky,
                                                 const float* __restrict__ dc, uint32_t nBins, const uint32_t* __restrict__ mx, u32x4* __restrict__ A,
                                                 uint8_t* __restrict__ nz) {
  const uint32_t lane = threadIdx.x & 63, g = blockIdx.y, ksEnd = min(nKS, (blockIdx.x + 1) * kSplitKs);
  const int e = H ? scale_exp(mx[0]) : 0;
  const uint32_t i = 32 * g + lane % 32;
  const bool live = i < nc;
  const float3 n = live ? N[i] : make_float3(0.f, 0.f, 0.f);
  for (uint32_t ks = blockIdx.x * kSplitKs + (threadIdx.x >> 6); ks < ksEnd; ks += 4) {
    const uint32_t d0 = 16 * ks + 8 * (lane / 32);
    uint32_t pk[NS][4];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int h = 0; h < 4; ++h) pk[s][h] = 0;
    bool any = false;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      const uint32_t d = d0 + j;
      const float c = !live ? 0.f : d < nSky ? coeff(bits, bytes, nc, i, d, n, D, wsky) : d - nSky < nBins ? dc[size_t(d - nSky) * nc + i] : 0.f;
      any |= c != 0.f;
      uint16_t p[NS];
      split<NS, H>(c, e, p);
#pragma unroll
      for (int s = 0; s < NS; ++s) pk[s][j / 2] |= uint32_t(p[s]) << (16 * (j & 1));
    }
    const size_t plane = size_t(nRG) * nKS * 64, at = (size_t(g) * nKS + ks) * 64 + lane;
#pragma unroll
    for (int s = 0; s < NS; ++s) A[s * plane + at] = u32x4{pk[s][0], pk[s][1], pk[s][2], pk[s][3]};
    const bool fragNz = __ballot(any) != 0;
    if (lane == 0) nz[size_t(ks) * nRG + g] = fragNz ? 1 : 0;
  }
}

// B's rows: S of each sky direction's patch, then (reflected light) the bins' radiance Sdc [b T + t]
template <int NS, bool H>
__global__ void k_split_s(const float* __restrict__ S, const int* __restrict__ row, uint32_t nSky, uint32_t Kp, int T, uint32_t Tp,
                          const float* __restrict__ Sdc, uint32_t nBins, const uint32_t* __restrict__ mx, uint16_t* __restrict__ B) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
  if (d >= Kp || t >= Tp) return;
  const float s = int(t) >= T ? 0.f : d < nSky ? S[size_t(row[d]) * T + t] : d - nSky < nBins ? Sdc[size_t(d - nSky) * T + t] : 0.f;
  uint16_t p[NS];
  split<NS, H>(s, H ? scale_exp(mx[1]) : 0, p);
  const size_t plane = size_t(Tp) * Kp, at = frag_index(t / 32, t % 32, d, Kp / 16);
#pragma unroll
  for (int k = 0; k < NS; ++k) B[k * plane + at] = p[k];
}

// v_mfma_f32_32x32x16 bf16 (gfx950): lane l's A operand is row l % 32, B operand column l % 32, both at the k slots of
// lane group l / 32 (the same slots for A and B, so the product sums every k once); acc[v] is C[8 (v / 4) + 4 (l / 32) +
// v % 4][l % 32].
//
// Block: 2 WM waves over 64 WM sensors x 128 steps, wave (wm, wn) 64 x 64 = 2 x 2 MFMA tiles (each A fragment serves two
// tiles, each B fragment two; WM = 4 shares each B fragment among four waves, at twice the block's A). Per slice of 32
// directions the block stages its fragments in LDS - NS x 2 WM row groups x 2 k-steps of A and NS x 4 column groups of
// B, laid out as in HBM so that both the copy and each wave's reads are contiguous - and loads the next slice into
// registers while the MFMAs of this one run. FP32 accumulation in blocks of kBlockSlices slices added to a running
// total, the epilogue (G g, the sun) in FP64. Every tile shape and order gives the same bits: each element's products
// are summed in the same order.
// Blocks are numbered XCD-major: the blocks an XCD runs (hardware id % 8) take consecutive tiles of a grouped order
// (`group` rows of sensor tiles per step column), so the tiles that share operand panels share that XCD's L2.
constexpr uint32_t kTileN = 128, kSliceK = 32, kMaxTileM = 256;
// Accumulation in blocks: the MFMAs of kBlockSlices slices (256 directions) accumulate in FP32 from zero, and each block's
// sum is added to a running FP32 total. A sum over K directions then carries about K / 256 + 6 x 16 roundings instead of
// the 6 K / 16 of one chain (6 products per k-step at NS = 3). Measured at P = 2305 on the urban block: BF16x6 1.9e-6 ->
// 3.9e-7, BF16x3 2.5e-6 -> 1.3e-6 relative to FP64. The totals take the registers of a third wave per SIMD (2 waves, 252
// VGPRs, no spills; 2 % slower than the single chain at 3 waves; 3 waves with the totals spilled: 30 % slower).
constexpr uint32_t kBlockSlices = 8;
// the slices a block's list holds (K up to 16,384 directions); a longer K multiplies every slice (nz null)
constexpr uint32_t kMaxListed = 512;
// one product of 16-byte operand fragments: bf16 pieces, or (H) fp16 ones
template <bool H>
__device__ __forceinline__ f32x16 mma(u32x4 a, u32x4 b, f32x16 c) {
  if constexpr (H)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
template <int NS, int WM, bool H>
__global__ __launch_bounds__(128 * WM) __attribute__((amdgpu_waves_per_eu(2))) void k_series_mfma(
    const u32x4* __restrict__ A, const u32x4* __restrict__ B, uint32_t nc, uint32_t nRG, uint32_t nCG, uint32_t nKS, int T,
    const uint32_t* __restrict__ bits, uint32_t bytes, const float3* __restrict__ N, const float3* __restrict__ D, const float* __restrict__ G,
    const float* __restrict__ g, uint32_t sun0, const int* __restrict__ sunOfCol, const float* __restrict__ sunW, const uint32_t* __restrict__ mx,
    const uint8_t* __restrict__ nz, float* __restrict__ E, uint32_t tilesT, uint32_t tiles, uint32_t group) {
  constexpr int kThreads = 128 * WM, kRows = 64 * WM, kGA = 2 * WM, kGB = 4;
  constexpr int kFragA = NS * kGA * 128, kFragB = NS * kGB * 128;   // uint4 per slice: planes x groups x 2 k-steps x 64 lanes
  constexpr int kPerA = kFragA / kThreads, kPerB = kFragB / kThreads;
  static_assert(kFragA % kThreads == 0 && kFragB % kThreads == 0, "each thread stages whole entries");
  __shared__ u32x4 sA[kFragA];
  __shared__ u32x4 sB[kFragB];
  static_assert(kFragA * 16 >= kRows * 128 && kFragB * 16 >= (kRows * 4 + 128 * 6) * 4, "the epilogue's staging fits the operand buffers");
  // XCD-major tile order
  const uint32_t id = blockIdx.x, per = (tiles + 7) / 8, xcd = id % 8, rank = id / 8;
  const uint32_t full = tiles % 8 == 0 ? 8 : tiles % 8;   // XCDs that take `per` tiles; the rest take per - 1
  const uint32_t tile = xcd < full ? xcd * per + rank : full * per + (xcd - full) * (per - 1) + rank;
  if (tile >= tiles || (xcd >= full && rank >= per - 1)) return;
  // grouped order: `group` sensor-tile rows walked down before the next step column
  const uint32_t tilesI = tiles / tilesT, grp = tile / (group * tilesT), first = grp * group;
  const uint32_t rows = tilesI - first < group ? tilesI - first : group, local = tile % (group * tilesT);
  const uint32_t ti = first + local % rows, tt = local / rows;
  const uint32_t rg0 = ti * kGA, cg0 = tt * kGB;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w % WM, wn = w / WM;
  // entry j = tid + kThreads m of a slice: plane j / (G 128), group (j / 128) % G, uint4 j % 128 of the group's two
  // k-steps (32-bit offsets: a plane holds fewer than 2^32 / 3 uint4 at any chunk size)
  const uint32_t gStride = nKS * 64, aPl = nRG * gStride, bPl = nCG * gStride;
  const auto aAt = [&](int m) {
    const uint32_t j = threadIdx.x + kThreads * m;
    return (j / (kGA * 128)) * aPl + (rg0 + (j / 128) % kGA) * gStride + j % 128;
  };
  const auto bAt = [&](int m) {
    const uint32_t j = threadIdx.x + kThreads * m;
    return (j / (kGB * 128)) * bPl + (cg0 + (j / 128) % kGB) * gStride + j % 128;
  };
  // The slices this block multiplies: those where some fragment of its kGA row groups holds a coefficient other than 0
  // (nz from k_split_a), in ascending order. The others add exact zeros: stepping over them leaves every sum's bits as
  // they are, and the accumulation blocks keep their bounds (kBlockSlices slices by index).
  const uint32_t nSlices = nKS / 2;
  __shared__ uint16_t list[kMaxListed];
  __shared__ uint32_t waveCount[kThreads / 64];
  const bool listed = nz != nullptr;   // else every slice, q = idx (the host passes no flags past kMaxListed slices)
  uint32_t nList = listed ? 0 : nSlices;
  for (uint32_t base = 0; listed && base < nSlices; base += kThreads) {
    const uint32_t q = base + threadIdx.x;
    bool keep = q < nSlices;
    if (keep) {
      const uint32_t* f0 = reinterpret_cast<const uint32_t*>(nz + size_t(2 * q) * nRG + rg0);
      const uint32_t* f1 = reinterpret_cast<const uint32_t*>(nz + size_t(2 * q + 1) * nRG + rg0);
      uint32_t any = 0;
#pragma unroll
      for (int k = 0; k < kGA / 4; ++k) any |= f0[k] | f1[k];
      keep = any != 0;
    }
    const uint64_t vote = __ballot(keep);
    if (lane == 0) waveCount[w] = uint32_t(__popcll(vote));
    __syncthreads();
    uint32_t before = nList;
    for (int v = 0; v < w; ++v) before += waveCount[v];
    if (keep) list[before + uint32_t(__popcll(vote & ((1ull << lane) - 1ull)))] = uint16_t(q);
    for (int v = 0; v < kThreads / 64; ++v) nList += waveCount[v];
    __syncthreads();
  }
  const auto sliceAt = [&](uint32_t idx) { return listed ? uint32_t(list[idx]) : idx; };
  u32x4 pa[kPerA], pb[kPerB];
  if (nList) {
    const uint32_t step = 128 * sliceAt(0);
#pragma unroll
    for (int m = 0; m < kPerA; ++m) pa[m] = A[aAt(m) + step];
#pragma unroll
    for (int m = 0; m < kPerB; ++m) pb[m] = B[bAt(m) + step];
  }
  // acc: the products of the current block of kBlockSlices slices; tot: the blocks summed
  f32x16 acc[2][2], tot[2][2];
#pragma unroll
  for (int fm = 0; fm < 2; ++fm)
#pragma unroll
    for (int fn = 0; fn < 2; ++fn) acc[fm][fn] = tot[fm][fn] = f32x16{};
  for (uint32_t idx = 0; idx < nList; ++idx) {
    const uint32_t q = sliceAt(idx), qNext = idx + 1 < nList ? sliceAt(idx + 1) : q;
#pragma unroll
    for (int m = 0; m < kPerA; ++m) sA[threadIdx.x + kThreads * m] = pa[m];
#pragma unroll
    for (int m = 0; m < kPerB; ++m) sB[threadIdx.x + kThreads * m] = pb[m];
    __syncthreads();
    {   // the next slice in flight while this one's MFMAs run (the last iteration reloads its own: no branch); the
        // scheduling barriers keep the loads here, ahead of the MFMAs, and each k-step's fragments after the last's
        // products, so that the two do not share registers
      const uint32_t step = 128 * qNext;
#pragma unroll
      for (int m = 0; m < kPerA; ++m) pa[m] = A[aAt(m) + step];
#pragma unroll
      for (int m = 0; m < kPerB; ++m) pb[m] = B[bAt(m) + step];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      u32x4 a[2][NS], b[2][NS];
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int f = 0; f < 2; ++f) {
          a[f][s] = sA[((s * kGA + 2 * wm + f) * 2 + kk) * 64 + lane];
          b[f][s] = sB[((s * kGB + 2 * wn + f) * 2 + kk) * 64 + lane];
        }
      // products whose exponent sum is at most NS - 1 pieces below the leading one, smallest first
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < 2; ++fn)
#pragma unroll
          for (int o = NS - 1; o >= 0; --o)
#pragma unroll
            for (int s = 0; s <= o; ++s) acc[fm][fn] = mma<H>(a[fm][s], b[fn][o - s], acc[fm][fn]);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    if (qNext / kBlockSlices != q / kBlockSlices || idx + 1 == nList) {   // a block of directions done: into the running sum
#pragma unroll
      for (int fm = 0; fm < 2; ++fm)
#pragma unroll
        for (int fn = 0; fn < 2; ++fn) { tot[fm][fn] += acc[fm][fn]; acc[fm][fn] = f32x16{}; }
    }
  }
  // Epilogue: the tile's rows (N, G), columns (g, the step's sun: D, weight) and the sun visibility of every (row,
  // column) staged in the operand buffers (free after the loop's last barrier), so that the 64 elements of a lane read
  // LDS instead of gathering from HBM; E = acc (H: times 2^-(eA + eS), exact) + G g + the sun's term, in FP64.
  const double unscale = H ? __builtin_ldexp(1.0, -(scale_exp(mx[0]) + scale_exp(mx[1]))) : 1.0;
  float* sRow = reinterpret_cast<float*>(sB);       // [kRows][4]: N x y z, G
  float* sCol = sRow + kRows * 4;                    // [128][6]: g, D x y z, sun weight, sun index (bits; -1: none)
  uint8_t* sVis = reinterpret_cast<uint8_t*>(sA);   // [kRows][128 columns]: the sun's transmittance byte
  const uint32_t i0 = 32 * rg0, t0 = 32 * cg0;
  if (int(threadIdx.x) < kRows) {
    const uint32_t i = i0 + threadIdx.x;
    const float3 n = i < nc ? N[i] : make_float3(0.f, 0.f, 0.f);
    float* r = sRow + 4 * threadIdx.x;
    r[0] = n.x; r[1] = n.y; r[2] = n.z; r[3] = i < nc ? G[i] : 0.f;
  } else if (threadIdx.x - kRows < 128) {
    const uint32_t c = threadIdx.x - kRows, t = t0 + c;
    const int j = int(t) < T ? sunOfCol[t] : -1;
    const float3 d = j >= 0 ? D[sun0 + uint32_t(j)] : make_float3(0.f, 0.f, 0.f);
    float* q = sCol + 6 * c;
    q[0] = int(t) < T ? g[t] : 0.f; q[1] = d.x; q[2] = d.y; q[3] = d.z; q[4] = j >= 0 ? sunW[j] : 0.f; q[5] = __int_as_float(j);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < uint32_t(kRows) * 128; k += kThreads) {   // consecutive threads: consecutive rows of one column
    const uint32_t r = k % kRows, c = k / kRows, i = i0 + r;
    const int j = __float_as_int(sCol[6 * c + 5]);
    sVis[r * 128 + c] = (j >= 0 && i < nc) ? uint8_t(vis_byte(bits, nc, i, sun0 + uint32_t(j), bytes)) : uint8_t(0);
  }
  __syncthreads();
#pragma unroll
  for (int fn = 0; fn < 2; ++fn) {
    const uint32_t c = 32 * (2 * wn + fn) + (lane & 31), t = t0 + c;
    if (int(t) >= T) continue;
    const float* q = sCol + 6 * c;
    const double gt = (double)q[0], sw = (double)q[4];
    const float3 dj = make_float3(q[1], q[2], q[3]);
#pragma unroll
    for (int fm = 0; fm < 2; ++fm)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const uint32_t r = 32 * (2 * wm + fm) + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4), i = i0 + r;
        if (i >= nc) continue;
        const float* rr = sRow + 4 * r;
        double e = (double)tot[fm][fn][v] * unscale + (double)rr[3] * gt;
        const uint32_t sv = sVis[r * 128 + c];
        if (sv) e += (double)(fmaxf(dlm::dot3(make_float3(rr[0], rr[1], rr[2]), dj), 0.f) * vis_T(sv)) * sw;
        E[size_t(i) * T + t] = (float)e;
      }
  }
}

// the ground view factor of each sensor of the chunk as the series reads it: float of the reduce's FP64 row
__global__ void k_ground_rows(const double* __restrict__ sums, uint32_t nc, float* __restrict__ G) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) G[i] = float(sums[i]);
}

// the MFMA methods' operand sizes of a chunk: Kp the directions (and bins) padded to a slice, the row groups padded to
// the tallest tile (so that one split serves every tile shape), the column groups to a tile
struct MfmaShape {
  uint32_t Kp, tilesT, nRG, nCG, Tp, nKS;
  size_t aBytes, bBytes, nzBytes;   // the split planes of A and B, then k_split_a's fragment flags [nKS x nRG]
  MfmaShape(uint32_t nc, uint32_t K, uint32_t T, int NS) {
    Kp = (K + kSliceK - 1) / kSliceK * kSliceK;
    tilesT = (T + kTileN - 1) / kTileN;
    nRG = (nc + kMaxTileM - 1) / kMaxTileM * (kMaxTileM / 32);
    nCG = tilesT * (kTileN / 32);
    Tp = nCG * 32;
    nKS = Kp / 16;
    aBytes = size_t(NS) * nRG * 32 * Kp * 2;
    bBytes = size_t(NS) * Tp * Kp * 2;
    nzBytes = (size_t(nKS) * nRG + 255) & ~size_t(255);
  }
  bool listed() const { return nKS / 2 <= kMaxListed; }
};

// The contraction's launch shapes: WM waves along the sensors (a tile of 64 WM sensors x 128 steps) and the tile-group
// order. A shape is picked per problem by timing each on its first chunk (the tuner below), or by DAYLIGHT_SERIES_TILE
// = "WMxGROUP".
struct TileShape { int wm; uint32_t group; };
constexpr TileShape kShapes[] = {{2, 4}, {2, 8}, {2, 16}, {4, 2}, {4, 4}, {4, 8}};
constexpr int kNumShapes = int(sizeof kShapes / sizeof kShapes[0]);

// the pieces each operand is split into by a method (1: BF16x3, 2: BF16x6, 3: FP16x3)
constexpr int pieces(int method) { return method == 2 ? 3 : 2; }

template <int NS, int WM, bool H>
void launch_mfma(dlg_gpu* g, const MfmaShape& sh, const u32x4* A, const u32x4* B, uint32_t group) {
  const dlg_series_plan& q = g->seriesPlan;
  const char* b = g->seriesIn.as<char>();
  const uint32_t tilesI = (q.nc + 64 * WM - 1) / (64 * WM), tiles = sh.tilesT * tilesI;
  k_series_mfma<NS, WM, H><<<(tiles + 7) / 8 * 8, 128 * WM, 0, g->stream>>>(
      A, B, q.nc, sh.nRG, sh.nCG, sh.nKS, q.T, g->bits.as<uint32_t>(), uint32_t(g->bytes), g->N.as<float3>() + g->i0, g->D.as<float3>(),
      reinterpret_cast<const float*>(b + q.oG), g->skyG.as<float>(), q.sun0, reinterpret_cast<const int*>(b + q.oCol),
      reinterpret_cast<const float*>(b + q.oSw), reinterpret_cast<const uint32_t*>(b + q.oMx),
      sh.listed() ? g->seriesWork.as<uint8_t>() + sh.aBytes + sh.bBytes : nullptr, g->seriesE.as<float>(), sh.tilesT, tiles, group);
}

template <int NS, bool H>
void launch_shape(dlg_gpu* g, const MfmaShape& sh, const u32x4* A, const u32x4* B, int shape) {
  const TileShape t = kShapes[shape];
  if (t.wm == 4) launch_mfma<NS, 4, H>(g, sh, A, B, t.group);
  else launch_mfma<NS, 2, H>(g, sh, A, B, t.group);
}

template <int NS, bool H>
void enqueue_mfma(dlg_gpu* g) {
  const dlg_series_plan& q = g->seriesPlan;
  const char* b = g->seriesIn.as<char>();
  const MfmaShape sh(q.nc, q.nSky + q.nBins, uint32_t(q.T), NS);
  uint16_t* A = g->seriesWork.as<uint16_t>();
  uint16_t* B = A + sh.aBytes / 2;
  const float3* N = g->N.as<float3>() + g->i0;
  const int* dRow = reinterpret_cast<const int*>(b + q.oRow);
  uint32_t* mx = reinterpret_cast<uint32_t*>(g->seriesIn.as<char>() + q.oMx);
  if (H) {   // the operands' largest magnitudes, for their scales (kernels only: the enqueue may be captured)
    const size_t nDc = size_t(q.nBins) * q.nc, nS = size_t(g->skyP) * size_t(q.T), nSdc = size_t(q.nBins) * size_t(q.T);
    k_scale_zero<<<1, 64, 0, g->stream>>>(mx);
    k_scale_max<<<1024, 256, 0, g->stream>>>(reinterpret_cast<const float*>(b + q.oW), q.nSky, g->dc.as<float>(), nDc, g->skyS.as<float>(), nS,
                                             g->skySdc.as<float>(), nSdc, mx);
  }
  k_split_a<NS, H><<<dim3((sh.nKS + kSplitKs - 1) / kSplitKs, sh.nRG), 256, 0, g->stream>>>(g->bits.as<uint32_t>(), uint32_t(g->bytes), q.nc, N, g->D.as<float3>(),
                                                                          q.nSky, sh.nKS, sh.nRG, reinterpret_cast<const float*>(b + q.oW),
                                                                          g->dc.as<float>(), q.nBins, mx, reinterpret_cast<u32x4*>(A),
                                                                          g->seriesWork.as<uint8_t>() + sh.aBytes + sh.bBytes);
  k_split_s<NS, H><<<dim3((sh.Kp + 255) / 256, sh.Tp), 256, 0, g->stream>>>(g->skyS.as<float>(), dRow, q.nSky, sh.Kp, q.T, sh.Tp,
                                                                           g->skySdc.as<float>(), q.nBins, mx, B);
  launch_shape<NS, H>(g, sh, reinterpret_cast<const u32x4*>(A), reinterpret_cast<const u32x4*>(B), q.shape);
}

// ------------------------------------------------------------------------------------------------ the tuner
// The shape chosen per problem (method, sensors to the next power of two, padded directions, steps), for the process.
struct TuneKey {
  int method;
  uint32_t ncPow2, Kp, T;
  bool operator==(const TuneKey& o) const { return method == o.method && ncPow2 == o.ncPow2 && Kp == o.Kp && T == o.T; }
};
std::mutex tuneMu;
std::vector<std::pair<TuneKey, int>> tuned;

TuneKey tune_key(const dlg_series_plan& q) {
  uint32_t p = 1;
  while (p < q.nc) p <<= 1;
  return {q.method, p, MfmaShape(q.nc, q.nSky + q.nBins, uint32_t(q.T), pieces(q.method)).Kp, uint32_t(q.T)};
}

// the forced shape of DAYLIGHT_SERIES_TILE ("WMxGROUP"), or -1
int forced_shape() {
  const char* e = std::getenv("DAYLIGHT_SERIES_TILE");
  if (!e) return -1;
  int wm = 0;
  unsigned grp = 0;
  if (std::sscanf(e, "%dx%u", &wm, &grp) != 2) return -1;
  for (int k = 0; k < kNumShapes; ++k)
    if (kShapes[k].wm == wm && kShapes[k].group == grp) return k;
  return -1;
}

// the tuned shape of the plan's problem, or -1 when it has not been timed
int known_shape(const dlg_series_plan& q) {
  if (const int f = forced_shape(); f >= 0) return f;
  const TuneKey k = tune_key(q);
  std::lock_guard<std::mutex> lk(tuneMu);
  for (const auto& [key, shape] : tuned)
    if (key == k) return shape;
  return -1;
}

// Times every shape on the chunk whose operands are split (each launch writes the same E) and records the fastest.
template <int NS, bool H>
int tune(dlg_gpu* g) {
  const dlg_series_plan& q = g->seriesPlan;
  const MfmaShape sh(q.nc, q.nSky + q.nBins, uint32_t(q.T), NS);
  const u32x4* A = g->seriesWork.as<u32x4>();
  const u32x4* B = reinterpret_cast<const u32x4*>(g->seriesWork.as<char>() + sh.aBytes);
  int best = 0;
  float bestMs = 0;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 0;
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return 0; }
  for (int k = 0; k < kNumShapes; ++k) {
    (void)hipEventRecord(e0, g->stream);
    launch_shape<NS, H>(g, sh, A, B, k);
    (void)hipEventRecord(e1, g->stream);
    float ms = 0;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
    if (k == 0 || ms < bestMs) { best = k; bestMs = ms; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(tuneMu);
  tuned.push_back({tune_key(q), best});
  return best;
}

}  // namespace

int32_t dlg_series_prepare(dlg_gpu* g, int64_t nc, const dlg_series_in* in, int32_t method, int32_t groundRow, char* err, int32_t errLen) {
  if (nc != g->nc) { dlg_fail(err, errLen, "dlg_series: no trace of this chunk"); return 1; }
  if (in->P != g->skyP || in->T != g->skyT) { dlg_fail(err, errLen, "dlg_series: S of another sky (run dlg_sky with series on)"); return 1; }
  if (in->reflect && (!g->skyReflect || g->dcN != nc || g->dcI0 != g->i0 || g->dcBins != uint32_t(g->skyP + 1))) {
    dlg_fail(err, errLen, "dlg_series: reflected light needs the chunk's dlg_bounce rows and a sky with reflect on");
    return 1;
  }
  if (method < 0 || method > 3) { dlg_fail(err, errLen, "dlg_series: method 0 (FP32), 1 (BF16x3), 2 (BF16x6) or 3 (FP16x3)"); return 1; }
  const int T = in->T;
  dlg_series_plan& q = g->seriesPlan;
  q = dlg_series_plan{};
  q.method = method;
  q.groundRow = groundRow;
  q.reflect = in->reflect;
  q.nc = uint32_t(nc);
  q.nSky = in->nSky;
  q.nBins = in->reflect ? g->dcBins : 0u;
  q.sun0 = in->sun0;
  q.T = T;
  // inputs: row | wsky | G | the sun of each column | sun weights, each 256-byte aligned
  size_t off = 0;
  const auto take = [&](size_t bytes) { const size_t at = off; off += (bytes + 255) & ~size_t(255); return at; };
  q.oRow = take(4 * size_t(in->nSky));
  q.oW = take(4 * size_t(in->nSky));
  q.oG = take(4 * size_t(nc));
  q.oCol = take(4 * size_t(T));
  q.oSw = take(4 * size_t(in->nSun ? in->nSun : 1));
  q.oMx = take(8);   // the fp16 operands' largest magnitudes (method 3)
  DLG_CHECK(g->seriesIn.grow(off));
  char* b = g->seriesIn.as<char>();
  std::vector<int> sunOfCol(static_cast<size_t>(T), -1);
  for (uint32_t j = 0; j < in->nSun; ++j)   // suns of night steps carry weight 0 and column 0: they add nothing
    if (in->sunW[j] != 0.f) sunOfCol[size_t(in->sunStep[j])] = int(j);
  g->seriesCols = sunOfCol;   // kept until the copies below have run (the caller waits before the next prepare)
  DLG_CHECK(hipMemcpyAsync(b + q.oRow, in->row, 4 * size_t(in->nSky), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipMemcpyAsync(b + q.oW, in->wsky, 4 * size_t(in->nSky), hipMemcpyHostToDevice, g->stream));
  if (groundRow < 0) DLG_CHECK(hipMemcpyAsync(b + q.oG, in->G, 4 * size_t(nc), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipMemcpyAsync(b + q.oCol, g->seriesCols.data(), 4 * size_t(T), hipMemcpyHostToDevice, g->stream));
  if (in->nSun) DLG_CHECK(hipMemcpyAsync(b + q.oSw, in->sunW, 4 * size_t(in->nSun), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(g->seriesE.grow(size_t(nc) * size_t(T) * 4));
  if (method) {
    const MfmaShape sh(uint32_t(nc), in->nSky + q.nBins, uint32_t(T), pieces(method));
    DLG_CHECK(g->seriesWork.grow(sh.aBytes + sh.bBytes + sh.nzBytes));
    q.shape = known_shape(q);
  }
  g->seriesN = nc;
  return 0;
}

void dlg_series_enqueue(dlg_gpu* g) {
  const dlg_series_plan& q = g->seriesPlan;
  const char* b = g->seriesIn.as<char>();
  float* G = reinterpret_cast<float*>(g->seriesIn.as<char>() + q.oG);
  if (q.groundRow >= 0) k_ground_rows<<<(q.nc + 255) / 256, 256, 0, g->stream>>>(g->out.as<double>() + size_t(q.groundRow) * q.nc, q.nc, G);
  if (q.method == 1) { enqueue_mfma<2, false>(g); return; }
  if (q.method == 2) { enqueue_mfma<3, false>(g); return; }
  if (q.method == 3) { enqueue_mfma<2, true>(g); return; }
  const float3* N = g->N.as<float3>() + g->i0;
  const int T = q.T;
  k_series_f32<<<dim3((T + kTile - 1) / kTile, (q.nc + kTile - 1) / kTile), 256, 0, g->stream>>>(
      g->bits.as<uint32_t>(), uint32_t(g->bytes), q.nc, N, g->D.as<float3>(), q.nSky, reinterpret_cast<const int*>(b + q.oRow),
      reinterpret_cast<const float*>(b + q.oW), g->skyS.as<float>(), T, G, g->skyG.as<float>(), q.sun0, reinterpret_cast<const int*>(b + q.oCol),
      reinterpret_cast<const float*>(b + q.oSw), g->seriesE.as<float>());
  if (q.reflect)
    k_series_dc_f32<<<dim3((T + kTile - 1) / kTile, (q.nc + kTile - 1) / kTile), 256, 0, g->stream>>>(g->dc.as<float>(), q.nc, g->dcBins,
                                                                                                       g->skySdc.as<float>(), T, g->seriesE.as<float>());
}

bool dlg_series_tuned(const dlg_gpu* g) { return g->seriesPlan.method == 0 || g->seriesPlan.shape >= 0; }

extern "C" int32_t dlg_series(dlg_gpu* g, int64_t nc, const dlg_series_in* in, int32_t method, float* E, char* err, int32_t errLen) {
  if (nc == 0 || in->T <= 0) return 0;
  if (dlg_series_prepare(g, nc, in, method, -1, err, errLen)) return 1;
  dlg_series_plan& q = g->seriesPlan;
  if (method && q.shape < 0) {   // the first chunk of this problem: split its operands and time every tile shape on it
    q.shape = 0;
    dlg_series_enqueue(g);
    q.shape = method == 1 ? tune<2, false>(g) : method == 2 ? tune<3, false>(g) : tune<2, true>(g);
  }
  dlg_tbegin(g);
  dlg_series_enqueue(g);
  DLG_CHECK(hipGetLastError());
  const float ms = dlg_tend(g);
  g->stats.seriesMs += ms;
  g->stats.lastSeriesMs = ms;
  if (E) {
    DLG_CHECK(hipMemcpyAsync(E, g->seriesE.p, size_t(nc) * size_t(in->T) * 4, hipMemcpyDeviceToHost, g->stream));
    DLG_CHECK(hipStreamSynchronize(g->stream));
  }
  return 0;
}
