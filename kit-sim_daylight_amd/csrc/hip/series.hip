// SPDX-License-Identifier: Apache-2.0
// The time-series contraction E = A S + G g + sun of a daylight chunk: replaces k_series / gpu_series (reference
// src/gpu.cu:119-149, 639-669). A[i, d] = max(0, dotf(n_i, D_d)) wsky_d on the visible sky directions (the daylight
// coefficients, exact 0 where a direction is blocked), S[row_d, t] the patch radiance of the non-night steps (dlg_sky).
//   method 0, FP32: each E[i, t] an fma chain over d in ascending order, then fma(G_i, g_t, .), then the sun's fma at its
//             step: the reference kernel's arithmetic (nvcc contracts it so), tiled 64 x 64 through LDS.
//   method 1, BF16x3: A and S split into bf16 hi + lo; hi.hi + hi.lo + lo.hi on MFMA (v_mfma_f32_32x32x8 bf16),
//             FP32 accumulation over blocks of 32 directions, the blocks summed in FP64; epilogue in FP64.
//   method 2, BF16x6: hi + mid + lo; the six products down to 2^-16 relative, as method 1 otherwise: FP32-level.
// Layout: A planes [nc x Kp] bf16 row-major (Kp = nSky rounded up to 32, zero padded); S^T planes [Tp x Kp] so that a lane's
// four k values of an MFMA operand are one 8-byte load for both operands.
#include <vector>

#include "dl_math.h"
#include "gpu_internal.h"

namespace {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ uint16_t bf16_rn(float f) {   // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(f);
  return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__device__ __forceinline__ bool vis_bit(const uint32_t* __restrict__ bits, uint32_t nc, uint32_t i, uint32_t d) {
  return (bits[size_t(d >> 5) * nc + i] >> (d & 31)) & 1u;
}

// the daylight coefficient of (sensor i, sky direction d), as k_series forms it (gpu.cu:137)
__device__ __forceinline__ float coeff(const uint32_t* __restrict__ bits, uint32_t nc, uint32_t i, uint32_t d, float3 n,
                                       const float3* __restrict__ D, const float* __restrict__ wsky) {
  if (!vis_bit(bits, nc, i, d)) return 0.f;
  return (fmaxf(dlm::dot3(n, D[d]), 0.f) * 1.f) * wsky[d];
}

// ---------------------------------------------------------------------------------------------- method 0: FP32
constexpr int kTile = 64, kKt = 16;
__global__ __launch_bounds__(256) void k_series_f32(const uint32_t* __restrict__ bits, uint32_t nc, const float3* __restrict__ N,
                                                    const float3* __restrict__ D, uint32_t nSky, const int* __restrict__ row,
                                                    const float* __restrict__ wsky, const float* __restrict__ S, int T, float* __restrict__ E) {
  __shared__ float As[kKt][kTile];
  __shared__ float Bs[kKt][kTile + 4];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const uint32_t i0 = blockIdx.y * kTile, t0 = blockIdx.x * kTile;
  float e[4][4] = {};
  for (uint32_t d0 = 0; d0 < nSky; d0 += kKt) {
    for (int k = threadIdx.x; k < kKt * kTile; k += 256) {   // A tile: 16 directions x 64 sensors, from the bits
      const int dd = k / kTile, ii = k % kTile;
      const uint32_t d = d0 + dd, i = i0 + ii;
      As[dd][ii] = (d < nSky && i < nc) ? coeff(bits, nc, i, d, N[i], D, wsky) : 0.f;
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
    for (int c = 0; c < 4; ++c) {
      const int t = int(t0) + tx * 4 + c;
      if (t < T) E[size_t(i) * T + t] = e[r][c];
    }
  }
}

// E += G g (fma), then the sun's term at its step (fma), per the reference's order
__global__ void k_series_f32_tail(const uint32_t* __restrict__ bits, uint32_t nc, const float3* __restrict__ N, const float3* __restrict__ D,
                                  const float* __restrict__ G, const float* __restrict__ g, int T, uint32_t sun0, uint32_t nSun,
                                  const int* __restrict__ sunStep, const float* __restrict__ sunW, float* __restrict__ E) {
  const uint32_t i = blockIdx.x;
  if (i >= nc) return;
  float* Ei = E + size_t(i) * T;
  for (int t = threadIdx.x; t < T; t += blockDim.x) Ei[t] = fmaf(G[i], g[t], Ei[t]);
  __syncthreads();
  if (threadIdx.x == 0) {
    const float3 n = N[i];
    for (uint32_t j = 0; j < nSun; ++j)
      if (vis_bit(bits, nc, i, sun0 + j)) Ei[sunStep[j]] = fmaf(fmaxf(dlm::dot3(n, D[sun0 + j]), 0.f) * 1.f, sunW[j], Ei[sunStep[j]]);
  }
}

// ---------------------------------------------------------------------------------------------- methods 1 / 2: MFMA
// split x into NS bf16 pieces, each the rounding of what the previous ones leave
template <int NS>
__device__ __forceinline__ void split(float x, uint16_t* out) {
  float r = x;
#pragma unroll
  for (int s = 0; s < NS; ++s) { out[s] = bf16_rn(r); r = r - bf16_f(out[s]); }
}

template <int NS>
__global__ void k_split_a(const uint32_t* __restrict__ bits, uint32_t nc, const float3* __restrict__ N, const float3* __restrict__ D,
                          uint32_t nSky, uint32_t Kp, const float* __restrict__ wsky, uint16_t* __restrict__ A) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (d >= Kp || i >= nc) return;
  const float c = d < nSky ? coeff(bits, nc, i, d, N[i], D, wsky) : 0.f;
  uint16_t p[NS];
  split<NS>(c, p);
  const size_t plane = size_t(nc) * Kp;
#pragma unroll
  for (int s = 0; s < NS; ++s) A[s * plane + size_t(i) * Kp + d] = p[s];
}

template <int NS>
__global__ void k_split_s(const float* __restrict__ S, const int* __restrict__ row, uint32_t nSky, uint32_t Kp, int T, uint32_t Tp,
                          uint16_t* __restrict__ B) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x, t = blockIdx.y;
  if (d >= Kp || t >= Tp) return;
  const float s = (d < nSky && int(t) < T) ? S[size_t(row[d]) * T + t] : 0.f;
  uint16_t p[NS];
  split<NS>(s, p);
  const size_t plane = size_t(Tp) * Kp;
#pragma unroll
  for (int k = 0; k < NS; ++k) B[k * plane + size_t(t) * Kp + d] = p[k];
}

// Block: 4 waves, 64 sensors x 64 steps; wave w: rows 32 (w & 1), columns 32 (w >> 1). v_mfma_f32_32x32x8 bf16 operand
// layout: lane l holds A[l % 32][4 (l / 32) + v] and B[4 (l / 32) + v][l % 32], v = 0..3; result acc[v] is
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
template <int NS>
__global__ __launch_bounds__(256) void k_series_mfma(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B, uint32_t nc, uint32_t Kp,
                                                     uint32_t Tp, int T, const uint32_t* __restrict__ bits, const float3* __restrict__ N,
                                                     const float3* __restrict__ D, const float* __restrict__ G, const float* __restrict__ g,
                                                     uint32_t sun0, const int* __restrict__ sunOfCol, const float* __restrict__ sunW,
                                                     float* __restrict__ E) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t i0 = blockIdx.y * 64 + 32 * (w & 1), t0 = blockIdx.x * 64 + 32 * (w >> 1);
  const uint32_t ia = i0 + (lane & 31), tb = t0 + (lane & 31), kq = 4 * (lane >> 5);
  const bool aOk = ia < nc;
  const size_t aPlane = size_t(nc) * Kp, bPlane = size_t(Tp) * Kp;
  const uint16_t* Ar = A + size_t(aOk ? ia : 0) * Kp + kq;
  const uint16_t* Br = B + size_t(tb) * Kp + kq;
  double acc64[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) acc64[v] = 0.0;
  for (uint32_t k0 = 0; k0 < Kp; k0 += 32) {
    f32x16 acc = {};
#pragma unroll
    for (uint32_t kk = 0; kk < 32; kk += 8) {
      bf16x4 a[NS], b[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        a[s] = aOk ? *reinterpret_cast<const bf16x4*>(Ar + s * aPlane + k0 + kk) : bf16x4{0, 0, 0, 0};
        b[s] = *reinterpret_cast<const bf16x4*>(Br + s * bPlane + k0 + kk);
      }
      // products whose exponent sum is at most NS - 1 pieces below the leading one, smallest first
#pragma unroll
      for (int o = NS - 1; o >= 0; --o)
#pragma unroll
        for (int s = 0; s <= o; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[s], b[o - s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) acc64[v] += (double)acc[v];
  }
  const uint32_t t = t0 + (lane & 31);
  if (int(t) >= T) return;
  const int j = sunOfCol[t];
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const uint32_t i = i0 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
    if (i >= nc) continue;
    double e = acc64[v] + (double)G[i] * (double)g[t];
    if (j >= 0 && vis_bit(bits, nc, i, sun0 + uint32_t(j))) e += (double)fmaxf(dlm::dot3(N[i], D[sun0 + uint32_t(j)]), 0.f) * (double)sunW[j];
    E[size_t(i) * T + t] = (float)e;
  }
}

template <int NS>
int32_t run_mfma(dlg_gpu* g, uint32_t nc, const dlg_series_in* in, const float* dG, const float* dg, const int* dRow, const float* dWsky,
                 const int* dSunOfCol, const float* dSunW, char* err, int32_t errLen) {
  const uint32_t Kp = (in->nSky + 31) / 32 * 32, T = uint32_t(in->T), Tp = (T + 63) / 64 * 64;
  const size_t aBytes = size_t(NS) * nc * Kp * 2, bBytes = size_t(NS) * Tp * Kp * 2;
  DLG_CHECK(g->seriesWork.grow(aBytes + bBytes));
  uint16_t* A = g->seriesWork.as<uint16_t>();
  uint16_t* B = A + aBytes / 2;
  const float3* N = g->N.as<float3>() + g->i0;
  k_split_a<NS><<<dim3((Kp + 255) / 256, nc), 256, 0, g->stream>>>(g->bits.as<uint32_t>(), nc, N, g->D.as<float3>(), in->nSky, Kp, dWsky, A);
  k_split_s<NS><<<dim3((Kp + 255) / 256, Tp), 256, 0, g->stream>>>(g->skyS.as<float>(), dRow, in->nSky, Kp, in->T, Tp, B);
  k_series_mfma<NS><<<dim3(Tp / 64, (nc + 63) / 64), 256, 0, g->stream>>>(A, B, nc, Kp, Tp, in->T, g->bits.as<uint32_t>(), N, g->D.as<float3>(),
                                                                          dG, dg, in->sun0, dSunOfCol, dSunW, g->seriesE.as<float>());
  DLG_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" int32_t dlg_series(dlg_gpu* g, int64_t nc, const dlg_series_in* in, int32_t method, float* E, char* err, int32_t errLen) {
  if (nc == 0 || in->T <= 0) return 0;
  if (nc != g->nc) { dlg_fail(err, errLen, "dlg_series: no trace of this chunk"); return 1; }
  if (in->P != g->skyP || in->T != g->skyT) { dlg_fail(err, errLen, "dlg_series: S of another sky (run dlg_sky with series on)"); return 1; }
  const int T = in->T;
  // inputs: row | wsky | G | sunStep-by-column | sunW (each 256-byte aligned), in the scratch buffer
  size_t off = 0;
  const auto take = [&](size_t bytes) { const size_t at = off; off += (bytes + 255) & ~size_t(255); return at; };
  const size_t oRow = take(4 * size_t(in->nSky)), oW = take(4 * size_t(in->nSky)), oG = take(4 * size_t(nc)), oCol = take(4 * size_t(T)),
               oSw = take(4 * size_t(in->nSun ? in->nSun : 1)), oSs = take(4 * size_t(in->nSun ? in->nSun : 1));
  DLG_CHECK(g->scratch.grow(off));
  char* b = g->scratch.as<char>();
  std::vector<int> sunOfCol(static_cast<size_t>(T), -1);
  for (uint32_t j = 0; j < in->nSun; ++j)   // suns of night steps carry weight 0 and column 0: they add nothing
    if (in->sunW[j] != 0.f) sunOfCol[size_t(in->sunStep[j])] = int(j);
  DLG_CHECK(hipMemcpyAsync(b + oRow, in->row, 4 * size_t(in->nSky), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipMemcpyAsync(b + oW, in->wsky, 4 * size_t(in->nSky), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipMemcpyAsync(b + oG, in->G, 4 * size_t(nc), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipMemcpyAsync(b + oCol, sunOfCol.data(), 4 * size_t(T), hipMemcpyHostToDevice, g->stream));
  if (in->nSun) {
    DLG_CHECK(hipMemcpyAsync(b + oSw, in->sunW, 4 * size_t(in->nSun), hipMemcpyHostToDevice, g->stream));
    DLG_CHECK(hipMemcpyAsync(b + oSs, in->sunStep, 4 * size_t(in->nSun), hipMemcpyHostToDevice, g->stream));
  }
  DLG_CHECK(g->seriesE.grow(size_t(nc) * size_t(T) * 4));
  g->seriesN = nc;
  const float* dg = g->skyG.as<float>();
  dlg_tbegin(g);
  if (method == 0) {
    const float3* N = g->N.as<float3>() + g->i0;
    k_series_f32<<<dim3((T + kTile - 1) / kTile, unsigned((nc + kTile - 1) / kTile)), 256, 0, g->stream>>>(
        g->bits.as<uint32_t>(), uint32_t(nc), N, g->D.as<float3>(), in->nSky, reinterpret_cast<const int*>(b + oRow),
        reinterpret_cast<const float*>(b + oW), g->skyS.as<float>(), T, g->seriesE.as<float>());
    k_series_f32_tail<<<unsigned(nc), 256, 0, g->stream>>>(g->bits.as<uint32_t>(), uint32_t(nc), N, g->D.as<float3>(),
                                                           reinterpret_cast<const float*>(b + oG), dg, T, in->sun0, in->nSun,
                                                           reinterpret_cast<const int*>(b + oSs), reinterpret_cast<const float*>(b + oSw),
                                                           g->seriesE.as<float>());
    DLG_CHECK(hipGetLastError());
  } else if (method == 1 || method == 2) {
    const int32_t rc = (method == 1 ? run_mfma<2> : run_mfma<3>)(g, uint32_t(nc), in, reinterpret_cast<const float*>(b + oG), dg,
                                                              reinterpret_cast<const int*>(b + oRow), reinterpret_cast<const float*>(b + oW),
                                                              reinterpret_cast<const int*>(b + oCol), reinterpret_cast<const float*>(b + oSw),
                                                              err, errLen);
    if (rc) return rc;
  } else { dlg_fail(err, errLen, "dlg_series: method 0 (FP32), 1 (BF16x3) or 2 (BF16x6)"); return 1; }
  const float ms = dlg_tend(g);
  g->stats.seriesMs += ms;
  g->stats.lastSeriesMs = ms;
  if (E) {
    DLG_CHECK(hipMemcpyAsync(E, g->seriesE.p, size_t(nc) * size_t(T) * 4, hipMemcpyDeviceToHost, g->stream));
    DLG_CHECK(hipStreamSynchronize(g->stream));
  }
  return 0;
}
