// SPDX-License-Identifier: Apache-2.0
// The cumulative reduces of a daylight run in one pass over the visibility bits: replaces k_reduce<Bits> and gpu_reduce
// (reference src/gpu.cu:99-114, 612-625), which ran once per quantity and re-read the bits each time.
// Each sensor's sum is the reference's: acc_q += double(float(max(0, dotf(n, D_d)))) x double(W[q][d]) over its visible
// directions in ascending order (sun hours: the float 1 in place of the cosine), in FP64, so the result is the same bits
// for any chunking. The order makes the sum a dependent chain; the work around it is spread over a wave:
//   a wave (one block) per sensor; per slice of 64 directions the lanes of the visible ones form their products and
//   store them compacted in direction order (position = visible lanes below), then lane q adds the slice's products of
//   quantity q in that order.
// HBM: 4 nWords bytes of bits + 24 (P, N) + 8 nQ per sensor; directions and weights stay in L2.
#include "dl_math.h"
#include "gpu_internal.h"

namespace {

constexpr int kMaxQ = 8, kWave = 64;

template <int NQ>
__global__ __launch_bounds__(kWave) void k_reduce(const uint32_t* __restrict__ bits, uint32_t nc, uint32_t nWords, const float3* __restrict__ N,
                                                  const float3* __restrict__ D, const float* __restrict__ W, uint32_t nDir,
                                                  uint32_t cosMask, double* __restrict__ out) {
  __shared__ double prod[NQ][kWave];
  const uint32_t lane = threadIdx.x, i = blockIdx.x;
  const float3 n = N[i];
  double acc = 0.0;          // lane q < NQ: quantity q
  for (uint32_t base = 0; base < nDir; base += kWave) {
    const uint32_t d = base + lane;
    const bool vis = d < nDir && ((bits[size_t(d >> 5) * nc + i] >> (d & 31)) & 1u);
    const uint64_t m = __ballot(vis);
    if (vis) {
      const uint32_t at = __popcll(m & ((uint64_t(1) << lane) - 1));
      const float c = fmaxf(dlm::dot3(n, D[d]), 0.f);
#pragma unroll
      for (int q = 0; q < NQ; ++q) prod[q][at] = (double)((cosMask >> q) & 1u ? c : 1.f) * (double)W[size_t(q) * nDir + d];
    }
    __syncthreads();
    const int cnt = __popcll(m);
    if (lane < uint32_t(NQ)) {
      const double* pq = prod[lane];
      for (int k = 0; k < cnt; ++k) acc += pq[k];
    }
    __syncthreads();
  }
  if (lane < uint32_t(NQ)) out[size_t(lane) * nc + i] = acc;
}

template <int NQ>
void launch(dlg_gpu* g, uint32_t nc, uint32_t mask) {
  k_reduce<NQ><<<nc, kWave, 0, g->stream>>>(g->bits.as<uint32_t>(), nc, uint32_t(g->nWords), g->N.as<float3>() + g->i0, g->D.as<float3>(),
                                           g->scratch.as<float>(), uint32_t(g->nDir), mask, g->out.as<double>());
}

}  // namespace

extern "C" int32_t dlg_reduce(dlg_gpu* g, int64_t nc, int32_t nQ, const float* W, const int32_t* useCos, double* out, char* err, int32_t errLen) {
  if (nc == 0 || nQ == 0) return 0;
  if (nc != g->nc) { dlg_fail(err, errLen, "dlg_reduce: no trace of this chunk"); return 1; }
  if (nQ < 0 || nQ > kMaxQ) { dlg_fail(err, errLen, "dlg_reduce: 1 to 8 quantities"); return 1; }
  DLG_CHECK(g->scratch.grow(size_t(nQ) * size_t(g->nDir) * 4));
  DLG_CHECK(g->out.grow(size_t(nQ) * size_t(nc) * 8));
  DLG_CHECK(hipMemcpyAsync(g->scratch.p, W, size_t(nQ) * size_t(g->nDir) * 4, hipMemcpyHostToDevice, g->stream));
  uint32_t mask = 0;
  for (int q = 0; q < nQ; ++q) mask |= useCos[q] ? 1u << q : 0u;
  dlg_tbegin(g);
  switch (nQ) {
    case 1: launch<1>(g, uint32_t(nc), mask); break;
    case 2: launch<2>(g, uint32_t(nc), mask); break;
    case 3: launch<3>(g, uint32_t(nc), mask); break;
    case 4: launch<4>(g, uint32_t(nc), mask); break;
    case 5: launch<5>(g, uint32_t(nc), mask); break;
    case 6: launch<6>(g, uint32_t(nc), mask); break;
    case 7: launch<7>(g, uint32_t(nc), mask); break;
    default: launch<8>(g, uint32_t(nc), mask); break;
  }
  DLG_CHECK(hipGetLastError());
  const float ms = dlg_tend(g);
  g->stats.reduceMs += ms;
  g->stats.lastReduceMs = ms;
  if (out) {
    DLG_CHECK(hipMemcpyAsync(out, g->out.p, size_t(nQ) * size_t(nc) * 8, hipMemcpyDeviceToHost, g->stream));
    DLG_CHECK(hipStreamSynchronize(g->stream));
  }
  return 0;
}
