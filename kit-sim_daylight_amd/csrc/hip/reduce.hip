// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The cumulative reduces of a daylight run in one pass over the visibility: replaces k_reduce<Bits> / k_reduce<Bytes> and
// gpu_reduce (reference src/gpu.cu:76-114, 612-625), which ran once per quantity and re-read the visibility each time.
// Each sensor's sum is the reference's: acc_q += double(float(max(0, dotf(n, D_d)) x T)) x double(W[q][d]) over its
// visible directions in ascending order (sun hours: the float 1 in place of the cosine; T = 1 for bits, byte / 255.f in
// byte mode), in FP64, so the result is the same bits for any chunking. The order makes each sum a dependent chain; the
// chains of different sensors run side by side:
//   a lane per sensor; every lane of a wave walks the directions together (the direction and its weights are the same
//   for all: staged in LDS per chunk of 256 directions, the weights widened to double there, and read as broadcasts),
//   reads its visibility word once per 32 directions and adds the product where its bit is set. Words that no lane of
//   the wave sees through are skipped. The lanes take the sensors in the trace's order (dlg_gpu::tileOrder, when the
//   chunk has one): a wave's 64 sensors are then a compact patch of one orientation, which see mostly the same
//   directions, so the wave steps over fewer directions that only some of its lanes see. Each sum is the same.
// HBM: 4 nWords bytes of bits (byte mode: 32 nWords) + 12 (N) + 8 nQ per sensor; directions and weights stay in L2.
#include "dl_math.h"
#include "gpu_internal.h"

namespace {

constexpr int kMaxQ = 8, kBlock = 256;

template <int NQ, bool BYTES>
__global__ __launch_bounds__(kBlock) void k_reduce(const uint32_t* __restrict__ vis, uint32_t nc, uint32_t nWords, const float3* __restrict__ N,
                                                   const float3* __restrict__ D, const float* __restrict__ W, uint32_t nDir,
                                                   uint32_t cosMask, const uint32_t* __restrict__ order, double* __restrict__ out) {
  // chunks of 8 words: their 256 directions and weights staged in LDS (every lane reads the same entry: a broadcast);
  // bits: the 8 visibility words of the lane loaded up front; bytes: per word its 8 quads, the fraction of a byte from a
  // table in LDS
  constexpr uint32_t kChunkW = 8, kChunkD = kChunkW * 32;
  __shared__ float sD[kChunkD * 3];
  __shared__ double sW[NQ][kChunkD];
  __shared__ float sT[BYTES ? 256 : 1];
  const uint32_t pos = blockIdx.x * kBlock + threadIdx.x;
  const bool live = pos < nc;
  const uint32_t i = live && order ? order[pos] : pos;
  const float3 n = live ? N[i] : make_float3(0.f, 0.f, 0.f);
  if (BYTES) sT[threadIdx.x] = dlm::byte_T(threadIdx.x);   // ordered before its reads by the chunk loop's barriers
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  const auto term = [&](uint32_t e, float T) {
    const float c = fmaxf(dlm::dot3(n, make_float3(sD[3 * e], sD[3 * e + 1], sD[3 * e + 2])), 0.f);
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] += (double)(((cosMask >> q) & 1u ? c : 1.f) * T) * sW[q][e];
  };
  for (uint32_t w0 = 0; w0 < nWords; w0 += kChunkW) {
    const uint32_t d0 = w0 * 32, dn = nDir - d0 < kChunkD ? nDir - d0 : kChunkD;
    uint32_t words[kChunkW];
    if (!BYTES) {
#pragma unroll
      for (uint32_t j = 0; j < kChunkW; ++j) words[j] = (live && w0 + j < nWords) ? vis[size_t(w0 + j) * nc + i] : 0u;
    }
    __syncthreads();   // the previous chunk's reads are done
    for (uint32_t k = threadIdx.x; k < dn; k += kBlock) {
      const float3 v = D[d0 + k];
      sD[3 * k] = v.x; sD[3 * k + 1] = v.y; sD[3 * k + 2] = v.z;
#pragma unroll
      for (int q = 0; q < NQ; ++q) sW[q][k] = (double)W[size_t(q) * nDir + d0 + k];
    }
    __syncthreads();
    if (BYTES) {
      for (uint32_t j = 0; j < kChunkW && 32 * j < dn; ++j) {
        uint32_t quad[8];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) quad[u] = live ? vis[size_t(8 * (w0 + j) + u) * nc + i] : 0u;
        uint32_t any = 0;
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) any |= quad[u];
        if (!__any(any != 0u)) continue;
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u)
#pragma unroll
          for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t v = (quad[u] >> (8 * b)) & 255u;   // 0 beyond the last direction
            if (v) term(32 * j + 4 * u + b, sT[v]);
          }
      }
      continue;
    }
#pragma unroll
    for (uint32_t j = 0; j < kChunkW; ++j) {
      const uint32_t word = words[j], base = 32 * j;
      if (base >= dn || !__any(word != 0u)) continue;
      const uint32_t cnt = dn - base < 32 ? dn - base : 32;
#pragma unroll 4
      for (uint32_t k = 0; k < cnt; ++k) {
        if (!((word >> k) & 1u)) continue;
        term(base + k, 1.f);
      }
    }
  }
  if (live) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[size_t(q) * nc + i] = acc[q];
  }
}

template <int NQ>
void launch(dlg_gpu* g, const uint32_t* vis, bool bytes, uint32_t nc, const float3* N, const float* W, uint32_t mask, const uint32_t* order,
            double* out) {
  const auto k = bytes ? k_reduce<NQ, true> : k_reduce<NQ, false>;
  k<<<(nc + kBlock - 1) / kBlock, kBlock, 0, g->stream>>>(vis, nc, uint32_t(g->nWords), N, g->D.as<float3>(), W, uint32_t(g->nDir), mask, order,
                                                          out);
}

void launch_nq(dlg_gpu* g, const uint32_t* vis, bool bytes, uint32_t nc, const float3* N, const float* W, int32_t nQ, uint32_t mask,
               const uint32_t* order, double* out) {
  switch (nQ) {
    case 1: launch<1>(g, vis, bytes, nc, N, W, mask, order, out); break;
    case 2: launch<2>(g, vis, bytes, nc, N, W, mask, order, out); break;
    case 3: launch<3>(g, vis, bytes, nc, N, W, mask, order, out); break;
    case 4: launch<4>(g, vis, bytes, nc, N, W, mask, order, out); break;
    case 5: launch<5>(g, vis, bytes, nc, N, W, mask, order, out); break;
    case 6: launch<6>(g, vis, bytes, nc, N, W, mask, order, out); break;
    case 7: launch<7>(g, vis, bytes, nc, N, W, mask, order, out); break;
    default: launch<8>(g, vis, bytes, nc, N, W, mask, order, out); break;
  }
}

}  // namespace

int32_t dlg_reduce_prepare(dlg_gpu* g, int64_t nc, int32_t nQ, const float* W, const int32_t* useCos, uint32_t& mask, char* err, int32_t errLen) {
  if (nc != g->nc) { dlg_fail(err, errLen, "dlg_reduce: no trace of this chunk"); return 1; }
  if (nQ < 0 || nQ > kMaxQ) { dlg_fail(err, errLen, "dlg_reduce: 1 to 8 quantities"); return 1; }
  DLG_CHECK(g->scratch.grow(size_t(nQ) * size_t(g->nDir) * 4));
  DLG_CHECK(g->out.grow(size_t(nQ) * size_t(nc) * 8));
  DLG_CHECK(hipMemcpyAsync(g->scratch.p, W, size_t(nQ) * size_t(g->nDir) * 4, hipMemcpyHostToDevice, g->stream));
  mask = 0;
  for (int q = 0; q < nQ; ++q) mask |= useCos[q] ? 1u << q : 0u;
  return 0;
}

void dlg_reduce_enqueue(dlg_gpu* g, uint32_t nc, int32_t nQ, uint32_t mask) {
  const bool ordered = g->orderGen == g->sensGen && g->orderI0 == g->i0 && g->orderN == int64_t(nc);   // the trace's order of this chunk
  launch_nq(g, g->bits.as<uint32_t>(), g->bytes != 0, nc, g->N.as<float3>() + g->i0, g->scratch.as<float>(), nQ, mask,
            ordered ? g->tileOrder.as<uint32_t>() : nullptr, g->out.as<double>());
}

void dlg_reduce_enqueue_from(dlg_gpu* g, const uint32_t* vis, bool bytes, uint32_t nc, const float* N, const float* W, int32_t nQ, uint32_t mask,
                             double* out) {
  launch_nq(g, vis, bytes, nc, reinterpret_cast<const float3*>(N), W, nQ, mask, nullptr, out);
}

extern "C" int32_t dlg_reduce(dlg_gpu* g, int64_t nc, int32_t nQ, const float* W, const int32_t* useCos, double* out, char* err, int32_t errLen) {
  if (nc == 0 || nQ == 0) return 0;
  uint32_t mask = 0;
  if (dlg_reduce_prepare(g, nc, nQ, W, useCos, mask, err, errLen)) return 1;
  dlg_tbegin(g);
  dlg_reduce_enqueue(g, uint32_t(nc), nQ, mask);
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
