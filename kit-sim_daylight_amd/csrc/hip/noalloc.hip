// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// A daylight chunk without a series - the trace, then the reduce of every quantity, the kernels dlg_chunk runs - over
// sensors, weights and sums that the caller holds in device memory, with every intermediate in one fixed device array
// of this file: the call allocates nothing. It has no counterpart in the reference; it is for callers that cannot
// allocate device memory (code under stream capture, serving loops that forbid allocation).
//
// The array holds the traced-ray count and the visibility (bits [nWords x nc], or byte mode's quads [8 nWords x nc]).
// There is one per device, held until the process ends, shared by every call on that device and keeping the last call's
// contents: the count is cleared by a kernel at each call and the visibility is written whole before it is read, so a
// call depends on nothing an earlier one left. Calls that overlap on one device race on the array: they must be
// ordered (one stream, or the caller's waits).
#include <mutex>
#include <string>

#include "gpu_internal.h"

namespace {

constexpr size_t kCountBytes = 256;              // the traced-ray count (8 bytes), padded to the visibility's alignment
constexpr size_t kVisBytes = size_t(64) << 20;   // the largest visibility of a call
constexpr int kMaxDevices = 64;

}  // namespace

// the fixed device array; host code uses the address that hipGetSymbolAddress gives for it on each device, never this name
static __device__ __align__(256) unsigned char scratch[kCountBytes + kVisBytes];

namespace {

std::mutex scratchMu;
void* scratchOf[kMaxDevices] = {};   // per device: the array's address there

__global__ void k_clear_count(unsigned long long* c) {
  if (threadIdx.x == 0) c[0] = 0ull;
}

}  // namespace

int32_t dlg_noalloc_setup(dlg_gpu* g, char* err, int32_t errLen) {
  int dev = 0;
  DLG_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) { dlg_fail(err, errLen, "dlg_init: device " + std::to_string(dev) + " beyond the array table"); return 1; }
  std::lock_guard<std::mutex> lk(scratchMu);
  if (!scratchOf[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      dlg_fail(err, errLen, std::string("dlg_init: the device array's address: ") + hipGetErrorString(e));
      return 1;
    }
    scratchOf[dev] = p;
  }
  g->noalloc = scratchOf[dev];
  return 0;
}

extern "C" {

int64_t dlg_chunk_noalloc_limit(void) { return int64_t(kVisBytes); }

int32_t dlg_chunk_noalloc(dlg_gpu* g, const float* P, const float* N, int64_t nc, uint32_t rayFlags, int32_t nQ, const float* W,
                          const int32_t* useCos, double* sums, int32_t wait, char* err, int32_t errLen) {
  if (!g->noalloc) { dlg_fail(err, errLen, "dlg_chunk_noalloc: the device array's address is unknown"); return 1; }
  if (nQ < 1 || nQ > 8) { dlg_fail(err, errLen, "dlg_chunk_noalloc: 1 to 8 rows of weights"); return 1; }
  if (nc < 0) { dlg_fail(err, errLen, "dlg_chunk_noalloc: a negative sensor count"); return 1; }
  if (nc == 0) return 0;
  if (g->nWords == 0) { dlg_fail(err, errLen, "dlg_chunk_noalloc: no directions (dlg_set_scene)"); return 1; }
  if (!P || !N || !W || !sums || !useCos) { dlg_fail(err, errLen, "dlg_chunk_noalloc: a null buffer"); return 1; }
  int stack = 0;
  if (dlg_trace_check(g, rayFlags, stack, err, errLen)) return 1;
  const bool bytes = !(rayFlags & 1u);
  const uint64_t perSensor = uint64_t(g->nWords) * (bytes ? 32 : 4);
  if (uint64_t(nc) > uint64_t(kVisBytes) / perSensor) {
    dlg_fail(err, errLen, "dlg_chunk_noalloc: " + std::to_string(nc) + " sensors x " + std::to_string(perSensor) + " bytes of visibility exceed the " +
                              std::to_string(kVisBytes) + "-byte device array (at most " + std::to_string(uint64_t(kVisBytes) / perSensor) +
                              " sensors a call)");
    return 1;
  }
  uint32_t mask = 0;
  for (int q = 0; q < nQ; ++q) mask |= useCos[q] ? 1u << q : 0u;
  unsigned long long* count = static_cast<unsigned long long*>(g->noalloc);
  uint32_t* vis = reinterpret_cast<uint32_t*>(static_cast<char*>(g->noalloc) + kCountBytes);
  k_clear_count<<<1, 64, 0, g->stream>>>(count);
  dlg_trace_enqueue_into(g, P, N, uint32_t(nc), rayFlags, stack, vis, count);
  dlg_reduce_enqueue_from(g, vis, bytes, uint32_t(nc), N, W, nQ, mask, sums);
  DLG_CHECK(hipGetLastError());
  if (wait) DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"
