// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// A daylight chunk without a series - the trace, then the reduce of every quantity, the kernels dlg_chunk runs - over
// sensors, weights and sums that the caller holds in device memory, with every intermediate in one fixed device array
// of this file: the call allocates nothing. It has no counterpart in the reference; it is for callers that cannot
// allocate device memory (code under stream capture, serving loops that forbid allocation).
//
// The array holds the traced-ray count and the trace's tile counters, the visibility (bits [nWords x nc], or byte mode's quads [8 nWords x nc]) and
// the traversal stacks' entries beyond their LDS part (one area per block of the persistent grid; the grid is cut to the
// areas that fit, which deep trees - many entries per area - pay for in waves, not in results). There is one array per
// device, held until the process ends, shared by every call on that device and keeping the last call's contents: the
// counts are cleared by kernels at each call, the visibility is written whole before it is read and a stack entry is read
// only after its lane wrote it, so a call depends on nothing an earlier one left. Calls that overlap on one device race
// on the array: they must be ordered (one stream, or the caller's waits).
#include <mutex>
#include <string>

#include "gpu_internal.h"

namespace {

constexpr size_t kCountBytes = 1024;             // the traced-ray count (8 bytes) at 0, the tile counters at 512
constexpr size_t kVisBytes = size_t(64) << 20;   // the largest visibility of a call
constexpr size_t kOvfBytes = size_t(32) << 20;   // the stacks' global areas
constexpr int kMaxDevices = 64;

}  // namespace

// the fixed device array; host code uses the address that hipGetSymbolAddress gives for it on each device, never this name
static __device__ __align__(256) unsigned char scratch[kCountBytes + kVisBytes + kOvfBytes];

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
  uint32_t* ovf = reinterpret_cast<uint32_t*>(static_cast<char*>(g->noalloc) + kCountBytes + kVisBytes);
  dlg_trace_shape sh = dlg_trace_grid(g, nc, stack, bytes, false);
  if (sh.ovfPerBlock && sh.grid * sh.ovfPerBlock > kOvfBytes) sh.grid = uint32_t(kOvfBytes / sh.ovfPerBlock / 8 * 8);
  k_clear_count<<<1, 64, 0, g->stream>>>(count);
  uint32_t* next = reinterpret_cast<uint32_t*>(static_cast<char*>(g->noalloc) + 512);
  static_assert(512 + dlg_tile_counter_bytes <= kCountBytes, "the counters fit their region");
  dlg_trace_enqueue_into(g, P, N, uint32_t(nc), rayFlags, stack, vis, count, ovf, sh.grid, next);
  dlg_reduce_enqueue_from(g, vis, bytes, uint32_t(nc), N, W, nQ, mask, sums);
  DLG_CHECK(hipGetLastError());
  if (wait) DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

}  // extern "C"
