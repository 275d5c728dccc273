// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The device state behind dl_device.h, shared by the HIP translation units of the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "dl_device.h"

#define DLG_CHECK(call)                                                                                             \
  do {                                                                                                              \
    hipError_t e_ = (call);                                                                                         \
    if (e_ != hipSuccess) { dlg_fail(err, errLen, std::string(#call) + ": " + hipGetErrorString(e_)); return 1; } \
  } while (0)

inline void dlg_fail(char* err, int32_t errLen, const std::string& msg) {
  if (err && errLen > 0) {
    const size_t n = msg.size() < size_t(errLen - 1) ? msg.size() : size_t(errLen - 1);
    for (size_t k = 0; k < n; ++k) err[k] = msg[k];
    err[n] = 0;
  }
}

// A device buffer that grows (never shrinks until trimmed).
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t grow(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e == hipSuccess) cap = bytes;
    return e;
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; cap = 0; }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// A built occluder scene: LBVH nodes (16 floats each), their 4-wide records (32 floats each, wide.h), leaf triangles
// (12 floats each: v0, v1, v2, the original index and the global slot as bits, 1 unused), its padding and height, and
// the largest slot its leaves name.
struct dlg_scene {
  DevBuf nodes, wide, leaves;
  int64_t nNodes = 0, nTris = 0;
  int32_t height = 0;
  float pad = 0.f;
  uint32_t slotMax = 0;
  bool borrowed = false;   // another context's arrays (dlg_borrow_scene): never freed here
  int64_t bytes() const { return int64_t(nodes.cap + wide.cap + leaves.cap); }
  void release() {
    if (!borrowed) { nodes.release(); wide.release(); leaves.release(); }
    *this = dlg_scene{};
  }
};

// the series of the current chunk as dlg_series_prepare set it up (offsets into dlg_gpu::seriesIn)
struct dlg_series_plan {
  int32_t method = 0, groundRow = -1, reflect = 0, T = 0;
  int32_t shape = -1;   // the MFMA methods' tile shape (series.hip kShapes), -1 until tuned
  uint32_t nc = 0, nSky = 0, nBins = 0, sun0 = 0;
  size_t oRow = 0, oW = 0, oG = 0, oCol = 0, oSw = 0, oMx = 0;
};

struct dlg_graph_cache;   // chunk.hip: the chunk's captured graph and its events

struct dlg_gpu {
  hipStream_t stream = nullptr;
  dlg_graph_cache* graph = nullptr;
  int32_t graphBroken = 0;   // a capture failed: chunks run their stages one by one
  double* hostSums = nullptr;   // pinned host memory for a chunk's sums (dlg_host_sums)
  size_t hostSumsBytes = 0;
  hipEvent_t ev[2] = {};
  dlg_scene scene;
  // resident sensors and directions; per direction its ray constants (dl_math.h dir_pack: 2 float4)
  DevBuf P, N, D, dirTab;
  int64_t nSens = 0, nDir = 0, nWords = 0;
  int64_t i0 = 0, nc = 0;          // the last traced chunk
  DevBuf bits;                     // [nWords x nc] bits, or byte mode's [8 nWords x nc] quads of transmittance bytes
  int32_t bytes = 0;               // the last trace wrote bytes
  DevBuf slots;                    // dlg_set_slots: [nSlots] dlg_slot
  int64_t nSlots = 0;
  // reflected light (dlg_bounce): rows dc[b nc + i] of the chunk, and the sky's bin radiance matrix Sdc [(P + 1) x Tup]
  DevBuf dc, skySdc;
  uint32_t dcBins = 0;
  int64_t dcN = 0, dcI0 = 0;
  DevBuf scratch, out;             // per-call inputs / outputs of the reduces
  // sky (dlg_sky): S [P x Tup] and g [Tup] of the last run, the tcos table
  DevBuf skyS, skyG, skyWork;
  int32_t skyP = 0, skyT = 0, skyReflect = 0;
  // series: the output, the split operands, the inputs (seriesPlan) and the host copy of the sun per column
  DevBuf seriesE, seriesWork, seriesIn;
  int64_t seriesN = 0;
  dlg_series_plan seriesPlan;
  std::vector<int> seriesCols;
  // views: classes | distances | row sines of the last dlg_view; the view rays' stack entries past the LDS part, one
  // area per block of their persistent grid
  DevBuf view, labels, viewOvf;
  int64_t viewN = 0, viewK = 0, viewI0 = 0;
  dlg_view_in viewIn{};
  dlg_stats_t stats{};
  int32_t cus = 0;
  // the traversal's counting variant (dlg_count_fetches) and its two counters
  int32_t countFetches = 0;
  DevBuf counts;
  // the rays traced past the hemisphere cull since the context was made (k_bits / k_bytes add to it)
  DevBuf traceCount;
  // the traversal stacks' entries past the LDS part, one area per block of the persistent grid, and that grid
  DevBuf stackOvf;
  uint32_t traceGrid = 0;
  DevBuf tileNext;   // the persistent trace's tile counters
  // The order the trace takes the sensors of a chunk in (dlg_trace_prepare): chunk-local indices sorted by the normal's
  // rounded components and the Morton code of the position's 1 m cell, so that a wave's 64 sensors are a compact patch
  // of one orientation. For the chunk [orderI0, orderI0 + orderN) of the sensors of generation orderGen; sensGen counts
  // every write of P / N.
  DevBuf tileOrder, orderWork;
  int64_t orderI0 = -1, orderN = -1, orderGen = -1, sensGen = 0;
  DevBuf sensWork;   // dlg_sensors: a sensor mesh on the device and its work arrays
  // the tree dlg_build makes (dlg_set_tree): DLG_TREE_SAH or DLG_TREE_LBVH
  int32_t tree = DLG_TREE_SAH;
  // mesh ingest (dlg_mesh_*): staged vertices [meshNv] float3, triangles [meshNt] uint3, their global slots, the
  // largest |coordinate| as float bits and a fault word; the per-mesh temporaries
  DevBuf meshV, meshI, meshS, meshWord, meshWork;
  int64_t meshNv = 0, meshNt = 0, meshCapV = 0, meshCapT = 0;
  int32_t meshSlotted = 0, meshOpen = 0;
  double meshC[3] = {0, 0, 0};
  // the device address of noalloc.hip's fixed array on this context's device (dlg_noalloc_setup)
  void* noalloc = nullptr;
};

// noalloc.hip: at context creation, with its device current: the address of the fixed array of dlg_chunk_noalloc
int32_t dlg_noalloc_setup(dlg_gpu* g, char* err, int32_t errLen);

// chunk.hip: frees the chunk's graph
void dlg_graph_free(dlg_gpu* g);

// lbvh.hip: the build over triangles already on the device (v [nv] float3, tri [nt] uint3), mx the largest |coordinate|
int32_t dlg_build_device(dlg_gpu* g, const float* v, int64_t nv, const uint32_t* tri, int64_t nt, float mx, char* err, int32_t errLen);
// gpu_ctx.hip: leaf slots from per-triangle slots (input order) already on the device, largest slot mx
int32_t dlg_leaf_slots_device(dlg_gpu* g, const uint32_t* slot, int64_t nt, uint32_t mx, char* err, int32_t errLen);

// sah.hip: the binned-SAH tree over nt >= 2 leaves in Morton order (nodes [nt - 1][16] written, leaves reordered in
// place), and the device bytes of its temporaries
int32_t dlg_sah_tree(dlg_gpu* g, int64_t nt, float* leaves, float* nodes, int32_t& height, char* err, int32_t errLen);
int64_t dlg_sah_bytes(int64_t nt);

// The stages of a chunk split into what must run outside a graph capture (checks, buffer growth, host uploads) and
// what a capture can hold (kernel launches on g->stream, no waits). trace.hip:
int32_t dlg_trace_prepare(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, int& stack, char* err, int32_t errLen);
void dlg_trace_enqueue(dlg_gpu* g, uint32_t rayFlags, int stack, unsigned long long* counts);
// the one-wave blocks of kernel fn a CU holds at once (the occupancy API, asked once per kernel)
uint32_t dlg_blocks_per_cu(const void* fn);
// the checks of a trace with these flags against the attached scene and the slots, and its stack size (no buffers)
int32_t dlg_trace_check(const dlg_gpu* g, uint32_t rayFlags, int& stack, char* err, int32_t errLen);
// the persistent grid of a trace (one-wave blocks, a multiple of 8) and the bytes of each block's global stack area
struct dlg_trace_shape { uint32_t grid; size_t ovfPerBlock; };
dlg_trace_shape dlg_trace_grid(const dlg_gpu* g, int64_t nc, int stack, bool bytes, bool count);
// the trace of nc sensors P, N ([nc] float3 on the device) into caller buffers: vis (bits [nWords x nc] or byte quads
// [8 nWords x nc]), the traced-ray count (added to), the stacks' global area (grid x ovfPerBlock bytes) and the tile
// counters (dlg_tile_counter_bytes; cleared by the launch)
void dlg_trace_enqueue_into(dlg_gpu* g, const float* P, const float* N, uint32_t nc, uint32_t rayFlags, int stack, uint32_t* vis,
                            unsigned long long* traced, uint32_t* ovf, uint32_t grid, uint32_t* next);
constexpr size_t dlg_tile_counter_bytes = 512;
int32_t dlg_trace_account(dlg_gpu* g, float ms, char* err, int32_t errLen);
// reduce.hip: W [nQ x nDir] uploaded, the output grown; the quantities' cosine mask
int32_t dlg_reduce_prepare(dlg_gpu* g, int64_t nc, int32_t nQ, const float* W, const int32_t* useCos, uint32_t& mask, char* err, int32_t errLen);
void dlg_reduce_enqueue(dlg_gpu* g, uint32_t nc, int32_t nQ, uint32_t mask);
// the reduce of caller buffers: vis as dlg_trace_enqueue_into wrote it, N [nc] float3, W [nQ x nDir], out [nQ x nc]
void dlg_reduce_enqueue_from(dlg_gpu* g, const uint32_t* vis, bool bytes, uint32_t nc, const float* N, const float* W, int32_t nQ, uint32_t mask,
                             double* out);
// series.hip: the series inputs uploaded (groundRow >= 0: G taken on the device from row groundRow of the reduce's
// output instead of in->G), the output and operands grown
int32_t dlg_series_prepare(dlg_gpu* g, int64_t nc, const dlg_series_in* in, int32_t method, int32_t groundRow, char* err, int32_t errLen);
void dlg_series_enqueue(dlg_gpu* g);
// whether the prepared series has its tile shape (the MFMA methods time every shape on a problem's first chunk)
bool dlg_series_tuned(const dlg_gpu* g);

// event timing on the library's stream: ms of the work enqueued between begin and end
inline void dlg_tbegin(dlg_gpu* g) { (void)hipEventRecord(g->ev[0], g->stream); }
inline float dlg_tend(dlg_gpu* g) {
  float ms = 0.f;
  (void)hipEventRecord(g->ev[1], g->stream);
  (void)hipEventSynchronize(g->ev[1]);
  (void)hipEventElapsedTime(&ms, g->ev[0], g->ev[1]);
  return ms;
}
