// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The device context of dl_device.h: one stream (default priority) and two timing events on device 0, resident sensors
// and directions, scene residency. Replaces the OptiX context / pipeline / SBT host code of reference src/gpu.cu:277-497.
#include <cstring>
#include <string>
#include <vector>

#include "dl_math.h"
#include "gpu_internal.h"

namespace {

// per direction the ray constants every ray along it shares (dl_math.h ray_dir, packed by dir_pack)
__global__ void k_dir_table(const float3* __restrict__ D, int64_t n, float4* __restrict__ tab) {
  const int64_t d = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (d >= n) return;
  dlm::Ray r;
  dlm::ray_dir(r, D[d]);
  dlm::dir_pack(r, tab[2 * d], tab[2 * d + 1]);
}

// leaf k's global slot (float 10, as bits) from the per-triangle slots in input order (its original index, float 9)
__global__ void k_leaf_slots(float* __restrict__ leaves, int64_t nt, const uint32_t* __restrict__ slot) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  leaves[12 * k + 10] = __uint_as_float(slot[__float_as_uint(leaves[12 * k + 9])]);
}

}  // namespace

int32_t dlg_leaf_slots_device(dlg_gpu* g, const uint32_t* slot, int64_t nt, uint32_t mx, char* err, int32_t errLen) {
  dlg_scene& s = g->scene;
  if (nt != s.nTris) { dlg_fail(err, errLen, "dlg_set_tri_slots: " + std::to_string(nt) + " slots for a scene of " + std::to_string(s.nTris) + " triangles"); return 1; }
  hipError_t e = hipSuccess;
  if (nt > 0) {
    k_leaf_slots<<<unsigned((nt + 255) / 256), 256, 0, g->stream>>>(s.leaves.as<float>(), nt, slot);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  }
  if (e != hipSuccess) { dlg_fail(err, errLen, std::string("dlg_set_tri_slots: ") + hipGetErrorString(e)); return 1; }
  s.slotMax = mx;
  return 0;
}

extern "C" {

int32_t dlg_init(dlg_gpu** out, char* err, int32_t errLen) {
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) { dlg_fail(err, errLen, "no HIP device"); return 1; }
  DLG_CHECK(hipSetDevice(0));
  dlg_gpu* g = new dlg_gpu;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess) g->cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&g->ev[0]) != hipSuccess ||
      hipEventCreate(&g->ev[1]) != hipSuccess) {
    dlg_fail(err, errLen, "hip stream / events");
    dlg_destroy(g);
    return 1;
  }
  if (dlg_noalloc_setup(g, err, errLen)) {
    dlg_destroy(g);
    return 1;
  }
  *out = g;
  return 0;
}

void dlg_destroy(dlg_gpu* g) {
  if (!g) return;
  dlg_graph_free(g);
  g->scene.release();
  for (DevBuf* b : {&g->P, &g->N, &g->D, &g->dirTab, &g->bits, &g->scratch, &g->out, &g->skyS,
                    &g->skyG, &g->skyWork, &g->seriesE, &g->seriesWork, &g->seriesIn, &g->view, &g->viewOvf, &g->labels, &g->counts, &g->traceCount, &g->stackOvf, &g->tileNext, &g->sensWork, &g->slots, &g->dc, &g->skySdc,
                    &g->meshV, &g->meshI, &g->meshS, &g->meshWord, &g->meshWork, &g->tileOrder, &g->orderWork})
    b->release();
  if (g->hostSums) (void)hipHostFree(g->hostSums);
  for (hipEvent_t e : g->ev) if (e) (void)hipEventDestroy(e);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

int32_t dlg_device_info(char* name, int32_t nameLen, int32_t* cus, int64_t* hbmBytes) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
  if (name && nameLen > 0) {
    std::snprintf(name, size_t(nameLen), "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (cus) *cus = prop.multiProcessorCount;
  if (hbmBytes) *hbmBytes = int64_t(prop.totalGlobalMem);
  return 0;
}

// Pinned host memory for n doubles of a chunk's sums, kept and grown across runs (null when it cannot be had: the
// caller uses its own memory). The downloads into it run at the link's rate, and a run touches no fresh pages.
double* dlg_host_sums(dlg_gpu* g, int64_t n) {
  const size_t bytes = size_t(n > 0 ? n : 1) * sizeof(double);
  if (bytes <= g->hostSumsBytes) return g->hostSums;
  if (g->hostSums) (void)hipHostFree(g->hostSums);
  g->hostSums = nullptr;
  g->hostSumsBytes = 0;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  g->hostSums = static_cast<double*>(p);
  g->hostSumsBytes = bytes;
  return g->hostSums;
}

void dlg_trim(dlg_gpu* g, int32_t sensors) {
  if (g->hostSums) (void)hipHostFree(g->hostSums);
  g->hostSums = nullptr;
  g->hostSumsBytes = 0;
  for (DevBuf* b : {&g->bits, &g->scratch, &g->out, &g->seriesE, &g->seriesWork, &g->seriesIn, &g->view, &g->viewOvf, &g->dc, &g->stackOvf, &g->sensWork, &g->meshV, &g->meshI, &g->meshS, &g->meshWork,
                    &g->tileOrder, &g->orderWork}) b->release();
  g->orderGen = -1;
  g->meshOpen = 0;
  g->meshNv = g->meshNt = 0;
  if (sensors) {
    g->P.release(); g->N.release(); g->labels.release();
    g->nSens = 0;
    g->sensGen += 1;
  }
}

dlg_scene* dlg_detach(dlg_gpu* g) {
  if (g->scene.borrowed) { g->scene = dlg_scene{}; return new dlg_scene; }   // a loan is dropped, not handed on
  dlg_scene* s = new dlg_scene(g->scene);
  g->scene = dlg_scene{};
  return s;
}

void dlg_attach(dlg_gpu* g, dlg_scene* s) {
  g->scene.release();
  g->scene = *s;
  delete s;
}

void dlg_borrow_scene(dlg_gpu* dst, const dlg_gpu* src) {
  dst->scene.release();
  dst->scene = src->scene;
  dst->scene.borrowed = true;
}

void dlg_free_scene(dlg_scene* s) {
  if (!s) return;
  s->release();
  delete s;
}

int64_t dlg_scene_bytes(const dlg_scene* s) { return s ? s->bytes() : 0; }

int64_t dlg_free_bytes(void) {
  size_t f = 0, t = 0;
  return hipMemGetInfo(&f, &t) == hipSuccess ? int64_t(f) : 0;
}

int32_t dlg_set_scene(dlg_gpu* g, const float* P, const float* N, int64_t nSens, const float* D, int64_t nDir, char* err, int32_t errLen) {
  g->nDir = nDir;
  g->nWords = (nDir + 31) / 32;
  DLG_CHECK(g->D.grow(size_t(nDir) * 12));
  if (nSens >= 0) {   // -1: the sensors dlg_sensors made stay
    g->nSens = nSens;
    DLG_CHECK(g->P.grow(size_t(nSens) * 12));
    DLG_CHECK(g->N.grow(size_t(nSens) * 12));
  }
  if (nSens > 0) {
    g->sensGen += 1;
    DLG_CHECK(hipMemcpyAsync(g->P.p, P, size_t(nSens) * 12, hipMemcpyHostToDevice, g->stream));
    DLG_CHECK(hipMemcpyAsync(g->N.p, N, size_t(nSens) * 12, hipMemcpyHostToDevice, g->stream));
  }
  if (nDir) {
    DLG_CHECK(hipMemcpyAsync(g->D.p, D, size_t(nDir) * 12, hipMemcpyHostToDevice, g->stream));
    DLG_CHECK(g->dirTab.grow(size_t(nDir) * 32));
    k_dir_table<<<unsigned((nDir + 255) / 256), 256, 0, g->stream>>>(g->D.as<float3>(), nDir, g->dirTab.as<float4>());
    DLG_CHECK(hipGetLastError());
  }
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

uint32_t dlg_ray_flags(int32_t cullBackfaces, int32_t translucent) { return 4u | (translucent ? 0u : 1u) | (cullBackfaces ? 16u : 0u); }

int32_t dlg_set_slots(dlg_gpu* g, const dlg_slot* slots, int64_t n, char* err, int32_t errLen) {
  g->nSlots = 0;
  if (n <= 0) return 0;
  DLG_CHECK(g->slots.grow(size_t(n) * sizeof(dlg_slot)));
  DLG_CHECK(hipMemcpyAsync(g->slots.p, slots, size_t(n) * sizeof(dlg_slot), hipMemcpyHostToDevice, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  g->nSlots = n;
  return 0;
}

int32_t dlg_set_tri_slots(dlg_gpu* g, const uint32_t* slot, int64_t nt, char* err, int32_t errLen) {
  dlg_scene& s = g->scene;
  if (nt != s.nTris) { dlg_fail(err, errLen, "dlg_set_tri_slots: " + std::to_string(nt) + " slots for a scene of " + std::to_string(s.nTris) + " triangles"); return 1; }
  uint32_t mx = 0;
  for (int64_t t = 0; t < nt; ++t) mx = slot[t] > mx ? slot[t] : mx;
  if (nt == 0) { s.slotMax = 0; return 0; }
  DevBuf d;
  hipError_t e = d.grow(size_t(nt) * 4);
  if (e == hipSuccess) e = hipMemcpyAsync(d.p, slot, size_t(nt) * 4, hipMemcpyHostToDevice, g->stream);
  if (e != hipSuccess) { d.release(); dlg_fail(err, errLen, std::string("dlg_set_tri_slots: ") + hipGetErrorString(e)); return 1; }
  const int32_t rc = dlg_leaf_slots_device(g, d.as<uint32_t>(), nt, mx, err, errLen);
  d.release();
  return rc;
}

int32_t dlg_download_bytes(dlg_gpu* g, int64_t nc, uint8_t* out, char* err, int32_t errLen) {
  if (nc != g->nc) { dlg_fail(err, errLen, "dlg_download_bytes: no trace of this chunk"); return 1; }
  if (!g->bytes) { dlg_fail(err, errLen, "dlg_download_bytes: the last trace wrote bits"); return 1; }
  const size_t nQuads = size_t(g->nWords) * 8, n = nQuads * size_t(nc);
  if (n == 0) return 0;
  std::vector<uint32_t> q(n);
  DLG_CHECK(hipMemcpyAsync(q.data(), g->bits.p, n * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  for (size_t d = 0; d < size_t(g->nDir); ++d) {   // quad d / 4, byte d % 4 -> direction-major [d nc + i]
    const uint32_t* row = q.data() + (d / 4) * size_t(nc);
    const uint32_t sh = 8 * uint32_t(d % 4);
    uint8_t* o = out + d * size_t(nc);
    for (size_t i = 0; i < size_t(nc); ++i) o[i] = uint8_t(row[i] >> sh);
  }
  return 0;
}

int32_t dlg_download_bits(dlg_gpu* g, int64_t nc, uint32_t* out, char* err, int32_t errLen) {
  if (nc != g->nc) { dlg_fail(err, errLen, "dlg_download_bits: no trace of this chunk"); return 1; }
  if (g->bytes) { dlg_fail(err, errLen, "dlg_download_bits: the last trace wrote bytes"); return 1; }
  const size_t n = size_t(g->nWords) * size_t(nc);
  if (n) DLG_CHECK(hipMemcpyAsync(out, g->bits.p, n * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int32_t dlg_scene_download(dlg_gpu* g, float* nodes, int64_t* nNodes, float* leaves, int64_t* nTris, int32_t* height) {
  if (nNodes) *nNodes = g->scene.nNodes;
  if (nTris) *nTris = g->scene.nTris;
  if (height) *height = g->scene.height;
  if (nodes && g->scene.nNodes && hipMemcpy(nodes, g->scene.nodes.p, size_t(g->scene.nNodes) * 64, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  if (leaves && g->scene.nTris && hipMemcpy(leaves, g->scene.leaves.p, size_t(g->scene.nTris) * 48, hipMemcpyDeviceToHost) != hipSuccess) return 1;
  return 0;
}

void dlg_count_fetches(dlg_gpu* g, int32_t on) { g->countFetches = on ? 1 : 0; }
void dlg_set_tree(dlg_gpu* g, int32_t tree) { g->tree = tree == DLG_TREE_LBVH ? DLG_TREE_LBVH : DLG_TREE_SAH; }

int32_t dlg_stats(const dlg_gpu* g, dlg_stats_t* out) {
  *out = g->stats;
  out->sceneBytes = g->scene.bytes();
  out->nodes = g->scene.nNodes;
  out->height = g->scene.height;
  return 0;
}

const float* dlg_series_device(dlg_gpu* g) { return g->seriesE.as<float>(); }

}  // extern "C"
