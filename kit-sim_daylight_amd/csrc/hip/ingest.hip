// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Mesh ingest on the GPU: the scene's points recentred and rounded to float, its faces fanned into triangles, where the
// scene is then built. The host's triangulate_into (reference mesh.cpp:137-168, called per mesh from run.cpp:120-147)
// restated as three passes per mesh: corner counts and triangles per face, two exclusive scans (each face's first
// corner and first triangle), one thread per face writing its triangles. Faces of 5 or more corners arrive
// ear-clipped by the host (rare in large meshes) and are placed by their face's scanned offset, so the triangle order is
// the host's: meshes in order, faces in order, a face's triangles in order.
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <string>
#include <vector>

#include "gpu_internal.h"

namespace {

constexpr int kBlock = 256;
unsigned grid_of(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

// fault bits in word[1]: a face with fewer than 3 corners, corners past the index array, an index out of range, a
// face of the host's list out of range, more triangles than staged
constexpr uint32_t kBadCount = 1, kBadCorner = 2, kBadIndex = 4, kBadNgon = 8, kBadRoom = 16;

// coordinate i of the mesh: (x - centre) in double rounded to float (the host's f3(p - c)); the largest |coordinate|
// into word[0] as float bits (non-negative floats order as their bits; NaN skipped, as fmaxf does)
__global__ void k_points(const double* __restrict__ xyz, int64_t n3, double cx, double cy, double cz, float* __restrict__ v,
                         uint32_t* __restrict__ word) {
  __shared__ uint32_t sMax;
  if (threadIdx.x == 0) sMax = 0;
  __syncthreads();
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n3) {
    const int k = int(i % 3);
    const float x = float(xyz[i] - (k == 0 ? cx : k == 1 ? cy : cz));
    v[i] = x;
    const float a = fabsf(x);
    if (a == a) atomicMax(&sMax, __float_as_uint(a));
  }
  __syncthreads();
  if (threadIdx.x == 0 && sMax) atomicMax(word, sMax);
}

// face f's corner count and triangles: 3 -> 1, 4 -> 2, 5 or more -> 0 here (the host's list sets them)
__global__ void k_face_counts(const int32_t* __restrict__ counts, int64_t nF, int64_t* __restrict__ cnt, int64_t* __restrict__ made,
                              uint32_t* __restrict__ word) {
  const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nF) return;
  const int32_t n = counts[f];
  cnt[f] = n;
  made[f] = n == 3 ? 1 : n == 4 ? 2 : 0;
  if (n < 3) atomicOr(word + 1, kBadCount);
}

__global__ void k_ngon_made(const int64_t* __restrict__ face, const int32_t* __restrict__ m, int64_t nNg, int64_t nF, int64_t* __restrict__ made,
                            uint32_t* __restrict__ word) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nNg) return;
  const int64_t f = face[k];
  if (f < 0 || f >= nF || m[k] < 0) { atomicOr(word + 1, kBadNgon); return; }
  made[f] = m[k];
}

// faces of 3 or 4 corners at their scanned offsets: (a, b, c), and (a, c, d) for a quad; vertex ids base + local, the
// face's global slot beside each triangle; the largest slot written into word[2]
__global__ void k_fan(const int32_t* __restrict__ counts, const int64_t* __restrict__ cOff, const int64_t* __restrict__ tOff, int64_t nF,
                      const int32_t* __restrict__ idx, int64_t nIdx, int64_t nPts, const uint16_t* __restrict__ faceSlot, uint32_t slotBase,
                      uint32_t base, uint32_t* __restrict__ tri, uint32_t* __restrict__ slot, int64_t tCap, uint32_t* __restrict__ word) {
  const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= nF) return;
  const int32_t n = counts[f];
  if (n != 3 && n != 4) return;
  const int64_t o = cOff[f], t = tOff[f];
  if (o < 0 || o + n > nIdx) { atomicOr(word + 1, kBadCorner); return; }
  if (t < 0 || t + (n - 2) > tCap) { atomicOr(word + 1, kBadRoom); return; }
  int32_t c[4] = {0, 0, 0, 0};
  for (int k = 0; k < n; ++k) {
    c[k] = idx[o + k];
    if (c[k] < 0 || int64_t(c[k]) >= nPts) { atomicOr(word + 1, kBadIndex); return; }
  }
  uint32_t* q = tri + 3 * t;
  q[0] = base + uint32_t(c[0]); q[1] = base + uint32_t(c[1]); q[2] = base + uint32_t(c[2]);
  if (n == 4) { q[3] = base + uint32_t(c[0]); q[4] = base + uint32_t(c[2]); q[5] = base + uint32_t(c[3]); }
  if (slot) {
    const uint32_t s = slotBase + (faceSlot ? uint32_t(faceSlot[f]) : 0u);
    slot[t] = s;
    if (n == 4) slot[t + 1] = s;
    atomicMax(word + 2, s);
  }
}

// the host's ear-clipped faces: face k of the list writes its m[k] triangles (corners ngTri[3 (start[k] + j) ..]) at its
// face's scanned offset
__global__ void k_ngon_fan(const int64_t* __restrict__ face, const int32_t* __restrict__ m, const int64_t* __restrict__ start, int64_t nNg,
                           const uint32_t* __restrict__ ngTri, int64_t nNgTri, const int64_t* __restrict__ tOff, int64_t nF, int64_t nPts,
                           const uint16_t* __restrict__ faceSlot, uint32_t slotBase, uint32_t base, uint32_t* __restrict__ tri,
                           uint32_t* __restrict__ slot, int64_t tCap, uint32_t* __restrict__ word) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= nNg) return;
  const int64_t f = face[k], s0 = start[k];
  const int32_t mk = m[k];
  if (f < 0 || f >= nF || mk < 0 || s0 < 0 || s0 + mk > nNgTri) { atomicOr(word + 1, kBadNgon); return; }
  const int64_t t = tOff[f];
  if (t < 0 || t + mk > tCap) { atomicOr(word + 1, kBadRoom); return; }
  const uint32_t s = slotBase + (faceSlot ? uint32_t(faceSlot[f]) : 0u);
  for (int32_t j = 0; j < mk; ++j) {
    for (int q = 0; q < 3; ++q) {
      const uint32_t c = ngTri[3 * (s0 + j) + q];
      if (int64_t(c) >= nPts) { atomicOr(word + 1, kBadIndex); return; }
      tri[3 * (t + j) + q] = base + c;
    }
    if (slot) slot[t + j] = s;
  }
  if (slot && mk > 0) atomicMax(word + 2, s);
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// the per-mesh temporaries' regions in g->meshWork
struct WorkLayout {
  size_t xyz, counts, idx, faceSlot, cnt, made, cOff, tOff, ngFace, ngMade, ngStart, ngTri, temp, total;
  WorkLayout(int64_t nPts, int64_t nF, int64_t nIdx, int64_t nNg, int64_t nNgTri, size_t tempBytes) {
    size_t at = 0;
    const auto take = [&](size_t b) { const size_t r = at; at += align256(b); return r; };
    xyz = take(size_t(nPts) * 24);
    counts = take(size_t(nF) * 4);
    idx = take(size_t(nIdx) * 4);
    faceSlot = take(size_t(nF) * 2);
    cnt = take(size_t(nF) * 8);
    made = take(size_t(nF) * 8);
    cOff = take(size_t(nF) * 8);
    tOff = take(size_t(nF) * 8);
    ngFace = take(size_t(nNg) * 8);
    ngMade = take(size_t(nNg) * 4);
    ngStart = take(size_t(nNg) * 8);
    ngTri = take(size_t(nNgTri) * 12);
    temp = take(tempBytes);
    total = at;
  }
};

size_t scan_temp_bytes(int64_t nF) {
  size_t b = 0;
  if (nF > 0) (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, static_cast<const int64_t*>(nullptr), static_cast<int64_t*>(nullptr), int(nF));
  return b;
}

std::string fault_text(uint32_t w) {
  std::string s;
  const auto add = [&](const char* t) { s += (s.empty() ? "" : ", ") + std::string(t); };
  if (w & kBadCount) add("a face with fewer than 3 corners");
  if (w & kBadCorner) add("a face's corners past the index array");
  if (w & kBadIndex) add("a vertex index out of range");
  if (w & kBadNgon) add("an ear-clipped face out of range");
  if (w & kBadRoom) add("more triangles than staged");
  return s;
}

void mesh_reset(dlg_gpu* g) {
  for (DevBuf* b : {&g->meshV, &g->meshI, &g->meshS, &g->meshWork}) b->release();
  g->meshNv = g->meshNt = g->meshCapV = g->meshCapT = 0;
  g->meshSlotted = g->meshOpen = 0;
}

}  // namespace

extern "C" {

int64_t dlg_mesh_bytes(int64_t nPts, int64_t nFaces, int64_t nIdx) {
  const WorkLayout w(nPts, nFaces, nIdx, 0, 0, 0);
  // + the scans' temporaries and the ear-clipped faces (bounded by the corners: a face of n corners makes n - 2)
  return int64_t(w.total) + int64_t(scan_temp_bytes(nFaces)) + nIdx * 24 + (int64_t(16) << 20);
}

int32_t dlg_mesh_begin(dlg_gpu* g, int64_t nPts, int64_t nTriCap, int32_t slotted, const double* centre, char* err, int32_t errLen) {
  mesh_reset(g);
  if (nPts < 0 || nTriCap < 0 || nPts >= (int64_t(1) << 32)) { dlg_fail(err, errLen, "dlg_mesh_begin: sizes out of range"); return 1; }
  DLG_CHECK(g->meshV.grow(size_t(nPts) * 12));
  DLG_CHECK(g->meshI.grow(size_t(nTriCap) * 12));
  if (slotted) DLG_CHECK(g->meshS.grow(size_t(nTriCap) * 4));
  DLG_CHECK(g->meshWord.grow(16));
  DLG_CHECK(hipMemsetAsync(g->meshWord.p, 0, 16, g->stream));
  g->meshCapV = nPts;
  g->meshCapT = nTriCap;
  g->meshSlotted = slotted ? 1 : 0;
  for (int k = 0; k < 3; ++k) g->meshC[k] = centre ? centre[k] : 0.0;
  g->meshOpen = 1;
  return 0;
}

int32_t dlg_mesh_add(dlg_gpu* g, const double* xyz, int64_t nPts, const int32_t* counts, int64_t nFaces, const int32_t* idx, int64_t nIdx,
                     const uint16_t* faceSlot, uint32_t slotBase, const int64_t* ngonFace, const int32_t* ngonMade, const uint32_t* ngonTri,
                     int64_t nNgon, int64_t* made, char* err, int32_t errLen) {
  if (made) *made = 0;
  if (!g->meshOpen) { dlg_fail(err, errLen, "dlg_mesh_add: no dlg_mesh_begin"); return 1; }
  if (nPts < 0 || nFaces < 0 || nIdx < 0 || nNgon < 0 || g->meshNv + nPts > g->meshCapV) {
    dlg_fail(err, errLen, "dlg_mesh_add: " + std::to_string(nPts) + " points past the " + std::to_string(g->meshCapV) + " staged");
    return 1;
  }
  if (nFaces >= (int64_t(1) << 31)) { dlg_fail(err, errLen, "dlg_mesh_add: more than 2^31 - 1 faces in one mesh"); return 1; }
  // the ear-clipped faces' triangle offsets, in list order
  std::vector<int64_t> start(size_t(nNgon), 0);
  int64_t nNgTri = 0;
  for (int64_t k = 0; k < nNgon; ++k) {
    if (ngonMade[k] < 0) { dlg_fail(err, errLen, "dlg_mesh_add: negative triangle count in the ear-clipped faces"); return 1; }
    start[size_t(k)] = nNgTri;
    nNgTri += ngonMade[k];
  }
  const size_t tempBytes = scan_temp_bytes(nFaces);
  const WorkLayout L(nPts, nFaces, nIdx, nNgon, nNgTri, tempBytes);
  DLG_CHECK(g->meshWork.grow(L.total));
  char* w = g->meshWork.as<char>();
  const auto at = [&](size_t off) { return static_cast<void*>(w + off); };
  hipStream_t st = g->stream;
  uint32_t* word = g->meshWord.as<uint32_t>();
  const uint32_t base = uint32_t(g->meshNv);
  float* v = g->meshV.as<float>() + 3 * g->meshNv;
  if (nPts > 0) {
    DLG_CHECK(hipMemcpyAsync(at(L.xyz), xyz, size_t(nPts) * 24, hipMemcpyHostToDevice, st));
    k_points<<<grid_of(3 * nPts), kBlock, 0, st>>>(static_cast<const double*>(at(L.xyz)), 3 * nPts, g->meshC[0], g->meshC[1], g->meshC[2], v, word);
    DLG_CHECK(hipGetLastError());
  }
  int64_t nMade = 0;
  if (nFaces > 0) {
    const bool slot = g->meshSlotted != 0;
    DLG_CHECK(hipMemcpyAsync(at(L.counts), counts, size_t(nFaces) * 4, hipMemcpyHostToDevice, st));
    if (nIdx > 0) DLG_CHECK(hipMemcpyAsync(at(L.idx), idx, size_t(nIdx) * 4, hipMemcpyHostToDevice, st));
    if (slot && faceSlot) DLG_CHECK(hipMemcpyAsync(at(L.faceSlot), faceSlot, size_t(nFaces) * 2, hipMemcpyHostToDevice, st));
    if (nNgon > 0) {
      DLG_CHECK(hipMemcpyAsync(at(L.ngFace), ngonFace, size_t(nNgon) * 8, hipMemcpyHostToDevice, st));
      DLG_CHECK(hipMemcpyAsync(at(L.ngMade), ngonMade, size_t(nNgon) * 4, hipMemcpyHostToDevice, st));
      DLG_CHECK(hipMemcpyAsync(at(L.ngStart), start.data(), size_t(nNgon) * 8, hipMemcpyHostToDevice, st));
      if (nNgTri > 0) DLG_CHECK(hipMemcpyAsync(at(L.ngTri), ngonTri, size_t(nNgTri) * 12, hipMemcpyHostToDevice, st));
    }
    int64_t* cnt = static_cast<int64_t*>(at(L.cnt));
    int64_t* mk = static_cast<int64_t*>(at(L.made));
    int64_t* cOff = static_cast<int64_t*>(at(L.cOff));
    int64_t* tOff = static_cast<int64_t*>(at(L.tOff));
    k_face_counts<<<grid_of(nFaces), kBlock, 0, st>>>(static_cast<const int32_t*>(at(L.counts)), nFaces, cnt, mk, word);
    if (nNgon > 0)
      k_ngon_made<<<grid_of(nNgon), kBlock, 0, st>>>(static_cast<const int64_t*>(at(L.ngFace)), static_cast<const int32_t*>(at(L.ngMade)), nNgon,
                                                     nFaces, mk, word);
    DLG_CHECK(hipGetLastError());
    size_t tb = tempBytes;
    DLG_CHECK(hipcub::DeviceScan::ExclusiveSum(at(L.temp), tb, cnt, cOff, int(nFaces), st));
    tb = tempBytes;
    DLG_CHECK(hipcub::DeviceScan::ExclusiveSum(at(L.temp), tb, mk, tOff, int(nFaces), st));
    const int64_t tCap = g->meshCapT - g->meshNt;
    uint32_t* tri = g->meshI.as<uint32_t>() + 3 * g->meshNt;
    uint32_t* ts = slot ? g->meshS.as<uint32_t>() + g->meshNt : nullptr;
    const uint16_t* fs = slot && faceSlot ? static_cast<const uint16_t*>(at(L.faceSlot)) : nullptr;
    k_fan<<<grid_of(nFaces), kBlock, 0, st>>>(static_cast<const int32_t*>(at(L.counts)), cOff, tOff, nFaces, static_cast<const int32_t*>(at(L.idx)),
                                             nIdx, nPts, fs, slotBase, base, tri, ts, tCap, word);
    if (nNgon > 0)
      k_ngon_fan<<<grid_of(nNgon), kBlock, 0, st>>>(static_cast<const int64_t*>(at(L.ngFace)), static_cast<const int32_t*>(at(L.ngMade)),
                                                    static_cast<const int64_t*>(at(L.ngStart)), nNgon, static_cast<const uint32_t*>(at(L.ngTri)),
                                                    nNgTri, tOff, nFaces, nPts, fs, slotBase, base, tri, ts, tCap, word);
    DLG_CHECK(hipGetLastError());
    int64_t last[2] = {0, 0};
    DLG_CHECK(hipMemcpyAsync(&last[0], tOff + (nFaces - 1), 8, hipMemcpyDeviceToHost, st));
    DLG_CHECK(hipMemcpyAsync(&last[1], mk + (nFaces - 1), 8, hipMemcpyDeviceToHost, st));
    DLG_CHECK(hipStreamSynchronize(st));
    nMade = last[0] + last[1];
  }
  uint32_t wd[2] = {0, 0};
  DLG_CHECK(hipMemcpyAsync(wd, word, 8, hipMemcpyDeviceToHost, st));
  DLG_CHECK(hipStreamSynchronize(st));
  if (wd[1]) {
    dlg_fail(err, errLen, "dlg_mesh_add: " + fault_text(wd[1]));
    mesh_reset(g);
    return 1;
  }
  if (nMade < 0 || g->meshNt + nMade > g->meshCapT) {
    dlg_fail(err, errLen, "dlg_mesh_add: more triangles than staged");
    mesh_reset(g);
    return 1;
  }
  g->meshNv += nPts;
  g->meshNt += nMade;
  if (made) *made = nMade;
  return 0;
}

int32_t dlg_mesh_staged(dlg_gpu* g, float* v, int64_t* nv, uint32_t* tri, int64_t* nt, uint32_t* slot, char* err, int32_t errLen) {
  if (!g->meshOpen) { dlg_fail(err, errLen, "dlg_mesh_staged: nothing staged"); return 1; }
  if (nv) *nv = g->meshNv;
  if (nt) *nt = g->meshNt;
  if (v && g->meshNv) DLG_CHECK(hipMemcpyAsync(v, g->meshV.p, size_t(g->meshNv) * 12, hipMemcpyDeviceToHost, g->stream));
  if (tri && g->meshNt) DLG_CHECK(hipMemcpyAsync(tri, g->meshI.p, size_t(g->meshNt) * 12, hipMemcpyDeviceToHost, g->stream));
  if (slot && g->meshSlotted && g->meshNt) DLG_CHECK(hipMemcpyAsync(slot, g->meshS.p, size_t(g->meshNt) * 4, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}

int32_t dlg_mesh_build(dlg_gpu* g, char* err, int32_t errLen) {
  if (!g->meshOpen) { dlg_fail(err, errLen, "dlg_mesh_build: no dlg_mesh_begin"); return 1; }
  uint32_t wd[3] = {0, 0, 0};
  DLG_CHECK(hipMemcpyAsync(wd, g->meshWord.p, 12, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  g->meshWork.release();   // the tree's room
  float mx = 0.f;
  std::memcpy(&mx, &wd[0], 4);
  int32_t rc = dlg_build_device(g, g->meshV.as<float>(), g->meshNv, g->meshI.as<uint32_t>(), g->meshNt, mx, err, errLen);
  if (rc == 0 && g->meshSlotted) rc = dlg_leaf_slots_device(g, g->meshS.as<uint32_t>(), g->meshNt, g->meshNt ? wd[2] : 0u, err, errLen);
  mesh_reset(g);
  return rc;
}

}  // extern "C"
