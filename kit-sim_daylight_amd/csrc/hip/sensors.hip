// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Sensor preparation on the GPU: the sensors of a sensor mesh - one per face at its centre along its normal, or one per
// point along its area-weighted normal - written straight into the context's resident sensor arrays (P, N), where the
// trace reads them. Replaces the host's face_geom / make_sensors_into / sensors_of (reference src/mesh.cpp:45-71,
// 315-349, src/run.cpp:192-213), with the same FP64 arithmetic in the same order, so every sensor is the host's to the
// bit (no fused multiply-adds: the library builds with -ffp-contract=off; division and square root are correctly
// rounded on both sides):
//   physical faces: Newell's vector N of the corners, area |N| / 2; the centre the fan's triangles' centroids weighted by
//     their signed areas along N / |N| (the corners' mean when that weight is under 1e-12); the normal N / |N| (0 for a
//     face under 1e-12 of area);
//   ladybug faces: the corners' mean, the normal of one triangle or of a quad's two halves summed;
//   points: the sum over their faces, in face order, of normal x area (per point a segment of the corners sorted by
//     point, stable, so each sum runs in the host's order), normalised, (0, 0, 1) when it is 0;
//   a sensor at p + offset n with normal n (front) and / or p - offset n with -n (back), recentred and rounded to float.
#include <hipcub/hipcub.hpp>

#include <string>

#include "gpu_internal.h"

namespace {

constexpr int kBlock = 256;
unsigned blocks(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

struct d3 { double x, y, z; };
__device__ __forceinline__ d3 add(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 mul(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double len(d3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ d3 unit(d3 a) {   // a times the reciprocal of its length, as the host's normalised
  const double l = len(a);
  return l > 0 ? mul(a, 1 / l) : d3{0, 0, 0};
}

struct FaceRef {
  const double* xyz;
  const int32_t* c;   // the face's corner indices (checked on the host; clamped here so that no read leaves the mesh)
  int n;
  int64_t nP;
  __device__ d3 at(int k) const {
    const int64_t i = c[k] < 0 ? 0 : c[k] >= nP ? nP - 1 : c[k];
    const double* p = xyz + 3 * size_t(i);
    return {p[0], p[1], p[2]};
  }
  __device__ d3 newell() const {
    d3 N{0, 0, 0};
    for (int i = 0; i < n; ++i) {
      const d3 a = at(i), b = at((i + 1) % n);
      N.x += (a.y - b.y) * (a.z + b.z);
      N.y += (a.z - b.z) * (a.x + b.x);
      N.z += (a.x - b.x) * (a.y + b.y);
    }
    return N;
  }
  __device__ d3 mean() const {
    d3 m{0, 0, 0};
    for (int i = 0; i < n; ++i) m = add(m, at(i));
    return mul(m, 1.0 / n);
  }
};

struct Geom { d3 normal, centre; double area; };

template <bool LB>
__device__ Geom face_geom(const FaceRef& f) {
  Geom g{{0, 0, 0}, {0, 0, 0}, 0};
  if (f.n < 3) return g;
  if (LB) {
    g.centre = f.mean();
    const d3 p0 = f.at(0);
    d3 N{0, 0, 0};
    if (f.n == 3) {
      N = cross(sub(f.at(1), p0), sub(f.at(2), p0));
      g.area = len(N) / 2;
    } else if (f.n == 4) {
      const d3 p1 = f.at(1), p2 = f.at(2), p3 = f.at(3);
      const d3 n1 = cross(sub(p1, p0), sub(p2, p0)), n2 = cross(sub(p3, p2), sub(p1, p2));
      N = add(n1, n2);
      g.area = (len(n1) + len(n2)) / 2;
    }
    g.normal = g.area < 1e-12 ? d3{0, 0, 0} : unit(N);
    return g;
  }
  const d3 N = f.newell(), u = unit(N), p0 = f.at(0);
  g.area = len(N) / 2;
  double wsum = 0;
  for (int i = 1; i + 1 < f.n; ++i) {
    const d3 a = f.at(i), b = f.at(i + 1);
    const double w = dot(cross(sub(a, p0), sub(b, p0)), u) / 2;
    g.centre = add(g.centre, mul(add(add(p0, a), b), w / 3));
    wsum += w;
  }
  g.centre = wsum > 1e-12 ? mul(g.centre, 1 / wsum) : f.mean();
  g.normal = g.area < 1e-12 ? d3{0, 0, 0} : unit(N);
  return g;
}

struct Out {
  float3* P;
  float3* N;
  int sides;   // 0 front, 1 back, 2 both
  double offset;
  d3 C;
  __device__ void put(int64_t j, d3 q, d3 nr) const {
    if (sides != 1) {
      const d3 p = sub(add(q, mul(nr, offset)), C);
      P[j] = make_float3(float(p.x), float(p.y), float(p.z));
      N[j] = make_float3(float(nr.x), float(nr.y), float(nr.z));
      ++j;
    }
    if (sides != 0) {
      const d3 p = sub(sub(q, mul(nr, offset)), C), m = mul(nr, -1);
      P[j] = make_float3(float(p.x), float(p.y), float(p.z));
      N[j] = make_float3(float(m.x), float(m.y), float(m.z));
    }
  }
};

// per face: its sensors (faces mode) or its normal x area (points mode); degenerate faces counted in cnt[0]
template <bool LB, bool VERTS>
__global__ void k_faces(const double* __restrict__ xyz, int64_t nP, const int32_t* __restrict__ counts, const int64_t* __restrict__ start,
                        const int32_t* __restrict__ idx, int64_t nC, int64_t nF, Out out, int64_t j0, int per, double* __restrict__ w,
                        int32_t* __restrict__ cnt) {
  const int64_t f = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (f >= nF) return;
  const int64_t c0 = start[f] < nC ? start[f] : nC;
  const int64_t n = counts[f] < nC - c0 ? counts[f] : nC - c0;
  const FaceRef face{xyz, idx + c0, int(n), nP};
  const Geom g = face_geom<LB>(face);
  if (g.area < 1e-12) atomicAdd(cnt, 1);
  if (VERTS) {
    w[3 * f] = g.normal.x * g.area;
    w[3 * f + 1] = g.normal.y * g.area;
    w[3 * f + 2] = g.normal.z * g.area;
  } else out.put(j0 + f * per, g.centre, g.normal);
}

// corner c's point (the sort key) and index (the value), and faceOf[c]
__global__ void k_corners(const int32_t* __restrict__ counts, const int64_t* __restrict__ start, const int32_t* __restrict__ idx, int64_t nC,
                          int64_t nF, int64_t nP, uint32_t* __restrict__ key, uint32_t* __restrict__ val, uint32_t* __restrict__ faceOf) {
  const int64_t f = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (f >= nF) return;
  for (int64_t c = start[f], e = c + counts[f]; c < e && c < nC; ++c) {
    key[c] = idx[c] < 0 ? 0u : idx[c] >= nP ? uint32_t(nP - 1) : uint32_t(idx[c]);
    val[c] = uint32_t(c);
    faceOf[c] = uint32_t(f);
  }
}

__global__ void k_widen(const int32_t* __restrict__ a, int64_t n, int64_t* __restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) b[i] = a[i] < 0 ? 0 : a[i];
}

__global__ void k_fill(int64_t* __restrict__ a, int64_t n, int64_t v) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) a[i] = v;
}

// first[v]: where point v's corners start among the sorted ones (-1: none)
__global__ void k_first(const uint32_t* __restrict__ key, int64_t nC, int64_t* __restrict__ first) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= nC) return;
  if (p == 0 || key[p] != key[p - 1]) first[key[p]] = p;
}

// per point: the sum of its faces' normal x area in corner order, its sensors; zero sums counted in cnt[1]
__global__ void k_points(const double* __restrict__ xyz, int64_t nP, const uint32_t* __restrict__ key, const uint32_t* __restrict__ val,
                         const uint32_t* __restrict__ faceOf, const double* __restrict__ w, const int64_t* __restrict__ first, int64_t nC, Out out,
                         int64_t j0, int per, int32_t* __restrict__ cnt) {
  const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= nP) return;
  d3 n{0, 0, 0};
  for (int64_t p = first[v]; p >= 0 && p < nC && key[p] == uint32_t(v); ++p) {
    const uint32_t f = faceOf[val[p]];
    n = add(n, d3{w[3 * f], w[3 * f + 1], w[3 * f + 2]});
  }
  if (len(n) == 0) {
    n = d3{0, 0, 1};
    atomicAdd(cnt + 1, 1);
  }
  out.put(j0 + v * per, d3{xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]}, unit(n));
}

}  // namespace

int64_t dlg_sensors_bytes(int64_t nPts, int64_t nFaces, int64_t nIdx) {
  size_t tmp = 0;
  if (nIdx > 0)
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                             static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), int(nIdx));
  size_t scan = 0;
  if (nFaces > 0) (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, static_cast<const int64_t*>(nullptr), static_cast<int64_t*>(nullptr), int(nFaces));
  return int64_t(24 * nPts + 4 * nFaces + 4 * nIdx + 8 * nFaces + 8 * nFaces + 24 * nFaces + 5 * 4 * nIdx + 8 * nPts + std::max(tmp, scan) + 1024);
}

extern "C" int32_t dlg_sensors_reserve(dlg_gpu* g, int64_t nSens, char* err, int32_t errLen) {
  if (nSens < 0) { dlg_fail(err, errLen, "dlg_sensors_reserve: a negative count"); return 1; }
  DLG_CHECK(g->P.grow(size_t(nSens) * 12));
  DLG_CHECK(g->N.grow(size_t(nSens) * 12));
  g->nSens = nSens;
  g->sensGen += 1;
  return 0;
}

extern "C" int32_t dlg_sensors(dlg_gpu* g, const dlg_sensor_mesh* m, int64_t j0, int32_t onVertices, int32_t sides, double offset, const double* centre,
                               int32_t ladybug, int64_t* made, int32_t* degenerate, int32_t* zeroNormals, char* err, int32_t errLen) {
  const int64_t nP = m->nPts, nF = m->nFaces, nC = m->nIdx;
  g->sensGen += 1;   // P / N are written: the trace's sensor order is made again
  const int per = sides == 2 ? 2 : 1;
  const int64_t n = (onVertices ? nP : nF) * per;
  *made = onVertices ? nP : nF;
  *degenerate = 0;
  *zeroNormals = 0;
  if (j0 < 0 || j0 + n > g->nSens) { dlg_fail(err, errLen, "dlg_sensors: the sensors run past dlg_sensors_reserve's count"); return 1; }
  if (nP < 0 || nF < 0 || nC < 0 || nP > int64_t(UINT32_MAX) || nC > int64_t(INT32_MAX)) { dlg_fail(err, errLen, "dlg_sensors: mesh sizes out of range"); return 1; }
  if (nF == 0 && !(onVertices && nP > 0)) return 0;
  if (nF > 0 && nP == 0) { dlg_fail(err, errLen, "dlg_sensors: faces without points"); return 1; }
  // the mesh on the device, then the work arrays, in one buffer
  size_t at = 0;
  const auto take = [&](size_t bytes) { const size_t o = at; at = (at + bytes + 255) / 256 * 256; return o; };
  const size_t oXyz = take(size_t(nP) * 24), oCnt = take(size_t(nF) * 4), oIdx = take(size_t(nC) * 4), oStart = take(size_t(nF) * 8),
               oStart64 = take(size_t(nF) * 8), oW = take(size_t(nF) * 24), oKey = take(size_t(nC) * 4), oVal = take(size_t(nC) * 4),
               oKey2 = take(size_t(nC) * 4), oVal2 = take(size_t(nC) * 4), oFace = take(size_t(nC) * 4), oFirst = take(size_t(nP) * 8),
               oCnt2 = take(8);
  size_t tmp = 0, scan = 0;
  if (nC > 0 && onVertices)
    DLG_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                                 static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), int(nC)));
  if (nF > 0) DLG_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan, static_cast<const int64_t*>(nullptr), static_cast<int64_t*>(nullptr), int(nF)));
  const size_t oTmp = take(std::max(tmp, scan));
  DLG_CHECK(g->sensWork.grow(at));
  char* b = g->sensWork.as<char>();
  const auto dp = [&](size_t o) { return static_cast<void*>(b + o); };
  double* xyz = static_cast<double*>(dp(oXyz));
  int32_t* cnts = static_cast<int32_t*>(dp(oCnt));
  int32_t* idx = static_cast<int32_t*>(dp(oIdx));
  int64_t* cnt64 = static_cast<int64_t*>(dp(oStart64));
  int64_t* start = static_cast<int64_t*>(dp(oStart));
  int32_t* counters = static_cast<int32_t*>(dp(oCnt2));
  hipStream_t st = g->stream;
  if (nP) DLG_CHECK(hipMemcpyAsync(xyz, m->xyz, size_t(nP) * 24, hipMemcpyHostToDevice, st));
  if (nF) DLG_CHECK(hipMemcpyAsync(cnts, m->counts, size_t(nF) * 4, hipMemcpyHostToDevice, st));
  if (nC) DLG_CHECK(hipMemcpyAsync(idx, m->idx, size_t(nC) * 4, hipMemcpyHostToDevice, st));
  DLG_CHECK(hipMemsetAsync(counters, 0, 8, st));
  const Out out{g->P.as<float3>(), g->N.as<float3>(), sides, offset, d3{centre[0], centre[1], centre[2]}};
  if (nF > 0) {   // each face's first corner: the exclusive scan of the counts, in 64 bits
    k_widen<<<blocks(nF), kBlock, 0, st>>>(cnts, nF, cnt64);
    size_t tb = scan;
    DLG_CHECK(hipcub::DeviceScan::ExclusiveSum(dp(oTmp), tb, cnt64, start, int(nF), st));
  }
  double* w = static_cast<double*>(dp(oW));
  if (!onVertices) {
    const auto k = ladybug ? k_faces<true, false> : k_faces<false, false>;
    k<<<blocks(nF), kBlock, 0, st>>>(xyz, nP, cnts, start, idx, nC, nF, out, j0, per, nullptr, counters);
  } else {
    uint32_t* key = static_cast<uint32_t*>(dp(oKey));
    uint32_t* val = static_cast<uint32_t*>(dp(oVal));
    uint32_t* key2 = static_cast<uint32_t*>(dp(oKey2));
    uint32_t* val2 = static_cast<uint32_t*>(dp(oVal2));
    uint32_t* faceOf = static_cast<uint32_t*>(dp(oFace));
    int64_t* first = static_cast<int64_t*>(dp(oFirst));
    if (nF > 0) {
      const auto k = ladybug ? k_faces<true, true> : k_faces<false, true>;
      k<<<blocks(nF), kBlock, 0, st>>>(xyz, nP, cnts, start, idx, nC, nF, out, j0, per, w, counters);
      k_corners<<<blocks(nF), kBlock, 0, st>>>(cnts, start, idx, nC, nF, nP, key, val, faceOf);
    }
    if (nC > 0) {   // the corners by point, stable: each point's corners in corner order (its faces in face order)
      int endBit = 1;
      while (endBit < 32 && (int64_t(1) << endBit) < nP) ++endBit;
      size_t tb = tmp;
      DLG_CHECK(hipcub::DeviceRadixSort::SortPairs(dp(oTmp), tb, key, key2, val, val2, int(nC), 0, endBit, st));
    }
    k_fill<<<blocks(nP), kBlock, 0, st>>>(first, nP, -1);
    if (nC > 0) k_first<<<blocks(nC), kBlock, 0, st>>>(key2, nC, first);
    k_points<<<blocks(nP), kBlock, 0, st>>>(xyz, nP, key2, val2, faceOf, w, first, nC, out, j0, per, counters);
  }
  DLG_CHECK(hipGetLastError());
  int32_t c[2] = {0, 0};
  DLG_CHECK(hipMemcpyAsync(c, counters, 8, hipMemcpyDeviceToHost, st));
  DLG_CHECK(hipStreamSynchronize(st));
  *degenerate = c[0];
  *zeroNormals = c[1];
  return 0;
}

extern "C" int32_t dlg_download_sensors(dlg_gpu* g, float* P, float* N, int64_t n, char* err, int32_t errLen) {
  if (n < 0 || n > g->nSens) { dlg_fail(err, errLen, "dlg_download_sensors: more sensors than resident"); return 1; }
  if (n == 0) return 0;
  DLG_CHECK(hipMemcpyAsync(P, g->P.p, size_t(n) * 12, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipMemcpyAsync(N, g->N.p, size_t(n) * 12, hipMemcpyDeviceToHost, g->stream));
  DLG_CHECK(hipStreamSynchronize(g->stream));
  return 0;
}
