// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Geometry of the meshes on the host threads: the bounding boxes every coordinate is recentred by, the triangles a scene
// draws (a quad as (0 1 2) (0 2 3); faces of 5 or more corners ear-clipped in their Newell plane), and the thin-solid
// test of a transmitting slot. The sensors of a mesh are made on the GPU (sensors.hip).
//
// The reference triangulates n-gons with earcut.hpp (v2.2.4, fetched at its build time, absent here); this library's ear
// clipping covers the same polygon, possibly with other diagonals.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <exception>
#include <mutex>
#include <thread>

#include "core.h"

namespace daylight {
namespace {

struct D3 {
  double x = 0, y = 0, z = 0;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(D3 a) { return std::sqrt(dot(a, a)); }
inline D3 normalised(D3 a) {
  const double l = norm(a);
  return l > 0 ? a * (1 / l) : D3{};
}
inline float3 to_f3(D3 a) { return make_float3(float(a.x), float(a.y), float(a.z)); }

// the corners of one face of a mesh
struct Face {
  const SimulationMesh& m;
  const int* idx;
  int n;
  D3 at(int k) const {
    const double* p = m.points.data() + 3 * size_t(idx[k]);
    return {p[0], p[1], p[2]};
  }
  // Newell's vector: twice the area, along the normal the winding gives
  D3 newell() const {
    D3 N;
    for (int i = 0; i < n; ++i) {
      const D3 a = at(i), b = at((i + 1) % n);
      N.x += (a.y - b.y) * (a.z + b.z);
      N.y += (a.z - b.z) * (a.x + b.x);
      N.z += (a.x - b.x) * (a.y + b.y);
    }
    return N;
  }
  D3 corner_mean() const {
    D3 c;
    for (int i = 0; i < n; ++i) c = c + at(i);
    return c * (1.0 / n);
  }
};

// Ear clipping of a simple polygon given counterclockwise in 2D: its triangles' corner indices (n - 2 of them; fewer,
// or none, when it is degenerate or no ear is found).
std::vector<int> clip_ears(const std::vector<std::array<double, 2>>& q) {
  std::vector<int> tri, ring(q.size());
  for (size_t k = 0; k < q.size(); ++k) ring[k] = int(k);
  const auto turn = [&](int a, int b, int c) {   // > 0: a b c turn left
    const auto &A = q[size_t(a)], &B = q[size_t(b)], &Cc = q[size_t(c)];
    return (B[0] - A[0]) * (Cc[1] - A[1]) - (B[1] - A[1]) * (Cc[0] - A[0]);
  };
  for (size_t budget = 4 * q.size() * q.size(); ring.size() > 3 && budget > 0;) {
    bool cut = false;
    for (size_t k = 0; k < ring.size() && budget > 0; ++k, --budget) {
      const size_t r = ring.size();
      const int a = ring[(k + r - 1) % r], b = ring[k], c = ring[(k + 1) % r];
      if (!(turn(a, b, c) > 0)) continue;   // reflex or flat
      const bool blocked = std::any_of(ring.begin(), ring.end(), [&](int p) {
        return p != a && p != b && p != c && turn(a, b, p) >= 0 && turn(b, c, p) >= 0 && turn(c, a, p) >= 0;
      });
      if (blocked) continue;
      tri.insert(tri.end(), {a, b, c});
      ring.erase(ring.begin() + std::ptrdiff_t(k));
      cut = true;
      break;
    }
    if (!cut) return tri;
  }
  if (ring.size() == 3 && turn(ring[0], ring[1], ring[2]) > 0) tri.insert(tri.end(), {ring[0], ring[1], ring[2]});
  return tri;
}

// A face of n >= 5 corners: its corners projected on its Newell plane and ear-clipped -> the mesh's corner indices of
// its triangles (3 per triangle, at most n - 2 triangles); empty when it has no normal or no ear is found.
std::vector<int> ngon_triangles(const Face& f) {
  const D3 N = f.newell();
  if (norm(N) == 0) return {};
  const D3 u3 = normalised(N);
  // the plane's axes (a right-handed u, v, N: the winding stays counterclockwise), from the axis least along N
  const D3 axis = std::fabs(u3.x) <= std::fabs(u3.y) && std::fabs(u3.x) <= std::fabs(u3.z) ? D3{1, 0, 0}
                  : std::fabs(u3.y) <= std::fabs(u3.z)                                       ? D3{0, 1, 0}
                                                                                               : D3{0, 0, 1};
  const D3 u = normalised(cross(u3, axis)), v = cross(u3, u);
  std::vector<std::array<double, 2>> poly(size_t(f.n));
  for (int k = 0; k < f.n; ++k) poly[size_t(k)] = {dot(f.at(k), u), dot(f.at(k), v)};
  std::vector<int> local = clip_ears(poly), out;
  for (size_t k = 0; k + 2 < local.size() && out.size() < 3 * size_t(f.n - 2); k += 3)
    for (int q = 0; q < 3; ++q) out.push_back(f.idx[local[k + size_t(q)]]);
  return out;
}

}  // namespace

void par_for(size_t n, const std::function<void(size_t)>& body) {
  const size_t workers = std::min<size_t>(n, std::max(1u, std::thread::hardware_concurrency()));
  if (workers <= 1) {
    for (size_t k = 0; k < n; ++k) body(k);
    return;
  }
  std::atomic<size_t> next{0};
  std::exception_ptr first;
  std::mutex mu;
  const auto work = [&] {
    for (size_t k; (k = next.fetch_add(1)) < n;) {
      try {
        body(k);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        if (!first) first = std::current_exception();
      }
    }
  };
  std::vector<std::thread> pool;
  pool.reserve(workers - 1);
  for (size_t t = 1; t < workers; ++t) pool.emplace_back(work);
  work();
  for (std::thread& t : pool) t.join();
  if (first) std::rethrow_exception(first);
}

void bbox_of(const std::vector<SimulationMeshView>& meshes, const std::function<bool(size_t)>& which, double box[6]) {
  // blocks of points bounded on the host threads (min and max are exact in any order), then merged
  constexpr size_t kBlock = size_t(1) << 18;
  struct Span { const double* p; size_t n; };
  std::vector<Span> spans;
  for (size_t k = 0; k < meshes.size(); ++k) {
    if (which && !which(k)) continue;
    const SimulationMesh& m = *meshes[k].mesh;
    for (size_t i = 0, n = m.numPoints(); i < n; i += kBlock) spans.push_back({m.points.data() + 3 * i, std::min(kBlock, n - i)});
  }
  std::vector<std::array<double, 6>> part(spans.size());
  par_for(spans.size(), [&](size_t s) {
    std::array<double, 6> b = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (const double *p = spans[s].p, *e = p + 3 * spans[s].n; p < e; p += 3)
      for (int a = 0; a < 3; ++a) { b[size_t(a)] = std::fmin(b[size_t(a)], p[a]); b[size_t(3 + a)] = std::fmax(b[size_t(3 + a)], p[a]); }
    part[s] = b;
  });
  for (size_t a = 0; a < 3; ++a) { box[a] = INFINITY; box[3 + a] = -INFINITY; }
  for (const auto& b : part) box_union(box, b.data());
}

void box_union(double box[6], const double other[6]) {
  for (size_t a = 0; a < 3; ++a) { box[a] = std::fmin(box[a], other[a]); box[3 + a] = std::fmax(box[3 + a], other[3 + a]); }
}

void box_centre(const double box[6], double c[3]) {
  for (size_t a = 0; a < 3; ++a) c[a] = box[a] <= box[3 + a] ? (box[a] + box[3 + a]) / 2 : 0;
}

size_t tri_cap(const SimulationMesh& m) {
  size_t n = 0;
  for (const int k : m.counts) n += k >= 3 ? size_t(k - 2) : 0;
  return n;
}

size_t slot_count(const SimulationMesh& m) { return m.slot.empty() ? 1 : size_t(*std::max_element(m.slot.begin(), m.slot.end())) + 1; }

size_t triangulate_into(const SimulationMesh& m, const double c[3], float3* vOut, uint3* tri, uint32_t* slots, unsigned base, uint32_t slotBase,
                        int& skipped) {
  const D3 C{c[0], c[1], c[2]};
  for (size_t i = 0; i < m.numPoints(); ++i) vOut[i] = to_f3(D3{m.points[3 * i], m.points[3 * i + 1], m.points[3 * i + 2]} - C);
  size_t made = 0;
  const int* idx = m.indices.data();
  for (size_t f = 0; f < m.counts.size(); idx += m.counts[f], ++f) {
    const int n = m.counts[f];
    const uint32_t slot = slotBase + (m.slot.empty() ? 0u : uint32_t(m.slot[f]));
    const auto emit = [&](int a, int b, int d) {
      if (slots) slots[made] = slot;
      tri[made++] = make_uint3(base + unsigned(a), base + unsigned(b), base + unsigned(d));
    };
    if (n == 3 || n == 4) {
      for (int t = 0; t + 2 < n; ++t) emit(idx[0], idx[t + 1], idx[t + 2]);
      continue;
    }
    const std::vector<int> nt = n >= 5 ? ngon_triangles(Face{m, idx, n}) : std::vector<int>{};
    if (nt.empty()) { ++skipped; continue; }
    for (size_t k = 0; k < nt.size(); k += 3) emit(nt[k], nt[k + 1], nt[k + 2]);
  }
  return made;
}

void ngon_faces(const SimulationMesh& m, std::vector<int64_t>& face, std::vector<int32_t>& made, std::vector<uint32_t>& tri, int& skipped) {
  face.clear();
  made.clear();
  tri.clear();
  const int* idx = m.indices.data();
  for (size_t f = 0; f < m.counts.size(); idx += m.counts[f], ++f) {
    const int n = m.counts[f];
    if (n < 3) ++skipped;
    if (n < 5) continue;
    const std::vector<int> nt = ngon_triangles(Face{m, idx, n});
    if (nt.empty()) ++skipped;
    face.push_back(int64_t(f));
    made.push_back(int32_t(nt.size() / 3));
    tri.insert(tri.end(), nt.begin(), nt.end());
  }
}

uint32_t thin_solid(const SimulationMesh& m, uint32_t s, const double c[3]) {
  const D3 C{c[0], c[1], c[2]};
  // the slot's faces: their corners welded on a 1e-5 m grid, their area, and the signed volume sum N . p0 / 6
  using Cell = std::array<long long, 3>;
  std::vector<Cell> cells;
  std::vector<int> sizes;
  double area = 0, volume = 0;
  const int* idx = m.indices.data();
  for (size_t f = 0; f < m.counts.size(); idx += m.counts[f], ++f) {
    if ((m.slot.empty() ? 0u : uint32_t(m.slot[f])) != s) continue;
    const Face face{m, idx, m.counts[f]};
    const D3 N = face.newell();
    area += norm(N) / 2;
    volume += dot(N, face.at(0) - C) / 6;
    for (int j = 0; j < face.n; ++j) {
      const D3 p = face.at(j) - C;
      cells.push_back({std::llround(p.x * 1e5), std::llround(p.y * 1e5), std::llround(p.z * 1e5)});
    }
    sizes.push_back(face.n);
  }
  if (!(area > 0)) return 0;
  const double thickness = 2 * std::fabs(volume) / area;
  if (thickness >= 0.1) return 0;
  // one id per welded position (the corners sorted by cell)
  std::vector<uint32_t> order(cells.size()), id(cells.size());
  for (uint32_t j = 0; j < order.size(); ++j) order[j] = j;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cells[a] < cells[b]; });
  uint32_t next = 0;
  for (size_t j = 0; j < order.size(); ++j) {
    if (j > 0 && cells[order[j]] != cells[order[j - 1]]) ++next;
    id[order[j]] = next;
  }
  // closed and consistently wound: every directed edge between welded ids has its reverse (degenerate edges dropped)
  std::vector<uint64_t> edges, reversed;
  for (size_t f = 0, j0 = 0; f < sizes.size(); j0 += size_t(sizes[f++]))
    for (int j = 0; j < sizes[f]; ++j) {
      const uint64_t a = id[j0 + size_t(j)], b = id[j0 + size_t((j + 1) % sizes[f])];
      if (a == b) continue;
      edges.push_back(a << 32 | b);
      reversed.push_back(b << 32 | a);
    }
  std::sort(edges.begin(), edges.end());
  std::sort(reversed.begin(), reversed.end());
  if (edges != reversed) return 0;
  return DLG_SLOT_THIN | (thickness < 1e-5 ? DLG_SLOT_TWIN : volume < 0 ? DLG_SLOT_INWARD : 0u);
}

}  // namespace daylight
