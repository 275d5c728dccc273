// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Mesh inputs of the path (reference src/mesh.cpp): OBJ / .dmesh readers, the occluder triangles (quads split
// (0,1,2)/(0,2,3); n-gons by ear clipping in the polygon's plane), sensor points and normals by the physical rules
// (Newell normal, signed-area fan centroid, area-weighted vertex normals, sides, offset, recentring), and the option rows
// that name meshes (--optics, --no-occlude, --target).
// The reference triangulates n-gons with earcut.hpp (v2.2.4, fetched at build time, not in the reference tree); this
// build's ear clipping gives the same area for simple polygons but may pick other diagonals.
#include <algorithm>
#include <atomic>
#include <array>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "core.h"
#include "dl_math.h"

namespace daylight {
namespace {

struct V3 { double x, y, z; };
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
double len(V3 a) { return std::sqrt(dot(a, a)); }
V3 unit(V3 a) { const double l = len(a); return l > 0 ? a * (1 / l) : V3{0, 0, 0}; }
V3 pt(const SimulationMesh& m, int i) { return {m.points[3 * size_t(i)], m.points[3 * size_t(i) + 1], m.points[3 * size_t(i) + 2]}; }
float3 f3(V3 a) { return make_float3(float(a.x), float(a.y), float(a.z)); }

// Newell's normal: length = twice the area, direction by the winding
V3 newell(const SimulationMesh& m, const int* idx, int n) {
  V3 N{0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const V3 a = pt(m, idx[i]), b = pt(m, idx[(i + 1) % n]);
    N.x += (a.y - b.y) * (a.z + b.z);
    N.y += (a.z - b.z) * (a.x + b.x);
    N.z += (a.x - b.x) * (a.y + b.y);
  }
  return N;
}

// Face geometry by the mode's rules (reference mesh.cpp:45-69). Physical: unit Newell normal (0 under 1e-12 area),
// signed-area-weighted fan centroid. Ladybug (RadiationStudy's, tri / quad faces): the corners' mean; a triangle's
// cross product; a quad's two triangle normals (0 1 2) and (2 3 1) summed, the area their half-lengths summed.
void face_geom(const SimulationMesh& m, const int* idx, int n, bool ladybug, V3& nrm, double& area, V3& cen) {
  nrm = {0, 0, 0}; area = 0; cen = {0, 0, 0};
  if (n < 3) return;
  const V3 P0 = pt(m, idx[0]);
  if (ladybug) {
    for (int i = 0; i < n; ++i) cen = cen + pt(m, idx[i]);
    cen = cen * (1.0 / n);
    V3 N{0, 0, 0};
    if (n == 3) { N = cross(pt(m, idx[1]) - P0, pt(m, idx[2]) - P0); area = len(N) / 2; }
    else if (n == 4) {
      const V3 P1 = pt(m, idx[1]), P2 = pt(m, idx[2]), P3 = pt(m, idx[3]);
      const V3 n1 = cross(P1 - P0, P2 - P0), n2 = cross(P3 - P2, P1 - P2);
      N = n1 + n2;
      area = (len(n1) + len(n2)) / 2;
    }
    nrm = area < 1e-12 ? V3{0, 0, 0} : unit(N);
    return;
  }
  const V3 N = newell(m, idx, n), Nu = unit(N);
  area = len(N) / 2;
  double sum = 0;
  for (int i = 1; i + 1 < n; ++i) {
    const V3 a = pt(m, idx[i]), b = pt(m, idx[i + 1]);
    const double w = dot(cross(a - P0, b - P0), Nu) / 2;
    cen = cen + (P0 + a + b) * (w / 3);
    sum += w;
  }
  if (sum > 1e-12) cen = cen * (1 / sum);
  else {
    cen = {0, 0, 0};
    for (int i = 0; i < n; ++i) cen = cen + pt(m, idx[i]);
    cen = cen * (1.0 / n);
  }
  nrm = area < 1e-12 ? V3{0, 0, 0} : unit(N);
}

// Ear clipping of a simple polygon given counterclockwise in 2D: n - 2 triangles (corner indices), fewer if it is degenerate.
void ear_clip(const std::vector<std::array<double, 2>>& q, std::vector<int>& tri) {
  std::vector<int> ring(q.size());
  for (size_t k = 0; k < q.size(); ++k) ring[k] = int(k);
  const auto cr = [&](int a, int b, int c) {
    return (q[size_t(b)][0] - q[size_t(a)][0]) * (q[size_t(c)][1] - q[size_t(a)][1]) - (q[size_t(b)][1] - q[size_t(a)][1]) * (q[size_t(c)][0] - q[size_t(a)][0]);
  };
  size_t guard = 0;
  while (ring.size() > 3 && guard < 4 * q.size() * q.size()) {
    bool cut = false;
    for (size_t k = 0; k < ring.size(); ++k, ++guard) {
      const int a = ring[(k + ring.size() - 1) % ring.size()], b = ring[k], c = ring[(k + 1) % ring.size()];
      if (!(cr(a, b, c) > 0)) continue;   // reflex or flat: not an ear
      bool inside = false;
      for (const int p : ring) {
        if (p == a || p == b || p == c) continue;
        if (cr(a, b, p) >= 0 && cr(b, c, p) >= 0 && cr(c, a, p) >= 0) { inside = true; break; }
      }
      if (inside) continue;
      tri.insert(tri.end(), {a, b, c});
      ring.erase(ring.begin() + std::ptrdiff_t(k));
      cut = true;
      break;
    }
    if (!cut) return;
  }
  if (ring.size() == 3 && cr(ring[0], ring[1], ring[2]) > 0) tri.insert(tri.end(), {ring[0], ring[1], ring[2]});
}

std::string stem(const std::string& path) {
  const size_t s = path.find_last_of("/\\"), b = s == std::string::npos ? 0 : s + 1, d = path.rfind('.');
  return path.substr(b, d == std::string::npos || d < b ? std::string::npos : d - b);
}

using Names = std::unordered_map<std::string, size_t>;
Names mesh_names(const std::vector<SimulationMeshView>& meshes) {
  Names byName;
  for (size_t k = 0; k < meshes.size(); ++k) if (!byName.emplace(meshes[k].name, k).second) byName[meshes[k].name] = SIZE_MAX;
  return byName;
}
size_t find_mesh(const Names& byName, const std::string& name, const std::string& row, std::string& err) {
  const auto it = byName.find(name);
  if (it == byName.end()) err = row + "no mesh of that name";
  else if (it->second == SIZE_MAX) err = row + "several meshes have that name";
  else return it->second;
  return SIZE_MAX;
}

bool read_file(const std::string& path, std::string& buf) {
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::rewind(f);
  buf.assign(n > 0 ? size_t(n) : 0, '\0');
  const bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  return ok;
}

}  // namespace

void par_for(size_t n, const std::function<void(size_t)>& f) {
  const size_t nt = std::min<size_t>(n, std::max(1u, std::thread::hardware_concurrency()));
  if (nt <= 1) { for (size_t k = 0; k < n; ++k) f(k); return; }
  std::exception_ptr first;
  std::mutex mu;
  std::atomic<size_t> next{0};
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (size_t k; (k = next++) < n;) {
        try { f(k); } catch (...) { std::lock_guard<std::mutex> l(mu); if (!first) first = std::current_exception(); }
      }
    });
  for (std::thread& t : th) t.join();
  if (first) std::rethrow_exception(first);
}

void bbox_centre(const std::vector<SimulationMeshView>& meshes, double c[3]) {
  // blocks of points on the host threads, each with its own bounds (min and max are exact in any order)
  constexpr size_t kBlock = size_t(1) << 18;
  struct Job { const double* p; size_t n; };
  std::vector<Job> jobs;
  for (const SimulationMeshView& v : meshes)
    for (size_t i = 0, n = v.mesh->numPoints(); i < n; i += kBlock) jobs.push_back({v.mesh->points.data() + 3 * i, std::min(kBlock, n - i)});
  std::vector<std::array<double, 6>> b(jobs.size());
  par_for(jobs.size(), [&](size_t j) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const double* p = jobs[j].p;
    for (size_t i = 0; i < jobs[j].n; ++i, p += 3)
      for (int k = 0; k < 3; ++k) { lo[k] = std::fmin(lo[k], p[k]); hi[k] = std::fmax(hi[k], p[k]); }
    b[j] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
  });
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (const std::array<double, 6>& x : b)
    for (int k = 0; k < 3; ++k) { lo[k] = std::fmin(lo[k], x[size_t(k)]); hi[k] = std::fmax(hi[k], x[size_t(3 + k)]); }
  for (int k = 0; k < 3; ++k) c[k] = lo[k] <= hi[k] ? (lo[k] + hi[k]) / 2 : 0;
}

size_t tri_cap(const SimulationMesh& m) {
  size_t n = 0;
  for (const int k : m.counts) n += k >= 3 ? size_t(k - 2) : 0;
  return n;
}

// A face of n >= 5 corners: its Newell plane, the corners projected on it and ear-clipped; out[3 k ..] the corner
// indices of its triangles (at most n - 2). -1: degenerate (no normal, or no ear found), the face is skipped.
static int ngon_into(const SimulationMesh& m, const int* idx, int n, std::vector<std::array<double, 2>>& poly, std::vector<int>& ear, int* out) {
  V3 N = newell(m, idx, n);
  if (len(N) == 0) return -1;
  N = unit(N);
  const V3 ax = std::fabs(N.x) <= std::fabs(N.y) && std::fabs(N.x) <= std::fabs(N.z) ? V3{1, 0, 0}
                : std::fabs(N.y) <= std::fabs(N.z)                                     ? V3{0, 1, 0}
                                                                                         : V3{0, 0, 1};
  const V3 u = unit(cross(N, ax)), v = cross(N, u);   // (u, v, N) right-handed: the winding is counterclockwise in (u, v)
  poly.clear();
  for (int k = 0; k < n; ++k) { const V3 p = pt(m, idx[k]); poly.push_back({dot(p, u), dot(p, v)}); }
  ear.clear();
  ear_clip(poly, ear);
  if (ear.empty()) return -1;
  int made = 0;
  for (size_t k = 0; k + 2 < ear.size() && k / 3 < size_t(n - 2); k += 3, ++made)
    for (int q = 0; q < 3; ++q) out[3 * made + q] = idx[ear[k + size_t(q)]];
  return made;
}

size_t triangulate_into(const SimulationMesh& m, const double c[3], float3* vOut, uint3* tri, uint32_t* slots, unsigned base, uint32_t slotBase,
                        int& skipped) {
  const V3 C{c[0], c[1], c[2]};
  for (size_t i = 0; i < m.numPoints(); ++i) vOut[i] = f3(pt(m, int(i)) - C);
  size_t made = 0, off = 0;
  uint32_t slot = slotBase;   // the current face's global slot
  const auto emit = [&](int a, int b, int d) {
    if (slots) slots[made] = slot;
    tri[made++] = make_uint3(base + unsigned(a), base + unsigned(b), base + unsigned(d));
  };
  std::vector<int> ear, out;
  std::vector<std::array<double, 2>> poly;
  for (size_t f = 0; f < m.counts.size(); ++f) {
    const int n = m.counts[f];
    const int* idx = m.indices.data() + off;
    off += size_t(n);
    slot = slotBase + (m.slot.empty() ? 0u : uint32_t(m.slot[f]));
    if (n == 3) emit(idx[0], idx[1], idx[2]);
    else if (n == 4) { emit(idx[0], idx[1], idx[2]); emit(idx[0], idx[2], idx[3]); }
    else if (n >= 5) {
      out.resize(3 * size_t(n - 2));
      const int k = ngon_into(m, idx, n, poly, ear, out.data());
      if (k < 0) { ++skipped; continue; }
      for (int j = 0; j < k; ++j) emit(out[3 * size_t(j)], out[3 * size_t(j) + 1], out[3 * size_t(j) + 2]);
    } else ++skipped;
  }
  return made;
}

void ngon_faces(const SimulationMesh& m, std::vector<int64_t>& face, std::vector<int32_t>& made, std::vector<uint32_t>& tri, int& skipped) {
  face.clear();
  made.clear();
  tri.clear();
  std::vector<int> ear, out;
  std::vector<std::array<double, 2>> poly;
  size_t off = 0;
  for (size_t f = 0; f < m.counts.size(); ++f) {
    const int n = m.counts[f];
    const int* idx = m.indices.data() + off;
    off += size_t(n);
    if (n < 5) { if (n < 3) ++skipped; continue; }
    out.resize(3 * size_t(n - 2));
    int k = ngon_into(m, idx, n, poly, ear, out.data());
    if (k < 0) { ++skipped; k = 0; }
    face.push_back(int64_t(f));
    made.push_back(k);
    for (int j = 0; j < 3 * k; ++j) tri.push_back(uint32_t(out[size_t(j)]));
  }
}

size_t slot_count(const SimulationMesh& m) { return m.slot.empty() ? 1 : size_t(*std::max_element(m.slot.begin(), m.slot.end())) + 1; }

uint32_t thin_solid(const SimulationMesh& m, uint32_t s, const double c[3]) {
  const V3 C{c[0], c[1], c[2]};
  // the corners of the slot's faces welded to a 1e-5 m grid, face after face; the area and the signed volume (sum N . p0
  // / 6 over planar faces, |N| = twice the area)
  using Key = std::array<long long, 3>;
  std::vector<Key> keys;
  std::vector<int> n;
  double area = 0, vol = 0;
  size_t off = 0;
  for (size_t f = 0; f < m.counts.size(); ++f) {
    const int k = m.counts[f];
    const int* idx = m.indices.data() + off;
    off += size_t(k);
    if ((m.slot.empty() ? 0u : uint32_t(m.slot[f])) != s) continue;
    const V3 N = newell(m, idx, k);
    area += len(N) / 2;
    vol += dot(N, pt(m, idx[0]) - C) / 6;
    for (int j = 0; j < k; ++j) {
      const V3 p = pt(m, idx[j]) - C;
      keys.push_back({std::llround(p.x * 1e5), std::llround(p.y * 1e5), std::llround(p.z * 1e5)});
    }
    n.push_back(k);
  }
  if (!(area > 0) || 2 * std::fabs(vol) / area >= 0.1) return 0;
  // ids of the welded positions: the corners sorted by key, equal keys one id
  std::vector<uint32_t> order(keys.size()), id(keys.size());
  for (uint32_t j = 0; j < uint32_t(order.size()); ++j) order[j] = j;
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return keys[x] < keys[y]; });
  for (size_t j = 0, cur = 0; j < order.size(); ++j) {
    if (j > 0 && keys[order[j]] != keys[order[j - 1]]) ++cur;
    id[order[j]] = uint32_t(cur);
  }
  // the faces' directed edges between welded ids (degenerate ones dropped) and the same reversed: equal multisets when
  // the slot is closed and consistently wound
  std::vector<uint64_t> fwd, rev;
  fwd.reserve(keys.size());
  rev.reserve(keys.size());
  for (size_t f = 0, j0 = 0; f < n.size(); j0 += size_t(n[f++]))
    for (int j = 0; j < n[f]; ++j) {
      const uint64_t a = id[j0 + size_t(j)], b = id[j0 + size_t((j + 1) % n[f])];
      if (a == b) continue;
      fwd.push_back(a << 32 | b);
      rev.push_back(b << 32 | a);
    }
  std::sort(fwd.begin(), fwd.end());
  std::sort(rev.begin(), rev.end());
  if (fwd != rev) return 0;   // open, or wound against itself
  const double th = 2 * std::fabs(vol) / area;
  return DLG_SLOT_THIN | (th < 1e-5 ? DLG_SLOT_TWIN : vol < 0 ? DLG_SLOT_INWARD : 0u);
}

bool mesh_ok(const SimulationMesh& m, const std::string& label, std::string& err) {
  uint64_t sum = 0;
  for (const int c : m.counts) {
    if (c < 3) { err = label + ": face with fewer than 3 vertices"; return false; }
    sum += uint64_t(c);
  }
  if (sum != m.indices.size()) { err = label + ": sum(counts) != n_indices"; return false; }
  if (!m.slot.empty() && m.slot.size() != m.counts.size()) { err = label + ": one optics slot per face needed"; return false; }
  const uint64_t n = m.numPoints();
  for (const int i : m.indices) if (i < 0 || uint64_t(i) >= n) { err = label + ": vertex index out of range"; return false; }
  return true;
}

bool optics_rows_ok(const std::vector<Optics>& rows, std::string& err) {
  for (const Optics& o : rows) {
    const std::string row = "--optics " + o.mesh + (o.slot >= 0 ? "/" + std::to_string(o.slot) : "") + ": ";
    if (o.mesh.empty() || o.slot < -1 || o.slot > 65535 || (o.mesh == "*" && o.slot >= 0)) { err = row + "NAME (or *) with an optional /SLOT 0..65535"; return false; }
    if (!(o.rho >= 0 && o.rho <= 1 && o.tau >= 0 && o.tau <= 1)) { err = row + "RHO and TAU must be in 0..1"; return false; }
    if (o.rho + o.tau > 1 + 1e-9) { err = row + "RHO + TAU must not exceed 1"; return false; }
    if (o.glass && o.rho != 0) { err = row + "glass reflects by angle: RHO must be 0"; return false; }
    if (o.glass && !(o.tau > 0 && o.tau <= 0.918)) { err = row + "glass TAU must be in (0, 0.918] (n 1.52 passes at most 0.918)"; return false; }
  }
  return true;
}

bool optics_ok(const Options& opt, std::string& err) {
  if (!optics_rows_ok(opt.optics, err)) return false;
  if (opt.transmission && opt.mode == Mode::Ladybug) { err = "--transmission on needs --mode physical (ladybug: every surface opaque and black)"; return false; }
  if (opt.bounces < 0 || opt.bounces > 20) { err = "--bounces must be 0..20"; return false; }
  if (opt.bounceSamples < 64 || opt.bounceSamples > 65536) { err = "--bounce-samples must be 64..65536"; return false; }
  if (opt.bounces > 0 && opt.mode == Mode::Ladybug) { err = "--bounces needs --mode physical (ladybug: every surface opaque and black)"; return false; }
  return true;
}

bool resolve_optics(const std::vector<SimulationMeshView>& meshes, const std::vector<Optics>& rows, std::vector<SlotOptics>& out, std::string& err) {
  std::vector<size_t> base(meshes.size() + 1, 0);
  for (size_t k = 0; k < meshes.size(); ++k) {
    const SimulationMesh& m = *meshes[k].mesh;
    if (!m.slot.empty() && m.slot.size() != m.numFaces()) {
      err = "mesh " + meshes[k].name + ": " + std::to_string(m.slot.size()) + " slots for " + std::to_string(m.numFaces()) + " faces";
      return false;
    }
    base[k + 1] = base[k] + slot_count(m);
  }
  const Names byName = mesh_names(meshes);
  std::vector<std::pair<std::string, int>> keys;
  for (const Optics& o : rows) keys.emplace_back(o.mesh, o.slot);
  std::sort(keys.begin(), keys.end());
  if (const auto d = std::adjacent_find(keys.begin(), keys.end()); d != keys.end()) {
    err = "--optics " + d->first + (d->second >= 0 ? "/" + std::to_string(d->second) : "") + " given twice";
    return false;
  }
  out.assign(base.back(), SlotOptics{});
  for (int rank = 0; rank < 3; ++rank)   // * < NAME < NAME/SLOT: later ranks overwrite
    for (const Optics& o : rows) {
      if ((o.mesh == "*" ? 0 : o.slot < 0 ? 1 : 2) != rank) continue;
      const std::string row = "--optics " + o.mesh + (o.slot >= 0 ? "/" + std::to_string(o.slot) : "") + ": ";
      size_t b = 0, e = out.size();
      if (rank) {
        const size_t k = find_mesh(byName, o.mesh, row, err);
        if (k == SIZE_MAX) return false;
        b = base[k];
        e = base[k + 1];
        if (rank == 2 && size_t(o.slot) >= e - b) { err = row + "slot out of range (the mesh has slots 0.." + std::to_string(e - b - 1) + ")"; return false; }
        if (rank == 2) { b += size_t(o.slot); e = b + 1; }
      }
      const SlotOptics so{float(o.rho), float(o.tau), o.glass ? float(dlm::tn_from_Tn(o.tau)) : 0.f, o.glass ? DLG_SLOT_GLASS : 0u};
      std::fill(out.begin() + std::ptrdiff_t(b), out.begin() + std::ptrdiff_t(e), so);
    }
  return true;
}

bool resolve_no_occlude(const std::vector<SimulationMeshView>& meshes, const std::vector<std::string>& names, std::vector<char>& open, std::string& err) {
  open.assign(meshes.size(), 0);
  if (names.empty()) return true;
  const Names byName = mesh_names(meshes);
  for (const std::string& n : names) {
    const std::string row = "--no-occlude " + n + ": ";
    const size_t k = find_mesh(byName, n, row, err);
    if (k == SIZE_MAX) return false;
    if (!meshes[k].sensor) { err = row + "a context mesh (leave it out to drop its shade)"; return false; }
    open[k] = 1;
  }
  return true;
}

bool resolve_targets(const std::vector<SimulationMeshView>& meshes, const std::vector<views::Target>& targets, bool selfOcclude,
                     const std::vector<char>& open, std::vector<std::string>& groups, std::vector<uint8_t>& cls, std::string& err) {
  groups.clear();
  cls.assign(meshes.size(), 2);
  if (targets.empty()) return true;
  const Names byName = mesh_names(meshes);
  for (const views::Target& t : targets) {
    const std::string row = "--target " + t.mesh + "=" + t.group + ": ";
    const std::string& g = t.group;
    const size_t sl = t.mesh.rfind('/');
    if (sl != std::string::npos && sl + 1 < t.mesh.size() && t.mesh.find_first_not_of("0123456789", sl + 1) == std::string::npos) {
      err = row + "face-slot targets (NAME/SLOT) are not supported, target the whole mesh";
      return false;
    }
    if (g.empty() || g.find_first_not_of("ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789_-") != std::string::npos) {
      err = row + "GROUP takes letters, digits, _ and - only";
      return false;
    }
    const auto ends = [&](const char* s) { const size_t n = std::strlen(s); return g.size() >= n && g.compare(g.size() - n, n, s) == 0; };
    if (g == "sky" || g == "ground" || g == "obstruction" || g == "glazing" || g == "volume" || g == "area" || g.rfind("distance", 0) == 0 ||
        g.rfind("reach", 0) == 0 || ends("_distance") || ends("_distance_min")) {
      err = row + "GROUP is reserved (sky, ground, obstruction, glazing, volume, area, distance*, reach*, *_distance, *_distance_min)";
      return false;
    }
    const size_t k = find_mesh(byName, t.mesh, row, err);
    if (k == SIZE_MAX) return false;
    if (cls[k] != 2) { err = row + "the mesh is a target twice"; return false; }
    if (meshes[k].sensor && (!selfOcclude || open[k])) { err = row + "the mesh blocks no rays (--no-self-occlude or --no-occlude): none can end on it"; return false; }
    const size_t gi = size_t(std::find(groups.begin(), groups.end(), g) - groups.begin());
    if (gi == groups.size() && groups.size() == 16) { err = row + "more than 16 target groups"; return false; }
    if (gi == groups.size()) groups.push_back(g);
    cls[k] = uint8_t(3 + gi);
  }
  return true;
}

size_t make_sensors_into(const SimulationMesh& m, const RunOptions& opt, const double c[3], float3* pOut, float3* nOut, std::string& warn,
                         bool ladybug) {
  const V3 C{c[0], c[1], c[2]};
  if (ladybug)   // RadiationStudy's rule: tri / quad faces only (nothing made)
    for (size_t f = 0; f < m.counts.size(); ++f)
      if (m.counts[f] > 4) {
        warn = "error: ladybug mode supports tri/quad faces only (mesh " + m.name + ", face " + std::to_string(f) + ")";
        return 0;
      }
  int degenerate = 0, zeroVn = 0;
  std::vector<V3> vn;
  if (opt.onVertices) vn.assign(m.numPoints(), V3{0, 0, 0});
  size_t j = 0;
  const auto push = [&](V3 q, V3 nr) {
    if (opt.sides != Sides::Back) { pOut[j] = f3(q + nr * opt.offset - C); nOut[j] = f3(nr); ++j; }
    if (opt.sides != Sides::Front) { pOut[j] = f3(q - nr * opt.offset - C); nOut[j] = f3(nr * -1); ++j; }
  };
  const size_t nF = m.counts.size();
  constexpr size_t kBlock = 4096;
  if (!opt.onVertices && nF > 2 * kBlock) {
    // faces are independent: blocks of faces on the host threads, each writing its own sensors (face f: f x sides)
    const size_t nS = opt.sides == Sides::Both ? 2 : 1, nB = (nF + kBlock - 1) / kBlock;
    std::vector<size_t> start(nB + 1, 0);
    for (size_t b = 0, off = 0, f = 0; b < nB; ++b) {
      start[b] = off;
      for (const size_t e = std::min(nF, f + kBlock); f < e; ++f) off += size_t(m.counts[f]);
    }
    std::atomic<int> deg{0};
    par_for(nB, [&](size_t b) {
      size_t off = start[b], jb = b * kBlock * nS;
      int bad = 0;
      for (size_t f = b * kBlock, e = std::min(nF, f + kBlock); f < e; ++f) {
        const int n = m.counts[f];
        const int* idx = m.indices.data() + off;
        off += size_t(n);
        V3 nrm, cen;
        double area;
        face_geom(m, idx, n, ladybug, nrm, area, cen);
        if (area < 1e-12) ++bad;
        if (opt.sides != Sides::Back) { pOut[jb] = f3(cen + nrm * opt.offset - C); nOut[jb] = f3(nrm); ++jb; }
        if (opt.sides != Sides::Front) { pOut[jb] = f3(cen - nrm * opt.offset - C); nOut[jb] = f3(nrm * -1); ++jb; }
      }
      deg += bad;
    });
    degenerate = deg;
    j = nF * nS;
  } else {
    size_t off = 0;
    for (size_t f = 0; f < nF; ++f) {
      const int n = m.counts[f];
      const int* idx = m.indices.data() + off;
      off += size_t(n);
      V3 nrm, cen;
      double area;
      face_geom(m, idx, n, ladybug, nrm, area, cen);
      if (area < 1e-12) ++degenerate;
      if (opt.onVertices) for (int k = 0; k < n; ++k) vn[size_t(idx[k])] = vn[size_t(idx[k])] + nrm * area;
      else push(cen, nrm);
    }
  }
  for (size_t v = 0; v < vn.size(); ++v) {
    V3 n = vn[v];
    if (len(n) == 0) { n = {0, 0, 1}; ++zeroVn; }
    push(pt(m, int(v)), unit(n));
  }
  if (degenerate) warn += "mesh " + m.name + ": " + std::to_string(degenerate) + " degenerate face(s)";
  if (zeroVn) warn += std::string(degenerate ? "; " : "mesh " + m.name + ": ") + std::to_string(zeroVn) + " zero-length vertex normal(s) set to (0,0,1)";
  return opt.onVertices ? m.numPoints() : m.numFaces();
}

bool read_obj(const std::string& path, std::vector<SimulationMesh>& out, std::string& err) {
  std::string buf;
  if (!read_file(path, buf)) { err = "cannot open " + path; return false; }
  std::vector<double> V3s;
  std::vector<int> remap, touched;
  SimulationMesh cur;
  cur.name = stem(path);
  const auto finish = [&]() {
    if (!cur.counts.empty()) {
      std::sort(touched.begin(), touched.end());   // the mesh keeps its points in file order
      cur.points.resize(3 * touched.size());
      for (size_t k = 0; k < touched.size(); ++k) {
        remap[size_t(touched[k])] = int(k);
        for (int a = 0; a < 3; ++a) cur.points[3 * k + size_t(a)] = V3s[3 * size_t(touched[k]) + size_t(a)];
      }
      for (int& k : cur.indices) k = remap[size_t(k)];
      out.push_back(std::move(cur));
    }
    for (const int g : touched) remap[size_t(g)] = -1;
    touched.clear();
    cur = SimulationMesh();
  };
  const char* s = buf.c_str();
  const char* const end = s + buf.size();
  const auto ws = [](const char*& p) { while (*p == ' ' || *p == '\t' || *p == '\r') ++p; };
  const auto tok = [](const char*& p) { while (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n') ++p; };
  for (size_t lineNo = 1; s < end; ++lineNo) {
    const char* nl = static_cast<const char*>(std::memchr(s, '\n', size_t(end - s)));
    const char* const e = nl ? nl : end;
    ws(s);
    if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
      ++s;
      for (int a = 0; a < 3; ++a) {
        ws(s);
        if (*s == '+') ++s;
        double q;
        const auto r = std::from_chars(s, e, q);
        if (r.ec != std::errc()) { err = path + ":" + std::to_string(lineNo) + ": bad vertex"; return false; }
        s = r.ptr;
        V3s.push_back(q);
      }
      remap.push_back(-1);
    } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
      ++s;
      int n = 0;
      for (;;) {
        ws(s);
        if (s >= e || *s == '#') break;
        if (*s == '+') ++s;
        long a;
        const auto r = std::from_chars(s, e, a);
        if (r.ec != std::errc()) { err = path + ":" + std::to_string(lineNo) + ": bad face token"; return false; }
        s = r.ptr;
        tok(s);
        const long nv = long(remap.size());
        a = a < 0 ? nv + a : a - 1;
        if (a < 0 || a >= nv) { err = path + ":" + std::to_string(lineNo) + ": vertex index out of range"; return false; }
        if (remap[size_t(a)] < 0) { remap[size_t(a)] = int(touched.size()); touched.push_back(int(a)); }
        cur.indices.push_back(int(a));
        ++n;
      }
      if (n) cur.counts.push_back(n);
    } else if ((s[0] == 'o' || s[0] == 'g') && (s[1] == ' ' || s[1] == '\t')) {
      finish();
      s += 2;
      ws(s);
      const char* t = s;
      tok(t);
      cur.name = t > s ? std::string(s, t) : stem(path);
    }
    s = e + 1;
  }
  finish();
  return true;
}

namespace {
bool parse_dmesh(const char* data, size_t size, const std::string& name, const std::string& label, std::vector<SimulationMesh>& out, std::string& err) {
  struct Header { char magic[4]; uint32_t flags; uint64_t nPoints, nFaces, nIndices; };
  static_assert(sizeof(Header) == 32, "dmesh header");
  Header h{};
  if (size < sizeof h) { err = label + ": truncated"; return false; }
  std::memcpy(&h, data, sizeof h);
  if (std::memcmp(h.magic, "DLM1", 4)) { err = label + ": not a .dmesh file (magic DLM1)"; return false; }
  if (h.flags & ~1u) { err = label + ": unsupported header flags " + std::to_string(h.flags) + " (only bit 0, the slot block, is defined)"; return false; }
  if (!(h.nPoints <= size / 24 && h.nFaces <= size / 4 && h.nIndices <= size / 4 && 32 + 24 * h.nPoints + 4 * h.nFaces + 4 * h.nIndices <= size)) {
    err = label + ": truncated";
    return false;
  }
  const uint64_t nSlots = (h.flags & 1) ? h.nFaces : 0;
  if (32 + 24 * h.nPoints + 4 * h.nFaces + 4 * h.nIndices + 2 * nSlots > size) { err = label + ": truncated slot block"; return false; }
  SimulationMesh m;
  m.name = name;
  m.points.resize(3 * h.nPoints); m.counts.resize(h.nFaces); m.indices.resize(h.nIndices); m.slot.resize(nSlots);
  const char* p = data + sizeof h;
  std::memcpy(m.points.data(), p, 24 * h.nPoints); p += 24 * h.nPoints;
  std::memcpy(m.counts.data(), p, 4 * h.nFaces); p += 4 * h.nFaces;
  std::memcpy(m.indices.data(), p, 4 * h.nIndices); p += 4 * h.nIndices;
  std::memcpy(m.slot.data(), p, 2 * nSlots);
  if (!mesh_ok(m, label, err)) return false;
  if (h.nFaces) out.push_back(std::move(m));
  return true;
}
}  // namespace

bool read_dmesh(const char* data, size_t size, const std::string& name, std::vector<SimulationMesh>& out, std::string& err) {
  return parse_dmesh(data, size, name, name, out, err);
}

bool read_dmesh(const std::string& path, std::vector<SimulationMesh>& out, std::string& err) {
  std::string buf;
  if (!read_file(path, buf)) { err = "cannot open " + path; return false; }
  return parse_dmesh(buf.data(), buf.size(), stem(path), path, out, err);
}

}  // namespace daylight
