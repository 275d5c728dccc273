// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// views::field and views::run (reference src/views.cpp is the contract: the field's checks and messages, the cell
// weights, the target and glass labels, the columns, --dump-rays). The run is a pipeline of stages on the engine's
// primary lane:
//   field    the grid (rows from HI down, columns clockwise) and the distance statistics
//   roles    which meshes block, their target groups, which of their faces rays pass (glass with --transmission on)
//   grid     each cell's weight (its solid angle, or its angle when flat), the field summing to 1
//   scene    the resident scene, each triangle's label (a mesh triangulated again where only some faces pass)
//   sensors  the lane's sensors and the labels uploaded, the room for the run
//   chunks   per chunk of sensors: dlg_view (classes, distances) -> dlg_reduce_view -> the columns (and the dump)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <new>
#include <unordered_map>

#include "engine.h"

namespace daylight::views {
namespace {

using Clk = std::chrono::steady_clock;
double ms_since(Clk::time_point t0) { return std::chrono::duration<double, std::milli>(Clk::now() - t0).count(); }
using File = std::unique_ptr<std::FILE, int (*)(std::FILE*)>;
constexpr double kDeg = 3.14159265358979323846 / 180;
constexpr uint8_t kPass = 255;   // a triangle rays pass (glass)

std::string reach_name(double D) {
  char s[48];
  std::snprintf(s, sizeof s, "reach_%g", D);
  return s;
}

std::string fmt(const char* f, double a, double b) {
  char s[256];
  std::snprintf(s, sizeof s, f, a, b);
  return s;
}

}  // namespace

bool field(const std::vector<SimulationMeshView>& meshes, const Options& opt, Field& f, std::string& err) {
  const auto no = [&](std::string msg) { err = std::move(msg); return false; };
  if (!(opt.fovH > 0 && opt.fovH <= 180) && opt.fovH != 360) return no("--fov H must be in (0, 180] (about the normal) or 360 (all around)");
  const bool flat = opt.elevLo == 0 && opt.elevHi == 0;
  if (!flat && !(opt.elevLo >= -90 && opt.elevLo < opt.elevHi && opt.elevHi <= 90))
    return no("--fov H,LO,HI needs -90 <= LO < HI <= 90, or LO = HI = 0 (horizontal rays only)");
  if (!(opt.cell > 0) || std::isinf(opt.cell)) return no("--cell must be positive and finite (degrees)");
  if (!std::isfinite(opt.north) || !std::isfinite(opt.offset)) return no("--north and --offset must be finite");
  if (!optics_rows_ok(opt.optics, err)) return false;
  const double cols = std::max(1.0, std::round(opt.fovH / opt.cell)), rows = std::max(1.0, std::round((opt.elevHi - opt.elevLo) / opt.cell));
  if (cols * rows > double(1 << 20)) return no(fmt("--fov / --cell: %.0f x %.0f rays per sensor exceed 2^20 (1048576); use a coarser --cell", rows, cols));
  if (!(opt.farDist > 0.1) || std::isinf(opt.farDist)) return no("--far must be more than 0.1 (metres) and finite");
  if (opt.reach.size() > 8) return no("--reach takes at most 8 distances");
  for (size_t k = 0; k < opt.reach.size(); ++k) {
    if (!(opt.reach[k] > 0) || std::isinf(opt.reach[k])) return no("--reach distances must be positive (metres) and finite");
    for (size_t j = 0; j < k; ++j)
      if (reach_name(opt.reach[j]) == reach_name(opt.reach[k])) return no("--reach distances must differ: two make column " + reach_name(opt.reach[k]));
  }
  if (std::isinf(opt.groundZ)) return no("--ground-z must be finite");
  const unsigned dist = opt.distance != DistAuto ? opt.distance : opt.fovH == 360 ? unsigned(DistIsovist) : unsigned(DistMean | DistMin);
  if ((dist & DistGroup) && !(dist & (DistMean | DistMin))) return no("--distance group adds per-group columns to mean and min: name one of them");
  if ((dist & DistIsovist) && (cols < 3 || (!flat && rows < 2))) {
    char s[256];
    std::snprintf(s, sizeof s, "--distance isovist needs at least 3 columns%s, the field has %.0f x %.0f: a finer --cell, --fov H,0,0 (the area) or --distance mean,min",
                  flat ? "" : " and 2 rows (a mesh)", rows, cols);
    return no(s);
  }
  f.nAz = unsigned(cols);
  f.nEl = unsigned(rows);
  f.dist = dist;
  f.flat = flat;
  f.groundZ = opt.groundZ;   // by default the lowest point of every mesh
  if (std::isnan(f.groundZ))
    for (const SimulationMeshView& v : meshes)
      for (size_t i = 2; i < v.mesh->points.size(); i += 3) f.groundZ = std::fmin(f.groundZ, v.mesh->points[i]);
  return true;
}

namespace {

class ViewsRun {
 public:
  ViewsRun(const std::vector<SimulationMeshView>& meshes, const Options& opt, Result& out, std::string& err, const char*& stage)
      : M(meshes), O(opt), out(out), err(err), stage(stage) {}

  bool go() {
    out = Result{};
    return field(M, O, fd, err) && roles() && grid() && open_dump() && scene() && sensors() && columns() && chunks();
  }
  bool dumping = false;   // the dump files were opened (a failed run removes them)

 private:
  const std::vector<SimulationMeshView>& M;
  const Options& O;
  Result& out;
  std::string& err;
  const char*& stage;

  Field fd;
  std::vector<char> open;
  std::vector<std::string> groups;
  std::vector<uint8_t> clsOf;       // per mesh: 2 obstruction, 3 + g target group g
  std::vector<SlotOptics> slots;
  std::vector<size_t> slotBase;
  std::vector<char> passAll, passAny;
  bool anyPass = false;
  std::string key;
  size_t nS = 1, nElem = 0, nSens = 0, K = 0, nClass = 3, NA = 0;
  bool iso = false, mean = false, least = false, perGroup = false;
  std::vector<float> rowSin;
  std::vector<double> wRow;
  File cls{nullptr, &std::fclose}, depth{nullptr, &std::fclose};
  Lane* lane = nullptr;
  dlg_gpu* g = nullptr;
  Scene sc;
  SceneLease lease;   // a run on a server's lane: its scene stays resident until the run ends
  std::vector<uint8_t> label;
  size_t chunkBytes = 0, perSensor = 1;
  dlg_stats_t st0{};

  bool fail(std::string msg) { err = std::move(msg); return false; }
  bool blocks(size_t k) const { return !M[k].sensor || (O.selfOcclude && !open[k]); }
  bool passes(size_t s) const { return O.transmission && (slots[s].flags & DLG_SLOT_GLASS) && slots[s].tau > 0; }

  // ---- roles (views.cpp:72-117)
  bool roles() {
    if (!resolve_no_occlude(M, O.noOcclude, open, err) || !resolve_targets(M, O.targets, O.selfOcclude, open, groups, clsOf, err) ||
        !resolve_optics(M, O.optics, slots, err))
      return false;
    // glass that rays pass: the global slots that are glass with tau > 0 under --transmission on (numbered as daylight
    // numbers them); per blocking mesh all of its faces, some, or none
    slotBase.assign(M.size() + 1, 0);
    for (size_t k = 0; k < M.size(); ++k) slotBase[k + 1] = slotBase[k] + slot_count(*M[k].mesh);
    passAll.assign(M.size(), 0);
    passAny.assign(M.size(), 0);
    for (size_t k = 0; O.transmission && k < M.size(); ++k) {
      if (!blocks(k)) continue;
      const SimulationMesh& m = *M[k].mesh;
      size_t n = 0;
      for (size_t f = 0; f < m.numFaces(); ++f) n += passes(slotBase[k] + (m.slot.empty() ? 0 : m.slot[f]));
      passAny[k] = n > 0;
      passAll[k] = n > 0 && n == m.numFaces();
      anyPass = anyPass || n > 0;
      if (clsOf[k] != 2 && passAll[k]) return fail("--target " + M[k].name + ": rays pass through it (glass with --transmission on), none can end on it");
    }
    if (!O.sceneKey.empty()) key = scene_key_of(O.sceneKey, O.selfOcclude) + open_key(open, O.selfOcclude);
    nS = O.sides == Sides::Both ? 2 : 1;
    for (const SimulationMeshView& v : M)
      if (v.sensor) nElem += O.onVertices ? v.mesh->numPoints() : v.mesh->numFaces();
    iso = fd.dist & DistIsovist;
    mean = fd.dist & DistMean;
    least = fd.dist & DistMin;
    perGroup = fd.dist & DistGroup;
    nSens = nElem * nS;
    K = size_t(fd.nAz) * fd.nEl;
    nClass = 3 + groups.size();
    NA = 3 * nClass + 1 + O.reach.size() + (O.transmission ? 1 : 0) + (iso ? 1 : 0);
    if (nSens == 0) return fail("no sensor elements (no sensor meshes, or none with faces / points)");
    if (!O.dumpPath.empty() && nSens * K * 5 > (size_t(4) << 30))
      return fail("--dump-rays: " + std::to_string(nSens) + " sensors x " + std::to_string(K) + " rays exceed 4 GB (fewer receivers or a coarser --cell)");
    return true;
  }

  // ---- grid: a row's cells weigh its solid angle (flat: its angle); the field sums to 1 (views.cpp:122-136)
  bool grid() {
    const double dAz = O.fovH / fd.nAz, dEl = (O.elevHi - O.elevLo) / fd.nEl;
    rowSin.assign(2 * size_t(fd.nEl), 0.f);
    wRow.assign(fd.nEl, 0.0);
    double total = 0;
    for (unsigned r = 0; r < fd.nEl; ++r) {
      const double top = std::sin((O.elevHi - r * dEl) * kDeg), bottom = std::sin((O.elevHi - (r + 1) * dEl) * kDeg);
      rowSin[2 * r] = float(bottom);
      rowSin[2 * r + 1] = float(top);
      wRow[r] = dAz * kDeg * (fd.flat ? 1 : top - bottom);
      total += wRow[r] * fd.nAz;
    }
    for (double& w : wRow) w /= total;
    return true;
  }

  bool open_dump() {
    if (O.dumpPath.empty()) return true;
    dumping = true;
    const std::string dpath = O.dumpPath + ".depth.npy";
    cls = File(std::fopen(O.dumpPath.c_str(), "wb"), &std::fclose);
    if (!cls || !npy_write_header(cls.get(), nSens, K, "<u1")) return fail("cannot write " + O.dumpPath);
    depth = File(std::fopen(dpath.c_str(), "wb"), &std::fclose);
    if (!depth || !npy_write_header(depth.get(), nSens, K, "<f4")) return fail("cannot write " + dpath);
    return true;
  }

  // ---- scene and labels: 2, 3 + a target mesh's group, or kPass for glass rays pass; a mixed mesh's triangles found by
  // triangulating it again as the scene did (same centre, same faces kept). A labelled mesh the resident scene lacks
  // (it was built under other names) forces a fresh build under the key.
  bool labels() {
    label.assign(sc.nTris, 2);
    std::unordered_map<std::string, const SceneMesh*> byName;
    for (const SceneMesh& s : sc.order)
      if (!byName.emplace(s.name, &s).second) byName[s.name] = nullptr;   // a name twice: ambiguous
    for (size_t k = 0; k < M.size(); ++k) {
      if (clsOf[k] == 2 && !passAny[k]) continue;
      const auto it = byName.find(M[k].name);
      if (it == byName.end() || !it->second || it->second->mesh != M[k].mesh) return false;
      const SceneMesh& s = *it->second;
      std::fill_n(label.begin() + std::ptrdiff_t(s.tri0), s.nTri, passAll[k] ? kPass : clsOf[k]);
      if (!passAny[k] || passAll[k]) continue;
      const SimulationMesh& m = *M[k].mesh;
      std::vector<float3> v(m.numPoints());
      std::vector<uint3> tri(tri_cap(m));
      std::vector<uint32_t> ts(tri.size());
      int skipped = 0;
      if (triangulate_into(m, sc.c, v.data(), tri.data(), ts.data(), 0, uint32_t(slotBase[k]), skipped) != s.nTri) return false;
      for (size_t t = 0; t < s.nTri; ++t)
        if (passes(ts[t])) label[s.tri0 + t] = kPass;
    }
    return true;
  }

  bool scene() {
    const auto blocksK = [this](size_t k) { return blocks(k); };
    lane = Engine::get().run_lane(err);
    if (!lane) return false;
    if (!Engine::get().acquire(lane, M, blocksK, false, key, O.residentScenes, O.cancel, &g, sc, out, stage, err, {}, {}, &lease)) return false;
    stage = "labelling the targets and the glass";
    if ((!groups.empty() || anyPass) && !labels()) {
      Engine::get().forget(key);
      if (!Engine::get().acquire(lane, M, blocksK, false, key, O.residentScenes, O.cancel, &g, sc, out, stage, err, {}, {}, &lease)) return false;
      if (!labels()) return fail("internal: a target or glass mesh is not in the scene as built");
    }
    dlg_stats(g, &st0);
    chunkBytes = O.chunkBytes;
    perSensor = std::max(4 * K, NA * sizeof(double));
    return Engine::get().admit(*lane, nSens, perSensor, 5 * K + NA * sizeof(double), label.size(), O.residentScenes, chunkBytes, err);
  }

  bool sensors() {
    stage = "making the sensors";
    const auto t0 = Clk::now();
    std::string warn;
    if (!lane_sensors(*lane, M, O, sc.c, nSens, out, warn, err)) return false;
    out.stats.prepMs += ms_since(t0);
    if (O.verbose) {
      char vg[48] = "off";
      if (O.virtualGround) std::snprintf(vg, sizeof vg, "at z %g", fd.groundZ);
      std::fprintf(stderr, "info: %zu sensors, %zu triangles%s, %u x %u = %zu rays per sensor, %zu target groups, virtual ground %s\n", nSens, sc.nTris,
                   out.sceneReused ? " (scene reused)" : "", fd.nEl, fd.nAz, K, groups.size(), vg);
      if (sc.skipped) std::fprintf(stderr, "warn: %d degenerate faces skipped in occluders\n", sc.skipped);
      if (!warn.empty()) std::fprintf(stderr, "warn: %s\n", warn.c_str());
      if (O.transmission) std::fprintf(stderr, "info: --transmission on: rays pass %zu glass triangles\n", size_t(std::count(label.begin(), label.end(), kPass)));
    }
    char e[512] = {};
    if (dlg_set_scene(g, nullptr, nullptr, -1, nullptr, 0, e, sizeof e) ||
        dlg_set_labels(g, label.data(), int64_t(label.size()), e, sizeof e))
      return fail(e);
    label = {};
    return true;
  }

  // ---- the columns (views.cpp:140-172) and a sensor's store of its sums (views.cpp:235-266)
  bool columns() {
    stage = "allocating the result";
    std::vector<std::string> names = {"sky", "ground", "obstruction"};
    names.insert(names.end(), groups.begin(), groups.end());
    if (O.transmission) names.push_back("glazing");
    if (iso) names.push_back(fd.flat ? "area" : "volume");
    if (mean) names.push_back("distance");
    if (least) names.push_back("distance_min");
    for (size_t k = 0; perGroup && k < groups.size(); ++k) {
      if (mean) names.push_back(groups[k] + "_distance");
      if (least) names.push_back(groups[k] + "_distance_min");
    }
    for (const double D : O.reach) names.push_back(reach_name(D));
    out.columns.clear();
    for (const std::string& n : names)
      for (size_t s = 0; s < nS; ++s) out.columns.push_back(n + ((O.sides == Sides::Back || (O.sides == Sides::Both && s == 1)) ? "_back" : "_front"));
    out.values.clear();
    out.f32.clear();
    if (O.f32Columns) out.f32.assign(out.columns.size() * nElem, 0.f);
    else out.values.assign(nElem * out.columns.size(), 0.0);
    return true;
  }

  // a[]: per class its weight, weight x ln d and least d; the own surface's weight; the reach sums; the glazing; the
  // isovist
  void store(size_t si, const double* a) {
    const double nearF = 0.1f, farF = float(O.farDist);
    const size_t el = si / nS, C = out.columns.size();
    size_t col = si % nS;
    const auto put = [&](double v) {
      if (O.f32Columns) out.f32[col * nElem + el] = float(v);
      else out.values[el * C + col] = v;
      col += nS;
    };
    // the log-mean distance (far: exactly the far distance) and the least one (clamped to [near, far]; none: NaN)
    const auto logMean = [&](double w, double wl, double m) { return !(w > 0) ? NAN : m >= farF ? O.farDist : std::exp(wl / w); };
    const auto minimum = [&](double m) { return std::isinf(m) ? NAN : m <= nearF ? 0.1 : m >= farF ? O.farDist : m; };
    const double own = a[3 * nClass];
    double w = 0, wl = 0, mn = INFINITY;
    for (size_t c = 0; c < nClass; ++c) {
      w += a[3 * c];
      wl += a[3 * c + 1];
      mn = std::min(mn, a[3 * c + 2]);
    }
    const double total = w + own;
    put(a[0] / total);
    put(a[3] / total);
    put((a[6] + own) / total);   // the own surface counts as obstruction
    for (size_t k = 0; k < groups.size(); ++k) put(a[3 * (3 + k)] / total);
    if (O.transmission) put(a[3 * nClass + 1 + O.reach.size()] / total);
    if (iso) put(a[NA - 1]);
    if (mean) put(logMean(w, wl, mn));
    if (least) put(minimum(mn));
    for (size_t k = 0; perGroup && k < groups.size(); ++k) {
      const double* gk = a + 3 * (3 + k);
      if (mean) put(logMean(gk[0], gk[1], gk[2]));
      if (least) put(minimum(gk[2]));
    }
    for (size_t r = 0; r < O.reach.size(); ++r) put(a[3 * nClass + 1 + r] / total);
  }

  bool chunks() {
    stage = "tracing";
    const bool world = O.fovH == 360;
    const size_t ncMax = std::max<size_t>(1, chunkBytes / perSensor);
    const dlg_view_in vin{fd.nAz, fd.nEl, float(world ? 0 : -O.fovH / 2), float(O.fovH / fd.nAz), rowSin.data(), world ? 1 : 0,
                          (world && O.offset == 0 && !fd.flat) ? 1 : 0, O.jitter ? 1 : 0, O.sides == Sides::Back ? 1 : 0, O.virtualGround ? 1 : 0,
                          uint32_t(nS), float(fd.groundZ - sc.c[2]), O.north};
    const uint32_t flags = (O.cullBackfaces ? 16u : 0u) | 1u;
    std::vector<float> reach(O.reach.begin(), O.reach.end());
    std::vector<double> sums;
    std::vector<uint8_t> cb;
    std::vector<float> db;
    char e[512] = {};
    for (size_t i0 = 0; i0 < nSens; i0 += ncMax) {
      if (O.cancel && O.cancel->load()) return fail("cancelled");
      const size_t nc = std::min(ncMax, nSens - i0);
      if (dlg_view(g, int64_t(i0), int64_t(nc), flags, &vin, e, sizeof e)) return fail(e);
      const auto t0 = Clk::now();
      sums.resize(nc * NA);
      if (dlg_reduce_view(g, int64_t(nc), wRow.data(), uint32_t(nClass), O.farDist, reach.data(), uint32_t(reach.size()), O.transmission ? 1 : 0,
                          iso ? (fd.flat ? 2u : 3u) : 0u, sums.data(), e, sizeof e))
        return fail(e);
      for (size_t i = 0; i < nc; ++i) store(i0 + i, sums.data() + i * NA);
      out.stats.reduceMs += ms_since(t0);
      if (O.progress) O.progress->store(double(i0 + nc) / double(nSens));
      if (!cls) continue;
      cb.resize(nc * K);
      db.resize(nc * K);
      if (dlg_download_view(g, int64_t(nc), cb.data(), db.data(), e, sizeof e)) return fail(e);
      if (std::fwrite(cb.data(), 1, cb.size(), cls.get()) != cb.size()) return fail("write failed: " + O.dumpPath);
      if (std::fwrite(db.data(), 4, db.size(), depth.get()) != db.size()) return fail("write failed: " + O.dumpPath + ".depth.npy");
    }
    dlg_stats_t s1{};
    dlg_stats(g, &s1);
    out.stats.traceMs = s1.traceMs - st0.traceMs;
    out.stats.rays = size_t(s1.rays - st0.rays);
    out.stats.launches = size_t(s1.launches - st0.launches);
    out.stats.gasBytes = size_t(s1.sceneBytes);
    if (O.verbose)
      std::fprintf(stderr, "info: scene %.1f MB build %.0f ms, %zu launches, %.3g rays in %.0f ms (%.2f GRays/s), reduce %.0f ms\n", s1.sceneBytes / 1048576.0,
                   s1.buildMs, out.stats.launches, double(out.stats.rays), out.stats.traceMs,
                   out.stats.traceMs > 0 ? double(out.stats.rays) / out.stats.traceMs / 1e6 : 0.0, out.stats.reduceMs);
    return true;
  }
};

}  // namespace

bool run(const std::vector<SimulationMeshView>& meshes, const Options& opt, Result& out, std::string& err) {
  const char* stage = "preparing the run";
  bool ok = false;
  ViewsRun r(meshes, opt, out, err, stage);
  try {
    ok = r.go();
  } catch (const std::bad_alloc&) {
    out = Result{};
    err = std::string("out of memory: host memory ran out while ") + stage;
  } catch (const std::exception& e) {
    out = Result{};
    err = std::string("internal error while ") + stage + ": " + e.what();
  }
  if (!ok && r.dumping) {   // only the files this run opened
    std::remove(opt.dumpPath.c_str());
    std::remove((opt.dumpPath + ".depth.npy").c_str());
  }
  return ok;
}

bool run(const std::vector<SimulationMesh>& meshes, const Options& opt, Result& out, std::string& err) { return run(views_of(meshes), opt, out, err); }

}  // namespace daylight::views
