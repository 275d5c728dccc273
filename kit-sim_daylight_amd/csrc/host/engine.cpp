// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The engine: lanes (device contexts) and the store of resident scenes (engine.h). Replaces the reference's process-wide
// Gpu and its parked GAS slots (reference src/run.cpp:30-43, 68-190) with an object that more than one lane can run
// against at a time.
#include "engine.h"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>

namespace daylight {
namespace {

constexpr size_t kVramMargin = size_t(128) << 20;   // left free beside every admission
constexpr size_t kMinChunk = size_t(32) << 20;
using Clk = std::chrono::steady_clock;
double ms_since(Clk::time_point t0) { return std::chrono::duration<double, std::milli>(Clk::now() - t0).count(); }
std::string gb(size_t b) { char s[32]; std::snprintf(s, sizeof s, "%.2f GB", b / 1073741824.0); return s; }
size_t free_bytes() { return size_t(dlg_free_bytes()); }

}  // namespace

Engine& Engine::get() {
  static Engine e;
  return e;
}

namespace {
thread_local Lane* boundLane = nullptr;
}

void Engine::bind(Lane* l) { boundLane = l; }

Lane* Engine::run_lane(std::string& err) { return boundLane ? boundLane : primary(err); }

Lane* Engine::primary(std::string& err) {
  if (!primary_.gpu) {
    char e[512] = {};
    if (dlg_init(&primary_.gpu, e, sizeof e)) { err = e; primary_.gpu = nullptr; return nullptr; }
  }
  return &primary_;
}

Lane* Engine::new_lane(std::string& err) {
  {
    std::lock_guard<std::mutex> pk(primaryMu_);
    if (!primary(err)) return nullptr;
  }
  auto* l = new Lane;
  char e[512] = {};
  if (dlg_init(&l->gpu, e, sizeof e)) { err = e; delete l; return nullptr; }
  std::lock_guard<std::mutex> lk(lanesMu_);
  lanes_.push_back(l);
  return l;
}

void Engine::release_lane(Lane* l) {
  if (!l) return;
  {
    std::lock_guard<std::mutex> lk(lanesMu_);
    lanes_.erase(std::remove(lanes_.begin(), lanes_.end(), l), lanes_.end());
  }
  dlg_destroy(l->gpu);   // a borrowed scene is left alone
  delete l;
}

bool Engine::room(size_t need) const { return free_bytes() >= need; }

bool Engine::pinned(uint64_t id) const {
  for (const auto& [k, n] : pins_)
    if (k == id) return n > 0;
  return false;
}

void Engine::unpin(uint64_t id) {
  std::lock_guard<std::mutex> sk(storeMu_);
  for (auto it = pins_.begin(); it != pins_.end(); ++it) {
    if (it->first != id) continue;
    if (--it->second > 0) return;
    pins_.erase(it);
    break;
  }
  for (auto it = forgotten_.begin(); it != forgotten_.end(); ++it)
    if (it->first == id) { dlg_free_scene(it->second); forgotten_.erase(it); return; }
}

void SceneLease::release() {
  if (id_) Engine::get().unpin(id_);
  id_ = 0;
}

bool Engine::evict_one() {
  for (auto it = parked_.end(); it != parked_.begin();) {
    --it;
    if (pinned(it->id)) continue;
    dlg_free_scene(it->parked);
    parked_.erase(it);
    return true;
  }
  return false;
}

// The installed scene leaves the primary lane: parked under its key when more scenes may stay resident or a lease pins
// it, else freed.
void Engine::park_or_free_installed(int residentScenes) {
  dlg_scene* s = dlg_detach(primary_.gpu);
  const bool keep = (!installed_.key.empty() && residentScenes > 1) || pinned(installed_.id);
  if (keep) {
    installed_.parked = s;
    parked_.push_front(installed_);
  } else dlg_free_scene(s);
  installed_ = Scene{};
}

bool Engine::acquire(Lane* lane, const std::vector<SimulationMeshView>& meshes, const std::function<bool(size_t)>& blocks, bool slotted,
                     const std::string& key, int residentScenes, const std::atomic<bool>* cancel, dlg_gpu** gpu, Scene& s, Result& out,
                     const char*& stage, std::string& err, const std::function<void(const char*)>& mark, const std::function<void(Scene&)>& fresh,
                     SceneLease* lease) {
  const bool other = lane && lane != &primary_;
  if (!gpu || !other) {   // the primary lane's runs hold its lock already
    std::lock_guard<std::mutex> sk(storeMu_);
    if (!acquire_locked(meshes, blocks, slotted, key, residentScenes, cancel, gpu != nullptr, s, out, stage, err, mark, fresh)) return false;
    if (gpu) *gpu = primary_.gpu;
    return true;
  }
  std::lock_guard<std::mutex> pk(primaryMu_);
  std::lock_guard<std::mutex> sk(storeMu_);
  if (!primary(err)) return false;
  if (!acquire_locked(meshes, blocks, slotted, key, residentScenes, cancel, true, s, out, stage, err, mark, fresh)) return false;
  dlg_borrow_scene(lane->gpu, primary_.gpu);
  *gpu = lane->gpu;
  if (lease) {
    lease->release();
    bool found = false;
    for (auto& [k, n] : pins_)
      if (k == installed_.id) { ++n; found = true; }
    if (!found) pins_.push_back({installed_.id, 1});
    lease->id_ = installed_.id;
  }
  return true;
}

bool Engine::acquire_locked(const std::vector<SimulationMeshView>& meshes, const std::function<bool(size_t)>& blocks, bool slotted,
                            const std::string& key, int residentScenes, const std::atomic<bool>* cancel, bool install, Scene& s, Result& out,
                            const char*& stage, std::string& err, const std::function<void(const char*)>& mark, const std::function<void(Scene&)>& fresh) {
  const auto note = [&](const char* what) { if (mark) mark(what); };
  // A resident scene under the key holds the blocking meshes; it serves this run when the run's centre (its box: the
  // scene's, grown by the meshes that do not block) is the scene's, bit for bit, so the recentred coordinates are the
  // ones a fresh build would make. Sensors that do not block can then change from run to run without a rebuild.
  const bool keyed = !key.empty() && primary_.gpu;
  double other[6];
  bool otherDone = false;
  const auto serves = [&](const Scene& sc) {
    if (!keyed || sc.key != key) return false;
    if (!otherDone) { bbox_of(meshes, [&](size_t k) { return !blocks(k); }, other); otherDone = true; }
    double box[6], c[3];
    std::copy(sc.box, sc.box + 6, box);
    box_union(box, other);
    box_centre(box, c);
    return c[0] == sc.c[0] && c[1] == sc.c[1] && c[2] == sc.c[2];
  };
  const bool isInstalled = serves(installed_);
  auto parkedAt = isInstalled ? parked_.end() : std::find_if(parked_.begin(), parked_.end(), serves);
  out.sceneReused = isInstalled || parkedAt != parked_.end();
  s = Scene{};
  if (out.sceneReused) { s = isInstalled ? installed_ : *parkedAt; s.parked = nullptr; }
  if (!install) return true;   // no rays: nothing to install
  if (cancel && cancel->load()) { err = "cancelled"; return false; }
  Lane* lane = primary(err);
  if (!lane) return false;
  if (isInstalled) return true;
  park_or_free_installed(residentScenes);
  if (parkedAt != parked_.end()) {   // back from the parked ones
    installed_ = *parkedAt;
    parked_.erase(parkedAt);
    dlg_attach(lane->gpu, installed_.parked);
    installed_.parked = nullptr;
    return true;
  }

  // a fresh build: the blocking meshes' points recentred and their faces fanned on the GPU (n-gons ear-clipped here)
  stage = "triangulating the scene";
  const auto tPrep = Clk::now();
  {
    double all[6];
    bbox_of(meshes, blocks, s.box);
    if (!otherDone) bbox_of(meshes, [&](size_t k) { return !blocks(k); }, other);
    std::copy(s.box, s.box + 6, all);
    box_union(all, other);
    box_centre(all, s.c);
  }
  const size_t n = meshes.size();
  std::vector<size_t> slotBase(n + 1, 0);
  size_t nPts = 0, nTri = 0;
  int64_t meshPeak = 0;   // the largest mesh's ingest temporaries
  for (size_t k = 0; k < n; ++k) {
    const SimulationMesh& m = *meshes[k].mesh;
    slotBase[k + 1] = slotBase[k] + slot_count(m);   // global slots number every mesh's, blocking or not
    if (!blocks(k)) continue;
    nPts += m.numPoints();
    nTri += tri_cap(m);
    meshPeak = std::max(meshPeak, dlg_mesh_bytes(int64_t(m.numPoints()), int64_t(m.counts.size()), int64_t(m.indices.size())));
  }
  note("centre");
  stage = "making room on the GPU";
  dlg_trim(lane->gpu, 1);
  lane->sensorKey.clear();
  const size_t staged = nPts * 12 + nTri * (slotted ? 16 : 12);
  const size_t build = size_t(dlg_build_bytes(int64_t(nPts), int64_t(nTri))) + (slotted ? nTri * 4 : 0);
  const size_t need = std::max(build, staged + size_t(meshPeak)) + kVramMargin;
  while (!room(need) && evict_one()) {   // the least recently used parked scenes go first
  }
  if (!room(need)) {
    err = "out of memory: building the scene (" + std::to_string(nTri) + " triangles) needs " + gb(need) + " of GPU memory, " + gb(free_bytes()) +
          " free (close other programs that use the GPU, or run fewer meshes)";
    return false;
  }
  stage = "triangulating the scene";
  struct Ngons { std::vector<int64_t> face; std::vector<int32_t> made; std::vector<uint32_t> tri; int skipped = 0; };
  std::vector<Ngons> ng(n);
  par_for(n, [&](size_t k) { if (blocks(k)) ngon_faces(*meshes[k].mesh, ng[k].face, ng[k].made, ng[k].tri, ng[k].skipped); });
  note("ear-clip");
  char e[512] = {};
  if (dlg_mesh_begin(lane->gpu, int64_t(nPts), int64_t(nTri), slotted ? 1 : 0, s.c, e, sizeof e)) { err = e; return false; }
  size_t tri0 = 0;
  for (size_t k = 0; k < n; ++k) {
    if (!blocks(k)) continue;
    const SimulationMesh& m = *meshes[k].mesh;
    int64_t made = 0;
    if (dlg_mesh_add(lane->gpu, m.points.data(), int64_t(m.numPoints()), m.counts.data(), int64_t(m.counts.size()), m.indices.data(),
                     int64_t(m.indices.size()), m.slot.empty() ? nullptr : m.slot.data(), uint32_t(slotBase[k]), ng[k].face.data(), ng[k].made.data(),
                     ng[k].tri.data(), int64_t(ng[k].face.size()), &made, e, sizeof e)) { err = e; return false; }
    s.order.push_back({meshes[k].name, tri0, size_t(made), meshes[k].mesh});
    tri0 += size_t(made);
    s.skipped += ng[k].skipped;
  }
  out.stats.prepMs += ms_since(tPrep);
  note("ingest");
  stage = "building the scene";
  const auto tBuild = Clk::now();
  if (dlg_mesh_build(lane->gpu, e, sizeof e)) { err = e; return false; }
  out.stats.buildMs += ms_since(tBuild);
  note("build");
  s.nTris = tri0;
  s.key = key;
  s.id = nextId_++;
  if (fresh) fresh(s);
  installed_ = s;
  return true;
}

bool Engine::admit(Lane& lane, size_t nSens, size_t perSensor, size_t chunkSensor, size_t fixedBytes, int residentScenes, size_t& chunkBytes,
                   std::string& err) {
  const auto need = [&] {
    const size_t perChunk = std::min(nSens, std::max<size_t>(1, chunkBytes / perSensor));
    return nSens * 24 + perChunk * chunkSensor + fixedBytes + kVramMargin;
  };
  if (!room(need())) {   // the lane's last buffers are kept for the next run unless the room is needed
    dlg_trim(lane.gpu, 1);
    lane.sensorKey.clear();
  }
  std::lock_guard<std::mutex> sk(storeMu_);
  const size_t keep = size_t(std::max(1, residentScenes)) - 1;
  size_t unpinned = 0;
  for (const Scene& p : parked_) unpinned += pinned(p.id) ? 0 : 1;
  while (unpinned > keep && evict_one()) --unpinned;
  while (!room(need()) && evict_one()) {
  }
  while (!room(need()) && chunkBytes > kMinChunk) chunkBytes /= 2;
  if (room(need())) return true;
  err = "out of memory: the run's " + std::to_string(nSens) + " sensors need " + gb(need()) + " of GPU memory beside the scene, " + gb(free_bytes()) +
        " free (close other programs that use the GPU, or use fewer receivers)";
  return false;
}

void Engine::forget(const std::string& key) {
  std::lock_guard<std::mutex> sk(storeMu_);
  if (installed_.key == key) installed_.key.clear();
  for (auto it = parked_.begin(); it != parked_.end();) {
    if (it->key != key) { ++it; continue; }
    if (pinned(it->id)) forgotten_.push_back({it->id, it->parked});
    else dlg_free_scene(it->parked);
    it = parked_.erase(it);
  }
}

bool lane_sensors(Lane& lane, const std::vector<SimulationMeshView>& meshes, const RunOptions& opt, const double c[3], size_t nSens, Result& out,
                  std::string& warn, std::string& err, bool ladybug) {
  std::string key;
  if (!opt.sceneKey.empty()) {   // the sensors depend on the meshes (named by the mesh key), these options and the centre
    char b[192];
    std::snprintf(b, sizeof b, "\nsensors v%d s%d o%a n%zu m%d c%a,%a,%a", opt.onVertices ? 1 : 0, int(opt.sides), opt.offset, nSens, ladybug ? 1 : 0,
                  c[0], c[1], c[2]);
    key = opt.sceneKey + b;
  }
  if (!key.empty() && key == lane.sensorKey) {
    out.ranges = lane.ranges;
    warn = lane.sensorWarn;
    return true;
  }
  lane.sensorKey.clear();
  const size_t per = opt.sides == Sides::Both ? 2 : 1;
  size_t total = 0;
  for (const SimulationMeshView& v : meshes)
    if (v.sensor) total += (opt.onVertices ? v.mesh->numPoints() : v.mesh->numFaces()) * per;
  if (total != nSens) { err = "internal: " + std::to_string(total) + " sensors for " + std::to_string(nSens); return false; }
  for (const SimulationMeshView& v : meshes) {   // ladybug mode's rule (RadiationStudy: tri / quad faces only)
    if (!v.sensor || !ladybug) continue;
    for (size_t f = 0; f < v.mesh->counts.size(); ++f)
      if (v.mesh->counts[f] > 4) { err = "error: ladybug mode supports tri/quad faces only (mesh " + v.mesh->name + ", face " + std::to_string(f) + ")"; return false; }
  }
  char e[512] = {};
  if (dlg_sensors_reserve(lane.gpu, int64_t(nSens), e, sizeof e)) { err = e; return false; }
  out.ranges.clear();
  warn.clear();
  size_t j = 0;
  for (const SimulationMeshView& v : meshes) {
    if (!v.sensor) continue;
    const SimulationMesh& m = *v.mesh;
    if (!mesh_ok(m, m.name, err)) return false;   // the device clamps indices; a mesh that needs it is refused here
    const dlg_sensor_mesh dm{m.points.data(), int64_t(m.numPoints()), m.counts.data(), int64_t(m.counts.size()), m.indices.data(),
                             int64_t(m.indices.size())};
    int64_t made = 0;
    int32_t degenerate = 0, zero = 0;
    if (dlg_sensors(lane.gpu, &dm, int64_t(j), opt.onVertices ? 1 : 0, opt.sides == Sides::Front ? 0 : opt.sides == Sides::Back ? 1 : 2, opt.offset, c,
                    ladybug ? 1 : 0, &made, &degenerate, &zero, e, sizeof e)) { err = e; return false; }
    j += size_t(made) * per;
    out.ranges.emplace_back(v.name, size_t(made));
    std::string note;
    if (degenerate) note = "mesh " + m.name + ": " + std::to_string(degenerate) + " degenerate face(s)";
    if (zero) note += std::string(degenerate ? "; " : "mesh " + m.name + ": ") + std::to_string(zero) + " zero-length vertex normal(s) set to (0,0,1)";
    if (!note.empty()) warn += (warn.empty() ? "" : "; ") + note;
  }
  if (key.empty()) return true;
  lane.ranges = out.ranges;
  lane.sensorWarn = warn;
  lane.sensorKey = key;
  return true;
}

std::string scene_key_of(const std::string& meshKey, bool selfOcclude) {
  if (selfOcclude) return meshKey;
  std::string k;
  for (size_t i = 0; i < meshKey.size();) {   // drop every " s<digits>" token
    if (meshKey.compare(i, 2, " s") == 0 && i + 2 < meshKey.size() && std::isdigit(static_cast<unsigned char>(meshKey[i + 2]))) {
      i += 2;
      while (i < meshKey.size() && std::isdigit(static_cast<unsigned char>(meshKey[i]))) ++i;
      continue;
    }
    k += meshKey[i++];
  }
  return k;
}

std::string open_key(const std::vector<char>& open, bool selfOcclude) {
  // sensor meshes left out of the scene by --no-occlude change the scene: their indices are part of its key
  std::string s;
  if (!selfOcclude) return s;
  for (size_t k = 0; k < open.size(); ++k)
    if (open[k]) s += (s.empty() ? "\nopen " : ",") + std::to_string(k);
  return s;
}

}  // namespace daylight
