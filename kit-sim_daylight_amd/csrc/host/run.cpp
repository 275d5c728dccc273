// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// daylight::run (reference src/run.cpp:216-758): validation, clocks, steps, the sky of every step on the GPU (dlg_sky),
// the direction table and per-quantity weights, the resident scene (LBVH in HBM, keyed), the sensors, and the chunk
// loop: trace -> every quantity reduced in one pass -> (series) the contraction on MFMA, streamed to .npy.
// Host work here is set-up of size O(steps + patches + directions) and the file I/O; everything per ray, per sensor
// or per (step, patch) runs on the GPU.
// Transmitting slots (--transmission on, tau > 0) trace in byte mode: a transmittance byte per direction instead of a bit,
// which the reduces and the contraction weight by. Reflected light (--bounces) adds the bounce pass's rows (dlg_bounce)
// weighted by the bins' radiance to the cumulative columns and the series.
// Not in this build: the Ladybug emulation mode. A run that asks for it fails with a message saying so.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>

#include "core.h"
#include "dl_math.h"

namespace daylight {
namespace {

using File = std::unique_ptr<std::FILE, decltype(&std::fclose)>;
using Clk = std::chrono::steady_clock;
double ms(Clk::time_point t0) { return std::chrono::duration<double, std::milli>(Clk::now() - t0).count(); }
constexpr double kPi = 3.14159265358979323846;
constexpr double kRf = 0.265074126, kGf = 0.670114631, kBf = 0.064811243;
constexpr double kNight = 1e-4;
constexpr size_t kVramMargin = size_t(128) << 20;

// The process's GPU and its resident scenes: `cur` installed, `parked` kept in HBM under their keys, most recent first.
struct Resident {
  dlg_gpu* gpu = nullptr;
  Scene cur;
  std::vector<Scene> parked;
};
Resident& resident() { static Resident r; return r; }
std::string gib(size_t b) { char s[32]; std::snprintf(s, sizeof s, "%.2f GB", b / 1073741824.0); return s; }

// The contraction's arithmetic for --timeseries: DAYLIGHT_SERIES = fp32 | bf16x3 | bf16x6 (default bf16x6, FP32-level).
int series_method() {
  const char* e = std::getenv("DAYLIGHT_SERIES");
  if (!e) return 2;
  if (!std::strcmp(e, "fp32")) return 0;
  if (!std::strcmp(e, "bf16x3")) return 1;
  return 2;
}

}  // namespace

bool npy_write_header(std::FILE* f, size_t rows, size_t cols, const char* descr) {
  char dict[256];
  const int n = std::snprintf(dict, sizeof dict, "{'descr': '%s', 'fortran_order': False, 'shape': (%zu, %zu), }", descr, rows, cols);
  if (n <= 0) return false;
  const size_t total = 10 + size_t(n) + 1, pad = (64 - total % 64) % 64;
  const uint16_t hlen = uint16_t(size_t(n) + pad + 1);
  if (std::fwrite("\x93NUMPY\x01\x00", 1, 8, f) != 8 || std::fwrite(&hlen, 2, 1, f) != 1) return false;
  if (std::fwrite(dict, 1, size_t(n), f) != size_t(n)) return false;
  for (size_t i = 0; i < pad; ++i) std::fputc(' ', f);
  std::fputc('\n', f);
  return true;
}

dlg_gpu* process_gpu(std::string& err) {
  Resident& R = resident();
  if (!R.gpu) {
    char e[512] = {};
    if (dlg_init(&R.gpu, e, sizeof e)) { err = e; R.gpu = nullptr; }
  }
  return R.gpu;
}

Scene& installed_scene() { return resident().cur; }

bool scene_acquire(const std::vector<SimulationMeshView>& meshes, const std::function<bool(size_t)>& blocks, bool slotted, const std::string& key,
                   int residentScenes, const std::atomic<bool>* cancel, dlg_gpu** gpu, Scene& s, Result& out, const char*& stage,
                   std::string& err, const std::function<void(const char*)>& mark) {
  const auto at_mark = [&](const char* what) { if (mark) mark(what); };
  Resident& R = resident();
  const bool installed = !key.empty() && R.gpu && key == R.cur.key;
  const auto parkedAt = [&] { return std::find_if(R.parked.begin(), R.parked.end(), [&](const Scene& p) { return p.key == key; }); };
  const bool reuse = installed || (!key.empty() && R.gpu && parkedAt() != R.parked.end());
  out.sceneReused = reuse;
  s = Scene{};
  if (reuse) { s = installed ? R.cur : *parkedAt(); s.parked = nullptr; }
  if (!gpu) return true;   // a run without rays: nothing to install
  if (cancel && cancel->load()) { err = "cancelled"; return false; }
  dlg_gpu* g = process_gpu(err);
  if (!g) return false;
  *gpu = g;
  if (installed) return true;
  if (!R.cur.key.empty() && residentScenes > 1) {   // park the installed keyed scene
    R.parked.insert(R.parked.begin(), R.cur);
    R.parked.front().parked = dlg_detach(g);
  } else dlg_free_scene(dlg_detach(g));
  R.cur = Scene{};
  if (auto it = parkedAt(); reuse && it != R.parked.end()) {
    Scene p = *it;
    R.parked.erase(it);
    dlg_attach(g, p.parked);
    p.parked = nullptr;
    R.cur = p;
    return true;
  }
  stage = "triangulating the scene";
  const auto tPrep = Clk::now();
  bbox_centre(meshes, s.c);
  std::vector<size_t> vBase(meshes.size() + 1, 0), tBase(meshes.size() + 1, 0), sBase(meshes.size() + 1, 0);
  for (size_t k = 0; k < meshes.size(); ++k) {
    vBase[k + 1] = vBase[k] + (blocks(k) ? meshes[k].mesh->numPoints() : 0);
    tBase[k + 1] = tBase[k] + (blocks(k) ? tri_cap(*meshes[k].mesh) : 0);
    sBase[k + 1] = sBase[k] + slot_count(*meshes[k].mesh);   // global slots: every mesh's, in input order
  }
  const size_t nPts = vBase.back(), nTri = tBase.back();
  at_mark("centre");
  stage = "making room on the GPU";
  dlg_trim(g, 1);
  // the peak: the staged triangles beside one mesh's ingest temporaries, or beside the tree's build
  int64_t work = 0;
  for (size_t k = 0; k < meshes.size(); ++k)
    if (blocks(k)) {
      const SimulationMesh& m = *meshes[k].mesh;
      work = std::max(work, dlg_mesh_bytes(int64_t(m.numPoints()), int64_t(m.counts.size()), int64_t(m.indices.size())));
    }
  const size_t staged = nPts * 12 + nTri * (slotted ? 16 : 12);
  const size_t need = std::max(size_t(dlg_build_bytes(int64_t(nPts), int64_t(nTri))) + (slotted ? nTri * 4 : 0), staged + size_t(work)) + kVramMargin;
  while (!R.parked.empty() && size_t(dlg_free_bytes()) < need) { dlg_free_scene(R.parked.back().parked); R.parked.pop_back(); }
  if (const size_t free = size_t(dlg_free_bytes()); free < need) {
    err = "out of memory: building the scene (" + std::to_string(nTri) + " triangles) needs " + gib(need) + " of GPU memory, " + gib(free) +
          " free (close other programs that use the GPU, or run fewer meshes)";
    return false;
  }
  // the faces of 5 or more corners ear-clipped on the host threads; the rest fanned on the GPU, mesh after mesh
  stage = "triangulating the scene";
  std::vector<std::vector<int64_t>> ngFace(meshes.size());
  std::vector<std::vector<int32_t>> ngMade(meshes.size());
  std::vector<std::vector<uint32_t>> ngTri(meshes.size());
  std::vector<int> skip(meshes.size(), 0);
  par_for(meshes.size(), [&](size_t k) { if (blocks(k)) ngon_faces(*meshes[k].mesh, ngFace[k], ngMade[k], ngTri[k], skip[k]); });
  at_mark("ear-clip");
  char e[512] = {};
  if (dlg_mesh_begin(g, int64_t(nPts), int64_t(nTri), slotted ? 1 : 0, s.c, e, sizeof e)) { err = e; return false; }
  size_t at = 0;
  for (size_t k = 0; k < meshes.size(); ++k) {
    if (!blocks(k)) continue;
    const SimulationMesh& m = *meshes[k].mesh;
    int64_t made = 0;
    if (dlg_mesh_add(g, m.points.data(), int64_t(m.numPoints()), m.counts.data(), int64_t(m.counts.size()), m.indices.data(), int64_t(m.indices.size()),
                     m.slot.empty() ? nullptr : m.slot.data(), uint32_t(sBase[k]), ngFace[k].data(), ngMade[k].data(), ngTri[k].data(),
                     int64_t(ngFace[k].size()), &made, e, sizeof e)) {
      err = e;
      return false;
    }
    s.order.push_back({meshes[k].name, at, size_t(made), meshes[k].mesh});
    at += size_t(made);
    s.skipped += skip[k];
  }
  out.stats.prepMs += ms(tPrep);
  at_mark("ingest");
  stage = "building the scene";
  const auto tBuild = Clk::now();
  if (dlg_mesh_build(g, e, sizeof e)) {
    err = e;
    return false;
  }
  out.stats.buildMs += ms(tBuild);
  at_mark("build");
  s.nTris = at;
  s.key = key;
  R.cur = s;
  return true;
}

bool run_admission(dlg_gpu* g, size_t nSens, size_t perSensor, size_t chunkSensor, size_t fixedBytes, int residentScenes, size_t& chunkBytes,
                   std::string& err) {
  Resident& R = resident();
  const auto need = [&] {
    const size_t nc = std::min(nSens, std::max<size_t>(1, chunkBytes / perSensor));
    return nSens * 2 * 12 + nc * chunkSensor + fixedBytes + kVramMargin;
  };
  // the last run's buffers are kept for this one (they grow in place) unless the room they hold is needed
  if (size_t(dlg_free_bytes()) < need()) dlg_trim(g, 1);
  while (!R.parked.empty() && (R.parked.size() + 1 > size_t(std::max(1, residentScenes)) || size_t(dlg_free_bytes()) < need())) {
    dlg_free_scene(R.parked.back().parked);
    R.parked.pop_back();
  }
  while (size_t(dlg_free_bytes()) < need() && chunkBytes > (size_t(32) << 20)) chunkBytes /= 2;
  if (const size_t free = size_t(dlg_free_bytes()); free < need()) {
    err = "out of memory: the run's " + std::to_string(nSens) + " sensors need " + gib(need()) + " of GPU memory beside the scene, " + gib(free) +
          " free (close other programs that use the GPU, or use fewer receivers)";
    return false;
  }
  return true;
}

std::string open_key(const std::vector<char>& open, bool selfOcclude) {
  std::string s;
  for (size_t k = 0; selfOcclude && k < open.size(); ++k) if (open[k]) s += (s.empty() ? "\nopen " : ",") + std::to_string(k);
  return s;
}

void scene_forget(const std::string& key) {
  Resident& R = resident();
  if (R.cur.key == key) R.cur.key.clear();
  for (auto it = R.parked.begin(); it != R.parked.end();)
    if (it->key == key) { dlg_free_scene(it->parked); it = R.parked.erase(it); } else ++it;
}

namespace {
struct SensorCache {
  std::string key;
  Sensors s;
  std::vector<std::pair<std::string, size_t>> ranges;
  std::string warn;
};
SensorCache& sensor_cache() {
  static SensorCache c;
  return c;
}
}  // namespace

const Sensors* sensors_cached(const std::vector<SimulationMeshView>& meshes, const RunOptions& opt, const double c[3], size_t nSens,
                              Sensors& scratch, Result& out, std::string& warn, std::string& err, bool ladybug) {
  std::string key;
  if (!opt.sceneKey.empty()) {   // what shapes the sensors besides the meshes (their roles are part of the scene key)
    char b[96];
    std::snprintf(b, sizeof b, "\nsensors v%d s%d o%a n%zu m%d", opt.onVertices ? 1 : 0, int(opt.sides), opt.offset, nSens, ladybug ? 1 : 0);
    key = opt.sceneKey + b;
  }
  SensorCache& C = sensor_cache();
  if (!key.empty() && C.key == key) {
    out.ranges = C.ranges;
    warn = C.warn;
    return &C.s;
  }
  C.key.clear();
  if (!sensors_of(meshes, opt, c, nSens, scratch, out, warn, err, ladybug)) return nullptr;
  if (key.empty()) return &scratch;
  C.s = std::move(scratch);
  C.ranges = out.ranges;
  C.warn = warn;
  C.key = key;
  return &C.s;
}

bool sensors_of(const std::vector<SimulationMeshView>& meshes, const RunOptions& opt, const double c[3], size_t nSens, Sensors& s, Result& out,
                std::string& warn, std::string& err, bool ladybug) {
  const size_t nS = opt.sides == Sides::Both ? 2 : 1;
  s.p.resize(nSens);
  s.n.resize(nSens);
  std::vector<size_t> sAt(meshes.size() + 1, 0), elems(meshes.size(), 0);
  for (size_t k = 0; k < meshes.size(); ++k)
    sAt[k + 1] = sAt[k] + (meshes[k].sensor ? (opt.onVertices ? meshes[k].mesh->numPoints() : meshes[k].mesh->numFaces()) * nS : 0);
  std::vector<std::string> warns(meshes.size());
  par_for(meshes.size(), [&](size_t k) {
    if (meshes[k].sensor) elems[k] = make_sensors_into(*meshes[k].mesh, opt, c, s.p.data() + sAt[k], s.n.data() + sAt[k], warns[k], ladybug);
  });
  out.ranges.clear();
  for (size_t k = 0; k < meshes.size(); ++k) {
    if (!meshes[k].sensor) continue;
    if (warns[k].rfind("error:", 0) == 0) { err = warns[k]; return false; }   // the ladybug mode's n-gon error
    if (!warns[k].empty()) warn += (warn.empty() ? "" : "; ") + warns[k];
    out.ranges.emplace_back(meshes[k].name, elems[k]);
  }
  if (sAt.back() != nSens) { err = "internal: " + std::to_string(sAt.back()) + " sensors made for " + std::to_string(nSens); return false; }
  return true;
}

static bool run_daylight(const std::vector<SimulationMeshView>& meshes, const Weather* w, const Options& opt, Result& out, std::string& err,
                         const char*& stage, const SeriesSink* sink) {
  const auto cancelled = [&] { if (opt.cancel && opt.cancel->load()) { err = "cancelled"; return true; } return false; };
  out.stats = RunStats{};
  // -v: where the run's host time goes (stage boundaries, printed at the end)
  std::vector<std::pair<const char*, Clk::time_point>> marks{{"start", Clk::now()}};
  const auto mark = [&](const char* what) { if (opt.verbose) marks.emplace_back(what, Clk::now()); };
  // ---- validation (run.cpp:219-242)
  if (!optics_ok(opt, err)) return false;
  bool needWeather = false, wantVis = false, wantRad = false;
  for (const Quantity q : opt.quantities)
    if (q != Quantity::SunHours) { needWeather = true; if (q == Quantity::Illuminance) wantVis = true; else wantRad = true; }
  if (needWeather && !w) { err = "weather (--epw) required for radiation/irradiance/illuminance"; return false; }
  const bool ladybug = opt.mode == Mode::Ladybug;
  if (ladybug && wantVis) { err = "illuminance is not available in ladybug mode (RadiationStudy is radiometric)"; return false; }
  if (ladybug && opt.mf != 1 && opt.mf != 2) { err = "ladybug mode supports --sky-density 1 or 2 only"; return false; }
  if (!opt.dumpSky.empty() && !(ladybug && needWeather)) { err = "--dump-sky is ladybug-mode only (with weather)"; return false; }
  if (opt.mf != 1 && opt.mf != 2 && opt.mf != 4) { err = "--sky-density must be 1, 2 or 4"; return false; }
  const int K = ladybug ? 1 : std::max(1, opt.patchSamples);
  {
    const int m = int(std::lround(std::sqrt(double(K))));
    if (m * m != K) { err = "--patch-samples must be a square number (1, 4, 9, 16, ...)"; return false; }
  }
  const bool series = !opt.seriesPath.empty();
  bool wantDc = series && needWeather;
  for (const Quantity q : opt.quantities) wantDc = wantDc || (q != Quantity::Direct && q != Quantity::SunHours);
  const bool reflect = opt.bounces > 0 && wantDc;   // reflected light: radiation, diffuse, irradiance, illuminance, the series
  double lat = w ? w->lat : opt.lat, lon = w ? w->lon : opt.lon, tz = w ? w->tz : opt.tz;
  if (!std::isnan(opt.lat)) lat = opt.lat;
  if (!std::isnan(opt.lon)) lon = opt.lon;
  if (w) tz = w->tz + (lon - w->lon) / 15;
  const double shift = w && !std::isnan(opt.tz) ? opt.tz - tz : 0;
  if (std::isnan(lat) || std::isnan(lon) || std::isnan(tz)) { err = "location unknown: give --epw or --lat --lon --tz"; return false; }
  const bool leap = w ? w->leap : false;
  if (opt.verbose && w && w->shiftH != 0) std::fprintf(stderr, "info: PVGIS weather: irradiance moved %.2f h later to meet its sun\n", w->shiftH);

  // ---- steps
  std::vector<Step> steps;
  if (!make_steps(opt, w, steps, err, shift)) return false;
  if (steps.empty()) { err = "no timesteps selected"; return false; }
  const size_t T = steps.size();
  mark("steps");
  double durationHours = 0;
  for (const Step& s : steps) durationHours += s.dt;
  out.durationHours = durationHours;

  // ---- geometry roles and optics (run.cpp:263-296). --transmission off: every slot opaque (bits); --fresnel off: glass
  // passes tau at every angle. Only a slot with faces in the scene transmits (not a sensor mesh's under
  // --no-self-occlude or --no-occlude, nor an id no face uses).
  std::vector<SlotOptics> slots;
  std::vector<char> open;
  if (!resolve_optics(meshes, opt.optics, slots, err) || !resolve_no_occlude(meshes, opt.noOcclude, open, err)) return false;
  const auto blocks = [&](size_t k) { return !meshes[k].sensor || (opt.selfOcclude && !open[k]); };
  std::vector<char> inGas(opt.transmission || reflect ? slots.size() : 0, 0);
  for (size_t k = 0, b = 0; !inGas.empty() && k < meshes.size(); b += slot_count(*meshes[k++].mesh)) {
    const SimulationMesh& m = *meshes[k].mesh;
    if (!blocks(k)) continue;
    if (m.slot.empty()) inGas[b] = !m.counts.empty();
    else for (const auto f : m.slot) inGas[b + size_t(f)] = 1;
  }
  std::string lit;          // the transmitting slots: their triangles' slots are kept with the scene, so part of its identity
  size_t first = SIZE_MAX;  // reflect: the first slot in the scene
  bool differs = false;     // ... and another's rho or glass differs: the bounce pass reads each triangle's slot
  for (size_t s = 0; s < slots.size(); ++s) {
    if (!opt.transmission || !inGas[s]) slots[s].tau = 0;
    if (!opt.fresnel) slots[s].flags &= ~DLG_SLOT_GLASS;
    if (slots[s].tau > 0) lit += (lit.empty() ? "" : ",") + std::to_string(s);
    if (!reflect || !inGas[s]) continue;
    if (first == SIZE_MAX) first = s;
    const SlotOptics &a = slots[s], &b = slots[first];
    differs = differs || a.rho != b.rho || a.flags != b.flags || ((a.flags & DLG_SLOT_GLASS) && a.tn != b.tn);
  }
  const bool translucent = !lit.empty(), slotted = translucent || differs;
  const std::string key = opt.sceneKey.empty() ? "" : opt.sceneKey + (translucent ? "\ntransmitting " + lit : slotted ? "\nslot records" : "") +
                                                   open_key(open, opt.selfOcclude);
  const int nS = opt.sides == Sides::Both ? 2 : 1;
  size_t nElem = 0;
  for (const SimulationMeshView& v : meshes) if (v.sensor) nElem += opt.onVertices ? v.mesh->numPoints() : v.mesh->numFaces();
  const size_t nSens = nElem * size_t(nS);
  if (nSens == 0) { err = "no sensor elements (no sensor meshes, or all faces degenerate)"; return false; }
  if (ladybug)   // RadiationStudy's rule, checked before the scene is built (run.cpp:304-307)
    for (const SimulationMeshView& v : meshes)
      for (size_t f = 0; v.sensor && f < v.mesh->counts.size(); ++f)
        if (v.mesh->counts[f] > 4) { err = "error: ladybug mode supports tri/quad faces only (mesh " + v.name + ", face " + std::to_string(f) + ")"; return false; }

  // ---- the sky of every step, on the GPU (sun, Perez, cumulative weights, series matrix)
  if (cancelled()) return false;
  const auto tSky = Clk::now();
  stage = "computing the sky";
  dlg_gpu* g = process_gpu(err);
  if (!g) return false;
  mark("gpu");
  const Patches pa = make_patches(opt.mf);
  const int P = pa.P;
  const size_t Pz = size_t(P);
  std::vector<int> month(T), day(T), jd(T);
  std::vector<double> hour(T), dtv(T), alt(T), azi(T), sunRadT(T), sunVisT(T), skyRad(Pz, 0.0), skyVis(Pz, 0.0), sunBinRad(Pz, 0.0),
      sunBinVis(Pz, 0.0);
  std::vector<float> dni(T), dhi(T), vec(3 * T);
  std::vector<int> up(T);
  for (size_t t = 0; t < T; ++t) {
    month[t] = steps[t].month; day[t] = steps[t].day; jd[t] = steps[t].jd; hour[t] = steps[t].hour; dtv[t] = steps[t].dt;
    dni[t] = steps[t].dni; dhi[t] = steps[t].dhi;
  }
  const bool seriesVis = series && opt.quantities[0] == Quantity::Illuminance;
  const double grefl = opt.virtualGround ? opt.ground : 0;
  double gndRad = 0, gndVis = 0;
  int Tup = 0;
  {
    dlg_sky_in in{int32_t(T), month.data(), day.data(), jd.data(), hour.data(), dtv.data(), dni.data(), dhi.data(), lat, lon, tz, opt.north, grefl,
                  leap ? 1 : 0, opt.mf, needWeather && wantRad && !ladybug ? 1 : 0, needWeather && wantVis ? 1 : 0,
                  series && needWeather && !ladybug ? 1 : 0, seriesVis ? 1 : 0, tcos_table(), reflect ? 1 : 0};
    dlg_sky_out so{alt.data(), azi.data(), up.data(), vec.data(), skyRad.data(), skyVis.data(), &gndRad, &gndVis, sunRadT.data(), sunVisT.data(), &Tup,
                   reflect ? sunBinRad.data() : nullptr, reflect ? sunBinVis.data() : nullptr};
    char e[512] = {};
    if (dlg_sky(g, &in, &so, e, sizeof e)) { err = e; return false; }
  }
  // direction table [sky P K | ground = -sky | sun up-steps] (run.cpp:316-340)
  std::vector<float3> D;
  std::vector<int> rowOf;
  size_t nSky = 0;
  if (needWeather) {
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < K; ++k) {
        D.push_back(dlm::rotate_z(ladybug ? pa.lb[size_t(p)] : patch_sample(pa, p, k, K), opt.north));
        rowOf.push_back(p);
      }
    nSky = D.size();
    for (size_t d = 0; d < nSky; ++d) D.push_back(make_float3(-D[d].x, -D[d].y, -D[d].z));
  }
  const size_t sun0 = D.size();
  std::vector<int> sunStep;
  for (size_t t = 0; t < T; ++t)
    if (up[t]) { D.push_back(make_float3(vec[3 * t], vec[3 * t + 1], vec[3 * t + 2])); sunStep.push_back(int(t)); }
  const size_t nSun = sunStep.size(), nDir = D.size();
  out.nSky = nSky;
  out.nSun = nSun;
  std::vector<int> colOf(T, -1);
  for (size_t t = 0, c = 0; t < T; ++t) if (steps[t].dni + steps[t].dhi > kNight) colOf[t] = int(c++);
  if (!series) Tup = 0;
  // ladybug: gendaymtx -A on the host (its series columns replace the device sky's)
  LadybugSky lb;
  std::vector<float> lbS, lbG;
  if (ladybug && needWeather) {
    if (series) { lbS.assign(Pz * size_t(Tup), 0.f); lbG.assign(size_t(std::max(Tup, 0)), 0.f); }
    ladybug_sky(pa, steps, lat, lon, w ? w->tz : tz, grefl, durationHours, colOf, series ? lbS.data() : nullptr, series ? lbG.data() : nullptr, Tup, lb);
  }
  // ---- per-quantity weights (run.cpp:475-505); the ground visibility weights for the series
  const size_t nQ = opt.quantities.size();
  std::vector<float> Wq(nQ * nDir + (series ? nDir : 0), 0.f);
  std::vector<double> sunRad(nSun, 0.0), sunVis(nSun, 0.0);
  double droppedDni = 0;
  for (size_t j = 0; j < nSun; ++j) { sunRad[j] = needWeather ? sunRadT[size_t(sunStep[j])] : 0.0; sunVis[j] = needWeather ? sunVisT[size_t(sunStep[j])] : 0.0; }
  for (size_t t = 0; t < T && needWeather && !ladybug; ++t) if (!up[t] && steps[t].dni > 0) droppedDni += double(steps[t].dni) * steps[t].dt;
  for (size_t qi = 0; qi < nQ; ++qi) {
    const Quantity q = opt.quantities[qi];
    float* W = Wq.data() + qi * nDir;
    if (q == Quantity::SunHours) {
      for (size_t j = 0; j < nSun; ++j) W[sun0 + j] = float(steps[size_t(sunStep[j])].dt);
      continue;
    }
    const bool vis = q == Quantity::Illuminance, wantSky = q != Quantity::Direct, wantSun = q != Quantity::Diffuse;
    if (ladybug) {   // the sun is inside the patches (gendaymtx -d): sky and ground only (run.cpp:489-502)
      for (size_t d = 0; d < nSky; ++d) {
        const size_t p = size_t(rowOf[d]);
        W[d] = float(q == Quantity::Direct ? lb.dir[p] : q == Quantity::Diffuse ? lb.dif[p] : lb.rad[p]);
      }
      if (q != Quantity::Direct) for (size_t d = nSky; d < 2 * nSky; ++d) W[d] = float(lb.gnd);
      continue;
    }
    for (size_t d = 0; d < nSky; ++d) {
      const size_t p = size_t(rowOf[d]);
      W[d] = float(wantSky ? (vis ? skyVis[p] : skyRad[p]) * pa.omega[p] / K : 0.0);
    }
    if (q != Quantity::Direct)
      for (size_t d = nSky; d < 2 * nSky; ++d) W[d] = float((vis ? gndVis : gndRad) * pa.omega[size_t(rowOf[d - nSky])] / K);
    if (wantSun) for (size_t j = 0; j < nSun; ++j) W[sun0 + j] = float(vis ? sunVis[j] : sunRad[j]);
  }
  if (series) {
    float* Wg = Wq.data() + nQ * nDir;
    for (size_t d = nSky; d < 2 * nSky; ++d) Wg[d] = ladybug ? 1.f : pa.omega[size_t(rowOf[d - nSky])] / float(K);
  }
  // reflected light (run.cpp:506-515): radiance per bounce bin (0 the ground, 1 + p the sky and the sun binned), no cosine
  // (the rows hold it); radiation / diffuse / irradiance radiometric, illuminance visible, direct and sun hours none
  const size_t nBins = Pz + 1;
  std::vector<float> Wdc(reflect ? nQ * nBins : 0, 0.f);
  std::vector<char> hasDc(nQ, 0);
  for (size_t qi = 0; reflect && qi < nQ; ++qi) {
    const Quantity q = opt.quantities[qi];
    if (q == Quantity::Direct || q == Quantity::SunHours) continue;
    const bool vis = q == Quantity::Illuminance;
    float* W = Wdc.data() + qi * nBins;
    hasDc[qi] = 1;
    W[0] = float(vis ? gndVis : gndRad);
    for (size_t p = 0; p < Pz; ++p) W[p + 1] = float(vis ? skyVis[p] + sunBinVis[p] : skyRad[p] + sunBinRad[p]);
  }
  out.stats.skyMs = ms(tSky);
  mark("sky, directions, weights");
  if (opt.verbose && droppedDni > 0)
    std::fprintf(stderr, "warn: %.1f Wh/m2 of DNI at steps with the sun below the horizon dropped (gendaymtx would bin it)\n", droppedDni);

  // ---- the scene, the sensors, the result
  const size_t nWords = (nDir + 31) / 32;
  // per sensor of a chunk: its visibility (bits, or a byte per direction in whole words) or its series row
  const size_t perSensor =
      std::max<size_t>(1, std::max({(translucent ? 32 : 4) * nWords, series ? 4 * size_t(Tup) : size_t(0), reflect ? 4 * nBins : size_t(0)}));
  size_t chunkBytes = opt.chunkBytes;
  File sf(nullptr, std::fclose), bf(nullptr, std::fclose);
  std::vector<double> seriesHoys;
  if (series && !sink) {
    sf.reset(std::fopen(opt.seriesPath.c_str(), "wb"));
    if (!sf || !npy_write_header(sf.get(), nSens, size_t(Tup), "<f4")) { err = "cannot write " + opt.seriesPath; return false; }
    for (size_t t = 0; t < T; ++t) if (colOf[t] >= 0) seriesHoys.push_back(steps[t].hoy);
  }
  if (!opt.dumpBitsPath.empty()) {
    bf.reset(std::fopen(opt.dumpBitsPath.c_str(), "wb"));
    if (!bf || !npy_write_header(bf.get(), nSens, translucent ? nDir : nWords, translucent ? "<u1" : "<u4")) { err = "cannot write " + opt.dumpBitsPath; return false; }
  }
  dlg_gpu* gp = nullptr;
  Scene sc;
  if (!scene_acquire(meshes, blocks, slotted, key, opt.residentScenes, opt.cancel, nDir > 0 ? &gp : nullptr, sc, out, stage, err, mark)) return false;
  mark("scene");
  dlg_stats_t st0{};
  std::vector<uint32_t>& thin = sc.thin;
  if (gp) {
    dlg_stats(gp, &st0);
    if (translucent && !out.sceneReused) {   // a fresh build's thin solids, from the same meshes and centre, kept with it
      stage = "finding thin solids";
      const auto tThin = Clk::now();
      Scene& cur = installed_scene();
      cur.key.clear();   // no run reuses the scene before they are known
      thin.assign(slots.size(), 0);
      std::vector<size_t> sBase(meshes.size() + 1, 0);
      for (size_t k = 0; k < meshes.size(); ++k) sBase[k + 1] = sBase[k] + slot_count(*meshes[k].mesh);
      par_for(meshes.size(), [&](size_t k) {
        if (!blocks(k)) return;
        const SimulationMesh& m = *meshes[k].mesh;
        for (size_t s = 0; s < slot_count(m); ++s)
          if (slots[sBase[k] + s].tau > 0) thin[sBase[k] + s] = thin_solid(m, uint32_t(s), sc.c);
      });
      cur.thin = thin;
      cur.key = key;
      out.stats.prepMs += ms(tThin);
    }
    const size_t buffers = 1 + (reflect ? 1 : 0) + (series ? 3 : 0);   // the visibility; the bounce rows; the series' output and operands
    if (!run_admission(gp, nSens, perSensor, buffers * perSensor + 8 * (nQ + 1), 0, opt.residentScenes, chunkBytes, err)) return false;
  }
  // A thin solid's exit face passes everything. With --cull-backfaces a closed solid is crossed through one front face
  // anyway (the other is culled): tau once without the rule.
  for (size_t s = 0; !opt.cullBackfaces && s < slots.size() && s < thin.size(); ++s) slots[s].flags |= thin[s];
  stage = "making the sensors";
  const auto tSens = Clk::now();
  Sensors scratch;
  std::string warn;
  const Sensors* sens = sensors_cached(meshes, opt, sc.c, nSens, scratch, out, warn, err, ladybug);
  if (!sens) return false;
  out.stats.prepMs += ms(tSens);
  mark("sensors");
  if (opt.verbose) {
    std::fprintf(stderr, "info: %zu sensors, %zu triangles%s, %zu steps (%.1f h)\n", nSens, sc.nTris, out.sceneReused ? " (scene reused)" : "", T, durationHours);
    if (sc.skipped) std::fprintf(stderr, "warn: %d degenerate faces skipped in occluders\n", sc.skipped);
    if (!warn.empty()) std::fprintf(stderr, "warn: %s\n", warn.c_str());
    if (translucent) {
      size_t nLit = 0, nThin = 0;
      for (const SlotOptics& s : slots) { nLit += s.tau > 0; nThin += (s.flags & DLG_SLOT_THIN) != 0; }
      std::fprintf(stderr, "info: %zu of %zu optics slots transmit (%zu thin solids)\n", nLit, slots.size(), nThin);
    }
  }
  if (gp) {
    char e[512] = {};
    if (dlg_set_scene(gp, reinterpret_cast<const float*>(sens->p.data()), reinterpret_cast<const float*>(sens->n.data()), int64_t(nSens),
                      reinterpret_cast<const float*>(D.data()), int64_t(nDir), e, sizeof e)) { err = e; return false; }
    // a scene without per-triangle slots reads slot 0: the one record of its slots' common optics (the reference's one SBT record)
    const std::vector<SlotOptics> one{first == SIZE_MAX ? SlotOptics{} : slots[first]};
    const std::vector<SlotOptics>& up = slotted ? slots : one;
    if ((translucent || reflect) && dlg_set_slots(gp, reinterpret_cast<const dlg_slot*>(up.data()), int64_t(up.size()), e, sizeof e)) { err = e; return false; }
  }
  scratch = Sensors{};
  mark("upload");

  stage = "allocating the result";
  const size_t C = nQ * size_t(nS);
  out.columns.clear();
  for (size_t qi = 0; qi < nQ; ++qi)
    for (int s = 0; s < nS; ++s)
      out.columns.push_back(std::string(kQuantityNames[size_t(opt.quantities[qi])]) +
                            ((opt.sides == Sides::Back || (opt.sides == Sides::Both && s == 1)) ? "_back" : "_front"));
  out.values.clear();
  out.f32.clear();
  if (opt.f32Columns) out.f32.assign(C * nElem, 0.f);
  else out.values.assign(nElem * C, 0.0);
  const auto store = [&](size_t si, size_t qi, double v) {
    const Quantity q = opt.quantities[qi];
    if (q == Quantity::Irradiance) v = durationHours > 0 ? v * 1000.0 / durationHours : 0.0;
    else if (q == Quantity::Illuminance) v = durationHours > 0 ? v / durationHours : 0.0;
    const size_t e = si / size_t(nS), col = qi * size_t(nS) + si % size_t(nS);
    if (opt.f32Columns) out.f32[col * nElem + e] = float(v);
    else out.values[e * C + col] = v;
  };

  if (gp) {
    stage = "tracing";
    const size_t ncMax = std::max<size_t>(1, chunkBytes / perSensor);
    const uint32_t flags = dlg_ray_flags(opt.cullBackfaces ? 1 : 0, translucent ? 1 : 0);
    std::vector<int32_t> useCos(nQ + 1, 1);
    for (size_t qi = 0; qi < nQ; ++qi) useCos[qi] = opt.quantities[qi] != Quantity::SunHours;
    const int nR = int(nQ) + (series ? 1 : 0);
    std::vector<double> tmp, tmp2;
    std::vector<float> G, E;
    const dlg_bounce_in bin{uint32_t(opt.bounces), uint32_t(opt.bounceSamples), opt.mf, opt.north};
    std::vector<uint32_t> bitsBuf, bitRows;
    std::vector<uint8_t> byteBuf, byteRows;
    std::vector<float> wsky(nSky), sunW(nSun, 0.f);
    std::vector<int32_t> sunCol(nSun, 0);
    for (size_t d = 0; d < nSky; ++d) wsky[d] = pa.omega[size_t(rowOf[d])] / float(K);
    for (size_t j = 0; j < nSun; ++j) {
      const int c = colOf[size_t(sunStep[j])];
      sunCol[j] = c < 0 ? 0 : c;
      const double dtj = steps[size_t(sunStep[j])].dt;
      sunW[j] = (ladybug || c < 0 || dtj <= 0) ? 0.f : float((seriesVis ? sunVis[j] : sunRad[j] * 1000.0) / dtj);   // ladybug: in the patches
    }
    if (ladybug && series && Tup > 0) {   // the host's gendaymtx -A columns as the contraction's sky
      char e[512] = {};
      if (dlg_set_sky_series(gp, lbS.data(), lbG.data(), P, Tup, e, sizeof e)) { err = e; return false; }
    }
    const int method = series_method();
    for (size_t i0 = 0; i0 < nSens; i0 += ncMax) {
      if (cancelled()) return false;
      const size_t nc = std::min(ncMax, nSens - i0);
      char e[512] = {};
      if (dlg_trace(gp, int64_t(i0), int64_t(nc), flags, e, sizeof e)) { err = e; return false; }
      // the reflected light next: it writes only its own rows, so the reduces see the trace's visibility
      if (reflect && dlg_bounce(gp, int64_t(i0), int64_t(nc), dlg_bounce_flags(opt.cullBackfaces ? 1 : 0), &bin, e, sizeof e)) { err = e; return false; }
      const auto tRed = Clk::now();
      tmp.assign(size_t(nR) * nc, 0.0);
      if (dlg_reduce(gp, int64_t(nc), nR, Wq.data(), useCos.data(), tmp.data(), e, sizeof e)) { err = e; return false; }
      for (size_t qi = 0; qi < nQ; ++qi) {
        if (hasDc[qi]) {
          tmp2.assign(nc, 0.0);
          if (dlg_reduce_dc(gp, int64_t(nc), Wdc.data() + qi * nBins, tmp2.data(), e, sizeof e)) { err = e; return false; }
          for (size_t i = 0; i < nc; ++i) store(i0 + i, qi, tmp[qi * nc + i] + tmp2[i]);   // the same double sum as the reference's
        } else
          for (size_t i = 0; i < nc; ++i) store(i0 + i, qi, tmp[qi * nc + i]);
      }
      out.stats.reduceMs += ms(tRed);
      if (opt.progress) opt.progress->store(double(i0 + nc) / double(nSens));
      if (bf && translucent) {   // direction-major bytes -> row-major [nc x nDir] rows
        byteBuf.resize(nDir * nc);
        if (dlg_download_bytes(gp, int64_t(nc), byteBuf.data(), e, sizeof e)) { err = e; return false; }
        byteRows.resize(nc * nDir);
        for (size_t d = 0; d < nDir; ++d)
          for (size_t i = 0; i < nc; ++i) byteRows[i * nDir + d] = byteBuf[d * nc + i];
        if (std::fwrite(byteRows.data(), 1, byteRows.size(), bf.get()) != byteRows.size()) { err = "write failed: " + opt.dumpBitsPath; return false; }
      } else if (bf) {
        bitsBuf.resize(nWords * nc);
        if (dlg_download_bits(gp, int64_t(nc), bitsBuf.data(), e, sizeof e)) { err = e; return false; }
        bitRows.resize(nc * nWords);
        for (size_t wi = 0; wi < nWords; ++wi)
          for (size_t i = 0; i < nc; ++i) bitRows[i * nWords + wi] = bitsBuf[wi * nc + i];
        if (std::fwrite(bitRows.data(), 4, bitRows.size(), bf.get()) != bitRows.size()) { err = "write failed: " + opt.dumpBitsPath; return false; }
      }
      if (series) {
        G.resize(nc);
        for (size_t i = 0; i < nc; ++i) G[i] = float(tmp[nQ * nc + i]);
        const bool compute = Tup > 0 && needWeather;
        dlg_series_in in{uint32_t(nSky), rowOf.data(), wsky.data(), P, Tup, G.data(), uint32_t(sun0), uint32_t(nSun), sunCol.data(), sunW.data(),
                         reflect ? 1 : 0};
        if (sink) {   // the series stays on the device: the sink reads each chunk there
          if (compute && dlg_series(gp, int64_t(nc), &in, method, nullptr, e, sizeof e)) { err = e; return false; }
          if (*sink && !(*sink)(i0, nc, size_t(Tup), compute ? dlg_series_device(gp) : nullptr, err)) return false;
          continue;
        }
        E.assign(nc * size_t(Tup), 0.f);
        if (compute && dlg_series(gp, int64_t(nc), &in, method, E.data(), e, sizeof e)) { err = e; return false; }
        if (std::fwrite(E.data(), sizeof(float), E.size(), sf.get()) != E.size()) { err = "write failed: " + opt.seriesPath; return false; }
      }
    }
    mark("trace, reduce, series");
    dlg_stats_t st1{};
    dlg_stats(gp, &st1);
    out.stats.traceMs = st1.traceMs - st0.traceMs;
    out.stats.bounceMs = st1.bounceMs - st0.bounceMs;
    out.stats.rays = size_t(st1.rays - st0.rays);
    out.stats.launches = size_t(st1.launches - st0.launches);
    out.stats.gasBytes = size_t(st1.sceneBytes);
    if (opt.verbose) {
      std::fprintf(stderr, "info: scene %.1f MB (LBVH height %lld) build %.0f ms, %zu launches, %.3g rays in %.0f ms (%.2f GRays/s)\n",
                   st1.sceneBytes / 1048576.0, (long long)st1.height, st1.buildMs, out.stats.launches, double(out.stats.rays), out.stats.traceMs,
                   out.stats.traceMs > 0 ? double(out.stats.rays) / out.stats.traceMs / 1e6 : 0.0);
      if (reflect)
        std::fprintf(stderr, "info: reflected light: %d bounces, %d samples per sensor, %.3g paths in %.0f ms (%lld launches)%s\n", opt.bounces,
                     opt.bounceSamples, double(st1.paths - st0.paths), out.stats.bounceMs, (long long)(st1.bounceLaunches - st0.bounceLaunches),
                     differs ? ", a slot per triangle" : "");
    }
  }
  if (sf) {
    File hf(std::fopen((opt.seriesPath + ".hoys.txt").c_str(), "w"), std::fclose);
    if (hf) for (const double h : seriesHoys) std::fprintf(hf.get(), "%.4f\n", h);
  }
  mark("end");
  if (opt.verbose) {
    std::string line = "info: host stages (ms):";
    for (size_t k = 1; k < marks.size(); ++k)
      line += std::string(" ") + marks[k].first + " " + std::to_string(std::chrono::duration<double, std::milli>(marks[k].second - marks[k - 1].second).count()).substr(0, 6) + ";";
    std::fprintf(stderr, "%s\n", line.c_str());
  }
  if (!opt.dumpSun.empty()) {   // columns 7-9: Radiance sun.c at the exact location, for the gendaymtx -D comparison (run.cpp:713-723)
    File f(std::fopen(opt.dumpSun.c_str(), "w"), std::fclose);
    if (!f) { err = "cannot write " + opt.dumpSun; return false; }
    std::fprintf(f.get(), "# hoy month day hour alt_deg azi_deg up alt_radiance_deg azi_radiance_deg up_radiance dni dhi   (north %.3f)\n", opt.north);
    for (size_t t = 0; t < T; ++t) {
      double altR, aziR;
      sun_radiance(lat, -lon, -15.0 * tz, steps[t].jd, steps[t].hour, altR, aziR);
      std::fprintf(f.get(), "%.4f %d %d %.4f %.6f %.6f %d %.6f %.6f %d %.4f %.4f\n", steps[t].hoy, steps[t].month, steps[t].day, steps[t].hour, alt[t],
                   azi[t], up[t], altR / (kPi / 180.), aziR / (kPi / 180.), altR > -0.2665 * (kPi / 180.) ? 1 : 0, steps[t].dni, steps[t].dhi);
    }
  }
  if (!opt.dumpSky.empty()) {   // the two gendaymtx -A ASCII blocks, -s then -d, ground row first (run.cpp:725-736)
    File f(std::fopen(opt.dumpSky.c_str(), "w"), std::fclose);
    if (!f) { err = "cannot write " + opt.dumpSky; return false; }
    const char* head[2] = {"# gendaymtx -s -A ascii rows (ground row first), %.3g\n", "# gendaymtx -d -A ascii rows (ground row first), %.3g\n"};
    for (size_t b = 0; b < 2; ++b) {
      std::fputs(head[b], f.get());
      for (size_t r = 0; r <= Pz; ++r) {
        const float* v = &lb.dump[b * 3 * (Pz + 1) + 3 * r];
        std::fprintf(f.get(), "%.3g %.3g %.3g\n", v[0], v[1], v[2]);
      }
    }
  }
  (void)kPi; (void)kRf; (void)kGf; (void)kBf;
  return true;
}

bool run_series_device(const std::vector<SimulationMeshView>& meshes, const Weather* w, const Options& opt, const SeriesSink& sink, Result& out,
                       std::string& err) {
  const char* stage = "preparing the run";
  try {
    return run_daylight(meshes, w, opt, out, err, stage, &sink);
  } catch (const std::bad_alloc&) {
    out = Result{};
    err = std::string("out of memory: host memory ran out while ") + stage;
  } catch (const std::exception& e) {
    out = Result{};
    err = std::string("internal error while ") + stage + ": " + e.what();
  }
  return false;
}

bool run(const std::vector<SimulationMeshView>& meshes, const Weather* w, const Options& opt, Result& out, std::string& err) {
  const char* stage = "preparing the run";
  try {
    return run_daylight(meshes, w, opt, out, err, stage, nullptr);
  } catch (const std::bad_alloc&) {
    out = Result{};
    err = std::string("out of memory: host memory ran out while ") + stage;
  } catch (const std::exception& e) {
    out = Result{};
    err = std::string("internal error while ") + stage + ": " + e.what();
  }
  return false;
}

bool run(const std::vector<SimulationMesh>& meshes, const Weather* w, const Options& opt, Result& out, std::string& err) {
  return run(views_of(meshes), w, opt, out, err);
}

}  // namespace daylight
