// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// daylight::run as a pipeline of stages (reference src/run.cpp:216-758 is the contract: the same checks and messages,
// the same direction table, weights, columns, files and statistics):
//   check    the options and the weather the quantities need
//   steps    the site's clock and the evaluated steps
//   roles    which meshes block, which slots transmit, the scene's key, the sensor count
//   sky      the sun and the Perez sky of every step on the GPU (FP64), the direction table, the per-quantity weights
//   scene    the resident scene (engine.h), the thin solids of a fresh slotted build, the room for the run
//   sensors  the lane's sensors, uploaded with the directions and the slot optics
//   chunks   per chunk of sensors: trace -> (bounce) -> every quantity reduced in one pass -> (series) the contraction
//   dumps    --dump-sun, --dump-sky, the series' hours
// Every per-ray, per-sensor or per-(step, patch) loop runs on the GPU; the host work is O(steps + patches + directions).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <thread>

#include "dl_math.h"
#include "engine.h"

namespace daylight {
namespace {

using Clk = std::chrono::steady_clock;
double ms_since(Clk::time_point t0) { return std::chrono::duration<double, std::milli>(Clk::now() - t0).count(); }
using File = std::unique_ptr<std::FILE, int (*)(std::FILE*)>;
File open_file(const std::string& path, const char* mode) { return File(std::fopen(path.c_str(), mode), &std::fclose); }
constexpr double kPi = 3.14159265358979323846;
constexpr double kNight = 1e-4;   // dni + dhi at or below: a night step (no series column)

// The contraction's arithmetic for --timeseries, DAYLIGHT_SERIES = fp32 | bf16x3 | bf16x6 | fp16x3 (default fp16x3:
// three MFMA products at FP32-level accuracy; an unknown name takes the default)
int series_method() {
  const char* e = std::getenv("DAYLIGHT_SERIES");
  if (!e) return 3;
  return !std::strcmp(e, "fp32") ? 0 : !std::strcmp(e, "bf16x3") ? 1 : !std::strcmp(e, "bf16x6") ? 2 : 3;
}

std::string side_suffix(Sides sides, int s) { return (sides == Sides::Back || (sides == Sides::Both && s == 1)) ? "_back" : "_front"; }

class DaylightRun {
 public:
  DaylightRun(const std::vector<SimulationMeshView>& meshes, const Weather* w, const Options& opt, Result& out, std::string& err, const char*& stage,
              const SeriesSink* sink)
      : M(meshes), W(w), O(opt), out(out), err(err), stage(stage), sink(sink) {}

  bool go() {
    out.stats = RunStats{};
    marks.emplace_back("start", Clk::now());
    if (!check() || !steps_() || !roles() || !sky() || !scene() || !sensors() || !chunks()) return false;
    if (O.verbose) host_line();
    return dumps();
  }

 private:
  // ---- inputs
  const std::vector<SimulationMeshView>& M;
  const Weather* W;
  const Options& O;
  Result& out;
  std::string& err;
  const char*& stage;
  const SeriesSink* sink;
  std::vector<std::pair<const char*, Clk::time_point>> marks;   // -v: where the host time goes

  // ---- check / steps
  bool needWeather = false, wantRad = false, wantVis = false, ladybug = false, series = false, reflect = false;
  int K = 1;
  double lat = NAN, lon = NAN, tz = NAN;
  bool leap = false;
  std::vector<Step> st;
  double hours = 0;
  // ---- roles
  std::vector<SlotOptics> slots;
  std::vector<char> open, inScene;
  bool translucent = false, slotted = false, differs = false;   // differs: scene slots that reflect differently
  size_t firstSlot = SIZE_MAX;   // reflected light: the first slot in the scene
  std::string key;
  int nS = 1;
  size_t nElem = 0, nSens = 0;
  // ---- sky and directions
  Patches pa;
  std::vector<double> alt, azi, sunRadT, sunVisT, skyRad, skyVis, sunBinRad, sunBinVis;
  std::vector<int> up, colOf, rowOf, sunStep;
  double gndRad = 0, gndVis = 0;
  int Tup = 0;
  std::vector<float3> D;
  size_t nSky = 0, sun0 = 0, nSun = 0, nDir = 0;
  LadybugSky lb;
  std::vector<float> lbS, lbG;
  std::vector<float> Wq, Wdc;      // [nQ (+1 ground for the series)] x nDir; reflected light [nQ] x (P + 1)
  std::vector<char> hasDc;
  std::vector<double> sunRad, sunVis;
  // ---- scene and run
  Lane* lane = nullptr;
  dlg_gpu* g = nullptr;
  Scene sc;
  SceneLease lease;   // a run on a server's lane: its scene stays resident until the run ends
  size_t chunkBytes = 0, perSensor = 1;
  dlg_stats_t st0{};

  void mark(const char* what) { if (O.verbose) marks.emplace_back(what, Clk::now()); }
  bool fail(std::string msg) { err = std::move(msg); return false; }
  bool cancelled() { return O.cancel && O.cancel->load() && fail("cancelled"); }
  bool blocks(size_t k) const { return !M[k].sensor || (O.selfOcclude && !open[k]); }
  size_t nQ() const { return O.quantities.size(); }

  // ------------------------------------------------------------------------------------------------ check (run.cpp:219-242)
  bool check() {
    if (!optics_ok(O, err)) return false;
    for (const Quantity q : O.quantities) {
      if (q == Quantity::SunHours) continue;
      needWeather = true;
      (q == Quantity::Illuminance ? wantVis : wantRad) = true;
    }
    if (needWeather && !W) return fail("weather (--epw) required for radiation/irradiance/illuminance");
    ladybug = O.mode == Mode::Ladybug;
    if (ladybug && wantVis) return fail("illuminance is not available in ladybug mode (RadiationStudy is radiometric)");
    if (ladybug && O.mf != 1 && O.mf != 2) return fail("ladybug mode supports --sky-density 1 or 2 only");
    if (!O.dumpSky.empty() && !(ladybug && needWeather)) return fail("--dump-sky is ladybug-mode only (with weather)");
    if (O.mf != 1 && O.mf != 2 && O.mf != 4) return fail("--sky-density must be 1, 2 or 4");
    K = ladybug ? 1 : std::max(1, O.patchSamples);
    const int root = int(std::lround(std::sqrt(double(K))));
    if (root * root != K) return fail("--patch-samples must be a square number (1, 4, 9, 16, ...)");
    series = !O.seriesPath.empty();
    // reflected light feeds radiation, diffuse, irradiance, illuminance and the series
    bool feedsDc = series && needWeather;
    for (const Quantity q : O.quantities) feedsDc = feedsDc || (q != Quantity::Direct && q != Quantity::SunHours);
    reflect = O.bounces > 0 && feedsDc;
    return true;
  }

  // ------------------------------------------------------------------------------------------------ steps
  bool steps_() {
    // the site: the weather's, with --lat / --lon overriding it; the weather's clock moved with the longitude
    if (W) { lat = W->lat; lon = W->lon; tz = W->tz; }
    else { lat = O.lat; lon = O.lon; tz = O.tz; }
    if (!std::isnan(O.lat)) lat = O.lat;
    if (!std::isnan(O.lon)) lon = O.lon;
    if (W) tz = W->tz + (lon - W->lon) / 15;
    const double shift = W && !std::isnan(O.tz) ? O.tz - tz : 0;
    if (std::isnan(lat) || std::isnan(lon) || std::isnan(tz)) return fail("location unknown: give --epw or --lat --lon --tz");
    leap = W && W->leap;
    if (O.verbose && W && W->shiftH != 0) std::fprintf(stderr, "info: PVGIS weather: irradiance moved %.2f h later to meet its sun\n", W->shiftH);
    if (!make_steps(O, W, st, err, shift)) return false;
    if (st.empty()) return fail("no timesteps selected");
    mark("steps");
    for (const Step& s : st) hours += s.dt;
    out.durationHours = hours;
    return true;
  }

  // ------------------------------------------------------------------------------------------------ roles (run.cpp:263-296)
  bool roles() {
    if (!resolve_optics(M, O.optics, slots, err) || !resolve_no_occlude(M, O.noOcclude, open, err)) return false;
    // the slots that have faces in the scene (only those can transmit or reflect)
    inScene.assign(O.transmission || reflect ? slots.size() : 0, 0);
    for (size_t k = 0, base = 0; !inScene.empty() && k < M.size(); base += slot_count(*M[k].mesh), ++k) {
      const SimulationMesh& m = *M[k].mesh;
      if (!blocks(k)) continue;
      if (m.slot.empty()) inScene[base] = !m.counts.empty();
      else for (const uint16_t f : m.slot) inScene[base + f] = 1;
    }
    std::string lit;        // the transmitting slots (part of the scene's identity: their triangles keep their slots)
    for (size_t s = 0; s < slots.size(); ++s) {
      SlotOptics& so = slots[s];
      const bool here = s < inScene.size() && inScene[s];
      if (!O.transmission || !here) so.tau = 0;
      if (!O.fresnel) so.flags &= ~DLG_SLOT_GLASS;
      if (so.tau > 0) lit += (lit.empty() ? "" : ",") + std::to_string(s);
      if (!reflect || !here) continue;
      if (firstSlot == SIZE_MAX) firstSlot = s;
      const SlotOptics& f = slots[firstSlot];
      differs = differs || so.rho != f.rho || so.flags != f.flags || ((so.flags & DLG_SLOT_GLASS) && so.tn != f.tn);
    }
    translucent = !lit.empty();
    slotted = translucent || differs;
    if (!O.sceneKey.empty())
      key = scene_key_of(O.sceneKey, O.selfOcclude) + (translucent ? "\ntransmitting " + lit : slotted ? "\nslot records" : "") + open_key(open, O.selfOcclude);
    nS = O.sides == Sides::Both ? 2 : 1;
    for (const SimulationMeshView& v : M)
      if (v.sensor) nElem += O.onVertices ? v.mesh->numPoints() : v.mesh->numFaces();
    nSens = nElem * size_t(nS);
    if (nSens == 0) return fail("no sensor elements (no sensor meshes, or all faces degenerate)");
    if (ladybug)   // RadiationStudy's rule, before any scene is built (run.cpp:304-307)
      for (const SimulationMeshView& v : M)
        for (size_t f = 0; v.sensor && f < v.mesh->counts.size(); ++f)
          if (v.mesh->counts[f] > 4) return fail("error: ladybug mode supports tri/quad faces only (mesh " + v.name + ", face " + std::to_string(f) + ")");
    return true;
  }

  // ------------------------------------------------------------------------------------------------ sky (run.cpp:311-515)
  bool sky() {
    if (cancelled()) return false;
    const auto t0 = Clk::now();
    stage = "computing the sky";
    lane = Engine::get().run_lane(err);
    if (!lane) return false;
    g = lane->gpu;
    mark("gpu");
    pa = make_patches(O.mf);
    const size_t T = st.size(), P = size_t(pa.P);
    std::vector<int> month(T), day(T), jd(T);
    std::vector<double> hour(T), dt(T);
    std::vector<float> dni(T), dhi(T), vec(3 * T);
    for (size_t t = 0; t < T; ++t) {
      const Step& s = st[t];
      month[t] = s.month; day[t] = s.day; jd[t] = s.jd; hour[t] = s.hour; dt[t] = s.dt; dni[t] = s.dni; dhi[t] = s.dhi;
    }
    alt.assign(T, 0); azi.assign(T, 0); sunRadT.assign(T, 0); sunVisT.assign(T, 0); up.assign(T, 0);
    skyRad.assign(P, 0); skyVis.assign(P, 0); sunBinRad.assign(P, 0); sunBinVis.assign(P, 0);
    const bool seriesVis = series && O.quantities[0] == Quantity::Illuminance;
    const double grefl = O.virtualGround ? O.ground : 0;
    dlg_sky_in in{int32_t(T), month.data(), day.data(), jd.data(), hour.data(), dt.data(), dni.data(), dhi.data(), lat, lon, tz, O.north, grefl,
                  leap ? 1 : 0, O.mf, needWeather && wantRad && !ladybug ? 1 : 0, needWeather && wantVis ? 1 : 0,
                  series && needWeather && !ladybug ? 1 : 0, seriesVis ? 1 : 0, tcos_table(), reflect ? 1 : 0};
    dlg_sky_out so{alt.data(), azi.data(), up.data(), vec.data(), skyRad.data(), skyVis.data(), &gndRad, &gndVis, sunRadT.data(), sunVisT.data(), &Tup,
                   reflect ? sunBinRad.data() : nullptr, reflect ? sunBinVis.data() : nullptr};
    char e[512] = {};
    if (dlg_sky(g, &in, &so, e, sizeof e)) return fail(e);
    directions(vec);
    colOf.assign(T, -1);
    for (size_t t = 0, c = 0; t < T; ++t)
      if (st[t].dni + st[t].dhi > kNight) colOf[t] = int(c++);
    if (!series) Tup = 0;
    if (ladybug && needWeather) {   // gendaymtx -A on the host: its series columns replace the device sky's
      if (series) { lbS.assign(P * size_t(Tup), 0.f); lbG.assign(size_t(std::max(Tup, 0)), 0.f); }
      ladybug_sky(pa, st, lat, lon, W ? W->tz : tz, grefl, hours, colOf, series ? lbS.data() : nullptr, series ? lbG.data() : nullptr, Tup, lb);
    }
    weights();
    out.stats.skyMs = ms_since(t0);
    mark("sky, directions, weights");
    return true;
  }

  // the direction table [sky P K | ground = -sky | the sun at each up step] (run.cpp:316-340)
  void directions(const std::vector<float>& vec) {
    D.clear();
    rowOf.clear();
    sunStep.clear();
    if (needWeather) {
      for (int p = 0; p < pa.P; ++p)
        for (int k = 0; k < K; ++k) {
          D.push_back(dlm::rotate_z(ladybug ? pa.lb[size_t(p)] : patch_sample(pa, p, k, K), O.north));
          rowOf.push_back(p);
        }
      nSky = D.size();
      for (size_t d = 0; d < nSky; ++d) D.push_back(make_float3(-D[d].x, -D[d].y, -D[d].z));
    }
    sun0 = D.size();
    for (size_t t = 0; t < up.size(); ++t)
      if (up[t]) {
        D.push_back(make_float3(vec[3 * t], vec[3 * t + 1], vec[3 * t + 2]));
        sunStep.push_back(int(t));
      }
    nSun = sunStep.size();
    nDir = D.size();
    out.nSky = nSky;
    out.nSun = nSun;
  }

  // the weight of every direction per quantity (run.cpp:475-505), the ground's for the series, reflected light per bin
  void weights() {
    const size_t P = size_t(pa.P);
    Wq.assign(nQ() * nDir + (series ? nDir : 0), 0.f);
    sunRad.assign(nSun, 0.0);
    sunVis.assign(nSun, 0.0);
    for (size_t j = 0; j < nSun && needWeather; ++j) {
      sunRad[j] = sunRadT[size_t(sunStep[j])];
      sunVis[j] = sunVisT[size_t(sunStep[j])];
    }
    for (size_t qi = 0; qi < nQ(); ++qi) {
      const Quantity q = O.quantities[qi];
      float* w = Wq.data() + qi * nDir;
      if (q == Quantity::SunHours) {   // the step's length for every visible sun, no cosine
        for (size_t j = 0; j < nSun; ++j) w[sun0 + j] = float(st[size_t(sunStep[j])].dt);
        continue;
      }
      const bool vis = q == Quantity::Illuminance;
      if (ladybug) {   // the sun sits in the patches (gendaymtx -d): sky and ground only (run.cpp:489-502)
        for (size_t d = 0; d < nSky; ++d) {
          const size_t p = size_t(rowOf[d]);
          w[d] = float(q == Quantity::Direct ? lb.dir[p] : q == Quantity::Diffuse ? lb.dif[p] : lb.rad[p]);
        }
        if (q != Quantity::Direct) std::fill(w + nSky, w + 2 * nSky, float(lb.gnd));
        continue;
      }
      if (q != Quantity::Direct)   // (radiance x omega) / K in double, rounded to float
        for (size_t d = 0; d < nSky; ++d) {
          const size_t p = size_t(rowOf[d]);
          w[d] = float((vis ? skyVis[p] : skyRad[p]) * pa.omega[p] / K);
          w[nSky + d] = float((vis ? gndVis : gndRad) * pa.omega[p] / K);
        }
      if (q != Quantity::Diffuse)
        for (size_t j = 0; j < nSun; ++j) w[sun0 + j] = float(vis ? sunVis[j] : sunRad[j]);
    }
    if (series)   // the ground's view factor per sensor: the ground directions' solid angles
      for (size_t d = 0; d < nSky; ++d) Wq[nQ() * nDir + nSky + d] = ladybug ? 1.f : pa.omega[size_t(rowOf[d])] / float(K);
    // reflected light (run.cpp:506-515): bin 0 the ground, 1 + p the sky with the sun binned; no cosine (the rows hold it)
    Wdc.assign(reflect ? nQ() * (P + 1) : 0, 0.f);
    hasDc.assign(nQ(), 0);
    for (size_t qi = 0; reflect && qi < nQ(); ++qi) {
      const Quantity q = O.quantities[qi];
      if (q == Quantity::Direct || q == Quantity::SunHours) continue;
      const bool vis = q == Quantity::Illuminance;
      float* w = Wdc.data() + qi * (P + 1);
      hasDc[qi] = 1;
      w[0] = float(vis ? gndVis : gndRad);
      for (size_t p = 0; p < P; ++p) w[p + 1] = float(vis ? skyVis[p] + sunBinVis[p] : skyRad[p] + sunBinRad[p]);
    }
    if (O.verbose && needWeather && !ladybug) {
      double dropped = 0;
      for (size_t t = 0; t < st.size(); ++t)
        if (!up[t] && st[t].dni > 0) dropped += double(st[t].dni) * st[t].dt;
      if (dropped > 0) std::fprintf(stderr, "warn: %.1f Wh/m2 of DNI at steps with the sun below the horizon dropped (gendaymtx would bin it)\n", dropped);
    }
  }

  // ------------------------------------------------------------------------------------------------ scene
  bool scene() {
    const size_t nWords = (nDir + 31) / 32, nBins = size_t(pa.P) + 1;
    // per sensor of a chunk: its visibility (bits, or a byte per direction in whole words), its series row, its bounce row
    perSensor = std::max<size_t>(1, std::max({(translucent ? 32 : 4) * nWords, series ? 4 * size_t(Tup) : size_t(0), reflect ? 4 * nBins : size_t(0)}));
    chunkBytes = O.chunkBytes;
    const auto blocksK = [this](size_t k) { return blocks(k); };
    dlg_gpu* gp = nullptr;
    if (!Engine::get().acquire(lane, M, blocksK, slotted, key, O.residentScenes, O.cancel, nDir > 0 ? &gp : nullptr, sc, out, stage, err,
                               [this](const char* w) { mark(w); }, [this](Scene& fresh) { if (translucent) thin_solids(fresh); }, &lease))
      return false;
    mark("scene");
    if (!gp) { g = nullptr; return true; }
    g = gp;
    dlg_stats(g, &st0);
    // the visibility, the bounce rows, and the series' output and two operand sets per sensor of a chunk
    const size_t buffers = 1 + (reflect ? 1 : 0) + (series ? 3 : 0);
    return Engine::get().admit(*lane, nSens, perSensor, buffers * perSensor + 8 * (nQ() + 1), 0, O.residentScenes, chunkBytes, err);
  }

  // a fresh slotted build's thin solids (reference mesh.cpp:275-304 per transmitting slot), kept with the scene (the store
  // calls this before any other run can reuse the build)
  void thin_solids(Scene& cur) {
    stage = "finding thin solids";
    const auto t0 = Clk::now();
    std::vector<uint32_t>& thin = cur.thin;
    thin.assign(slots.size(), 0);
    std::vector<size_t> base(M.size() + 1, 0);
    for (size_t k = 0; k < M.size(); ++k) base[k + 1] = base[k] + slot_count(*M[k].mesh);
    par_for(M.size(), [&](size_t k) {
      if (!blocks(k)) return;
      for (size_t s = 0; s < slot_count(*M[k].mesh); ++s)
        if (slots[base[k] + s].tau > 0) thin[base[k] + s] = thin_solid(*M[k].mesh, uint32_t(s), cur.c);
    });
    out.stats.prepMs += ms_since(t0);
  }

  // ------------------------------------------------------------------------------------------------ sensors
  bool sensors() {
    // a thin solid's exit face passes everything; with --cull-backfaces a closed solid is crossed through one front
    // face anyway (the other is culled): tau once without the rule
    for (size_t s = 0; !O.cullBackfaces && s < slots.size() && s < sc.thin.size(); ++s) slots[s].flags |= sc.thin[s];
    stage = "making the sensors";
    const auto t0 = Clk::now();
    std::string warn;
    if (!lane) { lane = Engine::get().run_lane(err); if (!lane) return false; }
    if (!lane_sensors(*lane, M, O, sc.c, nSens, out, warn, err, ladybug)) return false;
    out.stats.prepMs += ms_since(t0);
    mark("sensors");
    if (O.verbose) {
      std::fprintf(stderr, "info: %zu sensors, %zu triangles%s, %zu steps (%.1f h)\n", nSens, sc.nTris, out.sceneReused ? " (scene reused)" : "",
                   st.size(), hours);
      if (sc.skipped) std::fprintf(stderr, "warn: %d degenerate faces skipped in occluders\n", sc.skipped);
      if (!warn.empty()) std::fprintf(stderr, "warn: %s\n", warn.c_str());
      if (translucent) {
        size_t lit = 0, thin = 0;
        for (const SlotOptics& s : slots) { lit += s.tau > 0; thin += (s.flags & DLG_SLOT_THIN) != 0; }
        std::fprintf(stderr, "info: %zu of %zu optics slots transmit (%zu thin solids)\n", lit, slots.size(), thin);
      }
    }
    if (g) {
      char e[512] = {};
      if (dlg_set_scene(g, nullptr, nullptr, -1, reinterpret_cast<const float*>(D.data()), int64_t(nDir), e, sizeof e)) return fail(e);
      // a scene without per-triangle slots reads slot 0: one record of the scene slots' common optics
      const std::vector<SlotOptics> one{firstSlot == SIZE_MAX ? SlotOptics{} : slots[firstSlot]};
      const std::vector<SlotOptics>& rows = slotted ? slots : one;
      if ((translucent || reflect) && dlg_set_slots(g, reinterpret_cast<const dlg_slot*>(rows.data()), int64_t(rows.size()), e, sizeof e)) return fail(e);
    }
    mark("upload");
    return true;
  }

  // ------------------------------------------------------------------------------------------------ chunks (run.cpp:607-690)
  // (also run on a helper thread while the first chunk is on the device: it touches only out.columns / values / f32)
  void columns() {
    out.columns.clear();
    for (size_t qi = 0; qi < nQ(); ++qi)
      for (int s = 0; s < nS; ++s) out.columns.push_back(std::string(kQuantityNames[size_t(O.quantities[qi])]) + side_suffix(O.sides, s));
    const size_t C = out.columns.size();
    out.values.clear();
    out.f32.clear();
    if (O.f32Columns) out.f32.assign(C * nElem, 0.f);
    else out.values.assign(nElem * C, 0.0);
  }

  // sensor si's sum of quantity qi into its element's column: irradiance as a mean W/m2, illuminance as a mean lux
  void store(size_t si, size_t qi, double v) {
    const Quantity q = O.quantities[qi];
    if (q == Quantity::Irradiance) v = hours > 0 ? v * 1000.0 / hours : 0.0;
    else if (q == Quantity::Illuminance) v = hours > 0 ? v / hours : 0.0;
    const size_t el = si / size_t(nS), col = qi * size_t(nS) + si % size_t(nS);
    if (O.f32Columns) out.f32[col * nElem + el] = float(v);
    else out.values[el * out.columns.size() + col] = v;
  }

  // store() of every quantity of sensors i0 .. i0 + nc from the chunk's sums s[qi nc + i]: sensor by sensor, so that an
  // element's columns are written together, in blocks over the host's threads for large chunks
  void store_rows(size_t i0, size_t nc, const double* s) {
    constexpr size_t kBlock = size_t(1) << 15;
    const auto rows = [&](size_t b) {
      for (size_t i = b * kBlock, end = std::min(nc, i + kBlock); i < end; ++i)
        for (size_t qi = 0; qi < nQ(); ++qi) store(i0 + i, qi, s[qi * nc + i]);
    };
    const size_t blocks = (nc + kBlock - 1) / kBlock;
    if (blocks > 1) par_for(blocks, rows);
    else if (blocks) rows(0);
  }

  File seriesFile{nullptr, &std::fclose}, bitsFile{nullptr, &std::fclose};
  std::vector<double> seriesHoys;
  std::thread alloc;   // columns() while the first chunk runs
  bool allocFailed = false;
  bool ready() {
    if (alloc.joinable()) alloc.join();
    return !allocFailed || fail("out of host memory for the result");
  }

  bool open_outputs() {
    const size_t nWords = (nDir + 31) / 32;
    if (series && !sink) {
      seriesFile = open_file(O.seriesPath, "wb");
      if (!seriesFile || !npy_write_header(seriesFile.get(), nSens, size_t(Tup), "<f4")) return fail("cannot write " + O.seriesPath);
      for (size_t t = 0; t < st.size(); ++t)
        if (colOf[t] >= 0) seriesHoys.push_back(st[t].hoy);
    }
    if (!O.dumpBitsPath.empty()) {
      bitsFile = open_file(O.dumpBitsPath, "wb");
      if (!bitsFile || !npy_write_header(bitsFile.get(), nSens, translucent ? nDir : nWords, translucent ? "<u1" : "<u4"))
        return fail("cannot write " + O.dumpBitsPath);
    }
    return true;
  }

  bool chunks() {
    if (!open_outputs()) return false;
    if (!g) {
      columns();
      return true;
    }
    // the result's columns are allocated (their fresh pages touched) on a helper thread while the device works; every
    // store waits for it (ready)
    alloc = std::thread([this] {
      try {
        columns();
      } catch (const std::bad_alloc&) {
        allocFailed = true;
      }
    });
    struct Join {
      std::thread& t;
      ~Join() { if (t.joinable()) t.join(); }
    } join{alloc};
    stage = "tracing";
    const size_t ncMax = std::max<size_t>(1, chunkBytes / perSensor), nBins = size_t(pa.P) + 1, nWords = (nDir + 31) / 32;
    const uint32_t flags = dlg_ray_flags(O.cullBackfaces ? 1 : 0, translucent ? 1 : 0);
    std::vector<int32_t> useCos(nQ() + 1, 1);   // sun hours count dt, the rest weigh by the cosine
    for (size_t qi = 0; qi < nQ(); ++qi) useCos[qi] = O.quantities[qi] != Quantity::SunHours;
    const int nR = int(nQ()) + (series ? 1 : 0);
    const dlg_bounce_in bounce{uint32_t(O.bounces), uint32_t(O.bounceSamples), O.mf, O.north};
    // the contraction's per-direction and per-sun inputs (run.cpp:634-641): W/m2 (or lux) of each sun at its column
    std::vector<float> wsky(nSky), sunW(nSun, 0.f);
    std::vector<int32_t> sunCol(nSun, 0);
    const bool seriesVis = series && O.quantities[0] == Quantity::Illuminance;
    for (size_t d = 0; d < nSky; ++d) wsky[d] = pa.omega[size_t(rowOf[d])] / float(K);
    for (size_t j = 0; j < nSun; ++j) {
      const int c = colOf[size_t(sunStep[j])];
      const double dt = st[size_t(sunStep[j])].dt;
      sunCol[j] = std::max(c, 0);
      if (!ladybug && c >= 0 && dt > 0) sunW[j] = float((seriesVis ? sunVis[j] : sunRad[j] * 1000.0) / dt);
    }
    char e[512] = {};
    if (ladybug && series && Tup > 0 && dlg_set_sky_series(g, lbS.data(), lbG.data(), pa.P, Tup, e, sizeof e)) return fail(e);
    const int method = series_method();
    std::vector<double> sums, dcSums;
    std::vector<float> G, E;
    for (size_t i0 = 0; i0 < nSens; i0 += ncMax) {
      if (cancelled()) return false;
      const size_t nc = std::min(ncMax, nSens - i0);
      if (!reflect) {   // one graph: trace, reduce, series (dlg_chunk)
        if (!graph_chunk(i0, nc, flags, useCos, nR, wsky, sunCol, sunW, method, sums, E)) return false;
        continue;
      }
      if (dlg_trace(g, int64_t(i0), int64_t(nc), flags, e, sizeof e)) return fail(e);
      // the bounce pass writes only its own rows: the reduces still see the trace's visibility
      if (reflect && dlg_bounce(g, int64_t(i0), int64_t(nc), dlg_bounce_flags(O.cullBackfaces ? 1 : 0), &bounce, e, sizeof e)) return fail(e);
      const auto tRed = Clk::now();
      sums.assign(size_t(nR) * nc, 0.0);
      if (dlg_reduce(g, int64_t(nc), nR, Wq.data(), useCos.data(), sums.data(), e, sizeof e)) return fail(e);
      if (!ready()) return false;
      for (size_t qi = 0; qi < nQ(); ++qi) {
        const double* s = sums.data() + qi * nc;
        if (!hasDc[qi]) { for (size_t i = 0; i < nc; ++i) store(i0 + i, qi, s[i]); continue; }
        dcSums.assign(nc, 0.0);
        if (dlg_reduce_dc(g, int64_t(nc), Wdc.data() + qi * nBins, dcSums.data(), e, sizeof e)) return fail(e);
        for (size_t i = 0; i < nc; ++i) store(i0 + i, qi, s[i] + dcSums[i]);   // the reference's one double sum
      }
      out.stats.reduceMs += ms_since(tRed);
      if (O.progress) O.progress->store(double(i0 + nc) / double(nSens));
      if (bitsFile && !dump_visibility(nc, nWords)) return false;
      if (!series) continue;
      G.resize(nc);
      for (size_t i = 0; i < nc; ++i) G[i] = float(sums[nQ() * nc + i]);
      const bool compute = Tup > 0 && needWeather;
      const dlg_series_in in{uint32_t(nSky), rowOf.data(), wsky.data(), pa.P, Tup, G.data(), uint32_t(sun0), uint32_t(nSun), sunCol.data(), sunW.data(),
                             reflect ? 1 : 0};
      if (sink) {   // the series stays on the device: the sink reads each chunk there
        if (compute && dlg_series(g, int64_t(nc), &in, method, nullptr, e, sizeof e)) return fail(e);
        if (*sink && !(*sink)(i0, nc, size_t(Tup), compute ? dlg_series_device(g) : nullptr, err)) return false;
        continue;
      }
      E.assign(nc * size_t(Tup), 0.f);
      if (compute && dlg_series(g, int64_t(nc), &in, method, E.data(), e, sizeof e)) return fail(e);
      if (std::fwrite(E.data(), sizeof(float), E.size(), seriesFile.get()) != E.size()) return fail("write failed: " + O.seriesPath);
    }
    if (!ready()) return false;
    mark("trace, reduce, series");
    stats();
    return true;
  }

  // A chunk as one graph on the device (dlg_chunk): the trace, every quantity's reduce and the series, whose ground
  // view factor comes from the reduce's last row on the device; then the columns, the dump and the series' output.
  bool graph_chunk(size_t i0, size_t nc, uint32_t flags, const std::vector<int32_t>& useCos, int nR, const std::vector<float>& wsky,
                   const std::vector<int32_t>& sunCol, const std::vector<float>& sunW, int method, std::vector<double>& sums, std::vector<float>& E) {
    const bool compute = series && Tup > 0 && needWeather;
    const dlg_series_in in{uint32_t(nSky), rowOf.data(), wsky.data(), pa.P, Tup, nullptr, uint32_t(sun0), uint32_t(nSun), sunCol.data(), sunW.data(), 0};
    const dlg_chunk_in ci{int64_t(i0), int64_t(nc), flags, nR, Wq.data(), useCos.data(), compute ? &in : nullptr, method};
    const auto t0 = Clk::now();
    dlg_stats_t a{}, b{};
    dlg_stats(g, &a);
    // the sums land in the context's pinned buffer (else in the run's own vector)
    double* s = dlg_host_sums(g, int64_t(nR) * int64_t(nc));
    if (!s) {
      sums.resize(size_t(nR) * nc);
      s = sums.data();
    }
    const bool toFile = series && !sink;
    E.assign(toFile ? nc * size_t(Tup) : 0, 0.f);
    char e[512] = {};
    if (dlg_chunk(g, &ci, s, toFile && compute ? E.data() : nullptr, e, sizeof e)) return fail(e);
    dlg_stats(g, &b);
    const auto tStore = Clk::now();
    if (!ready()) return false;
    store_rows(i0, nc, s);
    out.stats.reduceMs += (b.reduceMs - a.reduceMs) + ms_since(tStore);
    (void)t0;
    if (O.progress) O.progress->store(double(i0 + nc) / double(nSens));
    if (bitsFile && !dump_visibility(nc, (nDir + 31) / 32)) return false;
    if (!series) return true;
    if (sink) return !*sink || (*sink)(i0, nc, size_t(Tup), compute ? dlg_series_device(g) : nullptr, err);
    if (std::fwrite(E.data(), sizeof(float), E.size(), seriesFile.get()) != E.size()) return fail("write failed: " + O.seriesPath);
    return true;
  }

  // the chunk's visibility as rows [nc x nWords] of words, or byte mode's [nc x nDir] bytes (downloads are
  // direction-major)
  bool dump_visibility(size_t nc, size_t nWords) {
    char e[512] = {};
    if (translucent) {
      std::vector<uint8_t> dm(nDir * nc), rows(nc * nDir);
      if (dlg_download_bytes(g, int64_t(nc), dm.data(), e, sizeof e)) return fail(e);
      for (size_t d = 0; d < nDir; ++d)
        for (size_t i = 0; i < nc; ++i) rows[i * nDir + d] = dm[d * nc + i];
      if (std::fwrite(rows.data(), 1, rows.size(), bitsFile.get()) != rows.size()) return fail("write failed: " + O.dumpBitsPath);
      return true;
    }
    std::vector<uint32_t> wm(nWords * nc), rows(nc * nWords);
    if (dlg_download_bits(g, int64_t(nc), wm.data(), e, sizeof e)) return fail(e);
    for (size_t w = 0; w < nWords; ++w)
      for (size_t i = 0; i < nc; ++i) rows[i * nWords + w] = wm[w * nc + i];
    if (std::fwrite(rows.data(), 4, rows.size(), bitsFile.get()) != rows.size()) return fail("write failed: " + O.dumpBitsPath);
    return true;
  }

  void stats() {
    dlg_stats_t s1{};
    dlg_stats(g, &s1);
    out.stats.traceMs = s1.traceMs - st0.traceMs;
    out.stats.bounceMs = s1.bounceMs - st0.bounceMs;
    out.stats.rays = size_t(s1.rays - st0.rays);
    out.stats.launches = size_t(s1.launches - st0.launches);
    out.stats.gasBytes = size_t(s1.sceneBytes);
    if (!O.verbose) return;
    const double traced = double(s1.traced - st0.traced);
    std::fprintf(stderr, "info: scene %.1f MB (tree height %lld) build %.0f ms, %zu launches, %.3g rays in %.0f ms (%.2f GRays/s), %.3g traced (%.2f GRays/s traced)\n",
                 s1.sceneBytes / 1048576.0, (long long)s1.height, s1.buildMs, out.stats.launches, double(out.stats.rays), out.stats.traceMs,
                 out.stats.traceMs > 0 ? double(out.stats.rays) / out.stats.traceMs / 1e6 : 0.0, traced,
                 out.stats.traceMs > 0 ? traced / out.stats.traceMs / 1e6 : 0.0);
    if (reflect)
      std::fprintf(stderr, "info: reflected light: %d bounces, %d samples per sensor, %.3g paths in %.0f ms (%lld launches)%s\n", O.bounces, O.bounceSamples,
                   double(s1.paths - st0.paths), out.stats.bounceMs, (long long)(s1.bounceLaunches - st0.bounceLaunches),
                   differs ? ", a slot per triangle" : "");
  }

  void host_line() {
    mark("end");
    std::string line = "info: host stages (ms):";
    for (size_t k = 1; k < marks.size(); ++k) {
      char b[64];
      std::snprintf(b, sizeof b, " %s %.4g;", marks[k].first, std::chrono::duration<double, std::milli>(marks[k].second - marks[k - 1].second).count());
      line += b;
    }
    std::fprintf(stderr, "%s\n", line.c_str());
  }

  // ------------------------------------------------------------------------------------------------ dumps (run.cpp:708-736)
  bool dumps() {
    if (seriesFile) {
      seriesFile.reset();
      if (File h = open_file(O.seriesPath + ".hoys.txt", "w"))
        for (const double x : seriesHoys) std::fprintf(h.get(), "%.4f\n", x);
    }
    if (!O.dumpSun.empty()) {   // columns 7-9: Radiance sun.c at the exact site, for the gendaymtx -D comparison
      File f = open_file(O.dumpSun, "w");
      if (!f) return fail("cannot write " + O.dumpSun);
      std::fprintf(f.get(), "# hoy month day hour alt_deg azi_deg up alt_radiance_deg azi_radiance_deg up_radiance dni dhi   (north %.3f)\n", O.north);
      for (size_t t = 0; t < st.size(); ++t) {
        double altR, aziR;
        sun_radiance(lat, -lon, -15.0 * tz, st[t].jd, st[t].hour, altR, aziR);
        std::fprintf(f.get(), "%.4f %d %d %.4f %.6f %.6f %d %.6f %.6f %d %.4f %.4f\n", st[t].hoy, st[t].month, st[t].day, st[t].hour, alt[t], azi[t], up[t],
                     altR / (kPi / 180.), aziR / (kPi / 180.), altR > -0.2665 * (kPi / 180.) ? 1 : 0, st[t].dni, st[t].dhi);
      }
    }
    if (!O.dumpSky.empty()) {   // the two gendaymtx -A ASCII blocks, -s then -d, ground row first
      File f = open_file(O.dumpSky, "w");
      if (!f) return fail("cannot write " + O.dumpSky);
      const size_t P = size_t(pa.P);
      const char* head[2] = {"# gendaymtx -s -A ascii rows (ground row first), %.3g\n", "# gendaymtx -d -A ascii rows (ground row first), %.3g\n"};
      for (size_t b = 0; b < 2; ++b) {
        std::fputs(head[b], f.get());
        for (size_t r = 0; r <= P; ++r) {
          const float* v = &lb.dump[b * 3 * (P + 1) + 3 * r];
          std::fprintf(f.get(), "%.3g %.3g %.3g\n", v[0], v[1], v[2]);
        }
      }
    }
    return true;
  }
};

// the run under a guard: host memory running out, or any exception, becomes the run's message
template <class F> bool guarded(Result& out, std::string& err, const char*& stage, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    out = Result{};
    err = std::string("out of memory: host memory ran out while ") + stage;
  } catch (const std::exception& e) {
    out = Result{};
    err = std::string("internal error while ") + stage + ": " + e.what();
  }
  return false;
}

}  // namespace

// .npy v1.0 header: magic, version, header length, the dict padded with spaces to a multiple of 64 bytes with the
// newline
bool npy_write_header(std::FILE* f, size_t rows, size_t cols, const char* descr) {
  char dict[256];
  const int n = std::snprintf(dict, sizeof dict, "{'descr': '%s', 'fortran_order': False, 'shape': (%zu, %zu), }", descr, rows, cols);
  if (n <= 0) return false;
  std::string h(dict, size_t(n));
  h.append((64 - (10 + h.size() + 1) % 64) % 64, ' ');
  h += '\n';
  const uint16_t len = uint16_t(h.size());
  return std::fwrite("\x93NUMPY\x01\x00", 1, 8, f) == 8 && std::fwrite(&len, 2, 1, f) == 1 && std::fwrite(h.data(), 1, h.size(), f) == h.size();
}

bool run_series_device(const std::vector<SimulationMeshView>& meshes, const Weather* w, const Options& opt, const SeriesSink& sink, Result& out,
                       std::string& err) {
  const char* stage = "preparing the run";
  return guarded(out, err, stage, [&] { return DaylightRun(meshes, w, opt, out, err, stage, &sink).go(); });
}

bool run(const std::vector<SimulationMeshView>& meshes, const Weather* w, const Options& opt, Result& out, std::string& err) {
  const char* stage = "preparing the run";
  return guarded(out, err, stage, [&] { return DaylightRun(meshes, w, opt, out, err, stage, nullptr).go(); });
}

bool run(const std::vector<SimulationMesh>& meshes, const Weather* w, const Options& opt, Result& out, std::string& err) {
  return run(views_of(meshes), w, opt, out, err);
}

}  // namespace daylight
