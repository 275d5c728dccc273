// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The C ABI of include/daylight_c.h. Its contract is reference src/c_api.cpp's (the symbols, the status codes, the
// messages, the refused flags, immutable mesh handles with serials never reused, results that own their values, a
// per-thread message, no exception across the boundary); the implementation is this library's: a call is a Request
// (request.h) made from the flags and handles, then run on the engine's primary lane, whose runs take turns.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <locale.h>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "engine.h"
#include "request.h"

using namespace daylight;

namespace {

thread_local char tlsError[2048];   // a plain array: setting it never allocates (the out-of-memory path)
std::atomic<uint64_t> nextSerial{0};
std::atomic<int64_t> chunkCap{0};   // dlx_chunk_bytes

// records msg as this thread's last error and returns code
int32_t status(int32_t code, const std::string& msg) {
  std::snprintf(tlsError, sizeof tlsError, "%s", msg.c_str());
  return code;
}

// the body of an entry point with every exception turned into a status and a message
template <class F> int32_t boundary(F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return status(DL_RUNTIME, "out of memory: host memory ran out");
  } catch (const std::exception& e) {
    return status(DL_RUNTIME, e.what());
  } catch (...) {
    return status(DL_RUNTIME, "internal error");
  }
}

std::string extension(const std::string& path) {
  const size_t dot = path.find_last_of('.');
  std::string e = dot == std::string::npos ? std::string() : path.substr(dot);
  std::transform(e.begin(), e.end(), e.begin(), [](unsigned char c) { return char(std::tolower(c)); });
  return e;
}

double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

namespace daylight {

CLocaleScope::CLocaleScope() {
  locale_t c = newlocale(LC_ALL_MASK, "C", locale_t(0));
  c_ = c;
  old_ = c ? uselocale(c) : nullptr;
}
CLocaleScope::~CLocaleScope() {
  if (!c_) return;
  uselocale(static_cast<locale_t>(old_));
  freelocale(static_cast<locale_t>(c_));
}

int32_t make_request(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t nSensors, dl_mesh* const* context,
                     int32_t nContext, dl_control* control, bool seriesSink, Request& r, std::string& err) {
  const auto usage = [&](std::string msg) { err = std::move(msg); return int32_t(DL_USAGE); };
  if (tool != DL_DAYLIGHT && tool != DL_VIEWS) return usage("dl_run: unknown tool (DL_DAYLIGHT 0, DL_VIEWS 1)");
  if (argc < 0 || nSensors < 0 || nContext < 0 || (argc && !argv) || (nSensors && !sensors) || (nContext && !context))
    return usage("dl_run: a negative count or a NULL array");
  r.tool = tool;
  // argv as the tool sees it: its name first
  std::vector<std::string> words{tool == DL_VIEWS ? "views" : "daylight"};
  for (int32_t i = 0; i < argc; ++i) {
    if (!argv[i]) return usage("dl_run: a NULL flag");
    words.emplace_back(argv[i]);
  }
  std::vector<char*> av;
  for (std::string& w : words) av.push_back(w.data());
  // the meshes (sensor handles, then context handles) and their identity: "dl", then " s<serial>" / " c<serial>"
  std::string key = "dl";
  r.meshes.clear();
  for (const bool ctx : {false, true}) {
    dl_mesh* const* hs = ctx ? context : sensors;
    for (int32_t k = 0, n = ctx ? nContext : nSensors; k < n; ++k) {
      if (!hs[k]) return usage("dl_run: a NULL mesh handle");
      key += (ctx ? " c" : " s") + std::to_string(hs[k]->serial);
      for (const SimulationMesh& m : hs[k]->meshes) r.meshes.push_back({&m, m.name, !ctx});
    }
  }
  // flags the in-process form refuses: the CLI's file arguments and its benchmark
  const auto refused = [&](const cli::CommonArgs& a) -> int32_t {
    if (a.help) return usage("-h: dl_usage(tool) gives the usage text");
    if (!a.sensors.empty()) return usage("argument '" + a.sensors[0] + "': mesh files are not read in-process; dl_mesh_load(path) gives a handle");
    if (!a.context.empty()) return usage("--context: in-process the context meshes are dl_run's context handles");
    if (!a.out.empty()) return usage("--out: in-process the result holds the values and the manifest");
    if (a.bench) return usage("--bench: the CLIs' benchmark, not in-process");
    return nSensors ? int32_t(DL_OK) : usage("no mesh given");
  };
  std::vector<SlotOptics> slots;
  std::vector<char> open;
  if (tool == DL_DAYLIGHT) {
    if (cli::parse(int(av.size()), av.data(), r.day, err)) return DL_USAGE;
    if (const int32_t rc = refused(r.day)) return rc;
    if (const int rc = cli::prepare(r.day, r.weather, r.haveWeather, err)) return rc;
    if (!resolve_optics(r.meshes, r.day.opt.optics, slots, err) || !resolve_no_occlude(r.meshes, r.day.opt.noOcclude, open, err)) return DL_USAGE;
  } else {
    if (views::cli::parse(int(av.size()), av.data(), r.view, err)) return DL_USAGE;
    if (const int32_t rc = refused(r.view)) return rc;
    std::vector<std::string> groups;
    std::vector<uint8_t> cls;
    if (!resolve_no_occlude(r.meshes, r.view.opt.noOcclude, open, err) ||
        !resolve_targets(r.meshes, r.view.opt.targets, r.view.opt.selfOcclude, open, groups, cls, err) ||
        !resolve_optics(r.meshes, r.view.opt.optics, slots, err))
      return DL_USAGE;
  }
  RunOptions& o = r.run_options();
  o.sceneKey = key + (o.selfOcclude ? " self" : " noself");
  o.cancel = control ? &control->cancel : nullptr;
  o.progress = control ? &control->progress : nullptr;
  if (const int64_t cap = chunkCap.load()) o.chunkBytes = size_t(cap);
  if (seriesSink && (tool != DL_DAYLIGHT || r.day.opt.seriesPath.empty())) return usage("dlx_run_series_device: the flags name no --timeseries");
  return DL_OK;
}

int32_t abi_status(int32_t code, const std::string& msg) { return status(code, msg); }

std::unique_ptr<dl_result> make_result(const Request& r, Result& res) {
  auto out = std::make_unique<dl_result>();
  if (r.tool == DL_DAYLIGHT) out->manifest = cli::manifest_json(res);
  else {
    views::Field f;
    std::string ignored;
    views::field(r.meshes, r.view.opt, f, ignored);
    out->manifest = views::cli::manifest_json(res, r.view.opt, f);
  }
  out->cols = int64_t(res.columns.size());
  out->rows = out->cols ? int64_t(res.values.size()) / out->cols : 0;
  out->values = std::move(res.values);
  const RunStats& s = res.stats;
  const double row[DL_STATS] = {res.sceneReused ? 1.0 : 0.0, s.prepMs, s.skyMs, s.buildMs, s.traceMs, s.reduceMs, s.bounceMs,
                                double(s.rays), double(s.launches), double(s.gasBytes)};
  std::copy(row, row + DL_STATS, out->stats);
  return out;
}

}  // namespace daylight

namespace {

// dl_run and its device-series form (sink null: --timeseries goes to its file, as in the CLI)
int32_t run_call(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t nSensors, dl_mesh* const* context,
                 int32_t nContext, dl_control* control, const SeriesSink* sink, dl_result** out) {
  return boundary([&]() -> int32_t {
    if (!out) return status(DL_USAGE, "dl_run: out is NULL");
    *out = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const CLocaleScope cLocale;
    Request r;
    std::string err;
    if (const int32_t rc = make_request(tool, argc, argv, sensors, nSensors, context, nContext, control, sink != nullptr, r, err)) return status(rc, err);
    const auto t1 = std::chrono::steady_clock::now();
    Result res;
    {
      std::lock_guard<std::mutex> lk(Engine::get().primary_lock());
      const Weather* w = r.haveWeather ? &r.weather : nullptr;
      const bool ok = tool == DL_VIEWS ? views::run(r.meshes, r.view.opt, res, err)
                      : sink           ? run_series_device(r.meshes, w, r.day.opt, *sink, res, err)
                                       : run(r.meshes, w, r.day.opt, res, err);
      if (!ok) return status(DL_RUNTIME, err);
    }
    if (control) control->progress = 1;
    const auto t2 = std::chrono::steady_clock::now();
    *out = make_result(r, res).release();
    if (tool == DL_DAYLIGHT && r.verbose())
      std::fprintf(stderr, "info: dl_run (ms): flags and weather %.3f, run %.3f, result %.3f\n", ms_between(t0, t1), ms_between(t1, t2),
                   ms_between(t2, std::chrono::steady_clock::now()));
    return int32_t(DL_OK);
  });
}

}  // namespace

extern "C" {

int32_t dl_version(void) { return DL_ABI_VERSION; }
const char* dl_last_error(void) { return tlsError; }
const char* dl_usage(int32_t tool) {
  switch (tool) {
    case DL_DAYLIGHT: return cli::kUsage;
    case DL_VIEWS: return views::cli::kUsage;
    default: return nullptr;
  }
}

dl_mesh* dl_mesh_create(const char* name, const double* xyz, int64_t n_points, const int32_t* counts, int64_t n_faces, const int32_t* indices,
                        int64_t n_indices, const uint16_t* slots) {
  std::unique_ptr<dl_mesh> h;
  boundary([&]() -> int32_t {
    if (!name) return status(DL_USAGE, "dl_mesh_create: name is NULL");
    const bool badArrays = (n_points && !xyz) || (n_faces && !counts) || (n_indices && !indices);
    if (n_points < 0 || n_faces < 0 || n_indices < 0 || badArrays) return status(DL_USAGE, "dl_mesh_create: a negative count or a NULL array");
    SimulationMesh m;
    m.name = name;
    m.points.assign(xyz, xyz + 3 * n_points);
    m.counts.assign(counts, counts + n_faces);
    m.indices.assign(indices, indices + n_indices);
    if (slots) m.slot.assign(slots, slots + n_faces);
    std::string err;
    if (!mesh_ok(m, m.name, err)) return status(DL_USAGE, err);
    h = std::make_unique<dl_mesh>();
    if (m.numFaces()) h->meshes.push_back(std::move(m));   // a mesh without faces is a handle of no meshes
    h->serial = ++nextSerial;
    return int32_t(DL_OK);
  });
  return h.release();
}

dl_mesh* dl_mesh_load(const char* path) {
  std::unique_ptr<dl_mesh> h;
  boundary([&]() -> int32_t {
    if (!path) return status(DL_USAGE, "dl_mesh_load: path is NULL");
    const std::string p = path, ext = extension(p);
    auto made = std::make_unique<dl_mesh>();
    std::string err;
    const bool ok = ext == ".obj" ? read_obj(p, made->meshes, err) : ext == ".dmesh" ? read_dmesh(p, made->meshes, err) : false;
    if (!ok) return status(DL_RUNTIME, ext == ".obj" || ext == ".dmesh" ? err : p + ": unknown extension; .obj or .dmesh in-process");
    made->serial = ++nextSerial;
    h = std::move(made);
    return int32_t(DL_OK);
  });
  return h.release();
}

int32_t dl_mesh_count(const dl_mesh* mesh) { return mesh ? int32_t(mesh->meshes.size()) : -1; }
const char* dl_mesh_name(const dl_mesh* mesh, int32_t k) {
  if (!mesh || k < 0 || size_t(k) >= mesh->meshes.size()) return nullptr;
  return mesh->meshes[size_t(k)].name.c_str();
}
void dl_mesh_free(dl_mesh* mesh) { delete mesh; }

dl_control* dl_control_create(void) {
  dl_control* c = nullptr;
  boundary([&]() -> int32_t { c = new dl_control; return int32_t(DL_OK); });
  return c;
}
void dl_control_cancel(dl_control* control) { if (control) control->cancel = true; }
double dl_control_progress(const dl_control* control) { return control ? control->progress.load() : 0.0; }
void dl_control_free(dl_control* control) { delete control; }

int32_t dl_run(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors, dl_mesh* const* context,
               int32_t n_context, dl_control* control, dl_result** out) {
  return run_call(tool, argc, argv, sensors, n_sensors, context, n_context, control, nullptr, out);
}

int64_t dl_result_rows(const dl_result* result) { return result ? result->rows : -1; }
int64_t dl_result_cols(const dl_result* result) { return result ? result->cols : -1; }
const char* dl_result_manifest(const dl_result* result) { return result ? result->manifest.c_str() : nullptr; }

int32_t dl_result_values(const dl_result* result, double* out, int64_t n) {
  if (!result || !out) return status(DL_USAGE, "dl_result_values: a NULL argument");
  if (n < int64_t(result->values.size())) return status(DL_USAGE, "dl_result_values: out holds fewer than rows x cols values");
  std::copy(result->values.begin(), result->values.end(), out);
  return DL_OK;
}

int32_t dl_result_values_f32(const dl_result* result, float* out, int64_t n) {
  if (!result || !out) return status(DL_USAGE, "dl_result_values_f32: a NULL argument");
  if (n < int64_t(result->values.size())) return status(DL_USAGE, "dl_result_values_f32: out holds fewer than rows x cols values");
  std::transform(result->values.begin(), result->values.end(), out, [](double v) { return float(v); });
  return DL_OK;
}

int32_t dl_result_stats(const dl_result* result, double* out) {
  if (!result || !out) return status(DL_USAGE, "dl_result_stats: a NULL argument");
  std::copy(result->stats, result->stats + DL_STATS, out);
  return DL_OK;
}

void dl_result_free(dl_result* result) { delete result; }

// ------------------------------------------------------------------------------------------------ extensions (dlx_)

const double* dlx_result_data(const dl_result* result) { return result ? result->values.data() : nullptr; }

int32_t dlx_run_series_device(int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors, dl_mesh* const* context,
                              int32_t n_context, dl_control* control, dlx_series_sink sink, void* user, dl_result** out) {
  int32_t sinkRc = DL_OK;
  const SeriesSink each = [&](size_t i0, size_t nc, size_t T, const float* E, std::string& err) {
    if (!sink) return true;
    sinkRc = sink(user, int64_t(i0), int64_t(nc), int64_t(T), E);
    if (sinkRc != DL_OK) err = "dlx_run_series_device: the sink returned " + std::to_string(sinkRc);
    return sinkRc == DL_OK;
  };
  return run_call(DL_DAYLIGHT, argc, argv, sensors, n_sensors, context, n_context, control, &each, out);
}

int32_t dlx_gpu_counters(double* out, int32_t n) {
  if (!out || n < 0) return status(DL_USAGE, "dlx_gpu_counters: a NULL array or a negative count");
  std::lock_guard<std::mutex> lk(Engine::get().primary_lock());
  std::string err;
  Lane* l = Engine::get().primary(err);
  if (!l) return status(DL_RUNTIME, err);
  dlg_stats_t s{};
  dlg_stats(l->gpu, &s);
  const double c[DLX_COUNTERS] = {s.traceMs, double(s.launches), double(s.rays), double(s.nodeFetches), double(s.triFetches), double(s.countedRays),
                                  s.reduceMs, s.seriesMs, s.skyMs, s.buildMs, s.viewMs, s.lastTraceMs, s.lastSeriesMs, double(s.sceneBytes),
                                  double(s.height), double(s.waveIters), double(s.traced), double(s.graphLaunches), double(s.graphCaptures),
                                  double(s.leafIters)};
  std::copy(c, c + std::min<int32_t>(n, DLX_COUNTERS), out);
  return DL_OK;
}

void dlx_count_fetches(int32_t on) {
  std::lock_guard<std::mutex> lk(Engine::get().primary_lock());
  std::string err;
  if (Lane* l = Engine::get().primary(err)) dlg_count_fetches(l->gpu, on);
}

void dlx_chunk_bytes(int64_t bytes) { chunkCap.store(bytes > 0 ? bytes : 0); }

int32_t dlx_read_epw(const char* path, float* dni, float* dhi, int64_t cap, int64_t* rows, double* latLonTz) {
  return boundary([&]() -> int32_t {
    if (!path || !rows) return status(DL_USAGE, "dlx_read_epw: a NULL path or rows");
    const CLocaleScope cLocale;
    Weather w;
    std::string err;
    if (!read_epw(path, w, err)) return status(DL_RUNTIME, err);
    *rows = int64_t(w.dni.size());
    if (latLonTz) { latLonTz[0] = w.lat; latLonTz[1] = w.lon; latLonTz[2] = w.tz; }
    const size_t n = size_t(std::max<int64_t>(0, std::min(cap, *rows)));
    if (dni) std::copy_n(w.dni.begin(), n, dni);
    if (dhi) std::copy_n(w.dhi.begin(), n, dhi);
    return int32_t(DL_OK);
  });
}

int32_t dlx_thin_solid(const dl_mesh* mesh, int32_t slot, const double* centre, uint32_t* flags) {
  return boundary([&]() -> int32_t {
    if (!mesh || !flags || mesh->meshes.size() != 1 || slot < 0) return status(DL_USAGE, "dlx_thin_solid: one mesh, a slot >= 0 and flags");
    const double origin[3] = {0, 0, 0};
    *flags = thin_solid(mesh->meshes[0], uint32_t(slot), centre ? centre : origin);
    return int32_t(DL_OK);
  });
}

// the triangles the scene build draws from a one-mesh handle, with their global slots
namespace {
struct Drawn {
  std::vector<float3> v;
  std::vector<uint3> tri;
  std::vector<uint32_t> slot;
  int skipped = 0;
  size_t n = 0;
};
Drawn draw(const SimulationMesh& m, const double* centre, uint32_t slotBase, bool slots) {
  Drawn d;
  d.v.resize(m.numPoints());
  d.tri.resize(tri_cap(m));
  if (slots) d.slot.resize(d.tri.size());
  const double origin[3] = {0, 0, 0};
  d.n = triangulate_into(m, centre ? centre : origin, d.v.data(), d.tri.data(), slots ? d.slot.data() : nullptr, 0, slotBase, d.skipped);
  return d;
}
}  // namespace

int32_t dlx_triangle_slots(const dl_mesh* mesh, uint32_t slotBase, uint32_t* slots, int64_t cap, int64_t* n) {
  return boundary([&]() -> int32_t {
    if (!mesh || !n || mesh->meshes.size() != 1) return status(DL_USAGE, "dlx_triangle_slots: one mesh and n");
    const Drawn d = draw(mesh->meshes[0], nullptr, slotBase, true);
    *n = int64_t(d.n);
    if (slots) std::copy_n(d.slot.begin(), size_t(std::max<int64_t>(0, std::min<int64_t>(cap, *n))), slots);
    return int32_t(DL_OK);
  });
}

int32_t dlx_triangulate(const dl_mesh* mesh, const double* centre, float* v, uint32_t* tri, int64_t cap, int64_t* n, int32_t* skipped) {
  return boundary([&]() -> int32_t {
    if (!mesh || !n || mesh->meshes.size() != 1) return status(DL_USAGE, "dlx_triangulate: one mesh and n");
    const Drawn d = draw(mesh->meshes[0], centre, 0, false);
    *n = int64_t(d.n);
    if (skipped) *skipped = d.skipped;
    for (size_t i = 0; v && i < d.v.size(); ++i) { v[3 * i] = d.v[i].x; v[3 * i + 1] = d.v[i].y; v[3 * i + 2] = d.v[i].z; }
    for (size_t k = 0; tri && int64_t(k) < std::min(cap, *n); ++k) { tri[3 * k] = d.tri[k].x; tri[3 * k + 1] = d.tri[k].y; tri[3 * k + 2] = d.tri[k].z; }
    return int32_t(DL_OK);
  });
}

int32_t dlx_ladybug_sky(int32_t mf, int64_t T, const int32_t* jd, const double* hour, const float* dni, const float* dhi, double lat, double lon,
                        double fileTz, double grefl, double dur, double* dir, double* dif, double* rad, double* gnd, float* dump) {
  return boundary([&]() -> int32_t {
    if ((mf != 1 && mf != 2) || T <= 0 || !jd || !hour || !dni || !dhi || !dir || !dif || !rad || !gnd || !dump)
      return status(DL_USAGE, "dlx_ladybug_sky: sky density 1 or 2, T > 0 steps and every array");
    std::vector<Step> steps(static_cast<size_t>(T));
    for (size_t t = 0; t < steps.size(); ++t) { steps[t].jd = jd[t]; steps[t].hour = hour[t]; steps[t].dni = dni[t]; steps[t].dhi = dhi[t]; }
    const Patches pa = make_patches(mf);
    LadybugSky lb;
    ladybug_sky(pa, steps, lat, lon, fileTz, grefl, dur, {}, nullptr, nullptr, 0, lb);
    std::copy(lb.dir.begin(), lb.dir.end(), dir);
    std::copy(lb.dif.begin(), lb.dif.end(), dif);
    std::copy(lb.rad.begin(), lb.rad.end(), rad);
    *gnd = lb.gnd;
    std::copy(lb.dump.begin(), lb.dump.end(), dump);
    return int32_t(DL_OK);
  });
}

}  // extern "C"
