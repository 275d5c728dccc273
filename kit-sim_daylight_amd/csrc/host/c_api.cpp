// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The C ABI of include/daylight_c.h (reference src/c_api.cpp's contract): the CLI's own parser, rules and manifests on
// argv, meshes as immutable handles with never-reused serials (the resident scene's identity), one lock around the GPU
// runs, a per-thread error buffer, and no exception across the boundary.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <clocale>
#include <cstdio>
#include <cstring>
#include <locale.h>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "cli.h"
#include "core.h"
#include "daylight_c.h"

using namespace daylight;

struct dl_mesh { uint64_t serial = 0; std::vector<SimulationMesh> meshes; };
struct dl_control { std::atomic<bool> cancel{false}; std::atomic<double> progress{0}; };
struct dl_result { int64_t rows = 0, cols = 0; std::string manifest; std::vector<double> values; double stats[DL_STATS] = {}; };

namespace {

thread_local char lastError[2048];   // a plain array: setting it never allocates (the out-of-memory path)
std::mutex runMu;
std::atomic<uint64_t> serials{0};

int32_t fail(int32_t rc, const char* msg) {
  std::snprintf(lastError, sizeof lastError, "%s", msg);
  return rc;
}

template <class F> int32_t guard(F&& f) noexcept {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(DL_RUNTIME, "out of memory: host memory ran out");
  } catch (const std::exception& e) {
    return fail(DL_RUNTIME, e.what());
  } catch (...) {
    return fail(DL_RUNTIME, "internal error");
  }
}

std::string lower_ext(const std::string& path) {
  std::string ext = path.substr(std::min(path.find_last_of('.'), path.size()));
  for (char& c : ext) c = char(std::tolower(static_cast<unsigned char>(c)));
  return ext;
}

// The flags' and the EPW's numbers are read, and the manifest printed, in the "C" locale for the call (this thread only).
struct CLocale {
  locale_t c = newlocale(LC_ALL_MASK, "C", locale_t(0));
  locale_t old = c ? uselocale(c) : locale_t(0);
  ~CLocale() {
    if (c) { uselocale(old); freelocale(c); }
  }
};

int32_t refused(const cli::CommonArgs& a) {
  if (a.help) return fail(DL_USAGE, "-h: dl_usage(tool) gives the usage text");
  if (!a.sensors.empty())
    return fail(DL_USAGE, ("argument '" + a.sensors[0] + "': mesh files are not read in-process; dl_mesh_load(path) gives a handle").c_str());
  if (!a.context.empty()) return fail(DL_USAGE, "--context: in-process the context meshes are dl_run's context handles");
  if (!a.out.empty()) return fail(DL_USAGE, "--out: in-process the result holds the values and the manifest");
  if (a.bench) return fail(DL_USAGE, "--bench: the CLIs' benchmark, not in-process");
  return DL_OK;
}

std::atomic<int64_t> chunkOverride{0};   // dlx_chunk_bytes

void set_run(RunOptions& o, const std::string& key, dl_control* control) {
  o.sceneKey = key + (o.selfOcclude ? " self" : " noself");
  o.cancel = control ? &control->cancel : nullptr;
  o.progress = control ? &control->progress : nullptr;
  if (const int64_t c = chunkOverride.load()) o.chunkBytes = size_t(c);
}

int32_t finish(Result& r, std::string manifest, dl_result** out) {
  auto res = std::make_unique<dl_result>();
  res->cols = int64_t(r.columns.size());
  res->rows = res->cols ? int64_t(r.values.size()) / res->cols : 0;
  res->manifest = std::move(manifest);
  res->values = std::move(r.values);
  const RunStats& s = r.stats;
  const double st[DL_STATS] = {r.sceneReused ? 1.0 : 0.0, s.prepMs, s.skyMs, s.buildMs, s.traceMs, s.reduceMs, s.bounceMs,
                               double(s.rays), double(s.launches), double(s.gasBytes)};
  std::memcpy(res->stats, st, sizeof st);
  *out = res.release();
  return DL_OK;
}

}  // namespace

extern "C" {

int32_t dl_version(void) { return DL_ABI_VERSION; }
const char* dl_last_error(void) { return lastError; }
const char* dl_usage(int32_t tool) { return tool == DL_DAYLIGHT ? cli::kUsage : tool == DL_VIEWS ? views::cli::kUsage : nullptr; }

dl_mesh* dl_mesh_create(const char* name, const double* xyz, int64_t n_points, const int32_t* counts, int64_t n_faces, const int32_t* indices,
                        int64_t n_indices, const uint16_t* slots) {
  dl_mesh* made = nullptr;
  guard([&]() -> int32_t {
    if (!name) return fail(DL_USAGE, "dl_mesh_create: name is NULL");
    if (n_points < 0 || n_faces < 0 || n_indices < 0 || (n_points && !xyz) || (n_faces && !counts) || (n_indices && !indices))
      return fail(DL_USAGE, "dl_mesh_create: a negative count or a NULL array");
    SimulationMesh m;
    m.name = name;
    m.points.assign(xyz, xyz + 3 * n_points);
    m.counts.assign(counts, counts + n_faces);
    m.indices.assign(indices, indices + n_indices);
    if (slots) m.slot.assign(slots, slots + n_faces);
    std::string err;
    if (!mesh_ok(m, m.name, err)) return fail(DL_USAGE, err.c_str());
    auto h = std::make_unique<dl_mesh>();
    if (m.numFaces()) h->meshes.push_back(std::move(m));
    h->serial = ++serials;
    made = h.release();
    return DL_OK;
  });
  return made;
}

dl_mesh* dl_mesh_load(const char* path) {
  dl_mesh* made = nullptr;
  guard([&]() -> int32_t {
    if (!path) return fail(DL_USAGE, "dl_mesh_load: path is NULL");
    auto h = std::make_unique<dl_mesh>();
    const std::string p = path, ext = lower_ext(p);
    std::string err;
    bool ok = false;
    if (ext == ".obj") ok = read_obj(p, h->meshes, err);
    else if (ext == ".dmesh") ok = read_dmesh(p, h->meshes, err);
    else err = p + ": unknown extension; .obj or .dmesh in-process";
    if (!ok) return fail(DL_RUNTIME, err.c_str());
    h->serial = ++serials;
    made = h.release();
    return DL_OK;
  });
  return made;
}

int32_t dl_mesh_count(const dl_mesh* mesh) { return mesh ? int32_t(mesh->meshes.size()) : -1; }
const char* dl_mesh_name(const dl_mesh* mesh, int32_t k) {
  return mesh && k >= 0 && size_t(k) < mesh->meshes.size() ? mesh->meshes[size_t(k)].name.c_str() : nullptr;
}
void dl_mesh_free(dl_mesh* mesh) { delete mesh; }

dl_control* dl_control_create(void) {
  dl_control* made = nullptr;
  guard([&]() -> int32_t { made = new dl_control; return DL_OK; });
  return made;
}
void dl_control_cancel(dl_control* control) { if (control) control->cancel = true; }
double dl_control_progress(const dl_control* control) { return control ? control->progress.load() : 0.0; }
void dl_control_free(dl_control* control) { delete control; }

}  // extern "C"

namespace {

// dl_run and its device-series form: sink null writes --timeseries to its file, as the CLI does
int32_t run_tool(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors, dl_mesh* const* context,
                 int32_t n_context, dl_control* control, const SeriesSink* sink, dl_result** out) {
  return guard([&]() -> int32_t {
    if (!out) return fail(DL_USAGE, "dl_run: out is NULL");
    *out = nullptr;
    if (tool != DL_DAYLIGHT && tool != DL_VIEWS) return fail(DL_USAGE, "dl_run: unknown tool (DL_DAYLIGHT 0, DL_VIEWS 1)");
    if (argc < 0 || n_sensors < 0 || n_context < 0 || (argc && !argv) || (n_sensors && !sensors) || (n_context && !context))
      return fail(DL_USAGE, "dl_run: a negative count or a NULL array");
    std::vector<std::string> store;
    store.reserve(size_t(argc) + 1);
    store.emplace_back(tool == DL_VIEWS ? "views" : "daylight");
    for (int32_t i = 0; i < argc; ++i) {
      if (!argv[i]) return fail(DL_USAGE, "dl_run: a NULL flag");
      store.emplace_back(argv[i]);
    }
    std::vector<char*> av;
    for (std::string& s : store) av.push_back(s.data());
    std::vector<SimulationMeshView> mv;
    std::string key = "dl";
    for (int32_t pass = 0; pass < 2; ++pass) {
      dl_mesh* const* hs = pass ? context : sensors;
      const int32_t n = pass ? n_context : n_sensors;
      for (int32_t k = 0; k < n; ++k) {
        if (!hs[k]) return fail(DL_USAGE, "dl_run: a NULL mesh handle");
        key += (pass ? " c" : " s") + std::to_string(hs[k]->serial);
        for (const SimulationMesh& m : hs[k]->meshes) mv.push_back({&m, m.name, pass == 0});
      }
    }
    const CLocale cLocale;
    std::string err;
    Result r;
    if (tool == DL_DAYLIGHT) {
      const auto t0 = std::chrono::steady_clock::now();
      cli::Args a;
      if (cli::parse(int(av.size()), av.data(), a, err)) return fail(DL_USAGE, err.c_str());
      if (const int32_t rc = refused(a)) return rc;
      if (!n_sensors) return fail(DL_USAGE, "no mesh given");
      Weather w;
      bool haveWeather = false;
      if (const int rc = cli::prepare(a, w, haveWeather, err)) return fail(rc, err.c_str());
      std::vector<SlotOptics> slots;
      std::vector<char> open;
      if (!resolve_optics(mv, a.opt.optics, slots, err) || !resolve_no_occlude(mv, a.opt.noOcclude, open, err)) return fail(DL_USAGE, err.c_str());
      set_run(a.opt, key, control);
      if (sink && a.opt.seriesPath.empty()) return fail(DL_USAGE, "dlx_run_series_device: the flags name no --timeseries");
      const auto t1 = std::chrono::steady_clock::now();
      {
        std::lock_guard<std::mutex> lock(runMu);
        const bool ok = sink ? run_series_device(mv, haveWeather ? &w : nullptr, a.opt, *sink, r, err) : run(mv, haveWeather ? &w : nullptr, a.opt, r, err);
        if (!ok) return fail(DL_RUNTIME, err.c_str());
      }
      if (control) control->progress = 1;
      const auto t2 = std::chrono::steady_clock::now();
      const int32_t rc = finish(r, cli::manifest_json(r), out);
      if (a.opt.verbose) {
        const auto ms = [](auto x, auto y) { return std::chrono::duration<double, std::milli>(y - x).count(); };
        std::fprintf(stderr, "info: dl_run (ms): flags and weather %.3f, run %.3f, result %.3f\n", ms(t0, t1), ms(t1, t2),
                     ms(t2, std::chrono::steady_clock::now()));
      }
      return rc;
    }
    views::cli::Args a;
    if (views::cli::parse(int(av.size()), av.data(), a, err)) return fail(DL_USAGE, err.c_str());
    if (const int32_t rc = refused(a)) return rc;
    if (!n_sensors) return fail(DL_USAGE, "no mesh given");
    std::vector<char> open;
    std::vector<std::string> groups;
    std::vector<uint8_t> cls;
    std::vector<SlotOptics> slots;
    if (!resolve_no_occlude(mv, a.opt.noOcclude, open, err) || !resolve_targets(mv, a.opt.targets, a.opt.selfOcclude, open, groups, cls, err) ||
        !resolve_optics(mv, a.opt.optics, slots, err))
      return fail(DL_USAGE, err.c_str());
    set_run(a.opt, key, control);
    {
      std::lock_guard<std::mutex> lock(runMu);
      if (!views::run(mv, a.opt, r, err)) return fail(DL_RUNTIME, err.c_str());
    }
    if (control) control->progress = 1;
    views::Field f;
    views::field(mv, a.opt, f, err);
    return finish(r, views::cli::manifest_json(r, a.opt, f), out);
  });
}

}  // namespace

extern "C" {

int32_t dl_run(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors, dl_mesh* const* context,
               int32_t n_context, dl_control* control, dl_result** out) {
  return run_tool(tool, argc, argv, sensors, n_sensors, context, n_context, control, nullptr, out);
}

int32_t dlx_run_series_device(int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors, dl_mesh* const* context,
                              int32_t n_context, dl_control* control, dlx_series_sink sink, void* user, dl_result** out) {
  int32_t sinkRc = DL_OK;
  const SeriesSink s = [&](size_t i0, size_t nc, size_t T, const float* E, std::string& err) {
    if (!sink) return true;
    sinkRc = sink(user, int64_t(i0), int64_t(nc), int64_t(T), E);
    if (sinkRc != DL_OK) err = "dlx_run_series_device: the sink returned " + std::to_string(sinkRc);
    return sinkRc == DL_OK;
  };
  return run_tool(DL_DAYLIGHT, argc, argv, sensors, n_sensors, context, n_context, control, &s, out);
}

int32_t dlx_gpu_counters(double* out, int32_t n) {
  if (!out || n < 0) return fail(DL_USAGE, "dlx_gpu_counters: a NULL array or a negative count");
  std::lock_guard<std::mutex> lock(runMu);
  std::string err;
  dlg_gpu* g = process_gpu(err);
  if (!g) return fail(DL_RUNTIME, err.c_str());
  dlg_stats_t st{};
  dlg_stats(g, &st);
  const double c[DLX_COUNTERS] = {st.traceMs, double(st.launches), double(st.rays), double(st.nodeFetches), double(st.triFetches),
                                  double(st.countedRays), st.reduceMs, st.seriesMs, st.skyMs, st.buildMs, st.viewMs, st.lastTraceMs,
                                  st.lastSeriesMs, double(st.sceneBytes), double(st.height), double(st.waveIters), double(st.traced)};
  for (int32_t k = 0; k < n && k < DLX_COUNTERS; ++k) out[k] = c[k];
  return DL_OK;
}

void dlx_chunk_bytes(int64_t bytes) { chunkOverride.store(bytes > 0 ? bytes : 0); }

int32_t dlx_thin_solid(const dl_mesh* mesh, int32_t slot, const double* centre, uint32_t* flags) {
  return guard([&]() -> int32_t {
    if (!mesh || !flags || mesh->meshes.size() != 1 || slot < 0) return fail(DL_USAGE, "dlx_thin_solid: one mesh, a slot >= 0 and flags");
    const double zero[3] = {0, 0, 0};
    *flags = thin_solid(mesh->meshes[0], uint32_t(slot), centre ? centre : zero);
    return DL_OK;
  });
}

int32_t dlx_triangle_slots(const dl_mesh* mesh, uint32_t slotBase, uint32_t* slots, int64_t cap, int64_t* n) {
  return guard([&]() -> int32_t {
    if (!mesh || !n || mesh->meshes.size() != 1) return fail(DL_USAGE, "dlx_triangle_slots: one mesh and n");
    const SimulationMesh& m = mesh->meshes[0];
    std::vector<float3> v(m.numPoints());
    std::vector<uint3> tri(tri_cap(m));
    std::vector<uint32_t> ts(tri.size());
    const double zero[3] = {0, 0, 0};
    int skipped = 0;
    *n = int64_t(triangulate_into(m, zero, v.data(), tri.data(), ts.data(), 0, slotBase, skipped));
    for (int64_t k = 0; slots && k < *n && k < cap; ++k) slots[k] = ts[size_t(k)];
    return DL_OK;
  });
}

int32_t dlx_triangulate(const dl_mesh* mesh, const double* centre, float* v, uint32_t* tri, int64_t cap, int64_t* n, int32_t* skipped) {
  return guard([&]() -> int32_t {
    if (!mesh || !n || mesh->meshes.size() != 1) return fail(DL_USAGE, "dlx_triangulate: one mesh and n");
    const SimulationMesh& m = mesh->meshes[0];
    std::vector<float3> vv(m.numPoints());
    std::vector<uint3> tt(tri_cap(m));
    const double zero[3] = {0, 0, 0};
    int sk = 0;
    *n = int64_t(triangulate_into(m, centre ? centre : zero, vv.data(), tt.data(), nullptr, 0, 0, sk));
    if (skipped) *skipped = sk;
    for (size_t i = 0; v && i < vv.size(); ++i) { v[3 * i] = vv[i].x; v[3 * i + 1] = vv[i].y; v[3 * i + 2] = vv[i].z; }
    for (int64_t k = 0; tri && k < *n && k < cap; ++k) {
      tri[3 * k] = tt[size_t(k)].x; tri[3 * k + 1] = tt[size_t(k)].y; tri[3 * k + 2] = tt[size_t(k)].z;
    }
    return DL_OK;
  });
}

int32_t dlx_ladybug_sky(int32_t mf, int64_t T, const int32_t* jd, const double* hour, const float* dni, const float* dhi, double lat,
                        double lon, double fileTz, double grefl, double dur, double* dir, double* dif, double* rad, double* gnd, float* dump) {
  return guard([&]() -> int32_t {
    if ((mf != 1 && mf != 2) || T <= 0 || !jd || !hour || !dni || !dhi || !dir || !dif || !rad || !gnd || !dump)
      return fail(DL_USAGE, "dlx_ladybug_sky: sky density 1 or 2, T > 0 steps and every array");
    std::vector<Step> steps(static_cast<size_t>(T));
    for (size_t t = 0; t < steps.size(); ++t) { steps[t].jd = jd[t]; steps[t].hour = hour[t]; steps[t].dni = dni[t]; steps[t].dhi = dhi[t]; }
    const Patches pa = make_patches(mf);
    LadybugSky lb;
    ladybug_sky(pa, steps, lat, lon, fileTz, grefl, dur, {}, nullptr, nullptr, 0, lb);
    for (size_t p = 0; p < size_t(pa.P); ++p) { dir[p] = lb.dir[p]; dif[p] = lb.dif[p]; rad[p] = lb.rad[p]; }
    *gnd = lb.gnd;
    std::copy(lb.dump.begin(), lb.dump.end(), dump);
    return DL_OK;
  });
}

int32_t dlx_read_epw(const char* path, float* dni, float* dhi, int64_t cap, int64_t* rows, double* latLonTz) {
  return guard([&]() -> int32_t {
    if (!path || !rows) return fail(DL_USAGE, "dlx_read_epw: a NULL path or rows");
    const CLocale cLocale;
    Weather w;
    std::string err;
    if (!read_epw(path, w, err)) return fail(DL_RUNTIME, err.c_str());
    *rows = int64_t(w.dni.size());
    if (latLonTz) { latLonTz[0] = w.lat; latLonTz[1] = w.lon; latLonTz[2] = w.tz; }
    for (int64_t k = 0; k < *rows && k < cap; ++k) {
      if (dni) dni[k] = w.dni[size_t(k)];
      if (dhi) dhi[k] = w.dhi[size_t(k)];
    }
    return DL_OK;
  });
}

void dlx_count_fetches(int32_t on) {
  std::lock_guard<std::mutex> lock(runMu);
  std::string err;
  if (dlg_gpu* g = process_gpu(err)) dlg_count_fetches(g, on);
}

int64_t dl_result_rows(const dl_result* result) { return result ? result->rows : -1; }
int64_t dl_result_cols(const dl_result* result) { return result ? result->cols : -1; }
const char* dl_result_manifest(const dl_result* result) { return result ? result->manifest.c_str() : nullptr; }

int32_t dl_result_values(const dl_result* result, double* out, int64_t n) {
  if (!result || !out) return fail(DL_USAGE, "dl_result_values: a NULL argument");
  if (n < int64_t(result->values.size())) return fail(DL_USAGE, "dl_result_values: out holds fewer than rows x cols values");
  std::memcpy(out, result->values.data(), result->values.size() * sizeof(double));
  return DL_OK;
}

int32_t dl_result_values_f32(const dl_result* result, float* out, int64_t n) {
  if (!result || !out) return fail(DL_USAGE, "dl_result_values_f32: a NULL argument");
  if (n < int64_t(result->values.size())) return fail(DL_USAGE, "dl_result_values_f32: out holds fewer than rows x cols values");
  for (size_t i = 0; i < result->values.size(); ++i) out[i] = float(result->values[i]);
  return DL_OK;
}

int32_t dl_result_stats(const dl_result* result, double* out) {
  if (!result || !out) return fail(DL_USAGE, "dl_result_stats: a NULL argument");
  std::memcpy(out, result->stats, sizeof result->stats);
  return DL_OK;
}

void dl_result_free(dl_result* result) { delete result; }

}  // extern "C"
