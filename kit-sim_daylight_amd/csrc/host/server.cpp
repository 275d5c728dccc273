// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The serving loop of the C ABI (dlx_server_*, daylight_c.h). No counterpart in the reference, whose C ABI runs one call
// at a time behind one mutex (reference src/c_api.cpp:31, 231, 250) on one device context.
//
// Requests of both tools are parsed as dl_run parses them (make_request) and queued. Workers, one per lane (a device
// context with its own HIP stream: lanes overlap on the GPU), take them oldest first when the device memory they may
// need fits the budget beside the runs already going (admission). A lane that comes free also takes every queued
// daylight request it can merge with the oldest one - the same flags over the same context handles with the same
// recentring, without self-occlusion or file outputs - and serves them as one run over all their sensor meshes
// (continuous batching), then splits the rows. Scenes stay resident in the engine's store; a run pins its scene for its
// length. Results are the ones each request gives alone: per sensor the run's arithmetic depends on nothing but the
// sensor, the scene, the sky and the centre of the run, and merged requests share all of those.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.h"
#include "request.h"

namespace daylight {
namespace {

using Clk = std::chrono::steady_clock;
double ms_between(Clk::time_point a, Clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

constexpr size_t kFixedBytes = size_t(64) << 20;   // per run beside its sensors: direction tables, sky, weights, tuning

struct Job {
  int64_t id = 0;
  Request req;
  std::vector<dl_mesh*> sensors, context;   // the caller's handles (valid until the job is waited for)
  std::vector<std::string> argv;
  std::string batchKey;                     // empty: served alone
  size_t elems = 0;                         // sensor elements: the rows of its result
  size_t meshes = 0;                        // sensor meshes (its ranges in a merged result)
  size_t bytes = 0;                         // the admission estimate
  Clk::time_point submitted, started, finished;
  bool done = false;
  int32_t status = DL_OK;
  std::string err;
  std::unique_ptr<dl_result> result;
  size_t batch = 1;   // requests served by its run
};

// the device memory a request's run may take: per sensor its inputs and one chunk's buffers at the most directions
// and steps its flags allow, and a fixed part
size_t estimate(const Request& r, size_t elems) {
  if (r.tool == DL_DAYLIGHT) {
    const Options& o = r.day.opt;
    const size_t P = o.mf == 4 ? 2305 : o.mf == 2 ? 577 : 145, K = size_t(std::max(1, o.patchSamples));
    const size_t words = (2 * P * K + 8784 + 31) / 32;
    size_t per = 24 + 4 * words * (o.transmission ? 8 : 1) + 8 * (o.quantities.size() + 1);
    if (!o.seriesPath.empty()) per += 3 * 4 * 8784;
    if (o.bounces > 0) per += 4 * (P + 1);
    return elems * (o.sides == Sides::Both ? 2 : 1) * per + kFixedBytes;
  }
  const views::Options& o = r.view.opt;
  const double cell = o.cell > 0 ? o.cell : 1;
  const size_t rays = size_t(std::ceil(o.fovH / cell) * std::ceil((o.elevHi - o.elevLo) / cell)) + 1;
  return elems * (o.sides == Sides::Both ? 2 : 1) * (24 + 5 * rays + 8 * 32) + kFixedBytes;
}

}  // namespace

class Server {
 public:
  Server(int lanes, int64_t budget, int64_t batchSensors, std::string& err) : batchMax_(batchSensors > 0 ? size_t(batchSensors) : 4000000) {
    const size_t freeNow = size_t(std::max<int64_t>(0, dlg_free_bytes()));
    budget_ = budget > 0 ? size_t(budget) : freeNow / 10 * 6;   // default: 60 % of what is free now
    for (int k = 0; k < std::max(1, lanes); ++k) {
      Lane* l = Engine::get().new_lane(err);
      if (!l) { ok_ = false; break; }
      lanes_.push_back(l);
    }
    if (!ok_) return;
    for (Lane* l : lanes_) workers_.emplace_back([this, l] { work(l); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      closing_ = true;
    }
    cvWork_.notify_all();
    for (std::thread& t : workers_) t.join();
    for (Lane* l : lanes_) Engine::get().release_lane(l);
  }

  bool ok() const { return ok_; }

  int64_t submit(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t nSensors, dl_mesh* const* context,
                 int32_t nContext, std::string& err, int32_t& rc) {
    auto j = std::make_shared<Job>();
    j->submitted = Clk::now();   // the latency counts the parse (the weather file read) as dl_run's time does
    const CLocaleScope cLocale;
    rc = make_request(tool, argc, argv, sensors, nSensors, context, nContext, nullptr, false, j->req, err);
    if (rc) return -1;
    j->sensors.assign(sensors, sensors + nSensors);
    j->context.assign(context, context + nContext);
    for (int32_t i = 0; i < argc; ++i) j->argv.emplace_back(argv[i]);
    RunOptions& o = j->req.run_options();
    for (const SimulationMeshView& v : j->req.meshes)
      if (v.sensor) { j->elems += o.onVertices ? v.mesh->numPoints() : v.mesh->numFaces(); ++j->meshes; }
    j->bytes = estimate(j->req, j->elems);
    j->batchKey = batch_key(*j);
    std::lock_guard<std::mutex> lk(mu_);
    j->id = nextId_++;
    jobs_[j->id] = j;
    queue_.push_back(j);
    stats_.submitted += 1;
    stats_.peakQueue = std::max(stats_.peakQueue, double(queue_.size()));
    cvWork_.notify_one();
    return j->id;
  }

  int32_t wait(int64_t id, dl_result** out, double* times, std::string& err) {
    std::unique_lock<std::mutex> lk(mu_);
    const auto it = jobs_.find(id);
    if (it == jobs_.end()) { err = "dlx_server_wait: no request " + std::to_string(id) + " (unknown or already waited for)"; return DL_USAGE; }
    std::shared_ptr<Job> j = it->second;
    cvDone_.wait(lk, [&] { return j->done; });
    jobs_.erase(it);
    if (times) {
      times[0] = ms_between(j->submitted, j->started);
      times[1] = ms_between(j->started, j->finished);
      times[2] = double(j->batch);
    }
    if (j->status != DL_OK) { err = j->err; return j->status; }
    *out = j->result.release();
    return DL_OK;
  }

  void stats(double* out, int32_t n) {
    std::lock_guard<std::mutex> lk(mu_);
    const double s[DLX_SERVER_STATS] = {stats_.submitted, stats_.completed, stats_.failed, stats_.runs, stats_.merged, stats_.peakQueue,
                                        stats_.peakBytes, double(budget_), double(lanes_.size()), stats_.peakRunning};
    std::copy(s, s + std::min<int32_t>(n, DLX_SERVER_STATS), out);
  }

 private:
  struct Stats { double submitted = 0, completed = 0, failed = 0, runs = 0, merged = 0, peakQueue = 0, peakBytes = 0, peakRunning = 0; };

  // Requests that one run can serve: daylight, no self-occlusion (sensors cast no shade, so the scene is the context
  // alone), no file outputs, no meshes named by flags (names of merged meshes would mix), the same flags and context
  // handles, and the same centre of the run's meshes (so the recentring, and with it every coordinate, is the same).
  std::string batch_key(const Job& j) {
    if (j.req.tool != DL_DAYLIGHT) return {};
    const Options& o = j.req.day.opt;
    if (o.selfOcclude || !o.seriesPath.empty() || !o.dumpBitsPath.empty() || !o.dumpSky.empty() || !o.dumpSun.empty() || !o.noOcclude.empty())
      return {};
    for (const Optics& x : o.optics)
      if (!x.mesh.empty() && x.mesh != "*") return {};
    std::string k;
    for (const std::string& a : j.argv) k += a + '\x1f';
    for (dl_mesh* c : j.context) k += " c" + std::to_string(c->serial);
    double box[6], c[3];
    bbox_of(j.req.meshes, {}, box);
    box_centre(box, c);
    char b[80];
    std::snprintf(b, sizeof b, " @%a,%a,%a", c[0], c[1], c[2]);
    return k + b;
  }

  // the next run for a lane: the oldest admissible request and the queued ones that merge with it (lock held)
  std::vector<std::shared_ptr<Job>> take() {
    std::vector<std::shared_ptr<Job>> run;
    if (queue_.empty()) return run;
    const std::shared_ptr<Job> head = queue_.front();
    if (running_ > 0 && inflight_ + head->bytes > budget_) return run;   // the budget admits it once the device drains
    queue_.pop_front();
    run.push_back(head);
    size_t bytes = head->bytes, elems = head->elems;
    if (!head->batchKey.empty())
      for (auto it = queue_.begin(); it != queue_.end();) {
        const Job& o = **it;
        if (o.batchKey == head->batchKey && elems + o.elems <= batchMax_ && inflight_ + bytes + o.bytes - kFixedBytes <= budget_) {
          bytes += o.bytes - kFixedBytes;
          elems += o.elems;
          run.push_back(*it);
          it = queue_.erase(it);
        } else ++it;
      }
    inflight_ += bytes;
    running_ += 1;
    stats_.peakBytes = std::max(stats_.peakBytes, double(inflight_));
    stats_.peakRunning = std::max(stats_.peakRunning, double(running_));
    for (auto& j : run) j->bytes = 0;
    run.front()->bytes = bytes;   // the run's reservation, returned when it ends
    return run;
  }

  void work(Lane* lane) {
    Engine::bind(lane);
    for (;;) {
      std::vector<std::shared_ptr<Job>> run;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cvWork_.wait(lk, [&] {
          run = take();
          return !run.empty() || (closing_ && queue_.empty());
        });
        if (run.empty()) return;   // closing with nothing left
      }
      const Clk::time_point t0 = Clk::now();
      for (auto& j : run) j->started = t0;
      serve(run);
      const Clk::time_point t1 = Clk::now();
      {
        std::lock_guard<std::mutex> lk(mu_);
        inflight_ -= run.front()->bytes;
        running_ -= 1;
        stats_.runs += 1;
        stats_.merged += run.size() > 1 ? double(run.size()) : 0.0;
        for (auto& j : run) {
          j->finished = t1;
          j->batch = run.size();
          j->done = true;
          (j->status == DL_OK ? stats_.completed : stats_.failed) += 1;
        }
      }
      cvDone_.notify_all();
      cvWork_.notify_all();   // the budget freed may admit the next request
    }
  }

  // one run over the requests (one alone, or merged daylight requests): the result rows of each
  void serve(std::vector<std::shared_ptr<Job>>& run) {
    const CLocaleScope cLocale;
    Job& head = *run.front();
    std::string err;
    Result res;
    bool ok = false;
    try {
      if (run.size() == 1) {
        const Weather* w = head.req.haveWeather ? &head.req.weather : nullptr;
        ok = head.req.tool == DL_VIEWS ? views::run(head.req.meshes, head.req.view.opt, res, err) : daylight::run(head.req.meshes, w, head.req.day.opt, res, err);
      } else {
        // the merged run: every request's sensor meshes in order, then the context's; one mesh key over all the handles
        std::vector<SimulationMeshView> meshes;
        std::string key = "dl";
        for (auto& j : run)
          for (const SimulationMeshView& v : j->req.meshes)
            if (v.sensor) meshes.push_back(v);
        for (auto& j : run)
          for (dl_mesh* h : j->sensors) key += " s" + std::to_string(h->serial);
        for (const SimulationMeshView& v : head.req.meshes)
          if (!v.sensor) meshes.push_back(v);
        for (dl_mesh* h : head.context) key += " c" + std::to_string(h->serial);
        Options opt = head.req.day.opt;
        opt.sceneKey = key + " noself";
        const Weather* w = head.req.haveWeather ? &head.req.weather : nullptr;
        ok = daylight::run(meshes, w, opt, res, err);
      }
    } catch (const std::bad_alloc&) {
      err = "out of host memory";
    } catch (const std::exception& e) {
      err = std::string("internal: ") + e.what();
    }
    if (!ok) {
      for (auto& j : run) { j->status = DL_RUNTIME; j->err = err; }
      return;
    }
    if (run.size() == 1) {
      head.result = make_result(head.req, res);
      return;
    }
    const size_t C = res.columns.size();
    size_t row = 0, range = 0;
    for (auto& j : run) {
      Result part;
      part.columns = res.columns;
      part.durationHours = res.durationHours;
      part.nSky = res.nSky;
      part.nSun = res.nSun;
      part.sceneReused = res.sceneReused;
      part.stats = res.stats;
      size_t rows = 0;
      for (size_t m = 0; m < j->meshes; ++m, ++range) {
        part.ranges.push_back(res.ranges[range]);
        rows += res.ranges[range].second;
      }
      part.values.assign(res.values.begin() + std::ptrdiff_t(row * C), res.values.begin() + std::ptrdiff_t((row + rows) * C));
      row += rows;
      j->result = make_result(j->req, part);
    }
  }

  bool ok_ = true;
  size_t budget_ = 0, batchMax_ = 0, inflight_ = 0;
  int running_ = 0;
  std::vector<Lane*> lanes_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cvWork_, cvDone_;
  std::deque<std::shared_ptr<Job>> queue_;
  std::map<int64_t, std::shared_ptr<Job>> jobs_;
  int64_t nextId_ = 1;
  bool closing_ = false;
  Stats stats_;
};

}  // namespace daylight

struct dlx_server {
  std::unique_ptr<daylight::Server> impl;
};

namespace {
int32_t fail(int32_t code, const std::string& msg) { return daylight::abi_status(code, msg); }
}  // namespace

extern "C" {

dlx_server* dlx_server_open(int32_t lanes, int64_t hbm_budget, int64_t batch_sensors) {
  try {
    std::string err;
    if (lanes < 1 || lanes > 16) { fail(DL_USAGE, "dlx_server_open: 1 to 16 lanes"); return nullptr; }
    auto s = std::make_unique<dlx_server>();
    s->impl = std::make_unique<daylight::Server>(lanes, hbm_budget, batch_sensors, err);
    if (!s->impl->ok()) { fail(DL_RUNTIME, "dlx_server_open: " + err); return nullptr; }
    return s.release();
  } catch (const std::exception& e) {
    fail(DL_RUNTIME, std::string("dlx_server_open: ") + e.what());
    return nullptr;
  }
}

int64_t dlx_server_submit(dlx_server* s, int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors,
                          dl_mesh* const* context, int32_t n_context) {
  if (!s) return fail(DL_USAGE, "dlx_server_submit: no server"), -1;
  try {
    std::string err;
    int32_t rc = DL_OK;
    const int64_t id = s->impl->submit(tool, argc, argv, sensors, n_sensors, context, n_context, err, rc);
    if (id < 0) fail(rc, err);
    return id;
  } catch (const std::exception& e) {
    fail(DL_RUNTIME, std::string("dlx_server_submit: ") + e.what());
    return -1;
  }
}

int32_t dlx_server_wait(dlx_server* s, int64_t id, dl_result** out, double* times) {
  if (!s || !out) return fail(DL_USAGE, "dlx_server_wait: no server or out is NULL");
  *out = nullptr;
  std::string err;
  const int32_t rc = s->impl->wait(id, out, times, err);
  return rc ? fail(rc, err) : DL_OK;
}

int32_t dlx_server_stats(dlx_server* s, double* out, int32_t n) {
  if (!s || !out || n < 0) return fail(DL_USAGE, "dlx_server_stats: no server, a NULL array or a negative count");
  s->impl->stats(out, n);
  return DL_OK;
}

void dlx_server_close(dlx_server* s) { delete s; }

}  // extern "C"
