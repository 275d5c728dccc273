// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// One tool invocation of the C ABI, parsed and checked before it meets the GPU: the tool's flags (the CLI's own
// grammar, cli.h), the weather and time rules, the mesh handles as the run's meshes and the scene key they make.
// dl_run and the serving loop share it.
#pragma once
#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "cli.h"
#include "core.h"
#include "daylight_c.h"

struct dl_mesh {
  uint64_t serial = 0;                         // never reused: the identity of the meshes in a scene key
  std::vector<daylight::SimulationMesh> meshes;
};
struct dl_control {
  std::atomic<bool> cancel{false};
  std::atomic<double> progress{0};
};
struct dl_result {
  int64_t rows = 0, cols = 0;
  std::string manifest;
  std::vector<double> values;
  double stats[DL_STATS] = {};
};

namespace daylight {

struct Request {
  int32_t tool = DL_DAYLIGHT;
  std::vector<SimulationMeshView> meshes;   // the sensor handles' meshes, then the context handles'
  cli::Args day;                             // DL_DAYLIGHT
  views::cli::Args view;                     // DL_VIEWS
  Weather weather;
  bool haveWeather = false;
  bool verbose() const { return tool == DL_DAYLIGHT ? day.opt.verbose : view.opt.verbose; }
  RunOptions& run_options() { return tool == DL_DAYLIGHT ? static_cast<RunOptions&>(day.opt) : static_cast<RunOptions&>(view.opt); }
};

// Parses and checks a tool's flags over the handles into r: status DL_OK, or DL_USAGE / DL_RUNTIME with err (the
// messages of reference src/c_api.cpp). The run options get the scene key of the handles, the control's cancel flag
// and progress, and the chunk cap set by dlx_chunk_bytes.
int32_t make_request(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t nSensors, dl_mesh* const* context,
                     int32_t nContext, dl_control* control, bool seriesSink, Request& r, std::string& err);

// The result handle of a finished run: values, the tool's manifest and the statistics row (reference c_api.cpp's
// finish).
std::unique_ptr<dl_result> make_result(const Request& r, Result& res);

// records msg as the calling thread's dl_last_error() and returns code
int32_t abi_status(int32_t code, const std::string& msg);

// The "C" locale for the calling thread while it lives: the flags' and the weather file's numbers and the manifest
// do not depend on the caller's locale.
class CLocaleScope {
 public:
  CLocaleScope();
  ~CLocaleScope();
  CLocaleScope(const CLocaleScope&) = delete;
  CLocaleScope& operator=(const CLocaleScope&) = delete;

 private:
  void* c_ = nullptr;
  void* old_ = nullptr;
};

}  // namespace daylight
