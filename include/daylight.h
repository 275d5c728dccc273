// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// daylight: annual daylight / solar simulation on meshes, on MI355X. The public C++ API: the declarations of reference
// src/daylight.h (the same types, members, defaults and functions, so callers of the reference compile against this
// header unchanged); the implementation behind them is this library's (HIP kernels for gfx950, see dl_device.h).
#pragma once
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <iosfwd>
#include <string>
#include <utility>
#include <vector>

namespace daylight {

// A face-vertex mesh in world metres, Z up (reference daylight.h:16-25).
struct SimulationMesh {
  std::string name;                   // OBJ group / file stem
  std::vector<double> points;         // xyz interleaved
  std::vector<int> counts;            // vertices per face
  std::vector<int> indices;           // 0-based corner indices, right-handed winding
  std::vector<uint16_t> slot;         // optics slot per face; empty: every face slot 0
  bool sensor = true;                 // false: context, blocks rays only
  size_t numFaces() const { return counts.size(); }
  size_t numPoints() const { return points.size() / 3; }
};

// A mesh as one run names it and uses it, without a copy (daylight.h:30).
struct SimulationMeshView { const SimulationMesh* mesh = nullptr; std::string name; bool sensor = true; };

// Hourly weather, W/m2 averaged over the hour ending at each row, in the file's clock (daylight.h:34-40).
struct Weather {
  double lat = 0, lon = 0, tz = 0;
  bool leap = false;
  std::vector<float> dni, dhi;
  int missing = 0;
  double shiftH = 0;
};

// One evaluation step (daylight.h:47).
struct Step { int month = 1, day = 1, jd = 1; double hour = 0, hoy = 0, dt = 1; float dni = 0, dhi = 0; };

enum class Mode { Physical, Ladybug };
enum class Quantity { Radiation, Direct, Diffuse, Irradiance, Illuminance, SunHours };
inline constexpr const char* kQuantityNames[] = {"radiation", "direct", "diffuse", "irradiance", "illuminance", "sunhours"};
enum class Sides { Front, Back, Both };

// One --optics row (daylight.h:58).
struct Optics { std::string mesh; int slot = -1; double rho = 0.2, tau = 0; bool glass = false; };

// What every sensor-ray tool takes alike (daylight.h:61-88).
struct RunOptions {
  Sides sides = Sides::Front;
  bool cullBackfaces = false;
  bool onVertices = false;
  bool selfOcclude = true;
  std::vector<std::string> noOcclude;
  double north = 0;
  double offset = 0;
  bool virtualGround = true;
  std::vector<Optics> optics;
  bool transmission = false;
  std::string sceneKey;
  int residentScenes = 1;
  const std::atomic<bool>* cancel = nullptr;
  bool verbose = false;
  // caps each per-chunk device buffer (results do not depend on it). 64 GiB here, where the reference's 256 MiB
  // (daylight.h:85) was sized for 6 GB cards: on a 288 GB MI355X a run of a million sensors with its hourly series is
  // one chunk, so each kernel is one launch over every sensor mesh of the run; admission still halves the cap down to
  // 32 MiB when memory is short.
  size_t chunkBytes = size_t(64) << 30;
  bool f32Columns = false;
  std::atomic<double>* progress = nullptr;
};

// A daylight run (daylight.h:91-106).
struct Options : RunOptions {
  Mode mode = Mode::Physical;
  std::vector<Quantity> quantities = {Quantity::Radiation};
  int mf = 1;
  int patchSamples = 1;
  double ground = 0.2;
  bool fresnel = true;
  int bounces = 0;
  int bounceSamples = 1024;
  double lat = NAN, lon = NAN, tz = NAN;
  double t0 = NAN, t1 = NAN;
  int stepMin = 60;
  double dayFrom = 0, dayTo = 1440;
  std::string seriesPath;
  std::string dumpBitsPath, dumpSky, dumpSun;
};

// Where a run's time went (daylight.h:111-114). rays: sensor x direction slots (culled ones included, as the reference
// counts them).
struct RunStats {
  double prepMs = 0, skyMs = 0, buildMs = 0, traceMs = 0, reduceMs = 0, bounceMs = 0;
  size_t rays = 0, launches = 0, gasBytes = 0;
};

struct Result {
  std::vector<double> values;                            // N x C row-major
  std::vector<float> f32;                                // f32Columns: C x N column-major
  std::vector<std::string> columns;
  std::vector<std::pair<std::string, size_t>> ranges;    // (mesh name, elements) in input order
  double durationHours = 0;
  size_t nSky = 0, nSun = 0;
  bool sceneReused = false;
  RunStats stats;
};

bool read_obj(const std::string& path, std::vector<SimulationMesh>& out, std::string& err);
bool read_dmesh(const std::string& path, std::vector<SimulationMesh>& out, std::string& err);
bool read_dmesh(const char* data, size_t size, const std::string& name, std::vector<SimulationMesh>& out, std::string& err);
bool read_epw(const std::string& path, Weather& out, std::string& err);
bool read_epw(std::istream& in, const std::string& label, Weather& out, std::string& err);
bool clear_sky_weather(double lat, double lon, double tz, double turbidity, double altitude, Weather& out, std::string& err);
bool run(const std::vector<SimulationMesh>& meshes, const Weather* weather, const Options& opt, Result& out, std::string& err);
bool run(const std::vector<SimulationMeshView>& meshes, const Weather* weather, const Options& opt, Result& out, std::string& err);

}  // namespace daylight
