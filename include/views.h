// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// views: view analysis on the daylight library, on MI355X. The declarations of reference src/views.h (same types,
// defaults and functions); the rays and reductions run as HIP kernels (dl_device.h).
#pragma once
#include "daylight.h"

namespace daylight::views {

// Rays ending on mesh `mesh` count for `group` (views.h:13).
struct Target { std::string mesh, group; };

enum Dist : unsigned { DistMean = 1, DistMin = 2, DistGroup = 4, DistIsovist = 8, DistAuto = ~0u };

struct Options : RunOptions {
  double fovH = 120, elevLo = -70, elevHi = 50, cell = 2;
  bool jitter = true;
  std::vector<Target> targets;
  unsigned distance = DistAuto;
  double farDist = 1000;
  double groundZ = NAN;
  std::vector<double> reach;
  std::string dumpPath;
};

// The grid and resolved options of a run (views.h:40).
struct Field { unsigned nAz = 0, nEl = 0, dist = 0; bool flat = false; double groundZ = NAN; };
bool field(const std::vector<SimulationMeshView>& meshes, const Options& opt, Field& f, std::string& err);

bool run(const std::vector<SimulationMeshView>& meshes, const Options& opt, Result& out, std::string& err);
bool run(const std::vector<SimulationMesh>& meshes, const Options& opt, Result& out, std::string& err);

}  // namespace daylight::views
