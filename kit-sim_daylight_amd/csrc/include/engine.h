// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// The process's GPU as the host layer sees it: lanes and a store of resident scenes.
//
// A lane is one device context (dlg_gpu: a HIP stream, the per-run device buffers, the resident sensors and direction
// table) and the sensors it made last. Runs on one lane take turns; runs on different lanes overlap on the device. The
// primary lane serves daylight::run / views::run and the C ABI (reference src/run.cpp:30-35 keeps one Gpu for the
// process); a server adds lanes of its own.
//
// The scene store keeps built scenes in HBM under their keys (the mesh handles, the roles and the transmitting slots):
// the one installed in the primary lane and up to residentScenes - 1 parked ones (reference run.cpp:38-43, 68-156,
// 179-190). Other lanes borrow the installed scene's arrays for a run; the run's lease pins the scene, which the store
// then parks instead of freeing and never evicts until the lease ends (a forgotten pinned scene is freed then).
//
// Locks: the primary lane's (runs on it take turns; a run on another lane takes it only while it acquires its scene,
// which installs through the primary lane) before the store's (the scenes and their pins).
#pragma once
#include <atomic>
#include <condition_variable>
#include <functional>
#include <list>
#include <mutex>
#include <string>
#include <vector>

#include "core.h"

namespace daylight {

struct Lane {
  dlg_gpu* gpu = nullptr;
  // the sensors resident on the lane's context (dlg_sensors), made by the last run whose meshes, sensor options and
  // centre gave `sensorKey` (empty: none kept; cleared whenever the context's sensor arrays are freed)
  std::string sensorKey;
  std::vector<std::pair<std::string, size_t>> ranges;
  std::string sensorWarn;
};

// A run's hold on its scene (Engine::acquire on a lane other than the primary): the scene stays resident until it ends.
class SceneLease {
 public:
  SceneLease() = default;
  ~SceneLease() { release(); }
  SceneLease(const SceneLease&) = delete;
  SceneLease& operator=(const SceneLease&) = delete;
  void release();

 private:
  friend class Engine;
  uint64_t id_ = 0;
};

class Engine {
 public:
  static Engine& get();

  // the primary lane, its device context made on first use (null with err when there is no device)
  Lane* primary(std::string& err);
  // The lane this thread's runs use: the one bound by bind (a server's worker), else the primary lane.
  static void bind(Lane* l);
  Lane* run_lane(std::string& err);
  // a new lane with a context of its own borrowing the installed scene (null with err); release_lane gives it back
  Lane* new_lane(std::string& err);
  void release_lane(Lane* l);

  // The scene a run over `meshes` needs, installed in the primary lane: the installed one or a parked one under `key`,
  // else built from the meshes that block (slotted: each triangle's global slot kept with the leaves); fresh(scene) is
  // called on a fresh build before any other run can see it. On a lane other than the primary the lane then borrows it
  // under lease and *gpu is that lane's context. gpu null: a run without rays (nothing is installed). mark is called at
  // the build's stage boundaries.
  bool acquire(Lane* lane, const std::vector<SimulationMeshView>& meshes, const std::function<bool(size_t)>& blocks, bool slotted,
               const std::string& key, int residentScenes, const std::atomic<bool>* cancel, dlg_gpu** gpu, Scene& s, Result& out, const char*& stage,
               std::string& err, const std::function<void(const char*)>& mark, const std::function<void(Scene&)>& fresh, SceneLease* lease);
  // Room on the device for a run of nSens sensors beside the scene: per sensor 24 B resident plus chunkSensor bytes per
  // chunk of chunkBytes / perSensor sensors, plus fixedBytes. Frees the lane's last buffers and parked scenes as needed
  // and halves chunkBytes down to 32 MiB; false with err when even that does not fit.
  bool admit(Lane& lane, size_t nSens, size_t perSensor, size_t chunkSensor, size_t fixedBytes, int residentScenes, size_t& chunkBytes,
             std::string& err);
  void forget(const std::string& key);
  // the run lock of the primary lane (the C ABI's runs take turns on it)
  std::mutex& primary_lock() { return primaryMu_; }

 private:
  friend class SceneLease;
  Engine() = default;
  bool acquire_locked(const std::vector<SimulationMeshView>& meshes, const std::function<bool(size_t)>& blocks, bool slotted, const std::string& key,
                      int residentScenes, const std::atomic<bool>* cancel, bool install, Scene& s, Result& out, const char*& stage, std::string& err,
                      const std::function<void(const char*)>& mark, const std::function<void(Scene&)>& fresh);
  void park_or_free_installed(int residentScenes);
  bool evict_one();   // frees the least recently used parked scene that no lease pins (false: none)
  bool pinned(uint64_t id) const;
  void unpin(uint64_t id);
  bool room(size_t need) const;

  Lane primary_;
  Scene installed_;
  std::list<Scene> parked_;   // most recently used first
  std::vector<std::pair<uint64_t, int>> pins_;            // scene id -> leases
  std::vector<std::pair<uint64_t, dlg_scene*>> forgotten_;   // pinned scenes forgotten: freed when their last lease ends
  uint64_t nextId_ = 1;
  std::mutex primaryMu_;
  std::mutex storeMu_;        // installed_, parked_, pins_, forgotten_
  std::mutex lanesMu_;        // the lanes' bookkeeping below
  std::vector<Lane*> lanes_;
};

// The sensors of a run on a lane, made on the GPU into its context's resident sensor arrays (dlg_sensors) unless they
// are there from its last run (the same meshes - the mesh key names the handles - sensor options and centre); the
// run's ranges into out, the meshes' notes (degenerate faces, zero vertex normals) into warn.
bool lane_sensors(Lane& lane, const std::vector<SimulationMeshView>& meshes, const RunOptions& opt, const double c[3], size_t nSens, Result& out,
                  std::string& warn, std::string& err, bool ladybug = false);

}  // namespace daylight
