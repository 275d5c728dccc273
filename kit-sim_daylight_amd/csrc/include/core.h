// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Internal contracts of the host layer (the reference's src/internal.h, minus its OptiX seam, which dl_device.h
// replaces). The host prepares inputs (steps, patches, directions, triangles, sensors) and orchestrates the device;
// every per-ray, per-sensor or per-(step, patch) loop of the hot path runs on the GPU.
#pragma once
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>   // float3 / uint3

#include "daylight.h"
#include "dl_device.h"
#include "views.h"

namespace daylight {

// ---------------------------------------------------------------- sky (reference sky.cpp:240-294)
struct Patches {
  int P = 0;
  float alpha = 0;
  std::vector<float> alt, azi, omega;
  std::vector<int> row, ninrow;
  std::vector<float3> centre, lb;
};
Patches make_patches(int mf);
float3 patch_sample(const Patches& pa, int p, int k, int K);
const double* tcos_table();     // [1025] cos(k pi / 2048), Radiance tcos.c's table
// Ladybug mode's sky (ladybug.cpp; reference run.cpp:362-470): per patch the -d and -s gendaymtx -A averages printed
// %.3g as broadband x omega x hours / 1000 (dir, dif, their sum rad), the ground's gnd = ground x mean rad; dump the
// two averaged blocks (-s then -d, ground row first, rgb) for --dump-sky. S / gS (may be null): the series columns
// (bright(-s) + -d per patch, the ground's reflectance x the omega-weighted mean over P).
struct LadybugSky {
  std::vector<double> dir, dif, rad;
  double gnd = 0;
  std::vector<float> dump;
};
void ladybug_sky(const Patches& pa, const std::vector<Step>& steps, double lat, double lon, double fileTz, double grefl, double durationHours,
                 const std::vector<int>& colOf, float* S, float* gS, int Tup, LadybugSky& out);
// Radiance sun.c at a site (degrees; longitude and meridian west-positive): altitude and azimuth in radians
void sun_radiance(double latDeg, double lonWestDeg, double meridianWestDeg, int jd, double hour, double& alt, double& azi);
double round_fmt(double x, const char* fmt);
void gendaymtx_rgb(const Patches& pa, const double* tsinP, int jd, double alt, double azi, double dni, double dhi, double grefl, bool sun, bool sky,
                   float* out);
std::vector<double> patch_table_sine(const Patches& pa);

// ---------------------------------------------------------------- solar (reference solar.cpp)
bool make_steps(const Options& opt, const Weather* w, std::vector<Step>& out, std::string& err, double shift = 0);
bool parse_hoy(const std::string& s, bool leap, double& hoy, std::string& err);
void hoy_to_date(double hoy, bool leap, int& month, int& day, double& hour);
double extraterrestrial_irradiance(int doy);
void ineichen(double apparentZenithDeg, double turbidity, double altitudeM, double dniExtra, double& ghi, double& dni, double& dhi);

// ---------------------------------------------------------------- meshes (reference mesh.cpp)
struct Tris { std::vector<float3> v; std::vector<uint3> i; };
inline std::vector<SimulationMeshView> views_of(const std::vector<SimulationMesh>& meshes) {
  std::vector<SimulationMeshView> v;
  v.reserve(meshes.size());
  for (const SimulationMesh& m : meshes) v.push_back({&m, m.name, m.sensor});
  return v;
}
void par_for(size_t n, const std::function<void(size_t)>& f);
// The bounding box (lo xyz, hi xyz) of the points of the meshes k with which(k) (all when which is empty); empty: lo
// +inf, hi -inf. The scene and its sensors are recentred on the centre of the box of every mesh of the run (reference
// run.cpp:68-156), midway per axis (0 on an empty axis).
void bbox_of(const std::vector<SimulationMeshView>& meshes, const std::function<bool(size_t)>& which, double box[6]);
void box_union(double box[6], const double other[6]);
void box_centre(const double box[6], double c[3]);
size_t tri_cap(const SimulationMesh& m);
// m's triangles (points recentred by c, indices from base) into v / tri; with slots non-null, each triangle's global slot
// (slotBase + its face's slot) into slots. Returns the triangles made (degenerate n-gons skipped and counted).
size_t triangulate_into(const SimulationMesh& m, const double c[3], float3* v, uint3* tri, uint32_t* slots, unsigned base, uint32_t slotBase,
                        int& skipped);
// The faces of 5 or more corners of m, ear-clipped as triangulate_into does: face[k] their indices ascending, made[k]
// the triangles each made (0: skipped, counted in skipped), tri[3 ..] the corners (mesh-local) in face order. What the
// GPU ingest (dlg_mesh_add) takes beside the faces it fans itself.
void ngon_faces(const SimulationMesh& m, std::vector<int64_t>& face, std::vector<int32_t>& made, std::vector<uint32_t>& tri, int& skipped);
size_t slot_count(const SimulationMesh& m);
bool mesh_ok(const SimulationMesh& m, const std::string& label, std::string& err);
bool optics_ok(const Options& opt, std::string& err);
bool optics_rows_ok(const std::vector<Optics>& rows, std::string& err);
// per global slot its optics once resolved by precedence NAME/SLOT > NAME > *: the device's record (dl_device.h dlg_slot):
// rho, tau, tn (glass: tn_from_Tn(tau)), flags (DLG_SLOT_GLASS; a run adds DLG_SLOT_THIN / INWARD / TWIN)
struct SlotOptics { float rho = 0.2f, tau = 0.f, tn = 0.f; uint32_t flags = 0; };
static_assert(sizeof(SlotOptics) == sizeof(dlg_slot), "a slot uploads as the device's record");
// Faces of slot s of m (recentred by c) as one closed thin solid (reference mesh.cpp:275-304): DLG_SLOT_THIN when the
// corners welded on a 1e-5 m grid close it (every directed edge matched by its reverse), 2 |V| / S < 0.1 m, with
// DLG_SLOT_INWARD when its signed volume is negative or DLG_SLOT_TWIN instead when 2 |V| / S is under 1e-5 m (coincident
// face pairs: a double-faced sheet); else 0.
uint32_t thin_solid(const SimulationMesh& m, uint32_t s, const double c[3]);
bool resolve_optics(const std::vector<SimulationMeshView>& meshes, const std::vector<Optics>& rows, std::vector<SlotOptics>& out, std::string& err);
bool resolve_no_occlude(const std::vector<SimulationMeshView>& meshes, const std::vector<std::string>& names, std::vector<char>& open, std::string& err);
bool resolve_targets(const std::vector<SimulationMeshView>& meshes, const std::vector<views::Target>& targets, bool selfOcclude,
                     const std::vector<char>& open, std::vector<std::string>& groups, std::vector<uint8_t>& cls, std::string& err);

// ---------------------------------------------------------------- scene residency and runs (reference run.cpp:38-213)
bool npy_write_header(std::FILE* f, size_t rows, size_t cols, const char* descr);
struct SceneMesh { std::string name; size_t tri0 = 0, nTri = 0; const SimulationMesh* mesh = nullptr; };
struct Scene {
  std::string key;
  uint64_t id = 0;   // the store's number of the build (keys can repeat after a forget; ids do not)
  double c[3] = {0, 0, 0};
  double box[6] = {0, 0, 0, 0, 0, 0};   // the bounding box of the meshes it was built from (the blocking ones)
  size_t nTris = 0;
  int skipped = 0;
  std::vector<SceneMesh> order;
  std::vector<uint32_t> thin;   // per global slot: thin_solid of a transmitting slot (a slotted build's)
  dlg_scene* parked = nullptr;
};
std::string open_key(const std::vector<char>& open, bool selfOcclude);
// The scene key of a run from its mesh key (RunOptions::sceneKey, " s<serial>" per sensor handle and " c<serial>" per
// context handle): without self-occlusion the sensor meshes are not in the scene, so their handles leave the key and
// runs over different sensors share the context's scene.
std::string scene_key_of(const std::string& meshKey, bool selfOcclude);
double ms_since(const void* clockStart);
// daylight::run with the --timeseries output kept on the device: per chunk of sensors, sink(i0, nc, T, E) with E the
// device array float [nc x T] (null when the run has no series columns); an empty sink drops the series.
using SeriesSink = std::function<bool(size_t i0, size_t nc, size_t T, const float* deviceE, std::string& err)>;
bool run_series_device(const std::vector<SimulationMeshView>& meshes, const Weather* w, const Options& opt, const SeriesSink& sink, Result& out,
                       std::string& err);

}  // namespace daylight
