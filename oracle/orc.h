/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * Oracle: a plain-C, FP64-where-the-reference-is restatement of daylightcore's trace-and-accumulate path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (kit-sim_daylight_amd/) includes, links or calls this code; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker and as the reported CPU
 * baseline. Every function names the reference file:line it restates.
 *
 * Parity pins (what fixes each part independently of this code):
 *   sky     - tests/check_sky.cpp constants (patch counts, sum omega = 2 pi, Ladybug altitudes, Perez ratios, AddDirect
 *             energy, -O0 efficacy) and, where the reference sources build here, oracle/_ref (sky.cpp) bit for bit.
 *   solar   - tests/check_solar.cpp: 35 Ladybug sun goldens, 8 pvlib Ineichen goldens, Spencer, Boston clear-sky sums,
 *             step rules; oracle/_ref (solar.cpp) bit for bit.
 *   mesh    - tests/check_mesh.cpp closed forms (Newell normals, centroids, sides/offset, recentring).
 *   trace   - tests/check_gpu.cpp slab-test box scene, km-sized own planes, the 1 cm curb, the L foot; end to end the
 *             printed outputs of docs/daylight.md:133-150 and :163-180.
 *   views   - tests/check_views.cpp closed forms (kSky, far, isovist polygon/sphere) and docs/views.md:333-350.
 * Triangle intersection is this library's own watertight test (Woop, Benthin, Wald 2013), not OptiX's hardware test
 * (not reproducible off an NVIDIA RT core); it is pinned by the cases above, not by OptiX bit for bit.
 */
#ifndef ORC_H
#define ORC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- sky (reference src/sky.cpp) */
int orc_patch_count(int mf);
/* make_patches (sky.cpp:240-270): per patch alt, azi, omega (float), row, ninrow, centre[3], lb[3]; returns alpha. */
float orc_patches(int mf, float* alt, float* azi, float* omega, int* row, int* ninrow, float* centre, float* lb);
/* patch_sample (sky.cpp:272-283) */
void orc_patch_sample(int mf, int p, int k, int K, float out[3]);
double orc_tcos(double x);               /* tcos.c table cosine (sky.cpp:202-227) */
double orc_tsin(double x);
double orc_round_fmt(double x, const char* fmt);   /* sky.cpp:488-493 */
/* gendaymtx_step (sky.cpp:297-474): out[3 (P+1)], [0] ground row, [1..P] patches. */
void orc_gendaymtx_step(int mf, int jd, double alt, double azi, double dni, double dhi, double grefl, int sun, int sky,
                        int visible, float* out);
void orc_sun_radiance(double latDeg, double lonWestDeg, double meridianWestDeg, int jd, double hour, double* alt, double* azi);
void orc_dir_from_altaz(double altRad, double aziRad, float out[3]);
void orc_rotate_z(const float v[3], double deg, float out[3]);

/* ---------------------------------------------------------------- solar (reference src/solar.cpp) */
/* sun_noaa (solar.cpp:66-105): degrees, azimuth clockwise from north; vec towards the sun (true north). */
void orc_sun_noaa(double lat, double lon, double tz, int month, int day, double hourFrac, int leap, double* alt, double* azi,
                  int* up, float vec[3]);
double orc_extraterrestrial(int doy);                                      /* solar.cpp:158-161 */
void orc_ineichen(double z, double TL, double altitude, double dniExtra, double* ghi, double* dni, double* dhi);   /* 163-177 */
int orc_clear_sky(double lat, double lon, double tz, double TL, double altitude, float* dni, float* dhi);     /* 179-194 */
void orc_hoy_to_date(double hoy, int leap, int* month, int* day, double* hour);                              /* 196-202 */
int orc_parse_hoy(const char* s, int leap, double* hoy);                                                      /* 204-215 */

typedef struct { int month, day, jd; double hour, hoy, dt; float dni, dhi; } orc_step;   /* daylight.h:47 Step */
typedef struct {
  double t0, t1;            /* NaN: annual */
  int stepMin;
  double dayFrom, dayTo;
} orc_time;
/* make_steps (solar.cpp:217-259). Weather may be null (dni/dhi 0). Returns the count written (<= cap), -1 on error. */
long orc_make_steps(const orc_time* t, const float* dni, const float* dhi, long nRows, int leap, double shift, orc_step* out, long cap);

/* ---------------------------------------------------------------- meshes (reference src/mesh.cpp) */
typedef struct {
  const double* points; long nPoints;
  const int* counts; long nFaces;
  const int* indices;
  int sensor;                /* SimulationMesh::sensor (as the run names it) */
} orc_mesh;
void orc_bbox_centre(const orc_mesh* m, int nMesh, double c[3]);           /* mesh.cpp:107-118 */
/* make_sensors_into (mesh.cpp:315-349), physical rules; sides 0 front, 1 back, 2 both. Returns elements. */
long orc_sensors(const orc_mesh* m, int onVertices, int sides, double offset, const double c[3], float* p, float* n);
long orc_sensors_mode(const orc_mesh* m, int onVertices, int sides, double offset, int ladybug, const double c[3], float* p, float* n);
/* triangulate_into (mesh.cpp:137-168) for faces of 3 or 4 corners (n-gons: skipped and counted). Returns triangles. */
long orc_triangulate(const orc_mesh* m, const double c[3], float* v, uint32_t* tri, uint32_t base, int* skipped);

/* ---------------------------------------------------------------- trace (reference src/trace_dev.cu, src/params.h) */
typedef struct orc_scene orc_scene;
/* Triangles V[3 nv], I[3 nt]; brute force below a few hundred triangles, else a host BVH (same hit sets). */
orc_scene* orc_scene_new(const float* V, long nv, const uint32_t* I, long nt);
void orc_scene_free(orc_scene* s);
/* __raygen__bits semantics (trace_dev.cu:91-108): bits[w nc + i], word-major; cull = back faces culled. */
void orc_trace_bits(const orc_scene* s, const float* P, const float* N, long nc, const float* D, int nDir, int cull,
                    uint32_t* bits, int nthreads);
/* __raygen__bytes + __anyhit__tr semantics (trace_dev.cu:115-165): the transmittance byte of every (sensor, direction),
 * direction-major bytes[d nc + i]; triSlot: the global slot of each triangle (input order; NULL: all slot 0), slots
 * [nSlots][4] = rho, tau, tn, flags (as bits: 1 glass, 2 thin, 4 inward). The any-hit calls in the order of the sorted
 * hits, the transmittance a product in double. */
void orc_trace_bytes(const orc_scene* s, const float* P, const float* N, long nc, const float* D, int nDir, int cull, const uint32_t* triSlot,
                     const float* slots, long nSlots, uint8_t* bytes, int nthreads);
/* glass.h: tn_from_Tn, and glass_TR's transmittance and reflectance (in double, rounded to float) */
double orc_tn_from_Tn(double Tn);
float orc_glass_T(double c, double tn);
void orc_glass_TR(double c, double tn, float* T, float* R);
/* __raygen__bounce semantics (trace_dev.cu:260-311): reflected light of sensors i0 .. i0 + nc - 1 (global indices seed
 * the paths), M paths each with up to `bounces` diffuse reflections, rows dc[b nc + i] (bin 0 the ground, 1 + p the
 * Reinhart patches of sky density mf), slots as orc_trace_bytes; cull: back faces culled. */
void orc_bounce(const orc_scene* s, const float* P, const float* N, long nc, unsigned i0, const uint32_t* triSlot, const float* slots, long nSlots,
                unsigned bounces, unsigned M, int mf, double north, int cull, float* dc, int nthreads);
/* One ray as the bits pass classifies it: 1 visible, 0 blocked or culled. */
int orc_visible(const orc_scene* s, const float p[3], const float n[3], const float d[3], int cull);
/* The same against every triangle (no BVH): the hit set the traversal must reproduce. */
int orc_visible_brute(const float* V, const uint32_t* I, long nt, const float p[3], const float n[3], const float d[3], int cull);
/* The LBVH as built: nodes [n][16] (left box lo/hi, right box lo/hi, left, right child as bits; child >= 2^31 a leaf),
 * leaf triangles [nt][12] (v0, v1, v2, original index as bits), original index per leaf. Returns the node count. */
long orc_scene_nodes(const orc_scene* s, const float** node, const float** tri, const uint32_t** prim);
float orc_scene_pad(const orc_scene* s);
/* The binned-SAH tree of the MI355X build (sah.hip) replacing the scene's LBVH, leaves reordered: returns its height.
 * No reference counterpart (OptiX builds its own GAS, gpu.cu:386-457); traversals give the same results over either. */
int orc_scene_sah(orc_scene* s);
/* inner nodes on the longest root-to-leaf path (the LBVH's k_refit height) */
int orc_scene_height(const orc_scene* s);
/* Morton keys (21 bits per axis of the centroid in the centroid box) and Karras 2012 children of sorted keys. */
void orc_lbvh_keys(const float* V, const uint32_t* I, long nt, uint64_t* keys);
void orc_lbvh_topology(const uint64_t* sortedKeys, long n, uint32_t* left, uint32_t* right);
/* offset_ray (trace_dev.cu:16-25) */
void orc_offset_ray(const float p[3], const float n[3], float out[3]);
/* pcg3d / u01 / view_cell (params.h:64-89) */
void orc_pcg3d(uint32_t x, uint32_t y, uint32_t z, uint32_t out[3]);

typedef struct {
  unsigned nAz, nEl;
  float az0, dAz;
  const float* rowSin;      /* [2 nEl] bottom, top */
  int world, cull, jitter, back, vground;
  unsigned nS;
  float groundZ;
  float cosN, sinN;
  const uint8_t* label;     /* per triangle, null: every hit an obstruction */
} orc_view;
/* __raygen__view (trace_dev.cu:177-232): cls[d nc + i], dist[d nc + i]; i0 = the chunk's first global sensor. */
void orc_trace_view(const orc_scene* s, const float* P, const float* N, long nc, unsigned i0, const orc_view* v, int cull,
                    uint8_t* cls, float* dist, int nthreads);
void orc_view_cell(unsigned d, unsigned col, float rsLo, float rsHi, float az0, float dAz, int jitter, unsigned element,
                   unsigned side, float out[3]);

/* ---------------------------------------------------------------- accumulation (reference src/gpu.cu) */
/* k_reduce<Bits> (gpu.cu:99-114): out[i] = sum over set bits (ascending d) of double(float(max(0, dotf(n, D)))) * W[d]. */
void orc_reduce(const uint32_t* bits, long nc, int nDir, const float* N, const float* D, const float* W, int useCos, double* out);
/* The FP64 contraction: E[i T + t] = sum_{visible sky d} double(c_d) S[row_d, t] + double(G_i) g_t + sun, with
 * c_d = float(float(max(0, dotf(n, D_d))) * wsky[d]) as k_series (gpu.cu:119-149) forms it; all sums in double. */
void orc_series64(const uint32_t* bits, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                  const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                  const float* sunVec, const float* sunW, double* E, int nthreads);
/* k_series in FP32 (what the reference's CUDA kernel computes, its own rounding), for reference only. */
void orc_series32(const uint32_t* bits, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                  const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                  const float* sunVec, const float* sunW, float* E);
/* The same three over byte mode's direction-major bytes[d nc + i] (k_reduce<Bytes> / k_series<Bytes>, gpu.cu:83-92):
 * each product's first factor times byte / 255.f. */
void orc_reduce_bytes(const uint8_t* bytes, long nc, int nDir, const float* N, const float* D, const float* W, int useCos, double* out);
void orc_series64_bytes(const uint8_t* bytes, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                        const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                        const float* sunVec, const float* sunW, double* E, int nthreads);
void orc_series32_bytes(const uint8_t* bytes, long nc, const float* N, const float* D, unsigned nSky, const int* row, const float* wsky,
                        const float* S, int T, const float* g, const float* G, unsigned sun0, unsigned nSun, const int* sunStep,
                        const float* sunVec, const float* sunW, float* E);
/* k_series_dc (gpu.cu:153-169): E += sum_b dc(b, i) Sdc[b, t], E32 as the reference's FP32 kernel, E64 in double */
void orc_series_dc(const float* dc, long nc, unsigned nBins, const float* Sdc, int T, float* E32, double* E64);
/* k_view + k_iso (gpu.cu:171-274): out[i NA + a]. isoDim 0/2/3. */
void orc_reduce_view(const uint8_t* cls, const float* dist, long nc, unsigned K, unsigned nAz, const double* wRow, unsigned nClass,
                     double farDist, const float* reach, unsigned nReach, int glazing, unsigned isoDim, const orc_view* v,
                     unsigned i0, double* out);

/* ---------------------------------------------------------------- runs (reference src/run.cpp, src/views.cpp) */
enum { ORC_RADIATION, ORC_DIRECT, ORC_DIFFUSE, ORC_IRRADIANCE, ORC_ILLUMINANCE, ORC_SUNHOURS };
typedef struct {
  int quantities[6]; int nQ;
  int mf, patchSamples;
  double ground, north, offset;
  int sides, onVertices, selfOcclude, cullBackfaces, virtualGround;
  double lat, lon, tz;      /* NaN: from the weather */
  orc_time time;
  int wantSeries;           /* also the FP64 series (Tup columns) */
  int nthreads;
  int ladybug;              /* --mode ladybug */
} orc_daylight_opt;
typedef struct {
  long nElem, nSens, nCols;
  double* values;           /* [nElem x nCols], caller-sized: orc_daylight_size */
  double durationHours;
  long nSky, nSun, nDir;
  int Tup;
  double* series;           /* [nSens x Tup] when wantSeries, caller-sized */
} orc_daylight_out;
/* Sizes of a run before it runs: elements, columns, Tup, directions. 0 ok. */
int orc_daylight_size(const orc_mesh* m, int nMesh, const float* dni, const float* dhi, long nRows, double wlat, double wlon,
                      double wtz, int leap, const orc_daylight_opt* o, long* nElem, long* nCols, int* Tup, long* nDir);
/* Ladybug mode's sky (run.cpp:362-470) over T steps: dir / dif / rad [P], *gnd, dump [2 x 3 (P + 1)] (-s then -d
 * averages, %.3g), and when S: the series columns S [P x Tup], gS [Tup] of the steps with colOf >= 0. */
void orc_ladybug_sky(int mf, long T, const int* jd, const double* hour, const float* dni, const float* dhi, double lat, double lon,
                     double fileTz, double grefl, double dur, const int* colOf, int Tup, float* S, float* gS, double* dir, double* dif,
                     double* rad, double* gnd, float* dump);
/* run_daylight (run.cpp:216-738), physical or ladybug mode, opaque occluders. Weather arrays may be null (sun hours only). 0 ok. */
int orc_daylight(const orc_mesh* m, int nMesh, const float* dni, const float* dhi, long nRows, double wlat, double wlon,
                 double wtz, int leap, const orc_daylight_opt* o, orc_daylight_out* out, char* err, int errLen);

#ifdef __cplusplus
}
#endif
#endif
