/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * daylight_c.h - the C ABI of the MI355X build: the daylight and views tools in-process, with the same symbols, types,
 * constants and semantics as reference src/daylight_c.h (what bindings/python, bindings/csharp and bindings/ts bind), so
 * those bindings load libdaylight_amd.so in place of daylight_c.dll unchanged.
 *
 * A run takes the CLI's own flags (cli.h) and meshes as handles. Status codes are the CLI's exit codes (DL_OK, DL_USAGE,
 * DL_RUNTIME); after a failure dl_last_error() holds the message (per thread). No exception crosses the ABI. Any thread
 * may call; runs on the GPU take turns.
 */
#ifndef DAYLIGHT_C_H
#define DAYLIGHT_C_H

#include <stdint.h>

#define DL_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

#define DL_ABI_VERSION 1
#define DL_OK 0
#define DL_USAGE 1
#define DL_RUNTIME 2
#define DL_DAYLIGHT 0
#define DL_VIEWS 1
#define DL_STATS 10

typedef struct dl_mesh dl_mesh;
typedef struct dl_control dl_control;
typedef struct dl_result dl_result;

/* daylight_c.h:54 */
DL_API int32_t dl_version(void);
/* daylight_c.h:57: the last failed call's message on this thread ("" if none) */
DL_API const char* dl_last_error(void);
/* daylight_c.h:61: the tool's usage text; NULL for an unknown tool */
DL_API const char* dl_usage(int32_t tool);

/* daylight_c.h:72-73: a mesh from arrays, copied (the .dmesh arrays); NULL on error */
DL_API dl_mesh* dl_mesh_create(const char* name, const double* xyz, int64_t n_points, const int32_t* counts, int64_t n_faces,
                               const int32_t* indices, int64_t n_indices, const uint16_t* slots);
/* daylight_c.h:77: the meshes of an .obj or .dmesh file */
DL_API dl_mesh* dl_mesh_load(const char* path);
/* daylight_c.h:80 */
DL_API int32_t dl_mesh_count(const dl_mesh* mesh);
/* daylight_c.h:83 */
DL_API const char* dl_mesh_name(const dl_mesh* mesh, int32_t k);
/* daylight_c.h:87 */
DL_API void dl_mesh_free(dl_mesh* mesh);

/* daylight_c.h:92-104: cancel flag and progress of one run */
DL_API dl_control* dl_control_create(void);
DL_API void dl_control_cancel(dl_control* control);
DL_API double dl_control_progress(const dl_control* control);
DL_API void dl_control_free(dl_control* control);

/* daylight_c.h:119-120: one run of a tool with its CLI flags; sensors then context handles; *out a new result on DL_OK */
DL_API int32_t dl_run(int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors,
                      dl_mesh* const* context, int32_t n_context, dl_control* control, dl_result** out);

/* daylight_c.h:125-146: the result */
DL_API int64_t dl_result_rows(const dl_result* result);
DL_API int64_t dl_result_cols(const dl_result* result);
DL_API const char* dl_result_manifest(const dl_result* result);
DL_API int32_t dl_result_values(const dl_result* result, double* out, int64_t n);
DL_API int32_t dl_result_values_f32(const dl_result* result, float* out, int64_t n);
DL_API int32_t dl_result_stats(const dl_result* result, double* out);
DL_API void dl_result_free(dl_result* result);

/* ---- Extensions: no counterpart in the reference's daylight_c.h. For callers that consume the annual series on the
 * GPU (a serving loop, a benchmark) and for monitoring; the calls above are untouched by them. */

/* The result's rows x cols values (row-major doubles) where the result holds them, valid until dl_result_free: for
 * bindings that view them in place instead of copying them out with dl_result_values (NULL for a NULL result). */
DL_API const double* dlx_result_data(const dl_result* result);

/* A chunk of the series: rows [row0, row0 + rows) x cols float32 values, row-major, in device memory, valid until the
 * callback returns (NULL when the run has no series columns). Return DL_OK to go on; anything else stops the run. */
typedef int32_t (*dlx_series_sink)(void* user, int64_t row0, int64_t rows, int64_t cols, const float* device);
/* dl_run(DL_DAYLIGHT, ...) with the --timeseries output kept on the device and handed to sink chunk by chunk instead of
 * being written to the --timeseries file (which is not opened). sink NULL: the series is computed and dropped. */
DL_API int32_t dlx_run_series_device(int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors, dl_mesh* const* context,
                                     int32_t n_context, dl_control* control, dlx_series_sink sink, void* user, dl_result** out);
/* The process GPU's counters since it was made: [0] trace ms (HIP events on the library's stream), [1] trace launches,
 * [2] ray slots, [3] traversal node visits, [4] triangle tests, [5] ray slots counted (the counts come from the counting
 * variant, dlx_count_fetches), [6] reduce ms, [7] series ms, [8] sky ms, [9] build ms, [10] views ms, [11] last trace
 * launch ms, [12] last series ms, [13] resident scene bytes, [14] tree height, [15] traversal iterations executed by
 * the waves (counting variant), [16] rays traced (the ray slots past the hemisphere cull), [17] chunk graphs launched,
 * [18] chunk graphs captured, [19] triangle-test iterations executed by the waves (counting variant). Writes
 * min(n, DLX_COUNTERS) values. */
#define DLX_COUNTERS 20
DL_API int32_t dlx_gpu_counters(double* out, int32_t n);
/* on != 0: the traversal runs its counting variant (slower) and fills counters [3]-[5]. */
DL_API void dlx_count_fetches(int32_t on);
/* bytes > 0: the per-chunk device buffer cap (the C++ Options::chunkBytes) of every following dl_run; 0: the default.
 * Results do not depend on it; tests use it to run one case in many chunks. */
DL_API void dlx_chunk_bytes(int64_t bytes);
/* The weather file a run reads with --epw, as it reads it (no GPU involved): *rows data rows (8760 or 8784), their
 * DNI / DHI in W/m2 (missing values 0; the first min(rows, cap) written where the arrays are not NULL), and
 * latLonTz[3] = latitude, longitude, time zone from the LOCATION line when not NULL. DL_RUNTIME with the message on a
 * file it cannot use. */
DL_API int32_t dlx_read_epw(const char* path, float* dni, float* dhi, int64_t cap, int64_t* rows, double* latLonTz);

/* Host checks of the scene's optics (no GPU involved; the reference's tests call its internal functions the same way):
 * dlx_thin_solid: the thin-solid flags of slot `slot` of a one-mesh handle (0, or DLG_SLOT_THIN | DLG_SLOT_INWARD /
 * DLG_SLOT_TWIN of dl_device.h: 2 | 4 / 8) with the mesh recentred by centre[3] (NULL: the origin), reference
 * mesh.cpp:275-304. dlx_triangle_slots: the global slot (slotBase + the face's slot) of every triangle the mesh draws,
 * in build order, reference mesh.cpp:137-168 (Tris::slot); *n triangles, the first min(n, cap) written. */
DL_API int32_t dlx_thin_solid(const dl_mesh* mesh, int32_t slot, const double* centre, uint32_t* flags);
DL_API int32_t dlx_triangle_slots(const dl_mesh* mesh, uint32_t slotBase, uint32_t* slots, int64_t cap, int64_t* n);
/* The triangles the scene build draws from a one-mesh handle (reference mesh.cpp:137-168 triangulate_into: a triangle as
 * it is, a quad (0 1 2) (0 2 3), 5 or more corners ear-clipped in the face's Newell plane): v[3 nPoints] the points
 * recentred on centre[3] (NULL: the origin) as floats, tri[3 n] the corners, *n triangles (the first min(n, cap)
 * written), *skipped the degenerate faces left out. No GPU involved. */
DL_API int32_t dlx_triangulate(const dl_mesh* mesh, const double* centre, float* v, uint32_t* tri, int64_t cap, int64_t* n, int32_t* skipped);
/* Ladybug mode's sky, the host emulation of gendaymtx -A the reference runs inline (run.cpp:362-470; no GPU involved),
 * over T steps (jd, hour, dni, dhi) at a site (lat, lon east-positive, fileTz the weather file's zone in hours) with the
 * ground reflectance grefl and the run's hours dur: per patch of sky density mf (1 or 2) dir / dif / rad [P] (kWh/m2 x
 * sr-weights), the ground's *gnd, and dump [2 x 3 (P + 1)] the two averaged blocks --dump-sky prints (-s then -d, ground
 * row first). */
DL_API int32_t dlx_ladybug_sky(int32_t mf, int64_t T, const int32_t* jd, const double* hour, const float* dni, const float* dhi, double lat,
                               double lon, double fileTz, double grefl, double dur, double* dir, double* dif, double* rad, double* gnd,
                               float* dump);

/* The serving loop. No counterpart in the reference (its C ABI runs one call at a time behind one mutex, c_api.cpp:31,
 * 231, 250): a server with `lanes` device contexts (1..16; each a HIP stream of its own, so runs overlap on the GPU)
 * takes requests of both tools, parsed as dl_run parses them (the same flags, errors and results), and runs them as
 * lanes come free:
 *  - admission: a request starts when the device memory its run may need (an estimate from its sensors and flags) fits
 *    hbm_budget beside the runs already going (0: 60 % of the memory free at open); one that alone exceeds it runs
 *    when the device is otherwise idle;
 *  - resident scenes: the context scenes stay in HBM between requests (the engine's store), pinned while a run uses them;
 *  - continuous batching: a lane that comes free takes the oldest queued request and every queued daylight request
 *    that merges with it - the same flags and context handles and the same centre of the run's meshes, without
 *    self-occlusion, file outputs or flags naming meshes - up to batch_sensors sensor elements (0: 4,000,000), runs
 *    them as one run over all their sensor meshes and splits the rows; each request gets the values and manifest it
 *    gets alone.
 * dlx_server_submit returns the request's id (> 0), or -1 with dl_last_error() (the flags' usage errors are found
 * here); the mesh handles must stay valid until the request is waited for. dlx_server_wait blocks until the request is
 * done and gives its result (as dl_run; DL_RUNTIME with dl_last_error() when its run failed) and, when times is not
 * NULL, times[3] = {queued ms, run ms, requests served by its run}; each id is waited for once. dlx_server_stats gives
 * DLX_SERVER_STATS counters: submitted, completed, failed, runs, requests served by merged runs, peak queue length, peak
 * admitted bytes, the budget, lanes, peak runs at once. dlx_server_close finishes the queued requests and frees the
 * server. A process may run dl_run beside a server: it takes its turn on the primary context. */
typedef struct dlx_server dlx_server;
#define DLX_SERVER_STATS 10
DL_API dlx_server* dlx_server_open(int32_t lanes, int64_t hbm_budget, int64_t batch_sensors);
DL_API int64_t dlx_server_submit(dlx_server* server, int32_t tool, int32_t argc, const char* const* argv, dl_mesh* const* sensors, int32_t n_sensors,
                                 dl_mesh* const* context, int32_t n_context);
DL_API int32_t dlx_server_wait(dlx_server* server, int64_t id, dl_result** out, double* times);
DL_API int32_t dlx_server_stats(dlx_server* server, double* out, int32_t n);
DL_API void dlx_server_close(dlx_server* server);

#ifdef __cplusplus
}
#endif
#endif
