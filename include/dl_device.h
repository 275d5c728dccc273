/* SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0 */
/*
 * dl_device.h - the GPU seam of the MI355X build: plain C entry points over HIP, with host pointers and sizes only.
 *
 * Each function replaces one of the reference's internal `gpu_*` functions (reference src/internal.h:125-220, the OptiX
 * + CUDA layer of src/gpu.cu) or a host loop the reference runs beside them, and says which. The daylight / views
 * orchestration (the reference's run.cpp / views.cpp) calls these; so do the tests, through ctypes.
 *
 * Conventions: status 0 ok, nonzero failed with the message in err[0..errLen) (NUL-terminated); one dlg_gpu per process
 * (the reference's Gpu), made on first use; calls on one dlg_gpu run one at a time. Device buffers belong to the
 * dlg_gpu and grow as needed. Layouts: visibility bits word-major bits[w nc + i] (downloads transposed to row-major as
 * the reference's are); byte mode's transmittance in quads of directions, the 32-bit word [(d / 4) nc + i] holding
 * directions 4 (d / 4) .. + 3 of sensor i in its bytes (the download is direction-major, as the reference's); views
 * classes / distances sensor-major [i K + d], the download's layout.
 */
#ifndef DL_DEVICE_H
#define DL_DEVICE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dlg_gpu dlg_gpu;
typedef struct dlg_scene dlg_scene;

/* gpu_init / gpu_destroy (internal.h:127-128): the HIP context, the stream and events of device 0. */
int32_t dlg_init(dlg_gpu** out, char* err, int32_t errLen);
void dlg_destroy(dlg_gpu* g);
/* the device in use: name, compute units, HBM bytes (for logs and the bench) */
int32_t dlg_device_info(char* name, int32_t nameLen, int32_t* cus, int64_t* hbmBytes);

/* gpu_build (internal.h:133): the occluder triangles (vertices v[3 nv], indices tri[3 nt]) as a BVH installed in g:
 * 63-bit Morton codes of the centroids and a radix sort order the leaves; then either the binned-SAH tree over that
 * order (DLG_TREE_SAH, the default: 16 centroid bins per axis, surface-area heuristic down to single triangles) or the
 * LBVH of the codes (DLG_TREE_LBVH: Karras 2012 hierarchy, bottom-up boxes), as dlg_set_tree chose. Both are exact
 * functions of the input (the oracle builds the same arrays) and give the same ray results. nt == 0: an empty scene
 * (every ray misses). */
int32_t dlg_build(dlg_gpu* g, const float* v, int64_t nv, const uint32_t* tri, int64_t nt, char* err, int32_t errLen);
#define DLG_TREE_SAH 0
#define DLG_TREE_LBVH 1
/* The tree the next dlg_build makes (no reference counterpart: OptiX picks its own). */
void dlg_set_tree(dlg_gpu* g, int32_t tree);
/* gpu_build_bytes (internal.h:136): device bytes a build of nt triangles over nv vertices peaks at. */
int64_t dlg_build_bytes(int64_t nv, int64_t nt);
/* gpu_trim (internal.h:138): frees the per-run buffers (with sensors: the sensors too). */
void dlg_trim(dlg_gpu* g, int32_t sensors);
/* Pinned host memory for n doubles of a chunk's sums (dlg_chunk's `sums`), owned by g, kept and grown across runs,
   freed by dlg_trim and dlg_destroy; null when pinned memory cannot be had. No counterpart in the reference (its runs
   download into their own vectors). */
double* dlg_host_sums(dlg_gpu* g, int64_t n);
/* Mesh ingest on the GPU: the scene's triangles made where they are built (the reference triangulates on the host,
 * run.cpp:120-147 over mesh.cpp:137-168 triangulate_into, one mesh per host thread). dlg_mesh_begin stages room for
 * nPts vertices and up to nTriCap triangles (sum of n - 2 over the faces), with per-triangle global slots when slotted;
 * dlg_mesh_add appends one mesh: its points xyz[3 nPts] (double) recentred on centre and rounded to float, its faces
 * (counts[nFaces], corner indices idx[nIdx], mesh-local) fanned in face order — a triangle (a, b, c), a quad (a, b, c)
 * (a, c, d) — at the running triangle count, with the face's slot slotBase + faceSlot[f] (faceSlot may be null: 0).
 * Faces of 5 or more corners come triangulated by the host (the ear clipping of triangulate_into): ngonFace[nNgon]
 * their face indices ascending, ngonMade[nNgon] the triangles each made (0: skipped), ngonTri[3 sum(made)] the corners
 * (mesh-local) in face order. *made = the mesh's triangles. dlg_mesh_build builds the staged triangles as dlg_build
 * would the same arrays from the host (and the per-triangle slots as dlg_set_tri_slots), then frees the staging. The
 * points and faces must pass the host's mesh checks (counts >= 3 summing to nIdx, indices in range): a violation found
 * on the device fails the call and nothing out of range is read. */
int32_t dlg_mesh_begin(dlg_gpu* g, int64_t nPts, int64_t nTriCap, int32_t slotted, const double* centre, char* err, int32_t errLen);
int32_t dlg_mesh_add(dlg_gpu* g, const double* xyz, int64_t nPts, const int32_t* counts, int64_t nFaces, const int32_t* idx, int64_t nIdx,
                     const uint16_t* faceSlot, uint32_t slotBase, const int64_t* ngonFace, const int32_t* ngonMade, const uint32_t* ngonTri,
                     int64_t nNgon, int64_t* made, char* err, int32_t errLen);
int32_t dlg_mesh_build(dlg_gpu* g, char* err, int32_t errLen);
/* The staged arrays before dlg_mesh_build (for the check against the host triangulation): v[3 nv], tri[3 nt] (global
 * vertex ids), slot[nt] when slotted (may be null). Sizes in *nv / *nt when v and tri are null. */
int32_t dlg_mesh_staged(dlg_gpu* g, float* v, int64_t* nv, uint32_t* tri, int64_t* nt, uint32_t* slot, char* err, int32_t errLen);
/* device bytes one dlg_mesh_add of a mesh this large peaks at, beside the staging (for the scene's admission) */
int64_t dlg_mesh_bytes(int64_t nPts, int64_t nFaces, int64_t nIdx);
/* gpu_detach / gpu_attach / gpu_free (internal.h:141-144): a built scene moved out of g and kept resident. */
dlg_scene* dlg_detach(dlg_gpu* g);
void dlg_attach(dlg_gpu* g, dlg_scene* s);
void dlg_free_scene(dlg_scene* s);
int64_t dlg_scene_bytes(const dlg_scene* s);
/* No counterpart in the reference (one context per process there): dst's scene becomes src's installed scene, borrowed
 * (the arrays stay src's: dst never frees or replaces them; dlg_attach / dlg_detach on dst drop the loan). For lanes
 * that trace one resident scene from several contexts, each on its own stream. */
void dlg_borrow_scene(dlg_gpu* dst, const dlg_gpu* src);
/* gpu_free_bytes (internal.h:145) */
int64_t dlg_free_bytes(void);
/* The installed scene as built, for the bit-exact check against a host build: nodes [n][16] floats (left box lo xyz hi
 * xyz, right box, left child, right child as bits, 2 unused; child >= 2^31 = leaf), leaves [nt][12] floats (v0, v1, v2,
 * original index as bits, 2 unused), tree height. Sizes in *nNodes / *nTris when the pointers are null. */
int32_t dlg_scene_download(dlg_gpu* g, float* nodes, int64_t* nNodes, float* leaves, int64_t* nTris, int32_t* height);

/* Slot optics (the reference's Slot, params.h:7-19): rho the diffuse reflectance (the bounce pass), tau the transmittance at normal
 * incidence (> 0: the slot transmits), tn the glass transmissivity (tn_from_Tn(tau)), flags DLG_SLOT_*. */
typedef struct { float rho, tau, tn; uint32_t flags; } dlg_slot;
#define DLG_SLOT_GLASS 1u   /* transmittance by angle (Radiance glass); else tau at every angle */
#define DLG_SLOT_THIN 2u    /* a closed thin solid: its exit face passes everything */
#define DLG_SLOT_INWARD 4u  /* ... wound inward: entry is the back face */
#define DLG_SLOT_TWIN 8u    /* ... of coincident face pairs (the bounce pass only) */
/* gpu_set_slots (internal.h:139): the optics of every global slot, read by a byte-mode trace and the bounce pass. */
int32_t dlg_set_slots(dlg_gpu* g, const dlg_slot* slots, int64_t n, char* err, int32_t errLen);
/* The global slot of every triangle of the installed scene (input order, nt = the build's triangle count), kept with the
 * scene's leaves: the reference passes them with the build (Tris::slot, internal.h:58, gpu_build's per-triangle SBT
 * offsets). A scene built without them has every triangle in slot 0. */
int32_t dlg_set_tri_slots(dlg_gpu* g, const uint32_t* slot, int64_t nt, char* err, int32_t errLen);

/* gpu_set_scene (internal.h:146): resident sensors (points P[3 n], normals N[3 n], recentred) and directions D[3 nDir];
 * nSens = -1 keeps the resident sensors (those dlg_sensors made). */
int32_t dlg_set_scene(dlg_gpu* g, const float* P, const float* N, int64_t nSens, const float* D, int64_t nDir, char* err, int32_t errLen);
/* gpu_ray_flags (internal.h:218): the reference's OptiX flag values (TERMINATE_ON_FIRST_HIT 4 | DISABLE_ANYHIT 1 |
 * CULL_BACK_FACING_TRIANGLES 16), kept so callers and tests read the same numbers. */
uint32_t dlg_ray_flags(int32_t cullBackfaces, int32_t translucent);

/* Sensor preparation on the GPU, in place of the host's face_geom / make_sensors_into / sensors_of (reference
 * mesh.cpp:45-71, 315-349, run.cpp:192-213): dlg_sensors_reserve sizes the resident sensors to nSens (dlg_set_scene's
 * P / N, left as they are: each is then written by dlg_sensors); dlg_sensors writes one mesh's sensors from sensor
 * j0 on - per face (onVertices 0) or per point, sides 0 front / 1 back / 2 both per element, offset along the normal,
 * recentred on centre[3], ladybug's face rules when ladybug != 0 - with the host's FP64 arithmetic (bit-identical),
 * and gives the elements made, the faces under 1e-12 of area and the points whose normal sum is 0 (set to +z). The
 * mesh's indices must be in range (they are clamped, not checked). Directions come from dlg_set_scene with nSens = -1
 * (keep the sensors). */
typedef struct {
  const double* xyz;
  int64_t nPts;
  const int32_t* counts;
  int64_t nFaces;
  const int32_t* idx;
  int64_t nIdx;
} dlg_sensor_mesh;
int32_t dlg_sensors_reserve(dlg_gpu* g, int64_t nSens, char* err, int32_t errLen);
int32_t dlg_sensors(dlg_gpu* g, const dlg_sensor_mesh* m, int64_t j0, int32_t onVertices, int32_t sides, double offset, const double* centre,
                    int32_t ladybug, int64_t* made, int32_t* degenerate, int32_t* zeroNormals, char* err, int32_t errLen);
/* The first n resident sensors (points and normals, 3 floats each), for checks. */
int32_t dlg_download_sensors(dlg_gpu* g, float* P, float* N, int64_t n, char* err, int32_t errLen);
/* gpu_trace (internal.h:151): visibility bits of sensors [i0, i0 + nc) against every direction, on the device. Ray
 * flags without DISABLE_ANYHIT (dlg_ray_flags(., 1): a slot transmits): byte mode, the transmittance of every direction
 * as a byte (x 255, 0 = culled or blocked) instead of a bit, through the slots of dlg_set_slots / dlg_set_tri_slots
 * (the reference's __raygen__bytes + __anyhit__tr, trace_dev.cu:115-165). */
int32_t dlg_trace(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, char* err, int32_t errLen);
/* gpu_reduce (internal.h:154), every quantity of a run in one pass over the visibility: out[q nc + i] = sum over the
 * visible directions d (ascending) of double(float(useCos[q] ? max(0, N_i . D_d) : 1) * T) * double(W[q nDir + d]),
 * T = 1 (bits) or byte / 255.f (byte mode). */
int32_t dlg_reduce(dlg_gpu* g, int64_t nc, int32_t nQ, const float* W, const int32_t* useCos, double* out, char* err, int32_t errLen);
/* gpu_download_bits (internal.h:182): [nWords x nc] as on the device. */
int32_t dlg_download_bits(dlg_gpu* g, int64_t nc, uint32_t* out, char* err, int32_t errLen);
/* gpu_download_bytes (internal.h:183): byte mode's transmittance, direction-major [nDir x nc]. */
int32_t dlg_download_bytes(dlg_gpu* g, int64_t nc, uint8_t* out, char* err, int32_t errLen);

/* gpu_bounce (internal.h:175-178) + __raygen__bounce (trace_dev.cu:260-311): reflected light of sensors [i0, i0 + nc),
 * M paths per sensor with up to `bounces` diffuse reflections by the slots' rho (glass mirrors, porous and glass
 * transmit, thin solids' exit faces pass), binned where they escape into rows dc[b nc + i] kept on the device: bin 0 the
 * ground (below the horizon), 1 + p the Reinhart patches of sky density mf in gendaymtx order, true north = the model's
 * turned by -north; pi / M per path, so sum_b dc(b, i) L_b is the reflected irradiance under bin radiances L. The rows
 * do not depend on the chunking (seeds from the global sensor index, sums in sample order). rayFlags: dlg_bounce_flags. */
typedef struct {
  uint32_t bounces, M;          /* diffuse reflections a path may have; paths per sensor */
  int32_t mf;                   /* sky density 1, 2, 4 */
  double north;                 /* degrees */
} dlg_bounce_in;
/* gpu_bounce_flags (internal.h:220): DISABLE_ANYHIT (1) [| CULL_BACK_FACING_TRIANGLES (16)] */
uint32_t dlg_bounce_flags(int32_t cullBackfaces);
int32_t dlg_bounce(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, const dlg_bounce_in* b, char* err, int32_t errLen);
/* gpu_reduce_dc (internal.h:179): out[i] = sum over bins b (ascending) of double(dc(b, i)) * double(W[b]), W [P + 1] (P the sky patches). */
int32_t dlg_reduce_dc(dlg_gpu* g, int64_t nc, const float* W, double* out, char* err, int32_t errLen);
/* gpu_download_dc (internal.h:180): the rows [(P + 1) x nc] as on the device. */
int32_t dlg_download_dc(dlg_gpu* g, int64_t nc, float* out, char* err, int32_t errLen);

/* The sky of a run on the GPU, FP64 (replaces the host loops of reference run.cpp:311-446 and sky.cpp's
 * gendaymtx_step / solar.cpp's sun_noaa per step): the sun of every step, the Perez all-weather sky per step and patch,
 * the cumulative per-patch weights and, for a series, the patch radiance matrix S [P x Tup] kept on the device. */
typedef struct {
  int32_t T;
  const int32_t* month; const int32_t* day; const int32_t* jd;   /* [T] */
  const double* hour; const double* dt;                          /* [T] */
  const float* dni; const float* dhi;                            /* [T] */
  double lat, lon, tz, north, grefl;
  int32_t leap, mf;
  int32_t wantRad, wantVis;     /* radiometric / visible cumulative weights */
  int32_t series, seriesVis;    /* keep S (radiometric, or visible when seriesVis) and g for the non-night steps */
  const double* tcos;           /* [1025] the table cosine (Radiance tcos.c): cos(k pi / 2048) */
  int32_t reflect;              /* reflected light: the sun binned as gendaymtx -d (AddDirect) per patch, summed into
                                   sunBinRad / sunBinVis, and with series the bin radiance matrix Sdc kept on the device */
} dlg_sky_in;
typedef struct {
  double* alt; double* azi; int32_t* up; float* vec;             /* [T], vec [3 T] turned by north */
  double* skyRad; double* skyVis;                                /* [P] */
  double* gndRad; double* gndVis;                                /* [1] */
  double* sunRad; double* sunVis;                                /* [T]: the step's sun weight (used where up) */
  int32_t* Tup;                                                  /* non-night steps (series columns) */
  double* sunBinRad; double* sunBinVis;                          /* [P] reflect: the binned sun's cumulative weights, or null */
} dlg_sky_out;
int32_t dlg_sky(dlg_gpu* g, const dlg_sky_in* in, dlg_sky_out* out, char* err, int32_t errLen);
/* The last dlg_sky's series inputs: S [P x Tup] and g [Tup] (floats), as the contraction reads them, and with reflect the
 * bins' radiance Sdc [(P + 1) x Tup] (row 0 g, row 1 + p S plus the binned sun; null: not wanted). */
int32_t dlg_sky_series(dlg_gpu* g, float* S, float* gS, float* Sdc, char* err, int32_t errLen);
/* The series inputs from the host instead of dlg_sky: S [P x Tup], g [Tup] (ladybug mode's gendaymtx -A columns, the
 * host loop of reference run.cpp:377-396). */
int32_t dlg_set_sky_series(dlg_gpu* g, const float* S, const float* gS, int32_t P, int32_t Tup, char* err, int32_t errLen);

/* gpu_series (internal.h:155-171): E[i T + t] for the chunk's sensors, S from the last dlg_sky (P x T, T = Tup).
 * method 0: FP32 on the vector units, each sum in direction order by fma (the reference kernel's arithmetic);
 * method 1: BF16x3 on MFMA (operands split in bf16 hi + lo, three products, FP32 accumulation in blocks of 256
 * directions added to a running FP32 total, the epilogue in FP64); method 2: BF16x6 (hi + mid + lo, six products),
 * FP32-level accuracy; method 3: FP16x3 (operands scaled by powers of two into fp16's range, split in fp16 hi + lo,
 * three products on the fp16 MFMA, accumulated as method 1): method 1's cost, FP32-level accuracy.
 * E: host array [nc x T], or null to keep it on the device (dlg_series_device gives the pointer). */
typedef struct {
  uint32_t nSky;                /* sky directions 0 .. nSky - 1 (P K) */
  const int32_t* row;           /* [nSky] patch per direction */
  const float* wsky;            /* [nSky] omega / K */
  int32_t P, T;
  const float* G;               /* [nc] ground view factor per sensor (the ground term per step g comes with S) */
  uint32_t sun0, nSun;          /* sun directions */
  const int32_t* sunStep;       /* [nSun] column of each sun direction */
  const float* sunW;            /* [nSun] */
  int32_t reflect;              /* add the chunk's dlg_bounce rows times the last sky's Sdc: sum_b dc(b, i) Sdc[b, t] */
} dlg_series_in;
int32_t dlg_series(dlg_gpu* g, int64_t nc, const dlg_series_in* in, int32_t method, float* E, char* err, int32_t errLen);
/* the device copy of the last series (float [nc x T]) */
const float* dlg_series_device(dlg_gpu* g);

/* Views (ViewLaunch, internal.h:186-195): the grid of K = nAz x nEl cells per sensor, rows top-down, columns clockwise
 * from az0 (degrees); world: the 360 frame (true north, horizon); cull: directions behind the sensor are own surface;
 * jitter: a seeded point per cell (pcg3d of element, cell, side); vground: the virtual ground plane z = groundZ. */
typedef struct {
  uint32_t nAz, nEl;
  float az0, dAz;
  const float* rowSin;          /* [2 nEl] host: sin of each row's bottom and top elevation */
  int32_t world, cull, jitter, back, vground;
  uint32_t nS;                  /* sides per element (the seeds' element = sensor / nS) */
  float groundZ;                /* recentred */
  double north;                 /* degrees */
} dlg_view_in;
/* gpu_set_labels (internal.h:197): the class of a hit per triangle (input order; 2 obstruction, 3 + g target group g). */
int32_t dlg_set_labels(dlg_gpu* g, const uint8_t* label, int64_t n, char* err, int32_t errLen);
/* gpu_view (internal.h:201): classes and distances of sensors [i0, i0 + nc) on the device, sensor-major [nc x K]. */
int32_t dlg_view(dlg_gpu* g, int64_t i0, int64_t nc, uint32_t rayFlags, const dlg_view_in* v, char* err, int32_t errLen);
/* gpu_reduce_view (internal.h:208-209): out[i NA + a], NA = 3 nClass + 1 + nReach + glazing + (isoDim > 0): per class
 * weight, weight x ln d, min d; own surface; reach sums; glazing; isovist area (isoDim 2) or volume (3). */
int32_t dlg_reduce_view(dlg_gpu* g, int64_t nc, const double* wRow, uint32_t nClass, double farDist, const float* reach, uint32_t nReach,
                        int32_t glazing, uint32_t isoDim, double* out, char* err, int32_t errLen);
/* gpu_download_view (internal.h:210): classes and distances, row-major [nc x K] */
int32_t dlg_download_view(dlg_gpu* g, int64_t nc, uint8_t* cls, float* dist, char* err, int32_t errLen);

/* No counterpart in the reference (its chunk loop launches and waits kernel by kernel, run.cpp:625-688): the device
 * work of one daylight chunk as one HIP graph on the library's stream - dlg_trace of sensors [i0, i0 + nc), dlg_reduce
 * of nQ rows of weights W [nQ x nDir] (useCos as there), and with series (T > 0) dlg_series by `method` with the
 * ground view factor G taken on the device from the last row of the reduce (series->G is not read). The graph is
 * captured on the first chunk of a shape and replayed while the buffers, sizes and flags stay the same. sums [nQ x nc]
 * (as dlg_reduce's out), E [nc x T] (may be null: the series stays on the device, dlg_series_device). Not for the
 * counting variant or reflected light (use the calls one by one). */
typedef struct {
  int64_t i0, nc;
  uint32_t rayFlags;
  int32_t nQ;
  const float* W;
  const int32_t* useCos;
  const dlg_series_in* series;  /* null: no series */
  int32_t method;
} dlg_chunk_in;
int32_t dlg_chunk(dlg_gpu* g, const dlg_chunk_in* in, double* sums, float* E, char* err, int32_t errLen);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under stream capture, serving
 * loops that forbid allocation): dlg_chunk without a series, allocating nothing. Sensors P, N ([nc] x 3 floats), weights
 * W [nQ x nDir] and the output sums [nQ x nc] (as dlg_reduce's out) are the caller's device memory; directions, scene
 * and slots are those already set on g (dlg_set_scene, dlg_build / dlg_attach, dlg_set_slots); useCos [nQ] is host
 * memory. The visibility goes to a fixed device array of dlg_chunk_noalloc_limit() bytes per device, held until the
 * process ends: 4 nWords bytes per sensor (byte mode, rayFlags without bit 0: 32 nWords), and a call over more
 * sensors than fit fails with a message naming the limit. Kernels are enqueued on the library's stream; wait != 0
 * waits for them. Every call on a device shares that device's array: calls that overlap on one device race. */
int32_t dlg_chunk_noalloc(dlg_gpu* g, const float* P, const float* N, int64_t nc, uint32_t rayFlags, int32_t nQ, const float* W,
                          const int32_t* useCos, double* sums, int32_t wait, char* err, int32_t errLen);
int64_t dlg_chunk_noalloc_limit(void);

/* gpu_stats (internal.h:211-216), plus per-kernel device times of the last calls (HIP events on the library's stream). */
typedef struct {
  double buildMs, traceMs, reduceMs, seriesMs, skyMs, viewMs;
  int64_t rays, traced, launches, sceneBytes, nodes, height;
  double lastTraceMs, lastReduceMs, lastSeriesMs, lastSkyMs, lastViewMs, lastViewReduceMs;
  /* the counting variant's totals (dlg_count_fetches): node visits and triangle tests of the visibility traversal, the
   * ray slots they were counted over, and the traversal iterations the waves executed */
  int64_t nodeFetches, triFetches, countedRays, waveIters;
  /* reflected light: ms, paths, launches */
  double bounceMs;
  int64_t paths, bounceLaunches;
  /* dlg_chunk: graphs captured, graph launches */
  int64_t graphCaptures, graphLaunches;
  /* the counting variant's triangle-test iterations executed by the waves */
  int64_t leafIters;
} dlg_stats_t;
int32_t dlg_stats(const dlg_gpu* g, dlg_stats_t* out);
/* No counterpart in the reference (its OptiX traversal is opaque): on != 0 makes dlg_trace run the traversal's counting
 * variant, which adds every node visit and triangle test to dlg_stats_t (the algorithmic bytes of the traversal). */
void dlg_count_fetches(dlg_gpu* g, int32_t on);

#ifdef __cplusplus
}
#endif
#endif
