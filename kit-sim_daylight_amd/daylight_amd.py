# SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
"""ctypes binding of the MI355X library (lib/libdaylight_amd.so): the public C ABI of include/daylight_c.h (the binding the
reference's bindings/python/daylight/_native.py makes over daylight_c.dll) and the GPU seam of include/dl_device.h.

The library is the product: nothing here computes results; it moves arrays in and out. Loading fails loudly when the
built library is missing (run kit-sim_daylight_amd/build.py).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DAYLIGHT_AMD_LIB names another build of the library (as the reference binding's DAYLIGHT_DLL does)
LIB_PATH = os.environ.get("DAYLIGHT_AMD_LIB") or os.path.join(HERE, "lib", "libdaylight_amd.so")

F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")

DL_OK, DL_USAGE, DL_RUNTIME = 0, 1, 2
DL_DAYLIGHT, DL_VIEWS = 0, 1
DL_STATS = 10


class SkyIn(C.Structure):
    _fields_ = [("T", C.c_int32), ("month", C.c_void_p), ("day", C.c_void_p), ("jd", C.c_void_p), ("hour", C.c_void_p),
                ("dt", C.c_void_p), ("dni", C.c_void_p), ("dhi", C.c_void_p), ("lat", C.c_double), ("lon", C.c_double),
                ("tz", C.c_double), ("north", C.c_double), ("grefl", C.c_double), ("leap", C.c_int32), ("mf", C.c_int32),
                ("wantRad", C.c_int32), ("wantVis", C.c_int32), ("series", C.c_int32), ("seriesVis", C.c_int32), ("tcos", C.c_void_p),
                ("reflect", C.c_int32)]


class SkyOut(C.Structure):
    _fields_ = [("alt", C.c_void_p), ("azi", C.c_void_p), ("up", C.c_void_p), ("vec", C.c_void_p), ("skyRad", C.c_void_p),
                ("skyVis", C.c_void_p), ("gndRad", C.c_void_p), ("gndVis", C.c_void_p), ("sunRad", C.c_void_p),
                ("sunVis", C.c_void_p), ("Tup", C.c_void_p), ("sunBinRad", C.c_void_p), ("sunBinVis", C.c_void_p)]


class SeriesIn(C.Structure):
    _fields_ = [("nSky", C.c_uint32), ("row", C.c_void_p), ("wsky", C.c_void_p), ("P", C.c_int32), ("T", C.c_int32),
                ("G", C.c_void_p), ("sun0", C.c_uint32), ("nSun", C.c_uint32), ("sunStep", C.c_void_p),
                ("sunW", C.c_void_p), ("reflect", C.c_int32)]


class SensorMesh(C.Structure):
    _fields_ = [("xyz", C.c_void_p), ("nPts", C.c_int64), ("counts", C.c_void_p), ("nFaces", C.c_int64), ("idx", C.c_void_p), ("nIdx", C.c_int64)]


class ChunkIn(C.Structure):
    _fields_ = [("i0", C.c_int64), ("nc", C.c_int64), ("rayFlags", C.c_uint32), ("nQ", C.c_int32), ("W", C.c_void_p),
                ("useCos", C.c_void_p), ("series", C.c_void_p), ("method", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("buildMs", "traceMs", "reduceMs", "seriesMs", "skyMs", "viewMs")] + \
               [(n, C.c_int64) for n in ("rays", "traced", "launches", "sceneBytes", "nodes", "height")] + \
               [(n, C.c_double) for n in ("lastTraceMs", "lastReduceMs", "lastSeriesMs", "lastSkyMs", "lastViewMs", "lastViewReduceMs")] + \
               [(n, C.c_int64) for n in ("nodeFetches", "triFetches", "countedRays", "waveIters")] + \
               [("bounceMs", C.c_double), ("paths", C.c_int64), ("bounceLaunches", C.c_int64)] + \
               [("graphCaptures", C.c_int64), ("graphLaunches", C.c_int64), ("leafIters", C.c_int64)]


class BounceIn(C.Structure):
    _fields_ = [("bounces", C.c_uint32), ("M", C.c_uint32), ("mf", C.c_int32), ("north", C.c_double)]


class ViewIn(C.Structure):
    _fields_ = [("nAz", C.c_uint32), ("nEl", C.c_uint32), ("az0", C.c_float), ("dAz", C.c_float), ("rowSin", C.c_void_p),
                ("world", C.c_int32), ("cull", C.c_int32), ("jitter", C.c_int32), ("back", C.c_int32), ("vground", C.c_int32),
                ("nS", C.c_uint32), ("groundZ", C.c_float), ("north", C.c_double)]


SERIES_SINK = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_void_p)
DLX_COUNTERS = 20
DLX_SERVER_STATS = 10


# every function include/*.h declares: (name, restype, argtypes)
C_ABI = [
    ("dl_version", C.c_int32, []),
    ("dl_last_error", C.c_char_p, []),
    ("dl_usage", C.c_char_p, [C.c_int32]),
    ("dl_mesh_create", C.c_void_p, [C.c_char_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]),
    ("dl_mesh_load", C.c_void_p, [C.c_char_p]),
    ("dl_mesh_count", C.c_int32, [C.c_void_p]),
    ("dl_mesh_name", C.c_char_p, [C.c_void_p, C.c_int32]),
    ("dl_mesh_free", None, [C.c_void_p]),
    ("dl_control_create", C.c_void_p, []),
    ("dl_control_cancel", None, [C.c_void_p]),
    ("dl_control_progress", C.c_double, [C.c_void_p]),
    ("dl_control_free", None, [C.c_void_p]),
    ("dl_run", C.c_int32, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("dl_result_rows", C.c_int64, [C.c_void_p]),
    ("dl_result_cols", C.c_int64, [C.c_void_p]),
    ("dl_result_manifest", C.c_char_p, [C.c_void_p]),
    ("dl_result_values", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("dl_result_values_f32", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    ("dl_result_stats", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("dl_result_free", None, [C.c_void_p]),
    ("dlx_result_data", C.c_void_p, [C.c_void_p]),
    ("dlx_run_series_device", C.c_int32, [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    ("dlx_gpu_counters", C.c_int32, [C.c_void_p, C.c_int32]),
    ("dlx_server_open", C.c_void_p, [C.c_int32, C.c_int64, C.c_int64]),
    ("dlx_server_submit", C.c_int64, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    ("dlx_server_wait", C.c_int32, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("dlx_server_stats", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32]),
    ("dlx_server_close", None, [C.c_void_p]),
    ("dlx_count_fetches", None, [C.c_int32]),
    ("dlx_chunk_bytes", None, [C.c_int64]),
    ("dlx_read_epw", C.c_int32, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("dlx_thin_solid", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_uint32)]),
    ("dlx_triangle_slots", C.c_int32, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    ("dlx_triangulate", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    ("dlx_ladybug_sky", C.c_int32, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
]
SEAM = [
    ("dlg_init", C.c_int32, [C.POINTER(C.c_void_p), C.c_char_p, C.c_int32]),
    ("dlg_destroy", None, [C.c_void_p]),
    ("dlg_device_info", C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    ("dlg_build", C.c_int32, [C.c_void_p, F32P, C.c_int64, U32P, C.c_int64, C.c_char_p, C.c_int32]),
    ("dlg_build_bytes", C.c_int64, [C.c_int64, C.c_int64]),
    ("dlg_trim", None, [C.c_void_p, C.c_int32]),
    ("dlg_host_sums", C.c_void_p, [C.c_void_p, C.c_int64]),
    ("dlg_detach", C.c_void_p, [C.c_void_p]),
    ("dlg_attach", None, [C.c_void_p, C.c_void_p]),
    ("dlg_free_scene", None, [C.c_void_p]),
    ("dlg_scene_bytes", C.c_int64, [C.c_void_p]),
    ("dlg_borrow_scene", None, [C.c_void_p, C.c_void_p]),
    ("dlg_free_bytes", C.c_int64, []),
    ("dlg_scene_download", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    ("dlg_set_scene", C.c_int32, [C.c_void_p, F32P, F32P, C.c_int64, F32P, C.c_int64, C.c_char_p, C.c_int32]),
    ("dlg_ray_flags", C.c_uint32, [C.c_int32, C.c_int32]),
    ("dlg_trace", C.c_int32, [C.c_void_p, C.c_int64, C.c_int64, C.c_uint32, C.c_char_p, C.c_int32]),
    ("dlg_reduce", C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, F32P, I32P, F64P, C.c_char_p, C.c_int32]),
    ("dlg_download_bits", C.c_int32, [C.c_void_p, C.c_int64, U32P, C.c_char_p, C.c_int32]),
    ("dlg_download_bytes", C.c_int32, [C.c_void_p, C.c_int64, U8P, C.c_char_p, C.c_int32]),
    ("dlg_set_slots", C.c_int32, [C.c_void_p, F32P, C.c_int64, C.c_char_p, C.c_int32]),
    ("dlg_set_tri_slots", C.c_int32, [C.c_void_p, U32P, C.c_int64, C.c_char_p, C.c_int32]),
    ("dlg_mesh_begin", C.c_int32, [C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_void_p, C.c_char_p, C.c_int32]),
    ("dlg_mesh_add", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_uint32,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_char_p, C.c_int32]),
    ("dlg_mesh_build", C.c_int32, [C.c_void_p, C.c_char_p, C.c_int32]),
    ("dlg_mesh_staged", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, C.POINTER(C.c_int64), C.c_void_p, C.c_char_p,
                                    C.c_int32]),
    ("dlg_mesh_bytes", C.c_int64, [C.c_int64, C.c_int64, C.c_int64]),
    ("dlg_bounce_flags", C.c_uint32, [C.c_int32]),
    ("dlg_bounce", C.c_int32, [C.c_void_p, C.c_int64, C.c_int64, C.c_uint32, C.POINTER(BounceIn), C.c_char_p, C.c_int32]),
    ("dlg_reduce_dc", C.c_int32, [C.c_void_p, C.c_int64, F32P, F64P, C.c_char_p, C.c_int32]),
    ("dlg_download_dc", C.c_int32, [C.c_void_p, C.c_int64, F32P, C.c_char_p, C.c_int32]),
    ("dlg_sky", C.c_int32, [C.c_void_p, C.POINTER(SkyIn), C.POINTER(SkyOut), C.c_char_p, C.c_int32]),
    ("dlg_set_sky_series", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int32]),
    ("dlg_sky_series", C.c_int32, [C.c_void_p, F32P, F32P, C.c_void_p, C.c_char_p, C.c_int32]),
    ("dlg_series", C.c_int32, [C.c_void_p, C.c_int64, C.POINTER(SeriesIn), C.c_int32, C.c_void_p, C.c_char_p, C.c_int32]),
    ("dlg_chunk", C.c_int32, [C.c_void_p, C.POINTER(ChunkIn), C.c_void_p, C.c_void_p, C.c_char_p, C.c_int32]),
    ("dlg_chunk_noalloc", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]),
    ("dlg_chunk_noalloc_limit", C.c_int64, []),
    ("dlg_sensors_reserve", C.c_int32, [C.c_void_p, C.c_int64, C.c_char_p, C.c_int32]),
    ("dlg_sensors", C.c_int32, [C.c_void_p, C.POINTER(SensorMesh), C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_void_p, C.c_int32,
                                C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_char_p, C.c_int32]),
    ("dlg_download_sensors", C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_char_p, C.c_int32]),
    ("dlg_series_device", C.c_void_p, [C.c_void_p]),
    ("dlg_stats", C.c_int32, [C.c_void_p, C.POINTER(Stats)]),
    ("dlg_count_fetches", None, [C.c_void_p, C.c_int32]),
    ("dlg_set_tree", None, [C.c_void_p, C.c_int32]),
    ("dlg_set_labels", C.c_int32, [C.c_void_p, U8P, C.c_int64, C.c_char_p, C.c_int32]),
    ("dlg_view", C.c_int32, [C.c_void_p, C.c_int64, C.c_int64, C.c_uint32, C.POINTER(ViewIn), C.c_char_p, C.c_int32]),
    ("dlg_reduce_view", C.c_int32, [C.c_void_p, C.c_int64, F64P, C.c_uint32, C.c_double, C.c_void_p, C.c_uint32, C.c_int32, C.c_uint32,
                                    F64P, C.c_char_p, C.c_int32]),
    ("dlg_download_view", C.c_int32, [C.c_void_p, C.c_int64, U8P, F32P, C.c_char_p, C.c_int32]),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdaylight_amd.so is not built: run python kit-sim_daylight_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, res, args in C_ABI + SEAM:
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported(names):
    """The names of `names` the library exports (no call is made)."""
    L = C.CDLL(LIB_PATH)
    return [n for n in names if hasattr(L, n)]


class GpuError(RuntimeError):
    pass


def _check(rc, err):
    if rc:
        raise GpuError(err.value.decode(errors="replace"))


class Gpu:
    """The seam (include/dl_device.h) over one dlg_gpu: arrays in, arrays out."""

    def __init__(self):
        self.L = lib()
        self.h = C.c_void_p()
        self.err = C.create_string_buffer(1024)
        _check(self.L.dlg_init(C.byref(self.h), self.err, 1024), self.err)
        self.nDir = 0

    def close(self):
        if self.h:
            self.L.dlg_destroy(self.h)
            self.h = C.c_void_p()

    def set_tree(self, tree):
        """'sah' (the default) or 'lbvh': the tree the next build makes."""
        self.L.dlg_set_tree(self.h, {"sah": 0, "lbvh": 1}[tree])

    def build(self, V, I):
        V = np.ascontiguousarray(V, np.float32).reshape(-1)
        I = np.ascontiguousarray(I, np.uint32).reshape(-1)
        if V.size == 0:
            V = np.zeros(3, np.float32)
        if I.size == 0:
            _check(self.L.dlg_build(self.h, V, 0, np.zeros(3, np.uint32), 0, self.err, 1024), self.err)
            return
        _check(self.L.dlg_build(self.h, V, V.size // 3, I, I.size // 3, self.err, 1024), self.err)

    def scene(self):
        nn, nt, h = C.c_int64(), C.c_int64(), C.c_int32()
        self.L.dlg_scene_download(self.h, None, C.byref(nn), None, C.byref(nt), C.byref(h))
        nodes = np.zeros(max(1, nn.value) * 16, np.float32)
        leaves = np.zeros(max(1, nt.value) * 12, np.float32)
        if self.L.dlg_scene_download(self.h, nodes.ctypes.data, C.byref(nn), leaves.ctypes.data, C.byref(nt), C.byref(h)):
            raise GpuError("scene download failed")
        return nodes[: nn.value * 16].reshape(-1, 16), leaves[: nt.value * 12].reshape(-1, 12), h.value

    def set_scene(self, P, N, D):
        P = np.ascontiguousarray(P, np.float32).reshape(-1)
        N = np.ascontiguousarray(N, np.float32).reshape(-1)
        D = np.ascontiguousarray(D, np.float32).reshape(-1)
        self.nDir = D.size // 3
        self.nSens = P.size // 3
        _check(self.L.dlg_set_scene(self.h, P, N, P.size // 3, D if D.size else np.zeros(3, np.float32), D.size // 3, self.err, 1024), self.err)

    def trace(self, i0, nc, cull=False, translucent=False):
        """translucent: byte mode (the slots of set_slots / set_tri_slots)."""
        _check(self.L.dlg_trace(self.h, i0, nc, self.L.dlg_ray_flags(int(cull), int(translucent)), self.err, 1024), self.err)

    def bits(self, nc):
        out = np.zeros(max(1, ((self.nDir + 31) // 32) * nc), np.uint32)
        _check(self.L.dlg_download_bits(self.h, nc, out, self.err, 1024), self.err)
        return out[: ((self.nDir + 31) // 32) * nc]

    def bytes(self, nc):
        """byte mode's transmittance, direction-major [nDir x nc]"""
        out = np.zeros(max(1, self.nDir * nc), np.uint8)
        _check(self.L.dlg_download_bytes(self.h, nc, out, self.err, 1024), self.err)
        return out[: self.nDir * nc].reshape(self.nDir, nc)

    def set_slots(self, slots):
        """slots: [n][4] rows rho, tau, tn, flags (flags as an integer: 1 glass, 2 thin, 4 inward, 8 twin)"""
        a = np.zeros((len(slots), 4), np.float32)
        for k, (rho, tau, tn, flags) in enumerate(slots):
            a[k, :3] = (rho, tau, tn)
            a[k, 3] = np.array([flags], np.uint32).view(np.float32)[0]
        _check(self.L.dlg_set_slots(self.h, a.reshape(-1), len(slots), self.err, 1024), self.err)

    def set_tri_slots(self, slot):
        slot = np.ascontiguousarray(slot, np.uint32)
        _check(self.L.dlg_set_tri_slots(self.h, slot if slot.size else np.zeros(1, np.uint32), slot.size, self.err, 1024), self.err)

    def mesh_begin(self, nPts, nTriCap, slotted, centre):
        c = np.ascontiguousarray(centre, np.float64)
        _check(self.L.dlg_mesh_begin(self.h, nPts, nTriCap, int(slotted), c.ctypes.data, self.err, 1024), self.err)

    def mesh_add(self, points, counts, indices, face_slot=None, slot_base=0, ngon=None):
        """One mesh into the staged scene; ngon = (faces int64, made int32, corners uint32 [3 sum made]) of its faces
        of 5 or more corners, ear-clipped by the host. Returns the triangles made."""
        keep = [np.ascontiguousarray(points, np.float64).reshape(-1), np.ascontiguousarray(counts, np.int32),
                np.ascontiguousarray(indices, np.int32)]
        fs = None if face_slot is None else np.ascontiguousarray(face_slot, np.uint16)
        ng = ngon or (np.zeros(0, np.int64), np.zeros(0, np.int32), np.zeros(0, np.uint32))
        ng = [np.ascontiguousarray(ng[0], np.int64), np.ascontiguousarray(ng[1], np.int32), np.ascontiguousarray(ng[2], np.uint32)]
        ptr = lambda a: a.ctypes.data if a is not None and a.size else None
        made = C.c_int64()
        _check(self.L.dlg_mesh_add(self.h, ptr(keep[0]), keep[0].size // 3, ptr(keep[1]), keep[1].size, ptr(keep[2]), keep[2].size, ptr(fs),
                                   slot_base, ptr(ng[0]), ptr(ng[1]), ptr(ng[2]), ng[0].size, C.byref(made), self.err, 1024), self.err)
        return made.value

    def mesh_staged(self, slotted=False):
        nv, nt = C.c_int64(), C.c_int64()
        _check(self.L.dlg_mesh_staged(self.h, None, C.byref(nv), None, C.byref(nt), None, self.err, 1024), self.err)
        V = np.zeros((max(nv.value, 1), 3), np.float32)
        I = np.zeros((max(nt.value, 1), 3), np.uint32)
        S = np.zeros(max(nt.value, 1), np.uint32) if slotted else None
        _check(self.L.dlg_mesh_staged(self.h, V.ctypes.data, C.byref(nv), I.ctypes.data, C.byref(nt), None if S is None else S.ctypes.data,
                                      self.err, 1024), self.err)
        return V[:nv.value], I[:nt.value], (S[:nt.value] if slotted else None)

    def mesh_build(self):
        _check(self.L.dlg_mesh_build(self.h, self.err, 1024), self.err)

    def reduce(self, nc, W, useCos):
        W = np.ascontiguousarray(W, np.float32).reshape(len(useCos), -1)
        out = np.zeros(len(useCos) * nc, np.float64)
        _check(self.L.dlg_reduce(self.h, nc, len(useCos), W.reshape(-1), np.array(useCos, np.int32), out, self.err, 1024), self.err)
        return out.reshape(len(useCos), nc)

    def chunk(self, i0, nc, W, useCos, cull=False, translucent=False):
        """dlg_chunk without a series: the sums [nQ x nc] of sensors [i0, i0 + nc) of set_scene"""
        W = np.ascontiguousarray(W, np.float32).reshape(len(useCos), -1)
        cos = np.array(useCos, np.int32)
        out = np.zeros(len(useCos) * nc, np.float64)
        ci = ChunkIn(i0, nc, self.L.dlg_ray_flags(int(cull), int(translucent)), len(useCos), W.ctypes.data, cos.ctypes.data, None, 0)
        _check(self.L.dlg_chunk(self.h, C.byref(ci), out.ctypes.data, None, self.err, 1024), self.err)
        return out.reshape(len(useCos), nc)

    def chunk_noalloc(self, P, N, nc, W, useCos, sums, cull=False, translucent=False, wait=True):
        """dlg_chunk_noalloc: P, N, W, sums are device addresses (ints); the directions are set_scene's"""
        cos = np.array(useCos, np.int32)
        _check(self.L.dlg_chunk_noalloc(self.h, P, N, nc, self.L.dlg_ray_flags(int(cull), int(translucent)), len(useCos), W, cos.ctypes.data,
                                        sums, int(wait), self.err, 1024), self.err)

    def sensors(self, meshes, onVertices=False, sides="front", offset=0.0, centre=(0.0, 0.0, 0.0), ladybug=False):
        """dlg_sensors of each (points, counts, indices) mesh in turn -> (P [n, 3], N [n, 3]) float32, the elements made per mesh,
        the degenerate faces and the zero vertex normals per mesh"""
        per = 2 if sides == "both" else 1
        ms = [(np.ascontiguousarray(p, np.float64).reshape(-1), np.ascontiguousarray(c, np.int32), np.ascontiguousarray(i, np.int32))
              for p, c, i in meshes]
        n = sum((len(p) // 3 if onVertices else len(c)) * per for p, c, i in ms)
        _check(self.L.dlg_sensors_reserve(self.h, n, self.err, 1024), self.err)
        cen = np.ascontiguousarray(centre, np.float64)
        j, made, deg, zero = 0, [], [], []
        for p, c, i in ms:
            sm = SensorMesh(p.ctypes.data, len(p) // 3, c.ctypes.data, len(c), i.ctypes.data, len(i))
            mk, dg, zr = C.c_int64(), C.c_int32(), C.c_int32()
            _check(self.L.dlg_sensors(self.h, C.byref(sm), j, int(onVertices), {"front": 0, "back": 1, "both": 2}[sides], float(offset),
                                      cen.ctypes.data, int(ladybug), C.byref(mk), C.byref(dg), C.byref(zr), self.err, 1024), self.err)
            j += mk.value * per
            made.append(mk.value)
            deg.append(dg.value)
            zero.append(zr.value)
        P, N = np.zeros((max(n, 1), 3), np.float32), np.zeros((max(n, 1), 3), np.float32)
        _check(self.L.dlg_download_sensors(self.h, P.ctypes.data, N.ctypes.data, n, self.err, 1024), self.err)
        return P[:n], N[:n], made, deg, zero

    def noalloc_limit(self):
        return self.L.dlg_chunk_noalloc_limit()

    def sky(self, steps, lat, lon, tz, leap, north, grefl, mf, tcos, wantRad=True, wantVis=False, series=False, seriesVis=False, reflect=False):
        T = len(steps["month"])
        arr = {k: np.ascontiguousarray(v) for k, v in steps.items()}
        keep = [arr, tcos]
        sin = SkyIn(T, arr["month"].ctypes.data, arr["day"].ctypes.data, arr["jd"].ctypes.data, arr["hour"].ctypes.data,
                    arr["dt"].ctypes.data, arr["dni"].ctypes.data, arr["dhi"].ctypes.data, lat, lon, tz, north, grefl, int(leap), mf,
                    int(wantRad), int(wantVis), int(series), int(seriesVis), tcos.ctypes.data, int(reflect))
        P = {1: 145, 2: 577, 4: 2305}[mf]
        out = dict(alt=np.zeros(T), azi=np.zeros(T), up=np.zeros(T, np.int32), vec=np.zeros(3 * T, np.float32), skyRad=np.zeros(P),
                   skyVis=np.zeros(P), gndRad=np.zeros(1), gndVis=np.zeros(1), sunRad=np.zeros(T), sunVis=np.zeros(T),
                   Tup=np.zeros(1, np.int32), sunBinRad=np.zeros(P), sunBinVis=np.zeros(P))
        so = SkyOut(*[out[k].ctypes.data for k in ("alt", "azi", "up", "vec", "skyRad", "skyVis", "gndRad", "gndVis", "sunRad", "sunVis", "Tup",
                                                   "sunBinRad", "sunBinVis")])
        _check(self.L.dlg_sky(self.h, C.byref(sin), C.byref(so), self.err, 1024), self.err)
        del keep
        out["vec"] = out["vec"].reshape(T, 3)
        out["Tup"] = int(out["Tup"][0])
        if series:
            S, g = np.zeros(max(1, P * out["Tup"]), np.float32), np.zeros(max(1, out["Tup"]), np.float32)
            Sdc = np.zeros(max(1, (P + 1) * out["Tup"]), np.float32) if reflect else None
            _check(self.L.dlg_sky_series(self.h, S, g, Sdc.ctypes.data if reflect else None, self.err, 1024), self.err)
            out["S"], out["g"] = S[: P * out["Tup"]].reshape(P, out["Tup"]), g[: out["Tup"]]
            if reflect:
                out["Sdc"] = Sdc[: (P + 1) * out["Tup"]].reshape(P + 1, out["Tup"])
        return out

    def set_sky_series(self, S, g):
        """the series inputs S [P x Tup] and g [Tup] from the host (dlg_set_sky_series) in place of the last sky's"""
        S = np.ascontiguousarray(S, np.float32)
        g = np.ascontiguousarray(g, np.float32)
        _check(self.L.dlg_set_sky_series(self.h, S.ctypes.data, g.ctypes.data, S.shape[0], S.shape[1], self.err, 1024), self.err)

    def series(self, nc, nSky, row, wsky, P, T, G, sun0, sunStep, sunW, method, download=True, reflect=False):
        row = np.ascontiguousarray(row, np.int32)
        wsky = np.ascontiguousarray(wsky, np.float32)
        G = np.ascontiguousarray(G, np.float32)
        sunStep = np.ascontiguousarray(sunStep, np.int32)
        sunW = np.ascontiguousarray(sunW, np.float32)
        si = SeriesIn(nSky, row.ctypes.data, wsky.ctypes.data, P, T, G.ctypes.data, sun0, len(sunW), sunStep.ctypes.data,
                      sunW.ctypes.data, int(reflect))
        E = np.zeros(nc * T, np.float32) if download else None
        _check(self.L.dlg_series(self.h, nc, C.byref(si), method, E.ctypes.data if download else None, self.err, 1024), self.err)
        return E.reshape(nc, T) if download else None

    def bounce(self, i0, nc, bounces, M, mf=1, north=0.0, cull=False):
        """reflected light of sensors [i0, i0 + nc): the rows [(P + 1) x nc] (bin 0 the ground, 1 + p the patches)"""
        b = BounceIn(bounces, M, mf, north)
        _check(self.L.dlg_bounce(self.h, i0, nc, self.L.dlg_bounce_flags(int(cull)), C.byref(b), self.err, 1024), self.err)
        nb = 144 * mf * mf + 2
        out = np.zeros(max(1, nb * nc), np.float32)
        _check(self.L.dlg_download_dc(self.h, nc, out, self.err, 1024), self.err)
        return out[: nb * nc].reshape(nb, nc)

    def reduce_dc(self, nc, W):
        out = np.zeros(nc, np.float64)
        _check(self.L.dlg_reduce_dc(self.h, nc, np.ascontiguousarray(W, np.float32), out, self.err, 1024), self.err)
        return out

    def stats(self):
        s = Stats()
        self.L.dlg_stats(self.h, C.byref(s))
        return {n: getattr(s, n) for n, _ in Stats._fields_}
