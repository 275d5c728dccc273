# SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
"""ctypes access to the oracle (liborc.so) and, where it was built, the reference's own sky / sun model (_ref/libdlref.so).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this module, as the checker
and the reported CPU baseline. The product never does.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
U64P = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")

QUANTITIES = ["radiation", "direct", "diffuse", "irradiance", "illuminance", "sunhours"]


def build():
    """Compile liborc.so (and _ref/libdlref.so when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class Mesh(C.Structure):
    _fields_ = [("points", C.c_void_p), ("nPoints", C.c_long), ("counts", C.c_void_p), ("nFaces", C.c_long),
                ("indices", C.c_void_p), ("sensor", C.c_int)]


class Time(C.Structure):
    _fields_ = [("t0", C.c_double), ("t1", C.c_double), ("stepMin", C.c_int), ("dayFrom", C.c_double), ("dayTo", C.c_double)]


class Step(C.Structure):
    _fields_ = [("month", C.c_int), ("day", C.c_int), ("jd", C.c_int), ("hour", C.c_double), ("hoy", C.c_double),
                ("dt", C.c_double), ("dni", C.c_float), ("dhi", C.c_float)]


class DaylightOpt(C.Structure):
    _fields_ = [("quantities", C.c_int * 6), ("nQ", C.c_int), ("mf", C.c_int), ("patchSamples", C.c_int),
                ("ground", C.c_double), ("north", C.c_double), ("offset", C.c_double), ("sides", C.c_int),
                ("onVertices", C.c_int), ("selfOcclude", C.c_int), ("cullBackfaces", C.c_int), ("virtualGround", C.c_int),
                ("lat", C.c_double), ("lon", C.c_double), ("tz", C.c_double), ("time", Time), ("wantSeries", C.c_int),
                ("nthreads", C.c_int), ("ladybug", C.c_int)]


class DaylightOut(C.Structure):
    _fields_ = [("nElem", C.c_long), ("nSens", C.c_long), ("nCols", C.c_long), ("values", C.c_void_p),
                ("durationHours", C.c_double), ("nSky", C.c_long), ("nSun", C.c_long), ("nDir", C.c_long), ("Tup", C.c_int),
                ("series", C.c_void_p)]


class View(C.Structure):
    _fields_ = [("nAz", C.c_uint), ("nEl", C.c_uint), ("az0", C.c_float), ("dAz", C.c_float), ("rowSin", C.c_void_p),
                ("world", C.c_int), ("cull", C.c_int), ("jitter", C.c_int), ("back", C.c_int), ("vground", C.c_int),
                ("nS", C.c_uint), ("groundZ", C.c_float), ("cosN", C.c_float), ("sinN", C.c_float), ("label", C.c_void_p)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_patch_count.restype = C.c_int
        L.orc_patches.restype = C.c_float
        L.orc_patches.argtypes = [C.c_int, F32P, F32P, F32P, I32P, I32P, F32P, F32P]
        L.orc_patch_sample.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, F32P]
        L.orc_tsin.restype = C.c_double
        L.orc_tsin.argtypes = [C.c_double]
        L.orc_tcos.restype = C.c_double
        L.orc_tcos.argtypes = [C.c_double]
        L.orc_round_fmt.restype = C.c_double
        L.orc_round_fmt.argtypes = [C.c_double, C.c_char_p]
        L.orc_gendaymtx_step.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_int, C.c_int, F32P]
        L.orc_sun_radiance.argtypes = [C.c_double] * 3 + [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_rotate_z.argtypes = [F32P, C.c_double, F32P]
        L.orc_dir_from_altaz.argtypes = [C.c_double, C.c_double, F32P]
        L.orc_sun_noaa.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), F32P]
        L.orc_extraterrestrial.restype = C.c_double
        L.orc_extraterrestrial.argtypes = [C.c_int]
        L.orc_ineichen.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)] * 3
        L.orc_clear_sky.restype = C.c_int
        L.orc_clear_sky.argtypes = [C.c_double] * 5 + [F32P, F32P]
        L.orc_hoy_to_date.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_parse_hoy.restype = C.c_int
        L.orc_parse_hoy.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_double)]
        L.orc_make_steps.restype = C.c_long
        L.orc_make_steps.argtypes = [C.POINTER(Time), C.c_void_p, C.c_void_p, C.c_long, C.c_int, C.c_double, C.c_void_p, C.c_long]
        L.orc_bbox_centre.argtypes = [C.c_void_p, C.c_int, F64P]
        L.orc_sensors.restype = C.c_long
        L.orc_sensors.argtypes = [C.POINTER(Mesh), C.c_int, C.c_int, C.c_double, F64P, F32P, F32P]
        L.orc_sensors_mode.restype = C.c_long
        L.orc_sensors_mode.argtypes = [C.POINTER(Mesh), C.c_int, C.c_int, C.c_double, C.c_int, F64P, F32P, F32P]
        L.orc_ladybug_sky.argtypes = [C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
        L.orc_triangulate.restype = C.c_long
        L.orc_triangulate.argtypes = [C.POINTER(Mesh), F64P, F32P, U32P, C.c_uint32, C.POINTER(C.c_int)]
        L.orc_scene_new.restype = C.c_void_p
        L.orc_scene_new.argtypes = [F32P, C.c_long, U32P, C.c_long]
        L.orc_scene_free.argtypes = [C.c_void_p]
        L.orc_scene_sah.restype = C.c_int
        L.orc_scene_sah.argtypes = [C.c_void_p]
        L.orc_scene_height.restype = C.c_int
        L.orc_scene_height.argtypes = [C.c_void_p]
        L.orc_scene_nodes.restype = C.c_long
        L.orc_scene_nodes.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
        L.orc_scene_pad.restype = C.c_float
        L.orc_scene_pad.argtypes = [C.c_void_p]
        L.orc_lbvh_keys.argtypes = [F32P, U32P, C.c_long, U64P]
        L.orc_lbvh_topology.argtypes = [U64P, C.c_long, U32P, U32P]
        L.orc_trace_bits.argtypes = [C.c_void_p, F32P, F32P, C.c_long, F32P, C.c_int, C.c_int, U32P, C.c_int]
        L.orc_visible.restype = C.c_int
        L.orc_visible.argtypes = [C.c_void_p, F32P, F32P, F32P, C.c_int]
        L.orc_visible_brute.restype = C.c_int
        L.orc_visible_brute.argtypes = [F32P, U32P, C.c_long, F32P, F32P, F32P, C.c_int]
        L.orc_offset_ray.argtypes = [F32P, F32P, F32P]
        L.orc_pcg3d.argtypes = [C.c_uint32] * 3 + [U32P]
        L.orc_view_cell.argtypes = [C.c_uint, C.c_uint, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_uint, C.c_uint, F32P]
        L.orc_trace_view.argtypes = [C.c_void_p, F32P, F32P, C.c_long, C.c_uint, C.POINTER(View), C.c_int, U8P, F32P, C.c_int]
        L.orc_reduce.argtypes = [U32P, C.c_long, C.c_int, F32P, F32P, F32P, C.c_int, F64P]
        L.orc_series64.argtypes = [U32P, C.c_long, F32P, F32P, C.c_uint, I32P, F32P, F32P, C.c_int, F32P, F32P, C.c_uint,
                                   C.c_uint, I32P, F32P, F32P, F64P, C.c_int]
        L.orc_series32.argtypes = [U32P, C.c_long, F32P, F32P, C.c_uint, I32P, F32P, F32P, C.c_int, F32P, F32P, C.c_uint,
                                   C.c_uint, I32P, F32P, F32P, F32P]
        L.orc_trace_bytes.argtypes = [C.c_void_p, F32P, F32P, C.c_long, F32P, C.c_int, C.c_int, C.c_void_p, F32P, C.c_long, U8P, C.c_int]
        L.orc_tn_from_Tn.restype = C.c_double
        L.orc_tn_from_Tn.argtypes = [C.c_double]
        L.orc_glass_T.restype = C.c_float
        L.orc_glass_T.argtypes = [C.c_double, C.c_double]
        L.orc_series_dc.argtypes = [F32P, C.c_long, C.c_uint, F32P, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_glass_TR.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_bounce.argtypes = [C.c_void_p, F32P, F32P, C.c_long, C.c_uint, C.c_void_p, F32P, C.c_long, C.c_uint, C.c_uint, C.c_int,
                                 C.c_double, C.c_int, F32P, C.c_int]
        L.orc_reduce_bytes.argtypes = [U8P, C.c_long, C.c_int, F32P, F32P, F32P, C.c_int, F64P]
        L.orc_series64_bytes.argtypes = [U8P, C.c_long, F32P, F32P, C.c_uint, I32P, F32P, F32P, C.c_int, F32P, F32P, C.c_uint,
                                         C.c_uint, I32P, F32P, F32P, F64P, C.c_int]
        L.orc_series32_bytes.argtypes = [U8P, C.c_long, F32P, F32P, C.c_uint, I32P, F32P, F32P, C.c_int, F32P, F32P, C.c_uint,
                                         C.c_uint, I32P, F32P, F32P, F32P]
        L.orc_reduce_view.argtypes = [U8P, F32P, C.c_long, C.c_uint, C.c_uint, F64P, C.c_uint, C.c_double, F32P, C.c_uint,
                                      C.c_int, C.c_uint, C.POINTER(View), C.c_uint, F64P]
        L.orc_daylight_size.restype = C.c_int
        L.orc_daylight_size.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_long, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.POINTER(DaylightOpt), C.POINTER(C.c_long), C.POINTER(C.c_long),
                                        C.POINTER(C.c_int), C.POINTER(C.c_long)]
        L.orc_daylight.restype = C.c_int
        L.orc_daylight.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_long, C.c_double, C.c_double, C.c_double,
                                   C.c_int, C.POINTER(DaylightOpt), C.POINTER(DaylightOut), C.c_char_p, C.c_int]
        _lib = L
    return _lib


def ref():
    """The reference's own sky / sun model, or None where it was not built (no reference sources)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libdlref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_patches.restype = C.c_int
        R.ref_patches.argtypes = [C.c_int, F32P, F32P, F32P, I32P, I32P, F32P, F32P, C.POINTER(C.c_float)]
        R.ref_patch_sample.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, F32P]
        R.ref_gendaymtx_step.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_int, C.c_int, C.c_int, F32P]
        R.ref_sun_noaa.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), F32P]
        R.ref_sun_radiance.argtypes = [C.c_double] * 3 + [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        R.ref_ineichen.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)] * 3
        R.ref_extraterrestrial.restype = C.c_double
        R.ref_extraterrestrial.argtypes = [C.c_int]
        R.ref_clear_sky.restype = C.c_int
        R.ref_clear_sky.argtypes = [C.c_double] * 5 + [F32P, F32P]
        R.ref_round_fmt.restype = C.c_double
        R.ref_round_fmt.argtypes = [C.c_double, C.c_char_p]
        R.ref_rotate_z.argtypes = [F32P, C.c_double, F32P]
        R.ref_make_steps.restype = C.c_long
        R.ref_make_steps.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                     C.c_long, C.c_int, C.c_double, F64P, C.c_long]
        _ref = R
    return _ref


# ---------------------------------------------------------------------------------------------------------------- helpers
def patches(mf):
    L = lib()
    P = L.orc_patch_count(mf)
    alt, azi, om = (np.zeros(P, np.float32) for _ in range(3))
    row, nin = np.zeros(P, np.int32), np.zeros(P, np.int32)
    cen, lb = np.zeros(3 * P, np.float32), np.zeros(3 * P, np.float32)
    alpha = L.orc_patches(mf, alt, azi, om, row, nin, cen, lb)
    return dict(P=P, alpha=alpha, alt=alt, azi=azi, omega=om, row=row, ninrow=nin, centre=cen.reshape(P, 3), lb=lb.reshape(P, 3))


def gendaymtx_step(mf, jd, alt, azi, dni, dhi, grefl, sun, sky, visible):
    P = lib().orc_patch_count(mf)
    out = np.zeros(3 * (P + 1), np.float32)
    lib().orc_gendaymtx_step(mf, jd, alt, azi, dni, dhi, grefl, int(sun), int(sky), int(visible), out)
    return out


def sun_noaa(lat, lon, tz, month, day, hour, leap=False):
    a, z, u, v = C.c_double(), C.c_double(), C.c_int(), np.zeros(3, np.float32)
    lib().orc_sun_noaa(lat, lon, tz, month, day, hour, int(leap), C.byref(a), C.byref(z), C.byref(u), v)
    return a.value, z.value, bool(u.value), v


def hoy_to_date(hoy, leap=False):
    m, d, h = C.c_int(), C.c_int(), C.c_double()
    lib().orc_hoy_to_date(hoy, int(leap), C.byref(m), C.byref(d), C.byref(h))
    return m.value, d.value, h.value


def clear_sky(lat, lon, tz, TL=3.0, altitude=0.0):
    dni, dhi = np.zeros(8760, np.float32), np.zeros(8760, np.float32)
    if lib().orc_clear_sky(lat, lon, tz, TL, altitude, dni, dhi):
        raise ValueError("bad clear-sky parameters")
    return dni, dhi


def make_steps(t0=math.nan, t1=math.nan, stepMin=60, dayFrom=0.0, dayTo=1440.0, dni=None, dhi=None, leap=False, shift=0.0):
    tm = Time(t0, t1, stepMin, dayFrom, dayTo)
    cap = 600000
    out = (Step * cap)()
    n = len(dni) if dni is not None else 0
    pd = dni.ctypes.data if dni is not None else None
    ph = dhi.ctypes.data if dhi is not None else None
    k = lib().orc_make_steps(C.byref(tm), pd, ph, n, int(leap), shift, C.cast(out, C.c_void_p), cap)
    if k < 0:
        raise ValueError("make_steps failed")
    return [out[i] for i in range(k)]


class MeshArrays:
    """A SimulationMesh as numpy arrays (float64 xyz, int32 counts / indices)."""

    def __init__(self, name, points, counts, indices, sensor=True):
        self.name = name
        self.points = np.ascontiguousarray(points, np.float64).reshape(-1)
        self.counts = np.ascontiguousarray(counts, np.int32)
        self.indices = np.ascontiguousarray(indices, np.int32)
        self.sensor = sensor

    def c(self, sensor=None):
        return Mesh(self.points.ctypes.data, len(self.points) // 3, self.counts.ctypes.data, len(self.counts),
                    self.indices.ctypes.data, int(self.sensor if sensor is None else sensor))


def read_obj(path):
    """OBJ groups as MeshArrays (v and f lines, 1-based or negative indices; each mesh keeps its used points in order)."""
    V, meshes = [], []
    name, faces = os.path.splitext(os.path.basename(path))[0], []

    def finish():
        if faces:
            used = sorted({i for f in faces for i in f})
            remap = {g: k for k, g in enumerate(used)}
            pts = np.array([V[g] for g in used], np.float64)
            meshes.append(MeshArrays(name, pts, [len(f) for f in faces], [remap[i] for f in faces for i in f]))

    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            V.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            idx = []
            for tok in t[1:]:
                if tok.startswith("#"):
                    break
                a = int(tok.split("/")[0])
                idx.append(a - 1 if a > 0 else len(V) + a)
            if idx:
                faces.append(idx)
        elif t[0] in ("o", "g"):
            finish()
            faces = []
            name = t[1] if len(t) > 1 else name
    finish()
    return meshes


def daylight(meshes, weather=None, quantities=("radiation",), mf=1, patchSamples=1, ground=0.2, north=0.0, offset=0.0,
             sides="front", onVertices=False, selfOcclude=True, cullBackfaces=False, virtualGround=True, lat=math.nan,
             lon=math.nan, tz=math.nan, t0=math.nan, t1=math.nan, stepMin=60, dayFrom=0.0, dayTo=1440.0, series=False,
             nthreads=None, ladybug=False):
    """The oracle's daylight run. weather: dict(lat, lon, tz, leap, dni, dhi) or None. Returns dict(values, series, ...)."""
    L = lib()
    q = [QUANTITIES.index(x) for x in quantities]
    o = DaylightOpt()
    for k, v in enumerate(q):
        o.quantities[k] = v
    o.nQ, o.mf, o.patchSamples, o.ground, o.north, o.offset = len(q), mf, patchSamples, ground, north, offset
    o.sides = {"front": 0, "back": 1, "both": 2}[sides]
    o.onVertices, o.selfOcclude, o.cullBackfaces, o.virtualGround = int(onVertices), int(selfOcclude), int(cullBackfaces), int(virtualGround)
    o.lat, o.lon, o.tz = lat, lon, tz
    o.time = Time(t0, t1, stepMin, dayFrom, dayTo)
    o.wantSeries = int(series)
    o.nthreads = nthreads or min(8, os.cpu_count() or 1)
    o.ladybug = int(ladybug)
    arr = (Mesh * len(meshes))(*[m.c() for m in meshes])
    if weather is not None:
        dni = np.ascontiguousarray(weather["dni"], np.float32)
        dhi = np.ascontiguousarray(weather["dhi"], np.float32)
        wp = (dni.ctypes.data, dhi.ctypes.data, len(dni), weather["lat"], weather["lon"], weather["tz"], int(weather.get("leap", False)))
    else:
        wp = (None, None, 0, math.nan, math.nan, math.nan, 0)
    ne, nc, tup, nd = C.c_long(), C.c_long(), C.c_int(), C.c_long()
    if L.orc_daylight_size(C.cast(arr, C.c_void_p), len(meshes), *wp, C.byref(o), C.byref(ne), C.byref(nc), C.byref(tup), C.byref(nd)):
        raise ValueError("oracle: bad run")
    nS = 2 if sides == "both" else 1
    values = np.zeros(ne.value * nc.value, np.float64)
    ser = np.zeros(max(1, ne.value * nS * tup.value) if series else 1, np.float64)
    out = DaylightOut()
    out.values = values.ctypes.data
    out.series = ser.ctypes.data if series else None
    err = C.create_string_buffer(512)
    if L.orc_daylight(C.cast(arr, C.c_void_p), len(meshes), *wp, C.byref(o), C.byref(out), err, 512):
        raise ValueError("oracle: " + err.value.decode())
    res = dict(values=values.reshape(ne.value, nc.value), durationHours=out.durationHours, nSky=out.nSky, nSun=out.nSun,
               nDir=out.nDir, Tup=out.Tup)
    if series:
        res["series"] = ser[: ne.value * nS * tup.value].reshape(ne.value * nS, tup.value)
    return res


def views(meshes, fov=(120.0, -70.0, 50.0), cell=2.0, jitter=True, far=1000.0, reach=(), distance=None, offset=0.0,
          sides="front", onVertices=False, selfOcclude=True, north=0.0, virtualGround=True, groundZ=math.nan,
          cullBackfaces=False, targets=None, nthreads=None, raw=False, transmission=False, glass=()):
    """The oracle's views run: the host steps of reference src/views.cpp (field :28-66, cell weights :126-136, labels,
    sensors, the launch frame :268-272, the columns :235-266) over the oracle's C kernels (orc_trace_view,
    orc_reduce_view). targets: {mesh name: group}. distance: a set of 'mean', 'min', 'group', 'isovist' or None for
    the default (isovist at 360, else mean and min). transmission: the glazing column, and rays pass the meshes named
    in glass (every face glass with tau > 0: label 255, views.cpp:86-106). Returns dict(values [N x C], columns, cls,
    dist when raw)."""
    L = lib()
    fovH, lo, hi = fov
    flat = lo == 0 and hi == 0
    nAz = max(1, round(fovH / cell))
    nEl = max(1, round((hi - lo) / cell))
    dist = set(distance) if distance is not None else ({"isovist"} if fovH == 360 else {"mean", "min"})
    if math.isnan(groundZ):
        groundZ = min(float(m.points[2::3].min()) for m in meshes if len(m.points))
    world = fovH == 360
    dAz = fovH / nAz
    dEl = (hi - lo) / nEl
    kDeg = math.pi / 180
    rowSin = np.zeros(2 * nEl, np.float32)
    wRow = np.zeros(nEl)
    tot = 0.0
    for k in range(nEl):
        top, bot = math.sin((hi - k * dEl) * kDeg), math.sin((hi - (k + 1) * dEl) * kDeg)
        rowSin[2 * k], rowSin[2 * k + 1] = bot, top
        wRow[k] = dAz * kDeg * (1 if flat else top - bot)
        tot += wRow[k] * nAz
    wRow /= tot
    # groups in first-appearance order; the label of a triangle: 2, or 3 + its mesh's group
    targets = targets or {}
    groups = []
    for m in meshes:
        g = targets.get(m.name)
        if g is not None and g not in groups:
            groups.append(g)
    nClass = 3 + len(groups)
    # scene: the blocking meshes recentred (as orc_run.c), labels per triangle in build order
    arr = (Mesh * len(meshes))(*[m.c() for m in meshes])
    c = np.zeros(3)
    L.orc_bbox_centre(C.cast(arr, C.c_void_p), len(meshes), c)
    blocks = [not m.sensor or selfOcclude for m in meshes]
    nv = sum(len(m.points) // 3 for m, b in zip(meshes, blocks) if b)
    ntCap = sum(int(np.maximum(m.counts - 2, 0).sum()) for m, b in zip(meshes, blocks) if b)
    V = np.zeros(3 * max(nv, 1), np.float32)
    I = np.zeros(3 * max(ntCap, 1), np.uint32)
    label = []
    vb = nt = 0
    skipped = C.c_int(0)
    for k, m in enumerate(meshes):
        if not blocks[k]:
            continue
        Vk = np.zeros(len(m.points), np.float32)
        Ik = np.zeros(3 * max(1, int(np.maximum(m.counts - 2, 0).sum())), np.uint32)
        n = L.orc_triangulate(C.byref(arr[k]), c, Vk if len(Vk) else np.zeros(3, np.float32), Ik, vb, C.byref(skipped))
        V[3 * vb:3 * vb + len(Vk)] = Vk
        I[3 * nt:3 * (nt + n)] = Ik[:3 * n]
        g = targets.get(m.name)
        label += [255 if transmission and m.name in glass else 2 if g is None else 3 + groups.index(g)] * n
        vb += len(m.points) // 3
        nt += n
    label = np.array(label if label else [2], np.uint8)
    nS = 2 if sides == "both" else 1
    sideCode = {"front": 0, "back": 1, "both": 2}[sides]
    nElem = sum((len(m.points) // 3 if onVertices else len(m.counts)) for m in meshes if m.sensor)
    nSens = nElem * nS
    P = np.zeros(3 * nSens, np.float32)
    N = np.zeros(3 * nSens, np.float32)
    at = 0
    for k, m in enumerate(meshes):
        if m.sensor:
            at += nS * L.orc_sensors(C.byref(arr[k]), int(onVertices), sideCode, offset, c, P[3 * at:], N[3 * at:])
    sc = L.orc_scene_new(V, nv, I, nt)
    rad = north * math.pi / 180
    v = View(nAz, nEl, -fovH / 2 if not world else 0.0, dAz, rowSin.ctypes.data, int(world), int(world and offset == 0 and not flat),
             int(jitter), int(sides == "back"), int(virtualGround), nS, groundZ - c[2], math.cos(rad), math.sin(rad),
             label.ctypes.data)
    K = nAz * nEl
    cls = np.zeros(K * nSens, np.uint8)
    dst = np.zeros(K * nSens, np.float32)
    L.orc_trace_view(sc, P, N, nSens, 0, C.byref(v), int(cullBackfaces), cls, dst, nthreads or min(8, os.cpu_count() or 1))
    L.orc_scene_free(sc)
    iso = "isovist" in dist
    nR = len(reach)
    gz = 1 if transmission else 0
    NA = 3 * nClass + 1 + nR + gz + (1 if iso else 0)
    sums = np.zeros(nSens * NA)
    reachF = np.array(reach if nR else [0.0], np.float32)
    L.orc_reduce_view(cls, dst, nSens, K, nAz, wRow, nClass, far, reachF, nR, gz, (2 if flat else 3) if iso else 0, C.byref(v), 0, sums)
    sums = sums.reshape(nSens, NA)
    # columns (views.cpp:235-266)
    names = ["sky", "ground", "obstruction"] + groups + (["glazing"] if transmission else [])
    if iso:
        names.append("area" if flat else "volume")
    mean, mn_, grp = "mean" in dist, "min" in dist, "group" in dist
    if mean:
        names.append("distance")
    if mn_:
        names.append("distance_min")
    for g in (groups if grp else []):
        if mean:
            names.append(g + "_distance")
        if mn_:
            names.append(g + "_distance_min")
    names += ["reach_%g" % D for D in reach]
    sfx = (["_front", "_back"] if nS == 2 else ["_back" if sides == "back" else "_front"])
    columns = [n + s for n in names for s in sfx]
    nearF, farF = float(np.float32(0.1)), float(np.float32(far))
    least = lambda m: math.nan if math.isinf(m) else 0.1 if m <= nearF else far if m >= farF else m
    logMean = lambda w, wl, m: math.nan if not w > 0 else far if m >= farF else math.exp(wl / w)
    out = np.zeros((nElem, len(columns)))
    for si in range(nSens):
        a = sums[si]
        e, col0 = si // nS, si % nS
        vals = []
        own = a[3 * nClass]
        w = sum(a[3 * k] for k in range(nClass))
        wl = sum(a[3 * k + 1] for k in range(nClass))
        mn = min(a[3 * k + 2] for k in range(nClass))
        total = w + own
        vals += [a[0] / total, a[3] / total, (a[6] + own) / total] + [a[3 * (3 + k)] / total for k in range(len(groups))]
        if transmission:
            vals.append(a[3 * nClass + 1 + nR] / total)
        if iso:
            vals.append(a[NA - 1])
        if mean:
            vals.append(logMean(w, wl, mn))
        if mn_:
            vals.append(least(mn))
        for k in (range(len(groups)) if grp else []):
            if mean:
                vals.append(logMean(a[3 * (3 + k)], a[3 * (3 + k) + 1], a[3 * (3 + k) + 2]))
            if mn_:
                vals.append(least(a[3 * (3 + k) + 2]))
        vals += [a[3 * nClass + 1 + r] / total for r in range(nR)]
        for j, x in enumerate(vals):
            out[e, col0 + j * nS] = x
    res = dict(values=out, columns=columns, nSens=nSens, K=K)
    if raw:
        res.update(cls=cls.reshape(K, nSens).T.copy(), dist=dst.reshape(K, nSens).T.copy(), sums=sums)
    return res


def sensors(mesh, ladybug=False, onVertices=False, sides="front", offset=0.0, centre=(0.0, 0.0, 0.0)):
    """make_sensors_into by the mode's rules: (P [n, 3], N [n, 3]) float32, or None for ladybug's n-gon error."""
    nS = 2 if sides == "both" else 1
    n = (len(mesh.points) // 3 if onVertices else len(mesh.counts)) * nS
    P = np.zeros(3 * max(n, 1), np.float32)
    N = np.zeros(3 * max(n, 1), np.float32)
    m = mesh.c()
    made = lib().orc_sensors_mode(C.byref(m), int(onVertices), {"front": 0, "back": 1, "both": 2}[sides], offset, int(ladybug),
                                  np.ascontiguousarray(centre, np.float64), P, N)
    if made < 0:
        return None
    return P[:3 * n].reshape(-1, 3), N[:3 * n].reshape(-1, 3)


def ladybug_sky(mf, jd, hour, dni, dhi, lat, lon, fileTz, grefl, dur, colOf=None, Tup=0):
    """Ladybug mode's gendaymtx -A emulation over the steps: dict(dir, dif, rad, gnd, dump [2, P + 1, 3], S, g)."""
    P = lib().orc_patch_count(mf)
    T = len(jd)
    jd = np.ascontiguousarray(jd, np.int32)
    hour = np.ascontiguousarray(hour, np.float64)
    dni = np.ascontiguousarray(dni, np.float32)
    dhi = np.ascontiguousarray(dhi, np.float32)
    d, f, r = np.zeros(P), np.zeros(P), np.zeros(P)
    g = C.c_double()
    dump = np.zeros(6 * (P + 1), np.float32)
    S = np.zeros(max(1, P * Tup), np.float32) if colOf is not None else None
    gS = np.zeros(max(1, Tup), np.float32) if colOf is not None else None
    col = np.ascontiguousarray(colOf, np.int32) if colOf is not None else None
    lib().orc_ladybug_sky(mf, T, jd.ctypes.data, hour.ctypes.data, dni.ctypes.data, dhi.ctypes.data, lat, lon, fileTz, grefl, dur,
                          None if col is None else col.ctypes.data, Tup, None if S is None else S.ctypes.data,
                          None if gS is None else gS.ctypes.data, d.ctypes.data, f.ctypes.data, r.ctypes.data, C.byref(g), dump.ctypes.data)
    out = dict(dir=d, dif=f, rad=r, gnd=g.value, dump=dump.reshape(2, P + 1, 3))
    if S is not None:
        out["S"] = S[:P * Tup].reshape(P, Tup)
        out["g"] = gS[:Tup]
    return out
