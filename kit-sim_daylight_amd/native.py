# SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
"""The daylight and views tools in-process through libdaylight_amd.so: the interface of the reference's Python binding
(bindings/python/daylight/native.py: SimulationMesh, Control, Result, run, usage; same names, argument meaning, errors
and stats keys) over the library's C ABI (include/daylight_c.h)::

    from native import SimulationMesh, run
    with SimulationMesh.load('tests/golden/box.obj') as box:
        r = run('daylight', ['--epw', 'site.epw', '--quantity', 'radiation'], box)
    print(r.manifest['columns'], r.values.shape)

run_series_device() is an addition of this library: a daylight run whose --timeseries output stays on the GPU; so is Server,
the serving loop (dlx_server_*): requests of both tools queued, admitted against an HBM budget, run on lanes of their
own over resident scenes, compatible daylight requests merged into one run.
"""
import ctypes as C
import json
import os
import weakref
from typing import NamedTuple

import numpy as np

import daylight_amd as _d

__all__ = ["Control", "DaylightError", "SimulationMesh", "Result", "Server", "run", "run_series_device", "usage", "gpu_counters",
           "count_fetches", "chunk_bytes", "read_epw"]

# dl_result_stats order (reference bindings/python/daylight/native.py:24-25)
_STATS = (("scene_reused", bool), ("prep_ms", float), ("sky_ms", float), ("build_ms", float), ("trace_ms", float),
          ("reduce_ms", float), ("bounce_ms", float), ("rays", int), ("launches", int), ("gas_bytes", int))
# dlx_gpu_counters order (include/daylight_c.h)
_COUNTERS = ("trace_ms", "trace_launches", "rays", "node_fetches", "tri_fetches", "counted_rays", "reduce_ms", "series_ms",
             "sky_ms", "build_ms", "views_ms", "last_trace_ms", "last_series_ms", "scene_bytes", "tree_height", "wave_iters",
             "traced", "graph_launches", "graph_captures", "leaf_iters")


class DaylightError(Exception):
    """A failed call: status 1 (usage: flags, names), 2 (runtime: a file, the GPU, 'cancelled'), and the core's message
    (reference bindings/python/daylight/_native.py:44-55)."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status

    def __reduce__(self):
        return DaylightError, (self.status, str(self))


def _check(status):
    if status != _d.DL_OK:
        raise DaylightError(status, (_d.lib().dl_last_error() or b"").decode(errors="replace"))


def _tool(tool):
    if tool not in ("daylight", "views"):
        raise ValueError(f"tool: 'daylight' or 'views', not {tool!r}")
    return _d.DL_DAYLIGHT if tool == "daylight" else _d.DL_VIEWS


class _Handle:
    def _own(self, handle, free, status):
        if not handle:
            _check(status)
        self._h = handle
        self._free = weakref.finalize(self, free, handle)
        self._free.atexit = False

    def close(self):
        self._free()
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class SimulationMesh(_Handle):
    """Meshes the core holds (dl_mesh), copied once and never changed; runs over the same objects reuse the GPU scene."""

    def __init__(self, name, points, counts, indices, slots=None):
        p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        c = np.ascontiguousarray(counts, np.int32).ravel()
        i = np.ascontiguousarray(indices, np.int32).ravel()
        s = None if slots is None else np.ascontiguousarray(slots, np.uint16).ravel()
        if s is not None and len(s) != len(c):
            raise ValueError(f"{name}: {len(s)} slots for {len(c)} faces")
        L = _d.lib()
        self._take(L.dl_mesh_create(name.encode(), p.ctypes.data, len(p), c.ctypes.data, len(c), i.ctypes.data, len(i),
                                    None if s is None else s.ctypes.data), _d.DL_USAGE)

    @classmethod
    def load(cls, path):
        m = cls.__new__(cls)
        m._take(_d.lib().dl_mesh_load(os.fsencode(path)), _d.DL_RUNTIME)
        return m

    def _take(self, handle, status):
        L = _d.lib()
        self._own(handle, L.dl_mesh_free, status)
        self.names = tuple(L.dl_mesh_name(handle, k).decode() for k in range(L.dl_mesh_count(handle)))


class Control(_Handle):
    """Cancel flag and progress of one run (dl_control)."""

    def __init__(self):
        self._own(_d.lib().dl_control_create(), _d.lib().dl_control_free, _d.DL_RUNTIME)

    def cancel(self):
        _d.lib().dl_control_cancel(self._h)

    @property
    def progress(self):
        return _d.lib().dl_control_progress(self._h)


class Result(NamedTuple):
    values: np.ndarray
    manifest: dict
    stats: dict


def usage(tool):
    return _d.lib().dl_usage(_tool(tool)).decode()


def _meshes(ms):
    ms = [ms] if isinstance(ms, SimulationMesh) else list(ms)
    if not all(isinstance(m, SimulationMesh) for m in ms):
        raise TypeError("sensors / context: a SimulationMesh or SimulationMesh objects")
    return ms


def _handles(ms):
    return (C.c_void_p * len(ms))(*[m._h for m in ms]) if ms else None


def _argv(args):
    argv = [os.fsencode(a) if isinstance(a, os.PathLike) else str(a).encode() for a in args]
    return len(argv), (C.c_char_p * max(1, len(argv)))(*argv)


class _ResultOwner:
    """Frees a dl_result once nothing views its values any more (the last reference to their buffer goes)."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        if self.handle:
            _d.lib().dl_result_free(self.handle)
            self.handle = None


def _take_result(out):
    """Result of a dl_result: its values viewed in place (dlx_result_data), the result freed with their last view."""
    L = _d.lib()
    owner = _ResultOwner(out)
    rows, cols = L.dl_result_rows(out), L.dl_result_cols(out)
    st = np.empty(_d.DL_STATS)
    _check(L.dl_result_stats(out, st.ctypes.data))
    manifest = L.dl_result_manifest(out).decode()
    ptr = L.dlx_result_data(out) if rows * cols else None
    if ptr:
        buf = (C.c_double * (rows * cols)).from_address(ptr)
        buf._owner = owner   # the values' buffer keeps the result alive
        values = np.frombuffer(buf, dtype=np.float64).reshape(rows, cols)
    else:
        values = np.empty((rows, cols))
    return Result(values, json.loads(manifest), {k: f(x) for (k, f), x in zip(_STATS, st.tolist())})


def run(tool, args, sensors, context=(), control=None):
    """One run of `tool` ('daylight' or 'views') with its CLI flags `args`, sensor and context meshes as SimulationMesh
    objects -> Result(values [N x C] float64, manifest dict, stats dict). DaylightError on failure."""
    t = _tool(tool)
    n, argv = _argv(args)
    sens, ctx = _meshes(sensors), _meshes(context)
    out = C.c_void_p()
    _check(_d.lib().dl_run(t, n, argv, _handles(sens), len(sens), _handles(ctx), len(ctx), control._h if control else None,
                           C.byref(out)))
    return _take_result(out)


def run_series_device(args, sensors, context=(), sink=None, control=None):
    """A daylight run with --timeseries kept on the GPU. sink(row0, rows, cols, device_ptr) is called per chunk of
    sensors (device_ptr: float32 [rows x cols] in device memory, valid during the call; 0 when there are no columns) and
    returns None / 0 to go on; sink None drops the series. -> Result as run()."""
    n, argv = _argv(args)
    sens, ctx = _meshes(sensors), _meshes(context)
    out = C.c_void_p()
    cb = None
    if sink is not None:
        def _cb(user, row0, rows, cols, ptr):
            try:
                rc = sink(row0, rows, cols, ptr or 0)
                return int(rc or 0)
            except Exception:   # an exception cannot cross the C callback: stop the run
                return 2
        cb = _d.SERIES_SINK(_cb)
    _check(_d.lib().dlx_run_series_device(n, argv, _handles(sens), len(sens), _handles(ctx), len(ctx),
                                          control._h if control else None, C.cast(cb, C.c_void_p) if cb else None, None,
                                          C.byref(out)))
    return _take_result(out)


def gpu_counters():
    """The process GPU's counters (dlx_gpu_counters) as a dict."""
    a = np.zeros(_d.DLX_COUNTERS)
    _check(_d.lib().dlx_gpu_counters(a.ctypes.data, len(a)))
    return dict(zip(_COUNTERS, a.tolist()))


def count_fetches(on):
    """Switch the traversal's counting variant on or off (dlx_count_fetches)."""
    _d.lib().dlx_count_fetches(1 if on else 0)


def chunk_bytes(n):
    """The per-chunk device buffer cap of the following runs (dlx_chunk_bytes; 0 restores the default)."""
    _d.lib().dlx_chunk_bytes(int(n))


def read_epw(path):
    """The weather file as a --epw run reads it (dlx_read_epw): (dni, dhi) float32 arrays and (lat, lon, tz)."""
    L = _d.lib()
    n = C.c_int64()
    _check(L.dlx_read_epw(os.fsencode(path), None, None, 0, C.byref(n), None))
    dni, dhi, loc = np.zeros(n.value, np.float32), np.zeros(n.value, np.float32), np.zeros(3)
    _check(L.dlx_read_epw(os.fsencode(path), dni.ctypes.data, dhi.ctypes.data, n.value, C.byref(n), loc.ctypes.data))
    return dni, dhi, tuple(loc.tolist())


# dlx_server_stats order (include/daylight_c.h)
_SERVER_STATS = ("submitted", "completed", "failed", "runs", "merged", "peak_queue", "peak_bytes", "budget", "lanes", "peak_running")


class Server(_Handle):
    """The serving loop (dlx_server_*): submit(tool, args, sensors, context) queues a request and returns its id at once
    (DaylightError for flag errors); wait(id) blocks until it is done -> (Result as run(), times dict: queued_ms, run_ms,
    batch). The meshes of a request must stay open until it is waited for."""

    def __init__(self, lanes=2, hbm_budget=0, batch_sensors=0):
        L = _d.lib()
        self._own(L.dlx_server_open(int(lanes), int(hbm_budget), int(batch_sensors)), L.dlx_server_close, _d.DL_RUNTIME)
        self._keep = {}

    def submit(self, tool, args, sensors, context=()):
        t = _tool(tool)
        n, argv = _argv(args)
        sens, ctx = _meshes(sensors), _meshes(context)
        rid = _d.lib().dlx_server_submit(self._h, t, n, argv, _handles(sens), len(sens), _handles(ctx), len(ctx))
        if rid < 0:
            msg = (_d.lib().dl_last_error() or b"").decode(errors="replace")
            raise DaylightError(_d.DL_USAGE, msg)
        self._keep[rid] = (sens, ctx)   # the handles stay alive until the request is waited for
        return rid

    def wait(self, rid):
        out = C.c_void_p()
        times = np.zeros(3)
        try:
            _check(_d.lib().dlx_server_wait(self._h, int(rid), C.byref(out), times.ctypes.data))
        finally:
            self._keep.pop(rid, None)
        return _take_result(out), {"queued_ms": float(times[0]), "run_ms": float(times[1]), "batch": int(times[2])}

    def stats(self):
        a = np.zeros(_d.DLX_SERVER_STATS)
        _check(_d.lib().dlx_server_stats(self._h, a.ctypes.data, len(a)))
        return dict(zip(_SERVER_STATS, a.tolist()))
