#!/usr/bin/env python3
# SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
"""Throughput of the daylight trace-and-accumulate path on one MI355X: sensor-hours/s of an annual run with its series.

Headline workload (the batched precision-form run): the seeded 3M-triangle city (250,000 extruded boxes on a 500 x 500
grid of 50 m plots, footprint sides U(10, 40) m, heights lognormal with median 15 m and sigma 0.5) as context, and three
sensor meshes of 1,000 / 150,000 / 850,000 faces (the roofs and facades of the central buildings in 1 m quads:
1,001,000 sensors) in ONE run, so one launch of each kernel covers all three. Weather: the library's clear sky
(Ineichen, turbidity 3) at lat 52.52 lon 13.41 tz 1 with an hourly cloud cover c ~ Beta(2, 3), written to an EPW. Run:
--sky-density 4 (2305 patches), 8760 hourly steps, radiation, direct, diffuse and illuminance, plus the radiometric
hourly series (1,001,000 x 4,456 non-night steps) by the default FP16x3 contraction on MFMA, kept on the GPU
(dlx_run_series_device; file output excluded from the timing). The sensors cast no shade of their own
(--no-self-occlude): they lie on the context boxes' faces.

A step is one such run through the C ABI with the context scene already resident on the GPU: the sky and sun of every
step (FP64), the visibility traversal, the per-sensor reductions and the series contraction, plus the host work of a
run. The timed region holds exactly --steps runs, bracketed by torch.cuda.synchronize().

The line also carries: the contraction methods side by side on the same workload (FP32 = the reference's arithmetic,
BF16x3, BF16x6, FP16x3) with their series error; the per-column error against the oracle's FP64 run of a sensor sample; the
traversal's roofline (HBM bytes from the committed rocprofv3 counters over the live launch time, and its limiter) and the
contraction's MFMA roofline (executed and algorithmic flops over the live launch time); traced and slot Grays/s; the
urban block (577 patches), the interior views and the byte-mode legs; the reference's own --bench commands through
this library's CLI; and the oracle's CPU time on a bounded sample of the headline workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import platform
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kit-sim_daylight_amd"))

import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3
NODE_BYTES, TRI_BYTES = 112, 48
PRODUCTS = {"bf16x3": 3, "bf16x6": 6, "fp16x3": 3}   # MFMA products per k-step and tile of each split method
DEFAULT_METHOD = "fp16x3"                              # the library's default contraction (DAYLIGHT_SERIES unset)   # a 4-wide record visit loads 7 x 16 bytes; a triangle 3 x 16
PROFILES = os.path.join(ROOT, "profiles")


def rel_err(x, want):
    """max |x - want| / max(|want|, 1e-3 max |want|) over the sample (the validation scale rule, reference
    validate/README.md:160)"""
    want = np.asarray(want, np.float64)
    scale = np.maximum(np.abs(want), 1e-3 * np.abs(want).max())
    return float(np.max(np.abs(np.asarray(x, np.float64) - want) / np.where(scale > 0, scale, 1.0)))


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return model, n


def profile_value(name, key):
    """A 'key: value' line of a committed profile summary, or None."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1].split()[0])
    return None


_hip = None


def device_rows(ptr, rows, cols):
    import ctypes as C
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    buf = np.empty((rows, cols), np.float32)
    if _hip.hipMemcpy(buf.ctypes.data, C.c_void_p(ptr), buf.nbytes, 2):
        raise RuntimeError("hipMemcpy of the series failed")
    return buf


def series_of(native, args, sensors, context, rows):
    """A run with the series on the device -> (Result, its first `rows` series rows)."""
    got = []

    def sink(row0, n, cols, ptr):
        take = min(n, rows - row0)
        if take > 0:
            got.append(device_rows(ptr, take, cols))
        return 0

    r = native.run_series_device(args, sensors=sensors, context=context, sink=sink)
    return r, np.concatenate(got)


def oracle_run(meshes, dni, dhi, nthreads, **kw):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    wx = dict(lat=W.LAT, lon=W.LON, tz=W.TZ, leap=False, dni=dni, dhi=dhi)
    return orc.daylight([orc.MeshArrays(n, *m, sensor=s) for n, m, s in meshes], weather=wx, selfOcclude=False, nthreads=nthreads, **kw)


def set_method(name):
    if name is None:
        os.environ.pop("DAYLIGHT_SERIES", None)
    else:
        os.environ["DAYLIGHT_SERIES"] = name


def timed_runs(native, torch, fn, n):
    torch.cuda.synchronize()
    c0 = native.gpu_counters()
    t0 = time.perf_counter()
    lat = []
    for _ in range(n):
        ts = time.perf_counter()
        r = fn()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c1 = native.gpu_counters()
    return r, wall, {k: c1[k] - c0[k] for k in c1}, lat


# ---------------------------------------------------------------- the CPU baseline (the oracle, on a bounded sample)

def cpu_baseline(ctx, mesh, dni, dhi, threads, model, target_s):
    """The oracle (host BVH, FP64 sums and series, the same ray rules) on the first faces of the 150,000-face mesh, the
    full context, 2305 patches and the full year; two probes give its fixed cost (the year's sky, the host BVH of the
    3M triangles) and its cost per sensor, and the sample fills target_s."""
    def sample(nf):
        sub = W.submesh(mesh, 0, nf)
        t0 = time.perf_counter()
        r = oracle_run([("sensors", sub, True), ("city", ctx, False)], dni, dhi, threads, quantities=("radiation",), mf=4, series=True)
        return time.perf_counter() - t0, r["durationHours"]

    n0, n1 = 256, 2048   # probes large enough that the per-sensor part stands out of the fixed part's noise
    t0, _ = sample(n0)
    t1, _ = sample(n1)
    per = max((t1 - t0) / (n1 - n0), t1 / n1 / 4, 1e-6)
    fixed = max(t0 - n0 * per, 0.0)
    nf = int(min(len(mesh[1]), max(n1, (target_s - fixed) / per)))
    t, hours = sample(nf)
    return {"value": round(nf * hours / t, 1), "unit": "sensor-hours/s", "cores": threads, "kind": "port",
            "cpu": model, "nproc": threads,
            "sample": "%d sensor faces of the 150,000-face mesh, the full 3M-triangle context, 2305 patches, full year (%.0f h); "
                      "the oracle: host BVH + trace + FP64 sums + FP64 series, %d threads (all the host's), %.1f s including "
                      "the host BVH build and the year's sky (%.1f s at %d sensors)" % (nf, hours, threads, t, t0, n0)}


# ---------------------------------------------------------------- legs

VIEW_ARGS = W.INTERIOR_VIEW_ARGS


def views_leg(native, torch, threads, steps=3, n_check=64):
    """The seeded interior (~2M triangles) seen all around from a 0.5 m floor grid of 100,000 sensors at eye height
    1.2 m, 2-degree cells (10,800 rays each): whole views runs through the C ABI on the resident scene; the first
    n_check sensors' columns checked against the oracle's run of them."""
    P, C, I = W.interior()
    gp, gc, gi = W.floor_grid()
    room = native.SimulationMesh("interior", P, C, I)
    grid = native.SimulationMesh("grid", gp, gc, gi)
    r = native.run("views", VIEW_ARGS, grid, context=room)   # builds and keeps the scene
    n = r.values.shape[0]
    r, wall, d, _ = timed_runs(native, torch, lambda: native.run("views", VIEW_ARGS, grid, context=room), steps)
    assert r.stats["scene_reused"]
    wall /= steps
    rays = n * 10800
    # accuracy: the first n_check sensors of the grid as a run of their own (same element indices: the jitter seeds agree)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    sub = W.submesh((gp, gc, gi), 0, n_check)
    rs = native.run("views", VIEW_ARGS, native.SimulationMesh("grid", *sub), context=room)
    o = orc.views([orc.MeshArrays("grid", *sub, sensor=True), orc.MeshArrays("interior", P, C, I, sensor=False)], fov=(360.0, -70.0, 50.0),
                  cell=2.0, jitter=True, far=1000.0, reach=(6.0, 20.0, 50.0), distance={"isovist", "mean", "min"}, offset=1.2,
                  selfOcclude=False, nthreads=threads)
    same_rows = bool(np.array_equal(rs.values, r.values[:n_check], equal_nan=True))
    errs = {}
    for c, name in enumerate(o["columns"]):
        g, w = rs.values[:, c], o["values"][:, c]
        ok = ~np.isnan(w)
        errs[name] = 0.0 if not ok.any() else float(np.max(np.abs(g[ok] - w[ok]) / np.maximum(np.abs(w[ok]), 1e-300)))
    return {"workload": "seeded interior, %d triangles (rooms U(3,8) m, door openings, clutter U(0.3,2) m), %d floor sensors at "
                        "0.5 m, eye 1.2 m, 360 x 120 degrees in 2-degree cells, jitter, isovist + mean/min + reach 6/20/50 m"
                        % (2 * len(C), n),
            "rays_per_run": rays, "runs": steps, "ms_per_run": round(1e3 * wall, 2),
            "grays_per_s": round(rays / wall / 1e9, 3),
            "trace_ms_per_run": round(d["views_ms"] / steps, 3),
            "reduce_ms_per_run": round(d["reduce_ms"] / steps, 3),
            "columns": r.manifest["columns"],
            "accuracy": {"sensors": n_check, "max_rel_err_per_column": errs, "rows_equal_in_the_full_run": same_rows,
                         "bar": "shares exact, distances 1e-12 (device vs host log)"}}


def urban_block_leg(native, torch, tmp, epw, dni, dhi, threads, steps=5):
    """The urban block (400 boxes, 4,800 triangles; the 40 central buildings' roofs and facades in 1 m quads: 107,344
    sensors), 577 patches, the same cloudy year: each contraction method, with the series error of 64 sensors."""
    boxes, central = W.city()
    ctx = W.box_mesh(boxes)
    sens = W.sensor_mesh(boxes[central])
    S = native.SimulationMesh("sensors", *sens)
    Cx = native.SimulationMesh("city", *ctx)
    args = ["--epw", epw, "--sky-density", "2", "--quantity", "radiation", "--timeseries", os.path.join(tmp, "block.npy"), "--no-self-occlude"]
    o = oracle_run([("sensors", W.submesh(sens, 0, 64), True), ("city", ctx, False)], dni, dhi, threads, quantities=("radiation",), mf=2,
                   series=True)
    out = {"workload": "seeded urban block: 400 boxes (4,800 triangles), roofs + facades of the 40 central buildings in <=1 m quads "
                       "(%d sensors), 577 patches, 8760 h, the cloudy year, radiation + device series" % len(sens[1]), "methods": {}}
    hours = 8760.0
    for name in ("fp32", "bf16x3", "bf16x6", "fp16x3"):
        set_method(name)
        native.run_series_device(args, sensors=S, context=Cx)
        _, wall, d, _ = timed_runs(native, torch, lambda: native.run_series_device(args, sensors=S, context=Cx), steps)
        _, E = series_of(native, args, S, Cx, 64)
        out["methods"][name] = {"ms_per_step": round(1e3 * wall / steps, 2), "series_ms_per_step": round(d["series_ms"] / steps, 3),
                                "sensor_hours_per_s": round(len(sens[1]) * hours * steps / wall, 1),
                                "series_max_rel_err": rel_err(E, o["series"])}
    set_method(None)
    m = out["methods"]
    for name in ("bf16x6", "fp16x3"):
        out[name + "_over_fp32"] = {"step": round(m["fp32"]["ms_per_step"] / m[name]["ms_per_step"], 2),
                                    "series": round(m["fp32"]["series_ms_per_step"] / m[name]["series_ms_per_step"], 2)}
    return out


def transmission_leg(native, torch, epw, tmp):
    """The urban block with every context face glass (tau 0.64, Radiance glass) under --transmission on: every ray walks
    every crossing (byte mode)."""
    boxes, central = W.city()
    S = native.SimulationMesh("sensors", *W.sensor_mesh(boxes[central]))
    Cx = native.SimulationMesh("city", *W.box_mesh(boxes))
    targs = ["--epw", epw, "--sky-density", "2", "--quantity", "radiation", "--timeseries", os.path.join(tmp, "t.npy"), "--no-self-occlude",
             "--transmission", "on", "--optics", "city=0,0.64,glass"]
    native.run_series_device(targs, sensors=S, context=Cx)
    rt, wall, d, _ = timed_runs(native, torch, lambda: native.run_series_device(targs, sensors=S, context=Cx), 2)
    assert rt.stats["scene_reused"]
    return {"workload": "the urban block with every context face glass (tau 0.64), --transmission on",
            "ms_per_step": round(1e3 * wall / 2, 2), "trace_ms_per_step": round(d["trace_ms"] / 2, 3),
            "grays_per_s": {"slots": round(d["rays"] / (d["trace_ms"] * 1e-3) / 1e9, 3),
                            "traced": round(d["traced"] / (d["trace_ms"] * 1e-3) / 1e9, 3)}}


SERVING_VIEW_ARGS = ["--fov", "360", "--cell", "2", "--far", "1000", "--jitter", "on", "--distance", "isovist,mean,min"]


def serving_leg(native, epw, n=100, lanes=3, seed=11, load_factor=0.8):
    """The serving mix over the resident 5M-triangle city (417,316 boxes): n requests, 70/30 daylight/views, sensor
    faces log-uniform in [1k, 500k] cut from the central buildings' faces, time span instant / 24 h / year and sky
    density 1 / 2 / 4 drawn uniformly (workloads.serving_load). First one request at a time through dl_run (its service
    times, and its latencies under the arrivals below, FIFO); then Poisson arrivals at load_factor / (its mean service
    time) replayed against a server of `lanes` lanes (admission by HBM budget, compatible daylight requests merged);
    then every request at once (capacity). Each server result is compared with the dl_run one."""
    import threading
    boxes, central = W.city_5m()
    ctx = W.box_mesh(boxes)
    pool = W.city_sensor_meshes(boxes, central, (("pool", 500_000),))[0][1]
    Cx = native.SimulationMesh("city5m", *ctx)
    load = W.serving_load(n, seed)
    meshes = [native.SimulationMesh("request%d" % k, *W.submesh(pool, r["f0"], r["f0"] + r["faces"])) for k, r in enumerate(load)]

    def args_of(r):
        if r["tool"] == "views":
            return SERVING_VIEW_ARGS
        return ["--epw", epw, "--sky-density", str(r["mf"]), "--quantity", "radiation,illuminance", "--no-self-occlude"] + W.SPANS[r["span"]]

    args = [args_of(r) for r in load]
    native.run("daylight", args_of(dict(load[0], tool="daylight")), meshes[0], Cx)   # the scene resident before either path
    service, want = [], []
    for k, r in enumerate(load):
        t0 = time.perf_counter()
        want.append(native.run(r["tool"], args[k], meshes[k], Cx).values)
        service.append(time.perf_counter() - t0)
    rate = load_factor / float(np.mean(service))
    arrive = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, n))
    fin, serial_lat = 0.0, []
    for a, sv in zip(arrive, service):
        fin = max(a, fin) + sv
        serial_lat.append(fin - a)
    srv = native.Server(lanes=lanes)
    try:
        got = [None] * n

        def waiter(k, rid):
            got[k] = srv.wait(rid)

        threads = []
        t0 = time.perf_counter()
        for k in range(n):
            d = arrive[k] - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            th = threading.Thread(target=waiter, args=(k, srv.submit(load[k]["tool"], args[k], meshes[k], Cx)))
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        replay = time.perf_counter() - t0
        st = srv.stats()
        server_lat = [(t["queued_ms"] + t["run_ms"]) / 1e3 for _, t in got]
        identical = all(np.array_equal(got[k][0].values, want[k], equal_nan=True) for k in range(n))
        t1 = time.perf_counter()
        ids = [srv.submit(load[k]["tool"], args[k], meshes[k], Cx) for k in range(n)]
        burst = [srv.wait(i) for i in ids]
        burst_s = time.perf_counter() - t1
        identical = identical and all(np.array_equal(burst[k][0].values, want[k], equal_nan=True) for k in range(n))
        st_burst = srv.stats()
    finally:
        srv.close()
    for m in meshes:
        m.close()
    pct = lambda x, q: round(1e3 * float(np.percentile(x, q)), 1)
    return {"workload": "seeded 5M-triangle city (417,316 boxes, %d triangles) resident; %d requests, %d daylight / %d views; sensor "
                        "faces log-uniform 1k-500k (mean %d); spans instant / 24 h / year and sky density 1 / 2 / 4 uniform; "
                        "views 360 x 120 degrees in 2-degree cells" % (2 * len(ctx[1]), n, sum(r["tool"] == "daylight" for r in load),
                                                                      sum(r["tool"] == "views" for r in load), int(np.mean([r["faces"] for r in load]))),
            "arrival_rate_per_s": round(rate, 3),
            "one_at_a_time": {"service_ms_mean": round(1e3 * float(np.mean(service)), 1), "capacity_requests_per_s": round(n / float(np.sum(service)), 3),
                              "p50_ms": pct(serial_lat, 50), "p99_ms": pct(serial_lat, 99)},
            "server": {"lanes": lanes, "p50_ms": pct(server_lat, 50), "p99_ms": pct(server_lat, 99), "replay_s": round(replay, 2),
                       "capacity_requests_per_s": round(n / burst_s, 3), "runs": st_burst["runs"], "merged_requests": st_burst["merged"],
                       "peak_running": st_burst["peak_running"], "budget_gb": round(st["budget"] / 1e9, 1)},
            "identical_to_one_at_a_time": identical,
            "note": "latency = submit to result (the flags and weather parse included, as in dl_run's time); one_at_a_time latencies "
                    "are its measured service times queued FIFO under the same arrivals"}


BOSTON = ["--clear-sky", "--lat", "42.35", "--lon", "-71.07", "--tz", "-5"]
# (label, tool arguments, the reference's published figure on an RTX 2060: docs/performance.md)
REFERENCE_BENCHES = [
    ("daylight --bench 100000 300000", ["daylight", "--bench", "100000", "300000"] + BOSTON,
     "2.0 GRays/s, trace 0.24 s, wall 0.6 s (performance.md:7)"),
    ("daylight --bench 1000000 3000000", ["daylight", "--bench", "1000000", "3000000"] + BOSTON,
     "2.8 GRays/s, trace 1.7 s, wall 2.6 s (performance.md:8)"),
    ("daylight --bench 1000000 12000000", ["daylight", "--bench", "1000000", "12000000"] + BOSTON,
     "1.9 GRays/s, trace 2.5 s, wall 4.2 s (performance.md:9)"),
    ("daylight --bench 1000000 12000000 --sky-density 2", ["daylight", "--bench", "1000000", "12000000", "--sky-density", "2"] + BOSTON,
     "1.9 GRays/s, trace 3.0 s, wall 4.8 s (performance.md:10)"),
    ("daylight --bench 1000000 12000000 --transmission on --optics bench_boxes/1=0.2,0.5,porous",
     ["daylight", "--bench", "1000000", "12000000", "--transmission", "on", "--optics", "bench_boxes/1=0.2,0.5,porous"] + BOSTON,
     "1.65 GRays/s, trace 2.9 s, wall 5.2 s (performance.md:15-19)"),
    ("daylight --bench 100000 300000 --transmission on --optics bench_boxes=0.2,0.5,porous",
     ["daylight", "--bench", "100000", "300000", "--transmission", "on", "--optics", "bench_boxes=0.2,0.5,porous"] + BOSTON,
     "0.5 GRays/s, every box porous (performance.md:19-20)"),
    ("daylight --bench 1000000 120000 --bounces 1", ["daylight", "--bench", "1000000", "120000", "--bounces", "1"] + BOSTON,
     "reflected-light pass 2.1 s, 494 M paths/s, wall 5.6 s (performance.md:29)"),
    ("daylight --bench 1000000 120000 --bounces 2", ["daylight", "--bench", "1000000", "120000", "--bounces", "2"] + BOSTON,
     "reflected-light pass 2.2 s, 455 M paths/s, wall 5.5 s (performance.md:30)"),
    ("daylight --bench 1000000 120000 --sky-density 4 --bounces 1",
     ["daylight", "--bench", "1000000", "120000", "--sky-density", "4", "--bounces", "1"] + BOSTON,
     "reflected-light pass 2.1 s, 485 M paths/s, wall 7.9 s (performance.md:31)"),
    ("views --bench 1000000 10000000", ["views", "--bench", "1000000", "10000000"],
     "1.98 GRays/s, trace 1.84 s, reduce 0.66 s, run 3.58-3.66 s (performance.md:105)"),
    ("views --bench 1000000 10000000 --fov 360", ["views", "--bench", "1000000", "10000000", "--fov", "360"],
     "trace 4.3-4.7 s, reduce 2.35 s, run 8.3 s (performance.md:122-129)"),
]


def reference_benches():
    """Each of the reference's published --bench runs with this library's CLI (a child process on the same GPU): its
    printed rays, trace time and run total. The annual Boston EPW of the reference's table is not available here; the
    Boston clear sky stands in (the same sun steps; radiation + sun hours)."""
    import re
    import subprocess
    lib = os.path.join(ROOT, "kit-sim_daylight_amd", "lib")
    out = []
    for label, argv, ref in REFERENCE_BENCHES:
        t0 = time.perf_counter()
        p = subprocess.run([os.path.join(lib, argv[0])] + argv[1:], capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        err = p.stderr
        row = {"command": label, "reference": ref, "exit": p.returncode, "process_s": round(wall, 2)}
        m = re.search(r"([0-9.e+]+) rays in ([0-9.]+) ms \(([0-9.]+) GRays/s\)", err)
        if m:
            row.update(rays=float(m.group(1)), trace_s=round(float(m.group(2)) / 1e3, 3), grays_per_s=float(m.group(3)))
        m = re.search(r"([0-9.e+]+) traced \(([0-9.]+) GRays/s traced\)", err)
        if m:
            row.update(traced=float(m.group(1)), traced_grays_per_s=float(m.group(2)))
        m = re.search(r"run ([0-9.]+) s total", err)
        if m:
            row["run_s"] = float(m.group(1))
        m = re.search(r"reduce ([0-9.]+) ms", err)
        if m:
            row["reduce_s"] = round(float(m.group(1)) / 1e3, 3)
        m = re.search(r"([0-9.e+]+) paths in ([0-9.]+) ms", err)
        if m:
            row.update(paths=float(m.group(1)), bounce_s=round(float(m.group(2)) / 1e3, 3),
                       paths_per_s=round(float(m.group(1)) / (float(m.group(2)) * 1e-3), 1))
        if p.returncode:
            row["stderr"] = err[-400:]
        print("bench: %s" % json.dumps(row), file=sys.stderr, flush=True)
        out.append(row)
    return out


# ---------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-methods", action="store_true", help="skip the contraction methods side by side")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the oracle's check of a sensor sample")
    ap.add_argument("--no-count", action="store_true", help="skip the traversal's counted run")
    ap.add_argument("--no-legs", action="store_true", help="skip the urban-block, views and byte-mode legs")
    ap.add_argument("--no-reference-benches", action="store_true", help="skip the reference's --bench runs through the CLI")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--only", choices=("headline", "views", "block", "transmission", "serving"), help="run one leg alone (profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (the HBM counters' scale) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported (the path shards by sensor rows, "
              "see the sharding notes)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        sys.exit("bench.py: no GPU visible")
    if a.calibrate_copy:
        src = torch.ones(a.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.calibrate_copy << 20, "copies": 5}), flush=True)
        return
    import native
    model, threads = cpu_info()
    tmp = tempfile.mkdtemp(prefix="dlbench")
    epw, dni, dhi = W.weather_epw(native, tmp)
    if a.only == "views":
        print(json.dumps({"views": views_leg(native, torch, threads)}), flush=True)
        return
    if a.only == "block":
        print(json.dumps({"urban_block": urban_block_leg(native, torch, tmp, epw, dni, dhi, threads, steps=a.steps)}), flush=True)
        return
    if a.only == "transmission":
        print(json.dumps({"transmission": transmission_leg(native, torch, epw, tmp)}), flush=True)
        return
    if a.only == "serving":
        print(json.dumps({"serving": serving_leg(native, epw)}), flush=True)
        return

    boxes, central = W.city_3m()
    ctx = W.box_mesh(boxes)
    meshes = W.city_sensor_meshes(boxes, central)
    n_sens = sum(len(m[1]) for _, m in meshes)
    quantities = ["radiation", "direct", "diffuse", "illuminance"]
    args = ["--epw", epw, "--sky-density", "4", "--quantity", ",".join(quantities), "--timeseries", os.path.join(tmp, "series.npy"),
            "--no-self-occlude"]
    print("bench: %d context triangles, sensor meshes %s; warmup %d, steps %d"
          % (2 * len(ctx[1]), [(n, len(m[1])) for n, m in meshes], a.warmup, a.steps), file=sys.stderr, flush=True)
    Cx = native.SimulationMesh("city", *ctx)
    S = [native.SimulationMesh(n, *m) for n, m in meshes]

    def step():
        return native.run_series_device(args, sensors=S, context=Cx)

    set_method(None)                  # the default contraction (FP16x3 on MFMA)
    r0 = step()
    assert r0.values.shape == (n_sens, len(quantities)), r0.values.shape
    hours = float(r0.manifest["duration_h"])
    for _ in range(a.warmup):
        step()
    r, wall, d, lat = timed_runs(native, torch, step, a.steps)
    assert r.stats["scene_reused"], "the context scene was rebuilt inside the timed region"
    sensor_hours = n_sens * hours
    value = sensor_hours * a.steps / wall
    launches = d["trace_launches"] / a.steps
    trace_launch_ms = d["trace_ms"] / max(d["trace_launches"], 1)
    series_ms = d["series_ms"] / a.steps
    print("bench: %.1f ms/step, %.4g sensor-hours/s, trace %.1f ms/step over %d launches, series %.1f ms/step"
          % (1e3 * wall / a.steps, value, d["trace_ms"] / a.steps, launches, series_ms), file=sys.stderr, flush=True)

    # the contraction's shape: K = 2305 sky directions (one sample per patch), T = the non-night steps
    _, E0 = series_of(native, args, S[:1], Cx, 1)
    Tup = E0.shape[1]
    P = 2305
    alg_flops = 2.0 * n_sens * P * Tup
    kp, tp, np_ = (P + 31) // 32 * 32, (Tup + 127) // 128 * 128, (n_sens + 127) // 128 * 128
    exec_flops = PRODUCTS[DEFAULT_METHOD] * 2.0 * np_ * kp * tp   # FP16x3: three products over the padded tiles
    contraction = {"kernel": "k_series_mfma<2, WM, fp16> (FP16x3, v_mfma_f32_32x32x16_f16)", "bound": "mfma",
                   "note": "executed = the three products over the padded shape, counting the slices of all-zero coefficients "
                           "that blocks step over as done (about 12 % of them on this workload; the counted MFMA ops are in "
                           "profiles/k_series_mfma_city3m_counters.txt); the series stage time includes the operand splits",
                   "executed_tflops": round(exec_flops / (series_ms * 1e-3) / 1e12, 1),
                   "achieved": round(exec_flops / (series_ms * 1e-3) / 1e12, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                   "frac": round(exec_flops / (series_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
                   "algorithmic_tflops": round(alg_flops / (series_ms * 1e-3) / 1e12, 1),
                   "algorithmic_frac": round(alg_flops / (series_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
                   "series_ms_per_step": round(series_ms, 3), "shape": [n_sens, P, Tup]}

    # the traversal's counted fetches (one untimed run of the counting variant): its cache-level traffic
    cache = None
    if not a.no_count:
        native.count_fetches(True)
        k0 = native.gpu_counters()
        step()
        k1 = native.gpu_counters()
        native.count_fetches(False)
        fetch_bytes = (k1["node_fetches"] - k0["node_fetches"]) * NODE_BYTES + (k1["tri_fetches"] - k0["tri_fetches"]) * TRI_BYTES
        waves = k1["wave_iters"] - k0["wave_iters"]
        leaf_waves = k1["leaf_iters"] - k0["leaf_iters"]
        cache = {"bytes_per_launch": round(fetch_bytes / launches), "gbs": round(fetch_bytes / launches / (trace_launch_ms * 1e-3) / 1e9, 1),
                 "simd_efficiency": None if not waves else round((k1["node_fetches"] - k0["node_fetches"]) / (64.0 * waves), 3),
                 "leaf_simd_efficiency": None if not leaf_waves else round((k1["tri_fetches"] - k0["tri_fetches"]) / (64.0 * leaf_waves), 3),
                 "leaf_iters_per_wave_iter": None if not waves else round(leaf_waves / waves, 3),
                 "note": "the counted 4-wide record (112 B) and triangle (48 B) fetches: served by L1/L2, not HBM"}
    traffic = profile_value("k_bits_city3m_traffic.txt", "traffic_bytes_per_launch")
    hbm = None if traffic is None else traffic / (trace_launch_ms * 1e-3) / 1e9
    roofline = {"kernel": "k_bits (visibility traversal, the step's largest kernel)", "bound": "valu",
                "achieved": None if hbm is None else round(hbm, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None if hbm is None else round(hbm / HBM_PEAK_GBS, 5), "traffic": traffic,
                "launch_ms": round(trace_launch_ms, 3),
                "limiter": {"valu_busy": profile_value("k_bits_city3m_counters.txt", "valu_busy"),
                            "valu_lanes_active": profile_value("k_bits_city3m_counters.txt", "valu_lanes_active"),
                            "source": "profiles/k_bits_city3m_counters.txt"},
                "note": "achieved = HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size: "
                        "profiles/k_bits_city3m_traffic.txt) / the live mean launch time; the tree is served from L2 / MALL, "
                        "so HBM is far from binding: the VALU issue is (counters above)",
                "cache": cache}

    # accuracy: the first sensors of the 1,000-face mesh as a run of their own, against the oracle's FP64 run
    acc = None
    methods = {}
    n_acc = 256
    sub = W.submesh(meshes[0][1], 0, n_acc)
    Ssub = native.SimulationMesh(meshes[0][0], *sub)
    if not a.no_accuracy:
        o = oracle_run([(meshes[0][0], sub, True), ("city", ctx, False)], dni, dhi, threads, quantities=tuple(quantities), mf=4, series=True)
        rs, Es = series_of(native, args, [Ssub], Cx, n_acc)
        acc = {"sensors": n_acc, "columns": {q: rel_err(rs.values[:, c], o["values"][:, c]) for c, q in enumerate(quantities)},
               "series_" + DEFAULT_METHOD: rel_err(Es, o["series"]), "bar": 2e-6}
        acc["within_bar"] = bool(max(list(acc["columns"].values()) + [acc["series_" + DEFAULT_METHOD]]) <= 2e-6)
        print("bench: accuracy %s" % json.dumps(acc), file=sys.stderr, flush=True)
    # the contraction methods side by side on the headline workload
    if not a.no_methods:
        for name in ("fp32", "bf16x3", "bf16x6", "fp16x3"):
            set_method(name)
            step()
            _, wm, dm, _ = timed_runs(native, torch, step, 2)
            sm = dm["series_ms"] / 2
            row = {"ms_per_step": round(1e3 * wm / 2, 2), "series_ms_per_step": round(sm, 3),
                   "sensor_hours_per_s": round(sensor_hours * 2 / wm, 1), "series_sensor_hours_per_s": round(sensor_hours / (sm * 1e-3), 1)}
            if name == "fp32":
                row["fp32_peak_frac"] = round(alg_flops / (sm * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4)
            else:
                row["mfma_executed_frac"] = round(PRODUCTS[name] * 2.0 * np_ * kp * tp / (sm * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4)
            if acc is not None:
                _, Em = series_of(native, args, [Ssub], Cx, n_acc)
                row["series_max_rel_err"] = rel_err(Em, o["series"])
            methods[name] = row
        set_method(None)
        for name in ("bf16x6", "fp16x3"):
            methods[name + "_over_fp32"] = {"step": round(methods["fp32"]["ms_per_step"] / methods[name]["ms_per_step"], 2),
                                            "series": round(methods["fp32"]["series_ms_per_step"] / methods[name]["series_ms_per_step"], 2)}
        print("bench: contraction methods %s" % json.dumps(methods), file=sys.stderr, flush=True)

    legs = {}
    if not a.no_legs:
        legs["urban_block"] = urban_block_leg(native, torch, tmp, epw, dni, dhi, threads)
        print("bench: urban block %s" % json.dumps(legs["urban_block"]), file=sys.stderr, flush=True)
        legs["views"] = views_leg(native, torch, threads)
        print("bench: views %s" % json.dumps({k: v for k, v in legs["views"].items() if k != "columns"}), file=sys.stderr, flush=True)
        legs["transmission"] = transmission_leg(native, torch, epw, tmp)
        print("bench: transmission %s" % json.dumps(legs["transmission"]), file=sys.stderr, flush=True)
    if not a.no_legs and not a.no_serving:
        legs["serving"] = serving_leg(native, epw)
        print("bench: serving %s" % json.dumps(legs["serving"]), file=sys.stderr, flush=True)
    refs = None if a.no_reference_benches else reference_benches()
    cpu = None if a.no_cpu else cpu_baseline(ctx, meshes[1][1], dni, dhi, threads, model, a.cpu_seconds)
    lat_ms = np.array(lat) * 1e3
    line = {
        "metric": "sensor-hours/s at <=2e-6 rel vs FP64 (annual radiation, direct, diffuse, illuminance + hourly series)",
        "value": round(value, 1), "unit": "sensor-hours/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 * wall / a.steps, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 trace, f64 sky and sums, fp16x3 mfma series (fp32 accumulation in 256-direction blocks)", "data": "synthetic",
        "config": {"workload": "batched precision-form daylight: seeded 3M-triangle city (250,000 boxes), three sensor meshes of "
                               "1,000 / 150,000 / 850,000 faces in one run, 2305 patches, 8760 h, Beta(2,3) cloudy clear-sky "
                               "weather, radiation + direct + diffuse + illuminance + the radiometric device series",
                   "sensors": n_sens, "context_triangles": 2 * len(ctx[1]), "patches": P, "hours": hours, "series_columns": Tup,
                   "contraction": DEFAULT_METHOD, "trace_launches_per_step": launches,
                   "graph_launches_per_step": d["graph_launches"] / a.steps},
        "grays_per_s": {"slots": round(d["rays"] / (d["trace_ms"] * 1e-3) / 1e9, 3),
                        "traced": round(d["traced"] / (d["trace_ms"] * 1e-3) / 1e9, 3),
                        "slots_per_step": d["rays"] / a.steps, "traced_per_step": d["traced"] / a.steps,
                        "note": "slots count every sensor x direction pair as the reference does (daylight.h:110); traced "
                                "counts the rays past the hemisphere cull"},
        "stages_ms_per_step": {k: round(d[k] / a.steps, 3) for k in ("trace_ms", "reduce_ms", "series_ms", "sky_ms")},
        "accuracy": acc,
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2), "p99": round(float(np.percentile(lat_ms, 99)), 2),
                       "note": "of back-to-back runs (one request at a time)"},
        "contraction_methods": methods,
        "roofline": roofline,
        "contraction_roofline": contraction,
        "legs": legs,
        "reference_benches": refs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
