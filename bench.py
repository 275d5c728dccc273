#!/usr/bin/env python3
"""Throughput of the daylight trace-and-accumulate path on one MI355X: sensor-hours/s of an annual run with its series.

Workload (seeded urban block): 400 extruded boxes (footprint sides U(10, 40) m, heights lognormal with median 15 m and
sigma 0.5) on a 20 x 20 grid of 50 m plots are the context; the roofs and facades of the 40 buildings nearest the centre,
cut into quads of at most 1 m, are the sensors (they sit on the context boxes' faces and do not block themselves,
--no-self-occlude). Weather: the library's own clear sky (Ineichen, turbidity 3) at lat 52.52 lon 13.41 tz 1, with an
hourly cloud cover c ~ Beta(2, 3) taking (1 - c) of the beam and moving half of what it took into the diffuse part,
written to an EPW. Run: --sky-density 2 (577 patches), 8760 hourly steps, radiation plus its --timeseries; the series
stays on the GPU (dlx_run_series_device), file output is excluded from the timing.

A step is one such run through the C ABI with the context scene already resident on the GPU: the sky and sun of every
step (FP64), the visibility traversal, the per-sensor reductions and the series contraction. The timed region holds
exactly --steps runs, bracketed by torch.cuda.synchronize(). The line also carries the traversal's roofline (HIP-event
times of its launches, counted node and triangle fetches), the three contraction methods side by side (FP32, BF16x3,
BF16x6) and the oracle's CPU time on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kit-sim_daylight_amd"))

LAT, LON, TZ = 52.52, 13.41, 1.0
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES = 112, 48   # a 4-wide record visit loads 7 x 16 bytes; a triangle 3 x 16


# ---------------------------------------------------------------- synthetic inputs

def city(seed=1, n_side=20, pitch=50.0, n_sensor_buildings=40):
    """Boxes (x0, y0, x1, y1, h) of the block and the indices of the central sensor buildings."""
    rng = np.random.default_rng(seed)
    n = n_side * n_side
    a = rng.uniform(10, 40, n)
    b = rng.uniform(10, 40, n)
    h = rng.lognormal(math.log(15.0), 0.5, n)
    gx, gy = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    cx = (gx.ravel() + 0.5 - n_side / 2) * pitch
    cy = (gy.ravel() + 0.5 - n_side / 2) * pitch
    # a 5 m street at least: the box moves within its plot
    cx = cx + rng.uniform(-1, 1, n) * (pitch - 5 - a) / 2
    cy = cy + rng.uniform(-1, 1, n) * (pitch - 5 - b) / 2
    boxes = np.stack([cx - a / 2, cy - b / 2, cx + a / 2, cy + b / 2, h], 1)
    central = np.argsort(np.hypot(cx, cy), kind="stable")[:n_sensor_buildings]
    return boxes, np.sort(central)


def box_mesh(boxes):
    """The boxes as quads (bottom, top, 4 sides), outward normals."""
    pts, idx = [], []
    for k, (x0, y0, x1, y1, h) in enumerate(boxes):
        v = np.array([[x0, y0, 0], [x1, y0, 0], [x1, y1, 0], [x0, y1, 0], [x0, y0, h], [x1, y0, h], [x1, y1, h], [x0, y1, h]])
        pts.append(v)
        q = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]) + 8 * k
        idx.append(q)
    pts, idx = np.concatenate(pts), np.concatenate(idx)
    return pts, np.full(len(idx), 4, np.int32), idx.ravel().astype(np.int32)


def grid_quads(origin, u, v, nu, nv):
    """nu x nv quads over the parallelogram origin + s u + t v (s, t in [0, 1]); normal u x v."""
    s = np.linspace(0, 1, nu + 1)
    t = np.linspace(0, 1, nv + 1)
    S, T = np.meshgrid(s, t, indexing="ij")
    P = origin + S[..., None] * u + T[..., None] * v
    pts = P.reshape(-1, 3)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * (nv + 1) + j
    quads = np.stack([a, a + nv + 1, a + nv + 2, a + 1], -1).reshape(-1, 4)
    return pts, quads


def sensor_mesh(boxes):
    """Roofs and facades of the boxes in quads of at most 1 m, facing out."""
    pts, idx, base = [], [], 0
    for x0, y0, x1, y1, h in boxes:
        a, b = x1 - x0, y1 - y0
        na, nb, nh = int(math.ceil(a)), int(math.ceil(b)), int(math.ceil(h))
        faces = [((x0, y0, h), (a, 0, 0), (0, b, 0), na, nb),           # roof, +z
                 ((x0, y0, 0), (a, 0, 0), (0, 0, h), na, nh),           # y0 side, -y
                 ((x1, y0, 0), (0, b, 0), (0, 0, h), nb, nh),           # x1 side, +x
                 ((x1, y1, 0), (-a, 0, 0), (0, 0, h), na, nh),          # y1 side, +y
                 ((x0, y1, 0), (0, -b, 0), (0, 0, h), nb, nh)]          # x0 side, -x
        for o, u, v, nu, nv in faces:
            p, q = grid_quads(np.array(o, float), np.array(u, float), np.array(v, float), nu, nv)
            pts.append(p)
            idx.append(q + base)
            base += len(p)
    pts, idx = np.concatenate(pts), np.concatenate(idx)
    return pts, np.full(len(idx), 4, np.int32), idx.ravel().astype(np.int32)


def interior(seed=1, side=158.0, wall_h=3.0, tri_budget=2_000_000):
    """The seeded interior: a rectilinear plan of rooms with sides U(3, 8) m, walls 3 m high tessellated into 0.25 m
    quads with a door opening (0.9-1.5 m wide, 2.1 m high) in 70 % of the wall segments, no ceiling, and clutter boxes
    with sides U(0.3, 2) m filling the rest of the triangle budget. -> (points, counts, indices) as quads."""
    rng = np.random.default_rng(seed)

    def cuts():
        c = [0.0]
        while c[-1] < side:
            c.append(min(side, c[-1] + rng.uniform(3, 8)))
        return np.array(c)

    xs, ys = cuts(), cuts()
    pts, quads = [], []
    base = 0

    def wall(p0, p1, lo, hi):   # vertical rectangle from p0 to p1 (xy), z in [lo, hi], 0.25 m quads
        nonlocal base
        length = math.hypot(p1[0] - p0[0], p1[1] - p0[1])
        nu, nv = max(1, int(math.ceil(length / 0.25))), max(1, int(math.ceil((hi - lo) / 0.25)))
        p, q = grid_quads(np.array([p0[0], p0[1], lo]), np.array([p1[0] - p0[0], p1[1] - p0[1], 0.0]), np.array([0, 0, hi - lo]), nu, nv)
        pts.append(p)
        quads.append(q + base)
        base += len(p)

    def segment(a, b, axis, at):   # a wall segment along x (axis 0) or y (axis 1) at the other coordinate `at`
        def xy(u):
            return (u, at) if axis == 0 else (at, u)
        if rng.uniform() < 0.7 and b - a > 2.0:
            w = rng.uniform(0.9, 1.5)
            d0 = rng.uniform(a + 0.3, b - 0.3 - w)
            wall(xy(a), xy(d0), 0.0, wall_h)
            wall(xy(d0 + w), xy(b), 0.0, wall_h)
            wall(xy(d0), xy(d0 + w), 2.1, wall_h)
        else:
            wall(xy(a), xy(b), 0.0, wall_h)

    for y in ys:
        for a, b in zip(xs[:-1], xs[1:]):
            segment(a, b, 0, y)
    for x in xs:
        for a, b in zip(ys[:-1], ys[1:]):
            segment(a, b, 1, x)
    nq = sum(len(q) for q in quads)
    nbox = max(0, (tri_budget - 2 * nq) // 12)
    sz = rng.uniform(0.3, 2.0, (nbox, 3))
    c = rng.uniform(0, side, (nbox, 2))
    boxes = np.stack([c[:, 0] - sz[:, 0] / 2, c[:, 1] - sz[:, 1] / 2, c[:, 0] + sz[:, 0] / 2, c[:, 1] + sz[:, 1] / 2, sz[:, 2]], 1)
    bp, bc, bi = box_mesh(boxes)
    pts.append(bp)
    quads.append(bi.reshape(-1, 4) + base)
    P = np.concatenate(pts)
    Q = np.concatenate(quads)
    return P, np.full(len(Q), 4, np.int32), Q.ravel().astype(np.int32)


def floor_grid(side=158.0, step=0.5, n=100_000):
    """Sensor quads of `step` m on the floor (z = 0, facing up), n of them from the plan's corner row by row."""
    per = int(side / step)
    rows = (n + per - 1) // per
    p, q = grid_quads(np.array([0.0, 0.0, 0.0]), np.array([per * step, 0, 0]), np.array([0, rows * step, 0]), per, rows)
    # grid_quads runs the first axis slowest; keep the first n quads
    q = q[:n]
    return p, np.full(len(q), 4, np.int32), q.ravel().astype(np.int32)


def write_epw(path, dni, dhi):
    rows = ["LOCATION,Synthetic,-,DEU,synthetic,000000,%.2f,%.2f,%.1f,0.0" % (LAT, LON, TZ),
            "DESIGN CONDITIONS,0", "TYPICAL/EXTREME PERIODS,0", "GROUND TEMPERATURES,0",
            "HOLIDAYS/DAYLIGHT SAVINGS,No,0,0,0", "COMMENTS 1,clear sky x Beta(2 3) cloud cover",
            "COMMENTS 2,", "DATA PERIODS,1,1,Data,Sunday, 1/ 1,12/31"]
    days = [31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    r = 0
    for m, nd in enumerate(days, 1):
        for d in range(1, nd + 1):
            for hr in range(1, 25):
                ghi = dni[r] + dhi[r]
                f = [1999, m, d, hr, 0, "?", 10.0, 5.0, 70, 101325, 0, 0, 300, "%.3f" % ghi, "%.3f" % dni[r], "%.3f" % dhi[r]]
                f += [0] * 19
                rows.append(",".join(str(x) for x in f))
                r += 1
    with open(path, "w") as fh:
        fh.write("\n".join(rows) + "\n")


def weather(native, tmp, seed=2):
    """Clear-sky DNI / DHI from the library (--clear-sky, --dump-sun), cloud-modulated; -> (dni, dhi) [8760] float32."""
    sun = os.path.join(tmp, "sun.txt")
    with native.SimulationMesh("probe", [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], [4], [0, 1, 2, 3]) as m:
        native.run("daylight", ["--clear-sky", "--lat", LAT, "--lon", LON, "--tz", TZ, "--quantity", "sunhours",
                                "--dump-sun", sun], m)
    d = np.loadtxt(sun, comments="#")
    alt, dni_cs, dhi_cs = d[:, 4], d[:, 10], d[:, 11]   # the reference's dump columns: ... up, sun.c alt azi up, dni, dhi
    assert len(d) == 8760, len(d)
    c = np.random.default_rng(seed).beta(2, 3, 8760)
    dni = dni_cs * (1 - c)
    dhi = dhi_cs + 0.5 * c * dni_cs * np.maximum(0, np.sin(np.radians(alt)))
    # multiples of 1/8 W/m2: exact in binary and in the EPW's decimal text, so the library (reading the file) and the
    # oracle (given the arrays) see the same numbers
    return (np.round(dni * 8) / 8).astype(np.float32), (np.round(dhi * 8) / 8).astype(np.float32)


# ---------------------------------------------------------------- the CPU baseline (the oracle, on a bounded sample)

def cpu_baseline(ctx, sens, dni, dhi, target_s=15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    threads = max(1, min(16, os.cpu_count() or 1))
    wx = dict(lat=LAT, lon=LON, tz=TZ, leap=False, dni=dni, dhi=dhi)
    ctx_m = orc.MeshArrays("city", ctx[0], ctx[1], ctx[2], sensor=False)

    def sample(nf):
        pts, counts, idx = sens
        sub = orc.MeshArrays("sensors", pts, counts[:nf], idx[: 4 * nf], sensor=True)
        t0 = time.perf_counter()
        r = orc.daylight([sub, ctx_m], weather=wx, quantities=("radiation",), mf=2, selfOcclude=False, series=True,
                         nthreads=threads)
        return time.perf_counter() - t0, r["durationHours"]

    # two probes give the fixed cost (the sky of the year) and the cost per sensor; the sample fills target_s
    n0, n1 = 64, 512
    t0, _ = sample(n0)
    t1, _ = sample(n1)
    per = max((t1 - t0) / (n1 - n0), 1e-6)
    nf = int(min(len(sens[1]), max(n1, (target_s - max(t0 - n0 * per, 0.0)) / per)))
    t, hours = sample(nf)
    return {"value": nf * hours / t, "unit": "sensor-hours/s", "cores": threads, "kind": "port",
            "sample": "%d of the sensor faces (the first in mesh order), full year (%.0f h), 577 patches, same context and "
                      "weather; oracle trace + FP64 sums + FP64 series, %d threads, %.1f s" % (nf, hours, threads, t)}


# ---------------------------------------------------------------- accuracy on the workload itself (the oracle, FP64)

def oracle_sample(ctx, sens, dni, dhi, n=64):
    """The oracle's FP64 run (annual radiation and hourly series) of the first n sensors of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.build()
    pts, counts, idx = sens
    sub = orc.MeshArrays("sensors", pts, counts[:n], idx[: 4 * n], sensor=True)
    ctx_m = orc.MeshArrays("city", ctx[0], ctx[1], ctx[2], sensor=False)
    o = orc.daylight([sub, ctx_m], weather=dict(lat=LAT, lon=LON, tz=TZ, leap=False, dni=dni, dhi=dhi), quantities=("radiation",),
                     mf=2, selfOcclude=False, series=True, nthreads=max(1, min(16, os.cpu_count() or 1)))
    return o["series"], o["values"][:, 0]


def accuracy(native, args, S, Cx, ref, n=64):
    """One more run of the workload: the device series of its first n sensors (copied out of the sink) and their annual
    radiation against the oracle's FP64 values. Error = max |x - ref| / max(|ref|, 1e-3 max |ref|) over the sample (the
    validation scale rule of the reference's validate/README.md:160)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    got = {}

    def sink(row0, rows, cols, ptr):
        if row0 == 0:
            buf = np.empty((min(n, rows), cols), np.float32)
            if hip.hipMemcpy(buf.ctypes.data, C.c_void_p(ptr), buf.nbytes, 2):
                return 2
            got["E"] = buf
        return 0

    r = native.run_series_device(args, sensors=S, context=Cx, sink=sink)

    def rel(x, want):
        scale = np.maximum(np.abs(want), 1e-3 * np.abs(want).max())
        return float(np.max(np.abs(x - want) / np.where(scale > 0, scale, 1.0)))

    e_series, e_rad = rel(got["E"].astype(np.float64), ref[0]), rel(r.values[:n, 0], ref[1])
    return {"sensors": n, "series_max_rel_err": e_series, "radiation_max_rel_err": e_rad, "bar": 2e-6,
            "within_bar": bool(e_series <= 2e-6 and e_rad <= 2e-6)}


# ---------------------------------------------------------------- the views leg (closest-hit rays and their reductions)

VIEW_ARGS = ["--fov", "360", "--offset", "1.2", "--cell", "2", "--far", "1000", "--jitter", "on", "--distance", "isovist,mean,min",
             "--reach", "en17037", "--no-occlude", "grid"]


def views_leg(native, torch, steps=3):
    """The seeded interior (~2M triangles) seen all around from a 0.5 m floor grid of 100,000 sensors at eye height
    1.2 m, 2-degree cells (10,800 rays each): whole views runs through the C ABI on the resident scene."""
    P, C, I = interior()
    gp, gc, gi = floor_grid()
    room = native.SimulationMesh("interior", P, C, I)
    grid = native.SimulationMesh("grid", gp, gc, gi)
    r = native.run("views", VIEW_ARGS, grid, context=room)   # builds and keeps the scene
    n = r.values.shape[0]
    c0 = native.gpu_counters()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = native.run("views", VIEW_ARGS, grid, context=room)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    c1 = native.gpu_counters()
    assert r.stats["scene_reused"]
    rays = n * 10800
    return {"workload": "seeded interior, %d triangles (rooms U(3,8) m, door openings, clutter U(0.3,2) m), %d floor sensors at "
                        "0.5 m, eye 1.2 m, 360 x 120 degrees in 2-degree cells, jitter, isovist + mean/min + reach 6/20/50 m"
                        % (2 * len(C), n),
            "rays_per_run": rays, "runs": steps, "ms_per_run": round(1e3 * wall, 2),
            "grays_per_s": round(rays / wall / 1e9, 3),
            "trace_ms_per_run": round((c1["views_ms"] - c0["views_ms"]) / steps, 3),
            "reduce_ms_per_run": round((c1["reduce_ms"] - c0["reduce_ms"]) / steps, 3),
            "columns": r.manifest["columns"]}


# ---------------------------------------------------------------- the reference's own --bench runs, through our CLI

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
    """Each of the reference's published --bench runs with this build's CLI (a child process on the same GPU): its
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


# ---------------------------------------------------------------- the HBM traffic of the traversal (rocprofv3 --pmc)

TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_bits_urban_block_traffic.txt")


def traffic_per_launch():
    """HBM bytes per traversal launch as the committed counter summary states them (FETCH_SIZE + WRITE_SIZE of the
    traversal, scaled by the counters' reading of a copy of known size), or None where there is none."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


# ---------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-methods", action="store_true", help="skip the side-by-side runs of the contraction methods")
    ap.add_argument("--no-count", action="store_true", help="skip the traversal's counted run (roofline achieved = null)")
    ap.add_argument("--no-views", action="store_true", help="skip the views leg")
    ap.add_argument("--only-views", action="store_true", help="run the views leg alone (prints its object as the line)")
    ap.add_argument("--no-reference-benches", action="store_true", help="skip the reference's --bench runs through the CLI")
    ap.add_argument("--no-transmission", action="store_true", help="skip the byte-mode leg (the context as glass)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (the HBM counters' scale) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the path shards by sensor rows, "
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
    if a.only_views:
        print(json.dumps({"views": views_leg(native, torch)}), flush=True)
        return

    boxes, central = city()
    ctx = box_mesh(boxes)
    sens = sensor_mesh(boxes[central])
    n_faces = len(sens[1])
    tmp = tempfile.mkdtemp(prefix="dlbench")
    dni, dhi = weather(native, tmp)
    epw = os.path.join(tmp, "block.epw")
    write_epw(epw, dni, dhi)
    args = ["--epw", epw, "--sky-density", "2", "--quantity", "radiation", "--timeseries", os.path.join(tmp, "series.npy"),
            "--no-self-occlude"]
    print("bench: %d context triangles, %d sensor faces; warmup %d, steps %d" % (2 * len(ctx[1]), n_faces, a.warmup, a.steps),
          file=sys.stderr, flush=True)

    def step(ms_series=None):
        r = native.run_series_device(args, sensors=S, context=Cx)
        return r

    S = native.SimulationMesh("sensors", *sens)
    Cx = native.SimulationMesh("city", *ctx)
    os.environ.pop("DAYLIGHT_SERIES", None)            # the default contraction (BF16x6 on MFMA)
    r0 = step()
    assert r0.values.shape == (n_faces, 1), r0.values.shape
    hours = float(r0.manifest["duration_h"])
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    c0 = native.gpu_counters()
    lat = []
    stages = dict.fromkeys(("prep_ms", "sky_ms", "trace_ms", "reduce_ms"), 0.0)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ts = time.perf_counter()
        r = step()
        lat.append(time.perf_counter() - ts)
        for k in stages:
            stages[k] += r.stats[k]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c1 = native.gpu_counters()
    assert r.stats["scene_reused"], "the context scene was rebuilt inside the timed region"
    d = {k: c1[k] - c0[k] for k in c1}
    sensor_hours = n_faces * hours
    value = sensor_hours * a.steps / wall
    rays = d["rays"] / a.steps
    print("bench: %.1f ms/step, %.3g sensor-hours/s, trace %.1f ms/step over %d launches, series %.1f ms/step"
          % (1e3 * wall / a.steps, value, d["trace_ms"] / a.steps, d["trace_launches"] / a.steps, d["series_ms"] / a.steps),
          file=sys.stderr, flush=True)

    # the traversal's algorithmic fetches: one counted run (untimed)
    launches = d["trace_launches"] / a.steps
    per_launch_ms = d["trace_ms"] / d["trace_launches"]
    alg_bytes_launch = achieved = simd_eff = None
    if not a.no_count:
        native.count_fetches(True)
        k0 = native.gpu_counters()
        step()
        k1 = native.gpu_counters()
        native.count_fetches(False)
        fetch_bytes = (k1["node_fetches"] - k0["node_fetches"]) * NODE_BYTES + (k1["tri_fetches"] - k0["tri_fetches"]) * TRI_BYTES
        n_words = math.ceil((k1["counted_rays"] - k0["counted_rays"]) / n_faces / 32)
        io_bytes = n_faces * n_words * (4 + 24)         # a visibility word written, the sensor's point + normal read per word
        alg_bytes_launch = (fetch_bytes + io_bytes) / launches
        achieved = alg_bytes_launch / (per_launch_ms * 1e-3) / 1e9
        waves = k1["wave_iters"] - k0["wave_iters"]
        simd_eff = (k1["node_fetches"] - k0["node_fetches"]) / (64.0 * waves) if waves else None

    ref = oracle_sample(ctx, sens, dni, dhi)
    acc = accuracy(native, args, S, Cx, ref)
    print("bench: accuracy %s" % json.dumps(acc), file=sys.stderr, flush=True)

    # the contraction methods side by side (two runs each)
    methods = {}
    for name in () if a.no_methods else ("fp32", "bf16x3", "bf16x6"):
        os.environ["DAYLIGHT_SERIES"] = name
        step()
        m0 = native.gpu_counters()
        tm = time.perf_counter()
        for _ in range(2):
            step()
        torch.cuda.synchronize()
        tm = (time.perf_counter() - tm) / 2
        m1 = native.gpu_counters()
        methods[name] = {"ms_per_step": round(1e3 * tm, 2), "series_ms_per_step": round((m1["series_ms"] - m0["series_ms"]) / 2, 3),
                         "sensor_hours_per_s": round(sensor_hours / tm, 1),
                         "series_max_rel_err": accuracy(native, args, S, Cx, ref)["series_max_rel_err"]}
    os.environ.pop("DAYLIGHT_SERIES", None)
    if methods:
        print("bench: contraction methods %s" % json.dumps(methods), file=sys.stderr, flush=True)

    # byte mode: the same workload with every context face glass (tau 0.64, Radiance glass) under --transmission on, so
    # every ray walks every crossing; the first run builds the slotted scene (its own key), two timed runs follow
    trans = None
    if not a.no_transmission:
        targs = args + ["--transmission", "on", "--optics", "city=0,0.64,glass"]
        native.run_series_device(targs, sensors=S, context=Cx)
        t0t = native.gpu_counters()
        tt = time.perf_counter()
        for _ in range(2):
            rt = native.run_series_device(targs, sensors=S, context=Cx)
        torch.cuda.synchronize()
        tt = (time.perf_counter() - tt) / 2
        t1t = native.gpu_counters()
        assert rt.stats["scene_reused"]
        trans = {"workload": "the urban block with every context face glass (tau 0.64), --transmission on",
                 "ms_per_step": round(1e3 * tt, 2), "sensor_hours_per_s": round(sensor_hours / tt, 1),
                 "trace_ms_per_step": round((t1t["trace_ms"] - t0t["trace_ms"]) / 2, 3),
                 "series_ms_per_step": round((t1t["series_ms"] - t0t["series_ms"]) / 2, 3),
                 "reduce_ms_per_step": round((t1t["reduce_ms"] - t0t["reduce_ms"]) / 2, 3),
                 "grays_per_s": round((t1t["rays"] - t0t["rays"]) / ((t1t["trace_ms"] - t0t["trace_ms"]) * 1e-3) / 1e9, 3),
                 "mean_radiation_vs_opaque": round(float(rt.values.mean() / r.values.mean()), 4)}
        print("bench: transmission %s" % json.dumps(trans), file=sys.stderr, flush=True)

    views = None if a.no_views else views_leg(native, torch)
    if views:
        print("bench: views %s" % json.dumps({k: v for k, v in views.items() if k != "columns"}), file=sys.stderr, flush=True)
    refs = None if a.no_reference_benches else reference_benches()
    cpu = None if a.no_cpu else cpu_baseline(ctx, sens, dni, dhi, a.cpu_seconds)
    lat_ms = np.array(lat) * 1e3
    line = {
        "metric": "sensor-hours/s at <=2e-6 rel vs FP64 (annual radiation + hourly series)",
        "value": round(value, 1), "unit": "sensor-hours/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1e3 * wall / a.steps, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 trace, f64 sky and sums, bf16x6 mfma series", "data": "synthetic",
        "config": {"workload": "seeded urban block: 400 boxes (4800 context triangles), roofs + facades of the 40 central "
                               "buildings in <=1 m quads, 577 patches, 8760 h, Beta(2,3) cloudy clear-sky weather, "
                               "radiation + device series",
                   "sensors": n_faces, "context_triangles": 2 * len(ctx[1]), "patches": 577, "hours": hours,
                   "series_columns": None, "contraction": "bf16x6"},
        "grays_per_s": {"slots": round(d["rays"] / (d["trace_ms"] * 1e-3) / 1e9, 3), "rays_per_step": rays},
        "stages_ms_per_step": {k: round(v / a.steps, 3) for k, v in stages.items()},
        "accuracy": acc,
        "latency_ms": {"p50": round(float(np.percentile(lat_ms, 50)), 2), "p99": round(float(np.percentile(lat_ms, 99)), 2)},
        "kernels_ms_per_step": {k: round(d[k] / a.steps, 3) for k in ("trace_ms", "reduce_ms", "series_ms", "sky_ms")},
        "contraction_methods": methods,
        "transmission": trans,
        "views": views,
        "reference_benches": refs,
        "roofline": {"kernel": "k_bits (visibility traversal)", "bound": "hbm",
                     "achieved": None if achieved is None else round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_per_launch(),
                     "launch_ms": round(per_launch_ms, 4), "simd_efficiency": None if simd_eff is None else round(simd_eff, 3),
                     "algorithmic_bytes_per_launch": None if alg_bytes_launch is None else round(alg_bytes_launch)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
