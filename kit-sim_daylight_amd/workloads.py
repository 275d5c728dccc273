# SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
"""Seeded synthetic inputs of the measured workloads, shared by bench.py and the tests.

- city(): extruded boxes (footprint sides U(10, 40) m, heights lognormal with median 15 m and sigma 0.5) on a square grid
  of 50 m plots, each moved within its plot so that a 5 m street remains; the buildings nearest the centre carry sensors.
  The urban block is the 20 x 20 grid (400 boxes, 4,800 triangles) with the 40 central buildings as sensors; the
  3M-triangle city is the 500 x 500 grid (250,000 boxes, 3,000,000 triangles).
- sensor_mesh(): roofs and facades of boxes in quads of at most `cell` metres, facing out.
- city_sensor_meshes(): the 1k / 150k / 850k-face sensor meshes over the 3M-triangle city's central buildings.
- city_5m(), serving_load(): the 5M-triangle city and the serving mix drawn over its central buildings' faces.
- interior(), floor_grid(): the seeded 2M-triangle interior and its 0.5 m floor grid of sensors.
- weather(), write_epw(): the cloud-modulated clear-sky year (Beta(2, 3) cover) at lat 52.52, lon 13.41, tz 1.

Meshes are (points [n x 3] float64, counts [f] int32, indices [sum counts] int32), the arrays a SimulationMesh takes.
"""
import math
import os

import numpy as np

LAT, LON, TZ = 52.52, 13.41, 1.0


def city(seed=1, n_side=20, pitch=50.0, n_sensor_buildings=40):
    """Boxes (x0, y0, x1, y1, h) of the city and the indices of the n_sensor_buildings nearest the centre (ascending)."""
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


_BOX_QUADS = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]], np.int64)


def box_mesh(boxes):
    """The boxes as quads (bottom, top, 4 sides), outward normals: 8 points and 6 quads per box."""
    x0, y0, x1, y1, h = (boxes[:, k] for k in range(5))
    z = np.zeros_like(h)
    pts = np.stack([np.stack(c, 1) for c in ((x0, y0, z), (x1, y0, z), (x1, y1, z), (x0, y1, z),
                                             (x0, y0, h), (x1, y0, h), (x1, y1, h), (x0, y1, h))], 1).reshape(-1, 3)
    idx = (_BOX_QUADS[None, :, :] + 8 * np.arange(len(boxes))[:, None, None]).reshape(-1, 4)
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


def _box_faces(box, cell):
    x0, y0, x1, y1, h = box
    a, b = x1 - x0, y1 - y0
    na, nb, nh = (max(1, int(math.ceil(x / cell))) for x in (a, b, h))
    return [((x0, y0, h), (a, 0, 0), (0, b, 0), na, nb),           # roof, +z
            ((x0, y0, 0), (a, 0, 0), (0, 0, h), na, nh),           # y0 side, -y
            ((x1, y0, 0), (0, b, 0), (0, 0, h), nb, nh),           # x1 side, +x
            ((x1, y1, 0), (-a, 0, 0), (0, 0, h), na, nh),          # y1 side, +y
            ((x0, y1, 0), (0, -b, 0), (0, 0, h), nb, nh)]          # x0 side, -x


def sensor_mesh(boxes, cell=1.0):
    """Roofs and facades of the boxes in quads of at most `cell` m, facing out."""
    pts, idx, base = [], [], 0
    for box in boxes:
        for o, u, v, nu, nv in _box_faces(box, cell):
            p, q = grid_quads(np.array(o, float), np.array(u, float), np.array(v, float), nu, nv)
            pts.append(p)
            idx.append(q + base)
            base += len(p)
    pts, idx = np.concatenate(pts), np.concatenate(idx)
    return pts, np.full(len(idx), 4, np.int32), idx.ravel().astype(np.int32)


def submesh(mesh, f0, f1):
    """Quads f0 .. f1 - 1 of a quad mesh with only the points they use (renumbered in first-use order)."""
    pts, _, idx = mesh
    q = idx.reshape(-1, 4)[f0:f1]
    used, inv = np.unique(q.ravel(), return_inverse=True)
    return pts[used], np.full(len(q), 4, np.int32), inv.astype(np.int32)


# ---------------------------------------------------------------- the 3M-triangle city and its three sensor meshes

CITY3M = dict(seed=4, n_side=500, pitch=50.0)
CITY3M_MESHES = (("district_s", 1_000), ("district_m", 150_000), ("district_l", 850_000))


def city_3m(n_sensor_buildings=600):
    """The 3M-triangle city: 250,000 boxes (12 triangles each) on a 500 x 500 grid of 50 m plots, seed 4."""
    return city(n_sensor_buildings=n_sensor_buildings, **CITY3M)


def city_sensor_meshes(boxes, central, sizes=CITY3M_MESHES):
    """Three sensor meshes of exactly the given face counts: the roofs and facades (1 m quads) of the central buildings,
    nearest first, cut in that order into consecutive meshes -> [(name, (points, counts, indices))]."""
    need = sum(n for _, n in sizes)
    order = central[np.argsort(np.hypot(boxes[central, 0] + boxes[central, 2], boxes[central, 1] + boxes[central, 3]), kind="stable")]
    parts, have = [], 0
    for k in order:
        m = sensor_mesh(boxes[k:k + 1])
        parts.append(m)
        have += len(m[1])
        if have >= need:
            break
    if have < need:
        raise ValueError("city_sensor_meshes: %d faces on the central buildings, %d wanted" % (have, need))
    base, pts, idx = 0, [], []
    for p, _, i in parts:
        pts.append(p)
        idx.append(i + base)
        base += len(p)
    allm = (np.concatenate(pts), None, np.concatenate(idx))
    out, f = [], 0
    for name, n in sizes:
        out.append((name, submesh(allm, f, f + n)))
        f += n
    return out


# ---------------------------------------------------------------- the serving load over the 5M-triangle city

CITY5M = dict(seed=5, n_side=646, pitch=50.0)


def city_5m(n_sensor_buildings=400):
    """The 5M-triangle city: 417,316 boxes (12 triangles each) on a 646 x 646 grid of 50 m plots, seed 5."""
    return city(n_sensor_buildings=n_sensor_buildings, **CITY5M)


SPANS = {"instant": ["--time", "06-21T12:00"], "day": ["--from", "06-21T00:00", "--to", "06-22T00:00", "--step", "60"], "year": []}


def serving_load(n, seed=11, pool_faces=500_000, views_share=0.3):
    """n requests of the serving mix: tool (daylight, or views with probability views_share), sensor faces log-uniform in
    [1,000, 500,000] cut at a random offset from a pool of pool_faces faces, a time span (instant / 24 h at 1 h / year)
    and a sky density (1 / 2 / 4) drawn uniformly -> [dict(tool, f0, faces, span, mf)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        faces = int(round(math.exp(rng.uniform(math.log(1_000), math.log(500_000)))))
        out.append(dict(tool="views" if rng.uniform() < views_share else "daylight", faces=faces,
                        f0=int(rng.integers(0, pool_faces - faces + 1)), span=str(rng.choice(list(SPANS))), mf=int(rng.choice([1, 2, 4]))))
    return out


# ---------------------------------------------------------------- the 2M-triangle interior

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


def floor_grid(side=158.0, step=0.5, n=100_000, x0=0.0, y0=0.0):
    """Sensor quads of `step` m on the floor (z = 0, facing up), n of them from (x0, y0) row by row."""
    per = int(side / step)
    rows = (n + per - 1) // per
    p, q = grid_quads(np.array([x0, y0, 0.0]), np.array([per * step, 0, 0]), np.array([0, rows * step, 0]), per, rows)
    # grid_quads runs the first axis slowest; keep the first n quads
    q = q[:n]
    return p, np.full(len(q), 4, np.int32), q.ravel().astype(np.int32)


INTERIOR_VIEW_ARGS = ["--fov", "360", "--offset", "1.2", "--cell", "2", "--far", "1000", "--jitter", "on", "--distance", "isovist,mean,min",
                      "--reach", "en17037", "--no-occlude", "grid"]


# ---------------------------------------------------------------- weather

def write_epw(path, dni, dhi, lat=LAT, lon=LON, tz=TZ):
    rows = ["LOCATION,Synthetic,-,DEU,synthetic,000000,%.2f,%.2f,%.1f,0.0" % (lat, lon, tz),
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
    """Clear-sky DNI / DHI from the library (--clear-sky, --dump-sun), cloud-modulated; -> (dni, dhi) [8760] float32.
    The values are multiples of 1/8 W/m2: exact in binary and in the EPW's decimal text, so the library (reading the
    file) and the oracle (given the arrays) see the same numbers."""
    sun = os.path.join(tmp, "sun.txt")
    with native.SimulationMesh("probe", [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], [4], [0, 1, 2, 3]) as m:
        native.run("daylight", ["--clear-sky", "--lat", LAT, "--lon", LON, "--tz", TZ, "--quantity", "sunhours",
                                "--dump-sun", sun], m)
    d = np.loadtxt(sun, comments="#")
    alt, dni_cs, dhi_cs = d[:, 4], d[:, 10], d[:, 11]   # the dump's columns: ... up, sun.c alt azi up, dni, dhi
    assert len(d) == 8760, len(d)
    c = np.random.default_rng(seed).beta(2, 3, 8760)
    dni = dni_cs * (1 - c)
    dhi = dhi_cs + 0.5 * c * dni_cs * np.maximum(0, np.sin(np.radians(alt)))
    return (np.round(dni * 8) / 8).astype(np.float32), (np.round(dhi * 8) / 8).astype(np.float32)


def weather_epw(native, tmp, seed=2):
    """weather() written to tmp/cloudy.epw -> (path, dni, dhi)."""
    dni, dhi = weather(native, tmp, seed)
    path = os.path.join(tmp, "cloudy.epw")
    write_epw(path, dni, dhi)
    return path, dni, dhi
