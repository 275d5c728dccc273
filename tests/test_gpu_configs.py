"""The HIP path on the measured workloads' own inputs, against the oracle.

- The urban block (400 boxes, 577 patches, the cloudy year): 2,000 of its sensor faces end to end; radiation within
  2e-6, sun hours exact, the series of every contraction method against the oracle's FP64 series.
- The 3M-triangle city at 2305 patches: its 1,000-face sensor mesh end to end (radiation, direct, diffuse,
  illuminance within 2e-6; the radiometric and the visible series of the default contraction within 2e-6 of FP64), and
  the same mesh inside the run of all three sensor meshes (1,001,000 sensors in one launch) giving the same rows bit
  for bit.
- Deep trees: the 2M-triangle interior (binned SAH, height 29: the 48-entry traversal stack) and the 3M-triangle city
  (LBVH, height 32: the 64-entry stack), visibility bits and transmittance bytes of a few hundred sensors looking every
  way, bit for bit.
- Deep stacks: a comb of 2^20 triangles whose boxes a ray along it crosses without a hit, and a wall across it found
  only through the entries the walk pops back from past the ones kept in LDS: bits and bytes (SAH and LBVH trees) and
  the views' classes and distances, bit for bit.
- The interior seen all around (360 x 120 degrees in 2-degree cells, jitter, isovist, reach) from 200 floor sensors:
  every ray's class and distance bit for bit, the columns as the views tests compare them.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import workloads as W
from test_gpu_core import bits_matrix
from test_transmission_cpu import GLASS, INWARD, THIN, slot_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


@pytest.fixture(scope="module")
def cloudy(native, tmp_path_factory):
    return W.weather_epw(native, str(tmp_path_factory.mktemp("weather")))


@pytest.fixture(scope="module")
def city3m():
    boxes, central = W.city_3m()
    return W.box_mesh(boxes), W.city_sensor_meshes(boxes, central)


def rel_err(x, want):
    """max |x - want| / max(|want|, 1e-3 max |want|) (the validation scale rule, reference validate/README.md:160)"""
    scale = np.maximum(np.abs(want), 1e-3 * np.abs(want).max())
    return float(np.max(np.abs(np.asarray(x, np.float64) - want) / np.where(scale > 0, scale, 1.0)))


_hip = None


def series_rows(native, args, sensors, context, rows):
    """A daylight run with its series kept on the GPU -> (Result, the series' first `rows` rows copied from the device)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    parts = []

    def sink(row0, n, cols, ptr):
        take = min(n, rows - row0)
        if take > 0:
            buf = np.empty((take, cols), np.float32)
            if _hip.hipMemcpy(buf.ctypes.data, C.c_void_p(ptr), buf.nbytes, 2):
                return 2
            parts.append(buf)
        return 0

    r = native.run_series_device(args, sensors=sensors, context=context, sink=sink)
    return r, np.concatenate(parts)


def with_method(name, f):
    old = os.environ.get("DAYLIGHT_SERIES")
    os.environ["DAYLIGHT_SERIES"] = name
    try:
        return f()
    finally:
        if old is None:
            os.environ.pop("DAYLIGHT_SERIES")
        else:
            os.environ["DAYLIGHT_SERIES"] = old


def oracle_run(orc, meshes, dni, dhi, **kw):
    wx = dict(lat=W.LAT, lon=W.LON, tz=W.TZ, leap=False, dni=dni, dhi=dhi)
    return orc.daylight(meshes, weather=wx, selfOcclude=False, nthreads=min(16, os.cpu_count() or 1), **kw)


# bars of the contraction methods against the FP64 series: BF16x6 and FP16x3 are the metric's 2e-6 (FP16x3's two fp16
# pieces keep 22 bits: 2^-21 per product at worst); BF16x3 keeps products to 2^-16 relative (its split's floor, about
# 1e-5 on small sensors); FP32 is the reference's own arithmetic (a fma chain over the patches: about sqrt(P) ulps)
SERIES_BARS = {"bf16x6": 2e-6, "fp16x3": 2e-6, "bf16x3": 2e-5, "fp32": 1e-5}


def test_urban_block_slice_vs_oracle(native, orc, cloudy):
    epw, dni, dhi = cloudy
    boxes, central = W.city()
    ctx = W.box_mesh(boxes)
    sens = W.submesh(W.sensor_mesh(boxes[central]), 0, 2000)
    print("case: urban block, 2000 sensors, 577 patches, every contraction method", flush=True)
    S = native.SimulationMesh("sensors", *sens)
    Cx = native.SimulationMesh("city", *ctx)
    args = ["--epw", epw, "--sky-density", "2", "--quantity", "radiation,sunhours", "--timeseries", "unused.npy", "--no-self-occlude"]
    o = oracle_run(orc, [orc.MeshArrays("sensors", *sens, sensor=True), orc.MeshArrays("city", *ctx, sensor=False)], dni, dhi,
                   quantities=("radiation", "sunhours"), mf=2, series=True)
    errs = {}
    for method, bar in SERIES_BARS.items():
        r, E = with_method(method, lambda: series_rows(native, args, S, Cx, 2000))
        assert r.values.shape == o["values"].shape
        assert rel_err(r.values[:, 0], o["values"][:, 0]) <= 2e-6
        assert np.array_equal(r.values[:, 1], o["values"][:, 1]), "sun hours"
        errs[method] = rel_err(E, o["series"])
        print("series %s: max rel error vs FP64 %.3e (bar %.0e)" % (method, errs[method], bar), flush=True)
        assert errs[method] <= bar, method


def test_city3m_p2305_vs_oracle(native, orc, cloudy, city3m):
    epw, dni, dhi = cloudy
    ctx, meshes = city3m
    name, small = meshes[0]
    print("case: 3M-triangle city, %s (%d faces), 2305 patches, radiation / direct / diffuse / illuminance" % (name, len(small[1])), flush=True)
    Cx = native.SimulationMesh("city", *ctx)
    handles = [native.SimulationMesh(n, *m) for n, m in meshes]
    q = ["radiation", "direct", "diffuse", "illuminance"]
    args = ["--epw", epw, "--sky-density", "4", "--quantity", ",".join(q), "--timeseries", "unused.npy", "--no-self-occlude"]
    n = len(small[1])
    alone, Ea = series_rows(native, args, [handles[0]], Cx, n)
    print("case: the three sensor meshes in one run (%d sensors)" % sum(len(m[1]) for _, m in meshes), flush=True)
    batch, Eb = series_rows(native, args, handles, Cx, n)
    assert batch.values.shape[0] == sum(len(m[1]) for _, m in meshes)
    assert np.array_equal(batch.values[:n], alone.values), "a mesh's rows depend on the run it is in"
    assert np.array_equal(Eb, Ea), "a mesh's series depends on the run it is in"
    meshes_o = [orc.MeshArrays(name, *small, sensor=True), orc.MeshArrays("city", *ctx, sensor=False)]
    o = oracle_run(orc, meshes_o, dni, dhi, quantities=tuple(q), mf=4, series=True)
    for c, col in enumerate(q):
        e = rel_err(alone.values[:, c], o["values"][:, c])
        print("%s: max rel error vs the oracle %.3e" % (col, e), flush=True)
        assert e <= 2e-6, col
    e = rel_err(Ea, o["series"])
    print("series (radiometric, default fp16x3): max rel error vs FP64 %.3e" % e, flush=True)
    assert e <= 2e-6
    # the visible series (illuminance first): its sky matrix and sun weights are the luminous ones
    args_v = ["--epw", epw, "--sky-density", "4", "--quantity", "illuminance", "--timeseries", "unused.npy", "--no-self-occlude"]
    _, Ev = series_rows(native, args_v, [handles[0]], Cx, n)
    ov = oracle_run(orc, meshes_o, dni, dhi, quantities=("illuminance",), mf=4, series=True)
    e = rel_err(Ev, ov["series"])
    print("series (visible, default fp16x3): max rel error vs FP64 %.3e" % e, flush=True)
    assert e <= 2e-6


def _soup(mesh):
    P, _, I = mesh
    q = I.reshape(-1, 4)
    tri = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], 1)
    return np.ascontiguousarray(P, np.float32).reshape(-1), np.ascontiguousarray(tri, np.uint32).reshape(-1)


def _looking_everywhere(rng, lo, hi, n):
    """n sensors at uniform points of the box [lo, hi] with uniform unit normals"""
    P = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    N = rng.normal(size=(n, 3))
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    return P, N.astype(np.float32)


def _bits_and_bytes_vs_oracle(orc, V, I, P, N, D, tree, check_stack=None):
    """Visibility bits, then byte mode with a porous, a glass and an opaque slot spread over the triangles, of sensors P, N
    along directions D over the tree `tree` of triangles V, I: every bit and byte against the oracle. check_stack(height,
    need) sees the tree's height and the any-hit stack it needs."""
    import daylight_amd
    nt, nc, nDir = len(I) // 3, len(P), len(D) // 3
    g = daylight_amd.Gpu()
    try:
        g.set_tree(tree)
        g.build(V, I)
        height = g.scene()[2]
        need = 3 * ((height + 1) // 2) + 1 + 2
        print("case: %d triangles, %s tree of height %d: stack of %d entries" % (nt, tree, height, need), flush=True)
        if check_stack:
            check_stack(height, need)
        g.set_scene(P.reshape(-1), N.reshape(-1), D)
        g.trace(0, nc)
        got = g.bits(nc)
        sc = orc.lib().orc_scene_new(V, len(V) // 3, I, nt)
        want = np.zeros_like(got)
        orc.lib().orc_trace_bits(sc, P.reshape(-1), N.reshape(-1), nc, D, nDir, 0, want, min(16, os.cpu_count() or 1))
        gm, wm = bits_matrix(got, nc, nDir), bits_matrix(want, nc, nDir)
        print("bits: %d open of %d rays" % (wm.sum(), wm.size), flush=True)
        assert np.array_equal(gm, wm), np.argwhere(gm != wm)[:8]
        tn = float(np.float32(orc.lib().orc_tn_from_Tn(0.64)))
        slots = [(0.2, 0.0, 0.0, 0), (0.2, 0.5, 0.0, 0), (0.0, 0.64, tn, GLASS)]
        ts = (np.arange(nt, dtype=np.uint32) // 7) % 3
        g.set_slots(slots)
        g.set_tri_slots(ts)
        g.trace(0, nc, translucent=True)
        by = g.bytes(nc)
        out = np.zeros(nDir * nc, np.uint8)
        orc.lib().orc_trace_bytes(sc, P.reshape(-1), N.reshape(-1), nc, D, nDir, 0, ts.ctypes.data, slot_rows(slots), len(slots), out,
                                  min(16, os.cpu_count() or 1))
        orc.lib().orc_scene_free(sc)
        want_b = out.reshape(nDir, nc)
        frac = (want_b > 0) & (want_b < 255)
        print("bytes: %d partly passed, %d open, %d blocked" % (frac.sum(), (want_b == 255).sum(), (want_b == 0).sum()), flush=True)
        bad = np.argwhere(by != want_b)
        assert len(bad) == 0, (len(bad), [(int(d), int(i), int(by[d, i]), int(want_b[d, i])) for d, i in bad[:8]])
    finally:
        g.close()


@pytest.mark.parametrize("scene,tree,stack", [("interior", "sah", 48), ("city3m", "lbvh", 64)])
def test_deep_tree_bits_and_bytes(orc, city3m, scene, tree, stack):
    rng = np.random.default_rng(7)
    if scene == "interior":
        V, I = _soup(W.interior())
        P, N = _looking_everywhere(rng, (5.0, 5.0, 0.2), (150.0, 150.0, 2.8), 256)
    else:
        V, I = _soup(city3m[0])
        P, N = _looking_everywhere(rng, (-400.0, -400.0, 0.5), (400.0, 400.0, 40.0), 256)
    pa = orc.patches(1)
    extra = rng.normal(size=(40, 3))
    extra[:, 2] = np.abs(extra[:, 2]) + 0.05
    extra /= np.linalg.norm(extra, axis=1, keepdims=True)
    D = np.concatenate([pa["centre"], -pa["centre"], extra]).astype(np.float32).reshape(-1)

    def check(height, need):
        assert need <= stack and (stack == 48 and need > 32 or stack == 64 and need > 48), (height, need)
    _bits_and_bytes_vs_oracle(orc, V, I, P, N, D, tree, check)


COMB_H = 1.0e7


def _comb(lg=20):
    """2^lg triangles along x, triangle k over [k, k + 0.9] x [0, H]^2 in the plane y = z (H = 10^7 m), and an opaque
    wall across them at x = 24.5 (two triangles, the last). A ray from x < 0 with a positive x component, started far
    from y = z, is inside every comb box and meets no comb triangle, so a nearest-first walk pushes the hit siblings at
    every record level on its way down to the first leaves (about 30 entries: past the 24 that each lane keeps in LDS,
    into the walk's global part), and then finds the wall only in the siblings it pops back from the deepest entries:
    a wrong global entry would lose the wall, and the ray would run on along the comb. (An instrumented build counted,
    over the SAH tree, about 9 global writes per ray along the comb, every one read back, and for the views of one
    sensor about 750 writes and 75 reads in 15-degree cells, 6,700 and 690 in 5-degree ones; the measured scenes' walks
    reach the global part a few times in all.)"""
    nt = 1 << lg
    k = np.arange(nt, dtype=np.float64)
    z = np.zeros(nt)
    V = np.stack([np.stack([k, z, z], 1), np.stack([k + 0.9, z, z], 1), np.stack([k, z + COMB_H, z + COMB_H], 1)], 1).reshape(-1, 3)
    h, w = COMB_H, 24.5
    wall = np.array([[w, 0, 0], [w, h, 0], [w, h, h], [w, 0, 0], [w, h, h], [w, 0, h]])
    V = np.concatenate([V, wall]).astype(np.float32).reshape(-1)
    return V, np.arange(len(V) // 3, dtype=np.uint32)


def _comb_sensors(rng, n):
    """n sensors facing +x: three quarters before the comb (their rays along it end at the wall), the rest beyond its
    end (open)"""
    P = np.stack([np.full(n, -5.0), rng.uniform(0.6, 0.9, n) * COMB_H, rng.uniform(0.1, 0.4, n) * COMB_H], 1).astype(np.float32)
    P[3 * n // 4:, 0] = (1 << 20) + 10.0
    return P, np.tile(np.array([1.0, 0.0, 0.0], np.float32), (n, 1))


@pytest.mark.parametrize("tree", ["sah", "lbvh"])
def test_deep_stacks_comb_bits_and_bytes(orc, tree):
    rng = np.random.default_rng(11)
    V, I = _comb()
    # few rays along the comb: the oracle's byte mode walks all of it, past the wall
    P, N = _comb_sensors(rng, 24)
    D = rng.normal(size=(16, 3))
    D[:8] = D[:8] * 0.2 + [1.0, 0.0, 0.0]   # 8 of them within a few degrees of +x
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    _bits_and_bytes_vs_oracle(orc, V, I, P, N, D.astype(np.float32).reshape(-1), tree)


def test_deep_stacks_comb_views(native, orc, tmp_path):
    V, I = _comb()
    nt = len(I) // 3
    pts = V.reshape(-1, 3).astype(np.float64)
    cnt, idx = np.full(nt, 3, np.int32), np.arange(3 * nt, dtype=np.int32)
    y0, z0 = 0.7 * COMB_H, 0.2 * COMB_H
    gp = np.array([[-5.0, y0, z0], [-5.0, y0 + 1, z0], [-5.0, y0 + 1, z0 + 1], [-5.0, y0, z0 + 1]])   # one quad facing +x
    gc, gi = np.array([4], np.int32), np.arange(4, dtype=np.int32)
    print("case: views from one quad along a comb of %d triangles, 360 x 120 degrees in 10-degree cells" % nt, flush=True)
    dump = str(tmp_path / "rays.npy")
    r = native.run("views", ["--fov", "360", "--cell", "10", "--dump-rays", dump], native.SimulationMesh("grid", gp, gc, gi),
                   context=native.SimulationMesh("comb", pts, cnt, idx))
    o = orc.views([orc.MeshArrays("grid", gp, gc, gi, sensor=True), orc.MeshArrays("comb", pts, cnt, idx, sensor=False)],
                  fov=(360.0, -70.0, 50.0), cell=10.0, raw=True, nthreads=min(16, os.cpu_count() or 1))
    assert r.manifest["columns"] == o["columns"]
    cls = np.load(dump)
    assert cls.shape == o["cls"].shape
    bad = np.argwhere(cls != o["cls"])
    assert len(bad) == 0, (len(bad), bad[:8])
    assert np.array_equal(np.load(dump + ".depth.npy").view(np.uint32), o["dist"].view(np.uint32))
    from test_gpu_views import compare
    compare(r.values, o["values"], o["columns"])
    print("classes: sky %d, ground %d, obstruction %d" % tuple((cls == c).sum() for c in range(3)), flush=True)


def test_interior_views_slice_vs_oracle(native, orc, tmp_path):
    from test_gpu_views import compare
    P, Cn, I = W.interior()
    gp, gc, gi = W.floor_grid(n=200, x0=40.0, y0=40.0)
    print("case: interior (%d triangles), 200 floor sensors, all around, every ray" % (2 * len(Cn)), flush=True)
    room = native.SimulationMesh("interior", P, Cn, I)
    grid = native.SimulationMesh("grid", gp, gc, gi)
    dump = str(tmp_path / "rays.npy")
    r = native.run("views", W.INTERIOR_VIEW_ARGS + ["--dump-rays", dump], grid, context=room)
    o = orc.views([orc.MeshArrays("grid", gp, gc, gi, sensor=True), orc.MeshArrays("interior", P, Cn, I, sensor=False)],
                  fov=(360.0, -70.0, 50.0), cell=2.0, jitter=True, far=1000.0, reach=(6.0, 20.0, 50.0), distance={"isovist", "mean", "min"},
                  offset=1.2, selfOcclude=False, raw=True, nthreads=min(16, os.cpu_count() or 1))
    assert r.manifest["columns"] == o["columns"]
    cls = np.load(dump)
    assert cls.shape == (200, 10800)
    bad = np.argwhere(cls != o["cls"])
    assert len(bad) == 0, (len(bad), bad[:8])
    assert np.array_equal(np.load(dump + ".depth.npy").view(np.uint32), o["dist"].view(np.uint32))
    compare(r.values, o["values"], o["columns"])
    print("classes: sky %d, ground %d, obstruction %d" % tuple((cls == c).sum() for c in range(3)), flush=True)
