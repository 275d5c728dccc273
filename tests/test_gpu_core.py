"""The HIP path through the seam (include/dl_device.h) against the oracle, on the GPU.

Bit-exact: the LBVH arrays (vs the oracle's host build of the same algorithm), visibility bits, the fused reduces, the
FP32 series. Within tolerance: the GPU sky (device libm vs host libm, FP64) and the split-BF16 MFMA series (vs FP64).
"""
import math

import numpy as np
import pytest

from scenes import Tris, box_scene_check_gpu, dirs_check_gpu, quad_dirs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import daylight_amd
    g = daylight_amd.Gpu()
    yield g
    g.close()


def oracle_scene(orc, V, I, tree="lbvh"):
    import ctypes as C
    L = orc.lib()
    sc = L.orc_scene_new(V, len(V) // 3, I, len(I) // 3)
    if tree == "sah":
        L.orc_scene_sah(sc)
    node, tri, prim = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nn = L.orc_scene_nodes(sc, C.byref(node), C.byref(tri), C.byref(prim))
    nt = len(I) // 3
    nodes = np.ctypeslib.as_array(C.cast(node, C.POINTER(C.c_float)), shape=(nn * 16,)).reshape(nn, 16).copy() if nt else np.zeros((0, 16), np.float32)
    leaves = np.ctypeslib.as_array(C.cast(tri, C.POINTER(C.c_float)), shape=(nt * 12,)).reshape(nt, 12).copy() if nt else np.zeros((0, 12), np.float32)
    return sc, nodes, leaves


def city(rng, nbox, extent):
    t = Tris()
    for _ in range(nbox):
        x0, y0 = rng.uniform(-extent, extent, 2)
        w, d = rng.uniform(5, 30, 2)
        h = rng.lognormal(math.log(15), 0.5)
        t.box(x0, y0, 0.0, x0 + w, y0 + d, h)
    e = 1.5 * extent
    t.quad((-e, -e, 0), (e, -e, 0), (e, e, 0), (-e, e, 0))
    return t


def bits_matrix(bits, nc, nDir):
    out = np.zeros((nc, nDir), np.uint8)
    for d in range(nDir):
        out[:, d] = (bits[(d // 32) * nc:(d // 32 + 1) * nc] >> (d % 32)) & 1
    return out


@pytest.mark.parametrize("nbox", [1, 7, 300, 20000])
def test_lbvh_bit_exact_vs_host_build(gpu, orc, nbox):
    print("case: LBVH, %d boxes" % nbox, flush=True)
    rng = np.random.default_rng(nbox)
    t = city(rng, nbox, 10.0 * math.sqrt(nbox) + 20)
    V, I = t.v, t.i
    gpu.set_tree("lbvh")
    try:
        gpu.build(V, I)
    finally:
        gpu.set_tree("sah")
    nodes, leaves, height = gpu.scene()
    sc, onodes, oleaves = oracle_scene(orc, V, I)
    oh = orc.lib().orc_scene_height(sc)
    orc.lib().orc_scene_free(sc)
    assert nodes.shape == onodes.shape and leaves.shape == oleaves.shape
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    assert np.array_equal(leaves[:, :10].view(np.uint32), oleaves[:, :10].view(np.uint32))
    assert height == oh and 1 <= height <= 160


def sah_scenes():
    """Scenes for each path of the SAH build: a small task only (<= 64 triangles), medium tasks (<= 4096), big tasks
    (chunked levels), and degenerate ranges (identical triangles: no valid plane, halved; a flat grid)."""
    out = []
    for nbox in (1, 5, 300, 2000):
        rng = np.random.default_rng(nbox)
        out.append(("city %d" % nbox, city(rng, nbox, 10.0 * math.sqrt(nbox) + 20)))
    t = Tris()
    for _ in range(150):
        t.quad((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))   # 300 identical-pair triangles
    out.append(("identical quads", t))
    t = Tris()
    for j in range(60):
        for i in range(60):
            t.quad((i, j, 0), (i + 1, j, 0), (i + 1, j + 1, 0), (i, j + 1, 0))
    rng = np.random.default_rng(3)
    for _ in range(40):
        x0, y0 = rng.uniform(0, 55, 2)
        t.box(x0, y0, 0.1, x0 + rng.uniform(1, 5), y0 + rng.uniform(1, 5), rng.uniform(1, 10))
    out.append(("grid + boxes", t))
    return out


@pytest.mark.parametrize("k", range(6))
def test_sah_bit_exact_vs_host_build(gpu, orc, k):
    """The binned-SAH tree (sah.hip) against the oracle's host restatement of the same rules: nodes, leaves and height
    bit for bit."""
    name, t = sah_scenes()[k]
    print("case: SAH, %s (%d triangles)" % (name, len(t.i) // 3), flush=True)
    V, I = t.v, t.i
    gpu.build(V, I)
    nodes, leaves, height = gpu.scene()
    sc, onodes, oleaves = oracle_scene(orc, V, I, "sah")
    oh = orc.lib().orc_scene_height(sc)
    orc.lib().orc_scene_free(sc)
    assert nodes.shape == onodes.shape and leaves.shape == oleaves.shape
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    assert np.array_equal(leaves[:, :10].view(np.uint32), oleaves[:, :10].view(np.uint32))
    assert height == oh


def test_sah_big_levels_bit_exact(gpu, orc):
    """More than 4096 triangles per task at the top: the chunked level kernels (bin, choose, count, offsets, scatter)."""
    print("case: SAH, 20000-box city", flush=True)
    rng = np.random.default_rng(20000)
    t = city(rng, 20000, 10.0 * math.sqrt(20000) + 20)
    gpu.build(t.v, t.i)
    nodes, leaves, height = gpu.scene()
    sc, onodes, oleaves = oracle_scene(orc, t.v, t.i, "sah")
    oh = orc.lib().orc_scene_height(sc)
    orc.lib().orc_scene_free(sc)
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    assert np.array_equal(leaves[:, :10].view(np.uint32), oleaves[:, :10].view(np.uint32))
    assert height == oh


def test_single_triangle_and_empty_scene(gpu, orc):
    t = Tris()
    t.quad((-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1))
    V, I = t.v, t.i[:3]
    gpu.build(V, I)
    nodes, leaves, _ = gpu.scene()
    sc, onodes, oleaves = oracle_scene(orc, V, I)
    orc.lib().orc_scene_free(sc)
    assert np.array_equal(nodes.view(np.uint32), onodes.view(np.uint32))
    D = dirs_check_gpu()
    P = np.array([1, 1, 0, -1, -1, 0, -1, -1, 0], np.float32)
    N = np.array([0, 0, 1, 0, 0, 1, 0, 0, -1], np.float32)
    gpu.build(np.zeros(0, np.float32), np.zeros(0, np.uint32))   # empty: every hemisphere direction open
    gpu.set_scene(P, N, D)
    gpu.trace(0, 3)
    b = bits_matrix(gpu.bits(3), 3, len(D) // 3)
    assert b[0].sum() == 147 and b[2].sum() == 1


def trace_both(gpu, orc, t, P, N, D, cull=False):
    gpu.build(t.v, t.i)
    gpu.set_scene(P, N, D)
    nc, nDir = len(P) // 3, len(D) // 3
    gpu.trace(0, nc, cull)
    got = gpu.bits(nc)
    sc = orc.lib().orc_scene_new(t.v, len(t.v) // 3, t.i, len(t.i) // 3)
    want = np.zeros_like(got)
    orc.lib().orc_trace_bits(sc, P, N, nc, D, nDir, int(cull), want, 8)
    orc.lib().orc_scene_free(sc)
    return got, want


def test_bits_box_scene(gpu, orc):
    P = np.array([1, 1, 0, -1, -1, 0, -1, -1, 0, 1, 1, 2, 1, 1, 1], np.float32)
    N = np.array([0, 0, 1, 0, 0, 1, 0, 0, -1, 0, 0, -1, 0, 0, 1], np.float32)
    for cull in (False, True):
        got, want = trace_both(gpu, orc, box_scene_check_gpu(), P, N, dirs_check_gpu(), cull)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("L", [1000.0, 10000.0])
def test_bits_km_own_planes(gpu, orc, L):
    h = L / 2
    t = Tris()
    t.quad((-h, -h, 0), (h, -h, 0), (h, h, 0), (-h, h, 0))
    P, N = [], []
    for p in ((0, 0, 0), (0.37, -0.21, 0), (0.3 * h, -0.4 * h, 0)):
        for nz in (1, -1):
            P += p; N += (0, 0, nz)
    P, N = np.array(P, np.float32), np.array(N, np.float32)
    D = quad_dirs()
    got, want = trace_both(gpu, orc, t, P, N, D)
    assert np.array_equal(got, want)
    b = bits_matrix(got, len(P) // 3, len(D) // 3)
    assert int(b.sum()) == 6 * (len(D) // 3) // 2   # every hemisphere direction open


def test_bits_city_random_sensors(gpu, orc):
    rng = np.random.default_rng(11)
    t = city(rng, 400, 200.0)
    nc = 3000
    P = np.zeros((nc, 3), np.float32)
    P[:, :2] = rng.uniform(-250, 250, (nc, 2))
    P[:, 2] = np.where(rng.uniform(size=nc) < 0.5, 0.0, rng.uniform(0, 30, nc))
    N = np.zeros((nc, 3), np.float32)
    N[:, 2] = 1
    k = rng.uniform(size=nc) < 0.3
    n = rng.normal(size=(int(k.sum()), 3))
    N[k] = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    sky = orc.patches(2)
    D = np.concatenate([sky["centre"], -sky["centre"]]).astype(np.float32).reshape(-1)
    got, want = trace_both(gpu, orc, t, P.reshape(-1), N.reshape(-1), D)
    assert np.array_equal(got, want)


def test_reduce_bit_exact(gpu, orc):
    rng = np.random.default_rng(2)
    t = city(rng, 100, 100.0)
    nc = 2000
    P = np.zeros((nc, 3), np.float32)
    P[:, :2] = rng.uniform(-120, 120, (nc, 2))
    N = np.tile(np.array([0, 0, 1], np.float32), (nc, 1))
    sky = orc.patches(1)
    D = np.concatenate([sky["centre"], -sky["centre"], rng.normal(size=(500, 3)).astype(np.float32)])
    D[290:] /= np.linalg.norm(D[290:], axis=1, keepdims=True)
    D = D.astype(np.float32).reshape(-1)
    got, want = trace_both(gpu, orc, t, P.reshape(-1), N.reshape(-1), D)
    assert np.array_equal(got, want)
    nDir = len(D) // 3
    W = rng.uniform(0, 2, (3, nDir)).astype(np.float32)
    useCos = [1, 0, 1]
    out = gpu.reduce(nc, W, useCos)
    for q in range(3):
        ref = np.zeros(nc)
        orc.lib().orc_reduce(want, nc, nDir, N.reshape(-1), D, W[q].copy(), useCos[q], ref)
        assert np.array_equal(out[q], ref)


def steps_of(orc, dni, dhi):
    st = orc.make_steps(dni=dni, dhi=dhi)
    return dict(month=np.array([s.month for s in st], np.int32), day=np.array([s.day for s in st], np.int32),
                jd=np.array([s.jd for s in st], np.int32), hour=np.array([s.hour for s in st]), dt=np.array([s.dt for s in st]),
                dni=np.array([s.dni for s in st], np.float32), dhi=np.array([s.dhi for s in st], np.float32)), st


def tcos_table():
    return np.array([math.cos((math.pi / 2.0 / 1024) * i) for i in range(1025)])


def oracle_rows(orc, mf, st, sky, t, grefl=0.2):
    """the oracle's gendaymtx row at the GPU's own sun angles (so only the Perez evaluation is compared)"""
    k = math.pi / 180.0   # degrees to radians as run.cpp:370 (alt x (pi / 180))
    r_ = orc.gendaymtx_step(mf, st[t].jd, sky["alt"][t] * k, sky["azi"][t] * k, st[t].dni, st[t].dhi, grefl, 0, 1, 0)
    r = np.asarray(r_, np.float64)   # each channel in double before the CIE weights (a float32 product would round them)
    return 0.265074126 * r[0::3] + 0.670114631 * r[1::3] + 0.064811243 * r[2::3]


def test_sky_suns_weights_and_series_matrix(gpu, orc):
    """GPU sun and Perez sky (FP64, device libm) against the oracle (host libm): suns to 1e-9 degrees, per-patch
    weights to 1e-12 relative, the series matrix S to one float rounding."""
    dni, dhi = orc.clear_sky(52.52, 13.41, 1)
    steps, st = steps_of(orc, dni, dhi)
    for mf in (1, 2, 4):
        print("case: GPU sky, sky density %d" % mf, flush=True)
        out = gpu.sky(steps, 52.52, 13.41, 1.0, False, 0.0, 0.2, mf, tcos_table(), wantRad=True, wantVis=True, series=True)
        pa = orc.patches(mf)
        for t in range(0, 8760, 7):
            a, z, u, v = orc.sun_noaa(52.52, 13.41, 1.0, st[t].month, st[t].day, st[t].hour)
            assert abs(out["alt"][t] - a) < 1e-9 and abs(out["azi"][t] - z) < 1e-9 and bool(out["up"][t]) == u
            assert np.max(np.abs(out["vec"][t] - v)) <= 1.2e-7
        skyRad = np.zeros(pa["P"])
        gnd = 0.0
        S = out["S"]
        col = 0
        for t in range(8760):
            if st[t].dni + st[t].dhi <= 1e-4:
                continue
            L = oracle_rows(orc, mf, st, out, t)
            skyRad += L[1:] * st[t].dt / 1000.0
            gnd += L[0] * st[t].dt / 1000.0
            assert np.all(np.abs(S[:, col] - L[1:].astype(np.float32)) <= 2.5e-7 * np.abs(L[1:])), t
            assert abs(out["g"][col] - np.float32(L[0])) <= 2.5e-7 * abs(L[0])
            col += 1
        assert col == out["Tup"]
        assert np.allclose(out["skyRad"], skyRad, rtol=1e-12, atol=0)
        assert abs(out["gndRad"][0] - gnd) <= 1e-12 * gnd


def series_case(gpu, orc, mf):
    """1,500 sensors among a random city, a cloudy clear-sky year at sky density mf: the traced chunk on the device, its
    inputs to the contraction and the oracle's arguments for the same contraction"""
    rng = np.random.default_rng(5)
    t = city(rng, 200, 120.0)
    nc = 1500
    P = np.zeros((nc, 3), np.float32)
    P[:, :2] = rng.uniform(-150, 150, (nc, 2))
    P[:, 2] = np.where(rng.uniform(size=nc) < 0.5, 0.0, rng.uniform(0, 20, nc))
    N = np.tile(np.array([0, 0, 1], np.float32), (nc, 1))
    pa = orc.patches(mf)
    Pp = pa["P"]
    dni, dhi = orc.clear_sky(52.52, 13.41, 1)
    dni[24 * 40:24 * 200] *= rng.uniform(0, 1, 24 * 160).astype(np.float32)   # some cloud
    steps, st = steps_of(orc, dni, dhi)
    sky = gpu.sky(steps, 52.52, 13.41, 1.0, False, 0.0, 0.2, mf, tcos_table(), wantRad=True, series=True)
    up = np.nonzero(sky["up"])[0]
    sunVec = sky["vec"][up]
    D = np.concatenate([pa["centre"], -pa["centre"], sunVec]).astype(np.float32).reshape(-1)
    nDir = len(D) // 3
    gpu.build(t.v, t.i)
    gpu.set_scene(P.reshape(-1), N.reshape(-1), D)
    gpu.trace(0, nc)
    bits = gpu.bits(nc)
    Tup = sky["Tup"]
    colOf = np.cumsum(dni + dhi > 1e-4) - 1
    colOf[~(dni + dhi > 1e-4)] = -1
    sunStep = np.array([max(0, colOf[t_]) for t_ in up], np.int32)
    sunW = np.array([0.0 if colOf[t_] < 0 else sky["sunRad"][t_] * 1000.0 / 1.0 for t_ in up], np.float32)
    wsky = (pa["omega"] / np.float32(1)).astype(np.float32)
    row = np.arange(Pp, dtype=np.int32)
    Wg = np.zeros(nDir, np.float32)
    Wg[Pp:2 * Pp] = wsky
    G = gpu.reduce(nc, Wg, [1])[0].astype(np.float32)
    S, g = sky["S"], sky["g"]   # the same S and g on both sides: the contraction alone is compared
    return dict(nc=nc, Pp=Pp, Tup=Tup, row=row, wsky=wsky, G=G, sunStep=sunStep, sunW=sunW, S=S, g=g,
                oargs=lambda w, S_, g_, G_=G: (bits, nc, N.reshape(-1), D, Pp, row, w, np.ascontiguousarray(S_, np.float32).reshape(-1), Tup,
                                               np.ascontiguousarray(g_, np.float32), np.ascontiguousarray(G_, np.float32), 2 * Pp, len(up), sunStep, sunVec.astype(np.float32).reshape(-1),
                                               sunW))


@pytest.mark.gpu
@pytest.mark.parametrize("mf", [2, 4])
def test_series_fp32_bit_exact_and_mfma_vs_fp64(gpu, orc, mf):
    print("case: series methods, sky density %d" % mf, flush=True)
    c = series_case(gpu, orc, mf)
    nc, Pp, Tup, row, wsky, G, sunStep, sunW = (c[k] for k in ("nc", "Pp", "Tup", "row", "wsky", "G", "sunStep", "sunW"))
    E32 = gpu.series(nc, Pp, row, wsky, Pp, Tup, G, 2 * Pp, sunStep, sunW, 0)
    E32o = np.zeros((nc, Tup), np.float32)
    args = c["oargs"](wsky, c["S"], c["g"])
    orc.lib().orc_series32(*args, E32o.reshape(-1))
    assert np.array_equal(E32, E32o)   # the reference kernel's fma order, bit for bit
    E64 = np.zeros((nc, Tup))
    orc.lib().orc_series64(*args, E64.reshape(-1), 8)
    scale = np.maximum(np.abs(E64), 1e-3 * np.abs(E64).max())
    for method, bar in ((1, 2e-5), (2, 2e-6), (3, 2e-6)):
        Em = gpu.series(nc, Pp, row, wsky, Pp, Tup, G, 2 * Pp, sunStep, sunW, method)
        err = np.max(np.abs(Em - E64) / scale)
        print("series method %d: max rel error vs FP64 %.3e" % (method, err))
        assert err <= bar


@pytest.mark.gpu
@pytest.mark.parametrize("k", [40, -40])
def test_series_split_scales_extreme_magnitudes(gpu, orc, k):
    """The sky weights and G times 2^k, S and g times 2^-k (the same E, the operands far outside fp16's range unscaled):
    every MFMA method against the oracle's FP64 series of the same inputs. FP16x3 holds its bar only by scaling each
    operand by a power of two into fp16's range and back (unscaled, 2^40 x the weights overflows fp16, 2^-40 flushes)."""
    print("case: series methods, weights x 2^%d, S x 2^%d" % (k, -k), flush=True)
    c = series_case(gpu, orc, 2)
    nc, Pp, Tup, row, sunStep, sunW = (c[k_] for k_ in ("nc", "Pp", "Tup", "row", "sunStep", "sunW"))
    w = (c["wsky"] * np.float32(2.0 ** k)).astype(np.float32)
    G = (c["G"] * np.float32(2.0 ** k)).astype(np.float32)
    S = (c["S"] * np.float32(2.0 ** -k)).astype(np.float32)
    g = (c["g"] * np.float32(2.0 ** -k)).astype(np.float32)
    gpu.set_sky_series(S, g)
    E64 = np.zeros((nc, Tup))
    orc.lib().orc_series64(*c["oargs"](w, S, g, G), E64.reshape(-1), 8)
    scale = np.maximum(np.abs(E64), 1e-3 * np.abs(E64).max())
    for method, bar in ((1, 2e-5), (2, 2e-6), (3, 2e-6)):
        Em = gpu.series(nc, Pp, row, w, Pp, Tup, G, 2 * Pp, sunStep, sunW, method)
        assert np.all(np.isfinite(Em)), method
        err = np.max(np.abs(Em - E64) / scale)
        print("series method %d: max rel error vs FP64 %.3e" % (method, err))
        assert err <= bar, method
