"""Transmission (byte mode) on the GPU: the HIP byte trace, its reduces and the contraction over bytes against the oracle,
and the daylight / views tools with --transmission on against the reference's own expectations (tests/check_gpu.cpp:
251-339 and 832-950, tests/check_views.cpp:457-500).

Bars: transmittance bytes identical to the oracle's; reduces bit for bit; the FP32 series bit for bit, the MFMA series
within their FP64 bars; run values as the reference's checks state them (exact where it checks ==).
"""
import math
import os

import numpy as np
import pytest

from scenes import Tris, dirs_check_gpu
from test_gpu_core import city, steps_of, tcos_table
from test_transmission_cpu import GLASS, INWARD, KGLASS_DEG, KGLASS_T, THIN, b255, porous_scene, slot_rows, trace_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import daylight_amd
    g = daylight_amd.Gpu()
    yield g
    g.close()


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


def gpu_bytes(gpu, tr, P, N, D, slots, cull=False):
    gpu.build(tr.v, tr.i)
    gpu.set_tri_slots(tr.s)
    gpu.set_slots(slots)
    gpu.set_scene(P, N, D)
    nc = len(P) // 3
    gpu.trace(0, nc, cull=cull, translucent=True)
    return gpu.bytes(nc)


PORO = [(0.2, 0.0, 0.0, 0), (0.2, 0.5, 0.0, 0), (0.2, 0.5, 0.0, 0)]


def test_porous_panes(gpu, orc):
    """check_gpu.cpp:251-320 at the seam: the bytes, the reduce and the series over them, a chunk at i0 = 1, 20000 copies
    of one sensor, and the same scene traced without any-hit giving an all-opaque build's bits."""
    print("case: porous panes, seam", flush=True)
    t, P, N = porous_scene()
    D = dirs_check_gpu()
    nDir, Z, NA = len(D) // 3, 145, 147
    by = gpu_bytes(gpu, t, P, N, D, PORO)
    assert by[Z, 0] == 128 and by[Z, 1] == 64 and by[Z, 2] == 0 and by[Z, 3] == 0 and by[NA, 3] == 255 and by[NA, 0] == 0
    assert np.array_equal(by, trace_bytes(orc, t, P, N, D, PORO))
    W = (np.float32(0.04) * (np.float32(1) + np.float32(0.01) * np.arange(nDir, dtype=np.float32))).astype(np.float32)
    got = gpu.reduce(4, np.stack([W, W]), [1, 0])
    for q, useCos in enumerate((1, 0)):
        want = np.zeros(4)
        orc.lib().orc_reduce_bytes(np.ascontiguousarray(by.reshape(-1)), 4, nDir, N, D, W, useCos, want)
        assert np.array_equal(got[q], want), useCos
    assert np.all(got[0] > 0)
    # series: three steps of a June sky, one sun (the zenith, direction Z) at step 1 with 100
    nSky, T = 145, 3
    # the contraction reads S and g from the last sky: the oracle gets the same S, g
    sky = gpu.sky(dict(month=np.array([6] * T, np.int32), day=np.array([21] * T, np.int32), jd=np.array([172] * T, np.int32),
                       hour=np.array([11.0, 12.0, 13.0]), dt=np.ones(T), dni=np.full(T, 500, np.float32), dhi=np.full(T, 100, np.float32)),
                  52.52, 13.41, 1.0, False, 0.0, 0.2, 1, tcos_table(), wantRad=True, series=True)
    gpu.set_scene(P, N, D)
    gpu.trace(0, 4, translucent=True)
    row = np.arange(nSky, dtype=np.int32)
    wsky = W[:nSky].copy()
    G = np.full(4, 0.25, np.float32)
    sunStep, sunW = np.array([1], np.int32), np.array([100.0], np.float32)
    S, g = np.ascontiguousarray(sky["S"]), np.ascontiguousarray(sky["g"])
    args = (np.ascontiguousarray(by.reshape(-1)), 4, N, D, nSky, row, wsky, S.reshape(-1), sky["Tup"], g, G, Z, 1, sunStep,
            np.array([0, 0, 1], np.float32), sunW)
    E32o = np.zeros((4, sky["Tup"]), np.float32)
    orc.lib().orc_series32_bytes(*args, E32o.reshape(-1))
    E64 = np.zeros((4, sky["Tup"]))
    orc.lib().orc_series64_bytes(*args, E64.reshape(-1), 2)
    assert np.array_equal(gpu.series(4, nSky, row, wsky, 145, sky["Tup"], G, Z, sunStep, sunW, 0), E32o)
    scale = np.maximum(np.abs(E64), 1e-3 * np.abs(E64).max())
    for method, bar in ((1, 2e-5), (2, 2e-6), (3, 2e-6)):
        Em = gpu.series(4, nSky, row, wsky, 145, sky["Tup"], G, Z, sunStep, sunW, method)
        assert np.max(np.abs(Em - E64) / scale) <= bar, method
    # a chunk at i0 = 1: the same bytes and sums
    gpu.trace(1, 2, translucent=True)
    assert np.array_equal(gpu.bytes(2), by[:, 1:3])
    assert np.array_equal(gpu.reduce(2, np.stack([W, W]), [1, 0]), got[:, 1:3])
    import daylight_amd
    with pytest.raises(daylight_amd.GpuError):
        gpu.bits(2)   # bytes were written, not bits
    # 20000 copies of s1 (several waves and words in flight): every byte equals s1's
    many = 20000
    gpu.set_scene(np.tile(P[3:6], many), np.tile(N[3:6], many), D)
    gpu.trace(0, many, translucent=True)
    cb = gpu.bytes(many)
    assert np.array_equal(cb, np.repeat(by[:, 1:2], many, axis=1))
    # transmission off: the same scene traced without any-hit gives the bits of an all-opaque build
    gpu.set_scene(P, N, D)
    gpu.trace(0, 4)
    off = gpu.bits(4)
    gpu.build(t.v, t.i)
    gpu.trace(0, 4)
    assert np.array_equal(off, gpu.bits(4))
    assert (off[(Z // 32) * 4] >> (Z % 32)) & 1 == 0


def test_glass_pane(gpu, orc):
    """check_gpu.cpp:323-338: bytes at the five rtrace angles within half a step; --fresnel off passes tau (163)."""
    print("case: glass pane, seam", flush=True)
    t = Tris()
    t.quad((-100, -100, 1), (100, -100, 1), (100, 100, 1), (-100, 100, 1))
    tn = float(np.float32(orc.lib().orc_tn_from_Tn(0.64)))
    D = np.array([[math.sin(math.radians(a)), 0, math.cos(math.radians(a))] for a in KGLASS_DEG], np.float32).reshape(-1)
    P, N = np.zeros(3, np.float32), np.array([0, 0, 1], np.float32)
    gb = gpu_bytes(gpu, t, P, N, D, [(0.0, 0.64, tn, GLASS)])[:, 0]
    print("glass bytes", list(gb), "(rtrace x 255: 163.2 159.3 137.6 94.3 29.4)")
    for k in range(5):
        assert abs(int(gb[k]) - 255 * KGLASS_T[k]) <= 0.5, (k, gb)
    assert np.array_equal(gb, trace_bytes(orc, t, P, N, D, [(0.0, 0.64, tn, GLASS)])[:, 0])
    assert list(gpu_bytes(gpu, t, P, N, D, [(0.0, 0.64, tn, 0)])[:, 0]) == [163] * 5


def glass_city(rng, nbox, extent):
    """A city of opaque, porous and glass blocks (slots 0-4) with thin glass panes (slot 5, 1 cm) and a porous canopy
    (slot 6) over an opaque ground (slot 0) the sensors lie in"""
    t = Tris()
    for _ in range(nbox):
        x0, y0 = rng.uniform(-extent, extent, 2)
        w, d = rng.uniform(5, 30, 2)
        h = rng.lognormal(math.log(15), 0.5)
        t.box(x0, y0, 0.0, x0 + w, y0 + d, h, slot=int(rng.integers(0, 5)))
    for _ in range(nbox // 2):
        x0, y0, z0 = rng.uniform(-extent, extent), rng.uniform(-extent, extent), rng.uniform(2, 20)
        if rng.uniform() < 0.5:
            t.box(x0, y0, z0, x0 + rng.uniform(3, 15), y0 + 0.01, z0 + rng.uniform(2, 6), slot=5)
        else:
            t.box(x0, y0, z0, x0 + 0.01, y0 + rng.uniform(3, 15), z0 + rng.uniform(2, 6), slot=5)
    t.quad((-extent, -extent, 25), (0, -extent, 25), (0, 0, 25), (-extent, 0, 25), 6)
    e = 1.5 * extent
    t.quad((-e, -e, 0), (e, -e, 0), (e, e, 0), (-e, e, 0), 0)
    return t


def city_slots(orc):
    tn = float(np.float32(orc.lib().orc_tn_from_Tn(0.64)))
    tn8 = float(np.float32(orc.lib().orc_tn_from_Tn(0.8)))
    return [(0.2, 0.0, 0.0, 0), (0.2, 0.5, 0.0, 0), (0.0, 0.64, tn, GLASS), (0.2, 0.3, 0.0, THIN), (0.0, 0.8, tn8, GLASS | THIN | INWARD),
            (0.0, 0.64, tn, GLASS | THIN), (0.3, 0.6, 0.0, 0)]


@pytest.mark.parametrize("cull", [False, True])
def test_city_bytes_reduce_series_vs_oracle(gpu, orc, cull):
    print("case: glass city bytes, cull %s" % cull, flush=True)
    rng = np.random.default_rng(21 + cull)
    t = glass_city(rng, 120, 90.0)
    slots = city_slots(orc)
    nc = 1200
    P = np.zeros((nc, 3), np.float32)
    P[:, :2] = rng.uniform(-120, 120, (nc, 2))
    P[:, 2] = np.where(rng.uniform(size=nc) < 0.6, 0.0, rng.uniform(0, 30, nc))   # most in the ground's plane
    N = np.tile(np.array([0, 0, 1], np.float32), (nc, 1))
    side = rng.uniform(size=nc) < 0.2   # some facing sideways
    ang = rng.uniform(0, 2 * np.pi, nc)
    N[side] = np.stack([np.cos(ang[side]), np.sin(ang[side]), np.zeros(side.sum())], 1).astype(np.float32)
    pa = orc.patches(1)
    Pp = pa["P"]
    extra = rng.normal(size=(40, 3))
    extra[:, 2] = np.abs(extra[:, 2]) + 0.05
    extra /= np.linalg.norm(extra, axis=1, keepdims=True)
    D = np.concatenate([pa["centre"], -pa["centre"], extra]).astype(np.float32).reshape(-1)
    nDir = len(D) // 3
    by = gpu_bytes(gpu, t, P.reshape(-1), N.reshape(-1), D, slots, cull=cull)
    want = trace_bytes(orc, t, P.reshape(-1), N.reshape(-1), D, slots, cull=cull)
    bad = np.argwhere(by != want)
    assert len(bad) == 0, (len(bad), [(int(d), int(i), int(by[d, i]), int(want[d, i])) for d, i in bad[:8]])
    frac = (by > 0) & (by < 255)
    print("rays %d, partly passed %d, blocked %d" % (by.size, frac.sum(), (by == 0).sum()))
    assert frac.sum() > 1000
    W = rng.uniform(0, 1, (3, nDir)).astype(np.float32)
    got = gpu.reduce(nc, W, [1, 0, 1])
    for q, useCos in enumerate((1, 0, 1)):
        ref = np.zeros(nc)
        orc.lib().orc_reduce_bytes(np.ascontiguousarray(want.reshape(-1)), nc, nDir, N.reshape(-1), D, W[q].copy(), useCos, ref)
        assert np.array_equal(got[q], ref), q


def test_city_series_bytes(gpu, orc):
    print("case: glass city series over bytes", flush=True)
    rng = np.random.default_rng(8)
    t = glass_city(rng, 100, 80.0)
    slots = city_slots(orc)
    nc = 900
    P = np.zeros((nc, 3), np.float32)
    P[:, :2] = rng.uniform(-100, 100, (nc, 2))
    N = np.tile(np.array([0, 0, 1], np.float32), (nc, 1))
    mf = 1
    pa = orc.patches(mf)
    Pp = pa["P"]
    dni, dhi = orc.clear_sky(52.52, 13.41, 1)
    steps, st = steps_of(orc, dni, dhi)
    sky = gpu.sky(steps, 52.52, 13.41, 1.0, False, 0.0, 0.2, mf, tcos_table(), wantRad=True, series=True)
    up = np.nonzero(sky["up"])[0]
    sunVec = sky["vec"][up]
    D = np.concatenate([pa["centre"], -pa["centre"], sunVec]).astype(np.float32).reshape(-1)
    nDir = len(D) // 3
    by = gpu_bytes(gpu, t, P.reshape(-1), N.reshape(-1), D, slots)
    Tup = sky["Tup"]
    colOf = np.cumsum(dni + dhi > 1e-4) - 1
    colOf[~(dni + dhi > 1e-4)] = -1
    sunStep = np.array([max(0, colOf[t_]) for t_ in up], np.int32)
    sunW = np.array([0.0 if colOf[t_] < 0 else sky["sunRad"][t_] * 1000.0 for t_ in up], np.float32)
    wsky = pa["omega"].astype(np.float32)
    row = np.arange(Pp, dtype=np.int32)
    Wg = np.zeros(nDir, np.float32)
    Wg[Pp:2 * Pp] = wsky
    G = gpu.reduce(nc, Wg, [1])[0].astype(np.float32)
    S, g = np.ascontiguousarray(sky["S"]), np.ascontiguousarray(sky["g"])
    args = (np.ascontiguousarray(by.reshape(-1)), nc, N.reshape(-1), D, Pp, row, wsky, S.reshape(-1), Tup, g, G, 2 * Pp, len(up), sunStep,
            sunVec.astype(np.float32).reshape(-1), sunW)
    E32o = np.zeros((nc, Tup), np.float32)
    orc.lib().orc_series32_bytes(*args, E32o.reshape(-1))
    assert np.array_equal(gpu.series(nc, Pp, row, wsky, Pp, Tup, G, 2 * Pp, sunStep, sunW, 0), E32o)
    E64 = np.zeros((nc, Tup))
    orc.lib().orc_series64_bytes(*args, E64.reshape(-1), 8)
    scale = np.maximum(np.abs(E64), 1e-3 * np.abs(E64).max())
    for method, bar in ((1, 2e-5), (2, 2e-6), (3, 2e-6)):
        Em = gpu.series(nc, Pp, row, wsky, Pp, Tup, G, 2 * Pp, sunStep, sunW, method)
        err = np.max(np.abs(Em - E64) / scale)
        print("series over bytes, method %d: max rel error vs FP64 %.3e" % (method, err))
        assert err <= bar


# ---------------------------------------------------------------------------------------------------- the tools


def squares(native, name, z, half, xs, slots=None):
    """check_gpu.cpp:35-44: one square face (facing +z) per x, centred at (x, 0, z)"""
    pts, idx = [], []
    for x in xs:
        b = len(pts) // 3
        pts += [x - half, -half, z, x + half, -half, z, x + half, half, z, x - half, half, z]
        idx += [b, b + 1, b + 2, b + 3]
    return native.SimulationMesh(name, np.array(pts, np.float64), np.array([4] * len(xs), np.int32), np.array(idx, np.int32), slots=slots)


def box_handle(native, name, x0, y0, z0, x1, y1, z1, soup, flip=False):
    from test_transmission_cpu import box_mesh, flipped
    m = box_mesh(native, name, x0, y0, z0, x1, y1, z1, soup)
    if flip:
        m = flipped(m)
    return native.SimulationMesh(name, np.array(m["points"], np.float64), np.array(m["counts"], np.int32), np.array(m["indices"], np.int32))


BOSTON = ["--quantity", "sunhours", "--lat", "42.35", "--lon", "-71.07", "--tz", "-5", "--transmission", "on"]
NOON = ["--time", "06-21T12:00"]


def test_run_optics_sun_hours(native, tmp_path):
    """check_gpu.cpp:836-860 and 897-921: two meshes x two slots (A/0 tau 1, A/1 tau 0.5, B/0 tau 0.25, B/1 opaque);
    --transmission off gives the values of a run without optics; --dump-bits writes uint8 rows [sensors x directions]
    with a transmitting slot, uint32 bits when the transmitting slot is not in the scene."""
    print("case: run, optics sun hours", flush=True)
    ground = squares(native, "ground", 0, 0.5, [0, 100, 200, 300])
    A = squares(native, "A", 1, 20, [0, 100], slots=[0, 1])
    B = squares(native, "B", 1, 20, [200, 300], slots=[0, 1])
    optics = "A/0=0,1,porous;A/1=0.2,0.5,porous;B=0.2,0;B/0=0.2,0.25,porous"
    r = native.run("daylight", BOSTON + NOON + ["--optics", optics], ground, [A, B])
    assert list(r.values[:, 0]) == [1, b255(128), b255(64), 0]
    off = native.run("daylight", [a for a in BOSTON if a not in ("--transmission", "on")] + NOON + ["--optics", optics], ground, [A, B])
    plain = native.run("daylight", [a for a in BOSTON if a not in ("--transmission", "on")] + NOON, ground, [A, B])
    assert list(off.values[:, 0]) == list(plain.values[:, 0]) == [0, 0, 0, 0]
    dump = str(tmp_path / "vis.npy")
    three = ["--from", "06-21T12:00", "--to", "06-21T15:00"]
    r = native.run("daylight", BOSTON + three + ["--optics", optics, "--dump-bits", dump], ground, [A, B])
    assert list(r.values[:, 0]) == [3, 3 * b255(128), 3 * b255(64), 0]
    d = np.load(dump)
    assert d.dtype == np.uint8 and d.shape == (4, 3)
    assert d.reshape(-1).tolist() == [255, 255, 255, 128, 128, 128, 64, 64, 64, 0, 0, 0]
    for extra in (["--no-self-occlude"], ["--no-occlude", "ground"]):
        os.remove(dump)
        r = native.run("daylight", BOSTON + three + ["--optics", "ground=0.2,0.5,porous", "--dump-bits", dump] + extra, ground, [A, B])
        assert list(r.values[:, 0]) == [0, 0, 0, 0]
        d = np.load(dump)
        assert d.dtype == np.uint32 and d.shape == (4, 1), extra


def test_run_glass_wall(native, orc):
    """check_gpu.cpp:862-877: a glass wall south of the sensor turned 30 degrees: --fresnel off passes tau like porous;
    on, glass at the sun's incidence (~75 degrees)."""
    print("case: run, glass wall", flush=True)
    s30, c30 = 0.5, math.sqrt(0.75)
    ground = squares(native, "ground", 0, 0.5, [0])
    wall = native.SimulationMesh("wall", np.array([-20 * c30, -1 - 20 * s30, 0, 20 * c30, -1 + 20 * s30, 0, 20 * c30, -1 + 20 * s30, 10,
                                                   -20 * c30, -1 - 20 * s30, 10], np.float64), np.array([4], np.int32), np.arange(4, dtype=np.int32))
    rg = native.run("daylight", BOSTON + NOON + ["--optics", "wall=0,0.64,glass"], ground, [wall])
    rgoff = native.run("daylight", BOSTON + NOON + ["--optics", "wall=0,0.64,glass", "--fresnel", "off"], ground, [wall])
    rpor = native.run("daylight", BOSTON + NOON + ["--optics", "wall=0,0.64,porous"], ground, [wall])
    assert list(rgoff.values[:, 0]) == list(rpor.values[:, 0]) == [b255(163)]
    sn = orc.sun_noaa(42.35, -71.07, -5, 6, 21, 12.0)
    vec = sn["vec"] if isinstance(sn, dict) else sn[3]
    cw = abs(-s30 * vec[0] + c30 * vec[1])
    Tw = orc.lib().orc_glass_T(cw, float(np.float32(orc.lib().orc_tn_from_Tn(0.64))))
    print("glass wall at %.1f deg incidence passes %.4f (glass %.4f, tau 0.64)" % (math.degrees(math.acos(cw)), rg.values[0, 0], Tw))
    assert abs(rg.values[0, 0] - b255(int(round(255 * Tw)))) <= 1.0 / 255 + 1e-6


def test_run_thin_solids_and_residency(native):
    """check_gpu.cpp:879-895: a pane box (indexed, a soup, wound inward) passes tau once, the 3 m cube twice, a
    double-faced sheet once; --cull-backfaces: every closed solid crossed through one front face; a resident scene keeps
    its thin solids (reused as built, with culling, after a tau edit)."""
    print("case: run, thin solids", flush=True)
    ground = squares(native, "ground", 0, 0.5, [0, 100, 200, 300, 400])
    pane = box_handle(native, "pane", -10, -10, 1, 10, 10, 1.01, False)
    soup = box_handle(native, "soup", 90, -10, 1, 110, 10, 1.01, True)
    cube = box_handle(native, "cube", 198.5, -1.5, 1, 201.5, 1.5, 4, False)
    flip = box_handle(native, "flip", 290, -10, 1, 310, 10, 1.01, False, flip=True)
    twin = native.SimulationMesh("twin", np.array([390, -10, 1, 410, -10, 1, 410, 10, 1, 390, 10, 1], np.float64), np.array([4, 4], np.int32),
                                 np.array([0, 1, 2, 3, 3, 2, 1, 0], np.int32))
    optics = ";".join("%s=0.2,0.5,porous" % n for n in ("pane", "soup", "cube", "flip", "twin"))
    ctx = [pane, soup, cube, flip, twin]
    thinV = [b255(128), b255(128), b255(64), b255(128), b255(128)]
    r = native.run("daylight", BOSTON + NOON + ["--optics", optics], ground, ctx)
    assert not r.stats["scene_reused"] and list(r.values[:, 0]) == thinV
    r = native.run("daylight", BOSTON + NOON + ["--optics", optics], ground, ctx)
    assert r.stats["scene_reused"] and list(r.values[:, 0]) == thinV
    r = native.run("daylight", BOSTON + NOON + ["--optics", optics, "--cull-backfaces"], ground, ctx)
    assert r.stats["scene_reused"] and list(r.values[:, 0]) == [b255(128)] * 5
    edited = optics.replace("pane=0.2,0.5", "pane=0.2,0.25")
    r = native.run("daylight", BOSTON + NOON + ["--optics", edited], ground, ctx)
    assert r.stats["scene_reused"] and list(r.values[:, 0]) == [b255(64), b255(128), b255(64), b255(128), b255(128)]


def test_run_residency_lid(native):
    """check_gpu.cpp:923-937: a tau edit reuses the scene; a slot that starts to transmit rebuilds."""
    print("case: run, lid residency", flush=True)
    ground = squares(native, "ground", 0, 0.5, [0])
    lid = squares(native, "lid", 1, 20, [0])
    run = lambda tau: native.run("daylight", BOSTON + NOON + ["--optics", "lid=0.2,%g,porous" % tau], ground, [lid])
    r = run(0.3)
    assert not r.stats["scene_reused"] and abs(r.values[0, 0] - 0.3) <= 0.5 / 255 + 1e-6
    r = run(0.5)
    assert r.stats["scene_reused"] and r.values[0, 0] == b255(128)
    ground2 = squares(native, "ground", 0, 0.5, [0])   # new handles: another scene key
    lid2 = squares(native, "lid", 1, 20, [0])
    run2 = lambda tau: native.run("daylight", BOSTON + NOON + ["--optics", "lid=0.2,%g,porous" % tau], ground2, [lid2])
    r = run2(0)
    assert not r.stats["scene_reused"] and r.values[0, 0] == 0
    r = run2(0.5)
    assert not r.stats["scene_reused"] and r.values[0, 0] == b255(128)
    r = run2(0.25)
    assert r.stats["scene_reused"] and r.values[0, 0] == b255(64)


def test_views_glass_vs_oracle(native, orc, tmp_path):
    """check_views.cpp:457-500: a facade looking through a glass pane at a park; --transmission off the pane blocks; on,
    rays pass it and the glazing share counts them; porous panes still block; every ray's class and distance as the
    oracle's."""
    print("case: views, glass pane", flush=True)

    def quad(name, pts, sensor):
        return (native.SimulationMesh(name, np.array(pts, np.float64), np.array([4], np.int32), np.arange(4, dtype=np.int32)),
                orc.MeshArrays(name, np.array(pts, np.float64), np.array([4], np.int32), np.arange(4, dtype=np.int32), sensor=sensor))

    facade = quad("facade", [0, 0, 0, 4, 0, 0, 4, 0, 3, 0, 0, 3], True)         # facing -y
    pane = quad("pane", [-1, -2, -0.5, 5, -2, -0.5, 5, -2, 4, -1, -2, 4], False)
    park = quad("park", [-30, -40, 0, 30, -40, 0, 30, -40, 8, -30, -40, 8], False)
    nat = [facade[0], pane[0], park[0]]
    ora = [facade[1], pane[1], park[1]]
    base = ["--target", "park=green", "--cell", "3"]
    kw = dict(targets={"park": "green"}, cell=3.0)
    off = native.run("views", base + ["--optics", "pane=0,0.64,glass"], [nat[0]], nat[1:])
    o = orc.views(ora, **kw)
    assert np.allclose(off.values, o["values"], rtol=1e-12, atol=0, equal_nan=True)
    dump = str(tmp_path / "rays.npy")
    on = native.run("views", base + ["--optics", "pane=0,0.64,glass", "--transmission", "on", "--dump-rays", dump], [nat[0]], nat[1:])
    o = orc.views(ora, transmission=True, glass=("pane",), raw=True, **kw)
    assert on.manifest["columns"] == o["columns"]
    cols = o["columns"]
    gz = cols.index("glazing_front")
    green = cols.index("green_front")
    assert np.all(on.values[:, gz] > 0.05) and np.all(on.values[:, green] > off.values[:, green])
    assert np.array_equal(np.load(dump), o["cls"])
    assert np.array_equal(np.load(dump + ".depth.npy").view(np.uint32), o["dist"].view(np.uint32))
    for c, name in enumerate(cols):
        g, w = on.values[:, c], o["values"][:, c]
        tol = 1e-12 * np.abs(w) if "distance" in name else 0
        assert np.all(np.abs(g - w) <= tol), name
    por = native.run("views", base + ["--optics", "pane=0,0.64,porous", "--transmission", "on"], [nat[0]], nat[1:])
    assert np.all(por.values[:, gz] == 0)
    assert np.array_equal(por.values[:, green], off.values[:, green])
