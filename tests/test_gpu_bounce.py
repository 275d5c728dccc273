"""Reflected light (the bounce pass) on the GPU: k_bounce's rows against the oracle's restatement (the same paths, bit for
bit), the reduce and the contraction over them, the sky's binned sun, and the daylight tool with --bounces against the
reference's own expectations (tests/check_gpu.cpp:456-667 and 951-1030).

Bars: rows identical to the oracle's; reduce_dc bit for bit; FP32 series bit for bit, the MFMA series within their FP64
bars; the sky's sun bins within host-vs-device libm; run values as the reference's checks state them.
"""
import math
import os

import numpy as np
import pytest

from scenes import Tris
from test_bounce_cpu import HALF, SA_N, SA_P, bounce as orc_bounce, glass_means, plane
from test_gpu_core import city, steps_of, tcos_table
from test_transmission_cpu import GLASS, INWARD, THIN, TWIN

pytestmark = pytest.mark.gpu
KPI = math.pi


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


def gpu_rows(gpu, tr, P, N, slots, bounces, M, mf=1, north=0.0, cull=False, i0=0):
    gpu.build(tr.v, tr.i)
    gpu.set_tri_slots(tr.s)
    gpu.set_slots(slots)
    gpu.set_scene(P, N, np.zeros(3, np.float32))
    nc = len(P) // 3 - i0
    return gpu.bounce(i0, nc, bounces, M, mf, north, cull)


def check_same(gpu, orc, tr, P, N, slots, bounces, M, mf=1, north=0.0, cull=False):
    got = gpu_rows(gpu, tr, P, N, slots, bounces, M, mf, north, cull)
    want = orc_bounce(orc, tr, P, N, slots, bounces, M, mf, north, cull)
    bad = np.argwhere(got != want)
    assert len(bad) == 0, (len(bad), [(int(b), int(i), float(got[b, i]), float(want[b, i])) for b, i in bad[:6]])
    return got


@pytest.mark.parametrize("mf", [1, 2, 4])
def test_plane_rows(gpu, orc, mf):
    """check_gpu.cpp:466-497 on the GPU rows, and the oracle's rows bit for bit"""
    print("case: bounce plane, sky density %d" % mf, flush=True)
    dc = check_same(gpu, orc, plane(), SA_P, SA_N, HALF, 1, 8192, mf)
    for i in range(2):
        assert abs(dc[:, i].astype(np.float64).sum() / (0.5 * KPI) - 1) <= 1e-4 and dc[0, i] == 0
    assert dc[:, 2].sum() == 0


def test_winding_culling_chunk(gpu, orc):
    print("case: bounce winding, culling, chunk", flush=True)
    rows = gpu_rows(gpu, plane(), SA_P, SA_N, HALF, 1, 4096)
    assert np.array_equal(gpu_rows(gpu, plane(away=True), SA_P, SA_N, HALF, 1, 4096), rows)
    assert not gpu_rows(gpu, plane(away=True), SA_P, SA_N, HALF, 1, 4096, cull=True).any()
    assert np.array_equal(gpu_rows(gpu, plane(), SA_P, SA_N, HALF, 1, 4096, cull=True), rows)
    P = np.concatenate([np.array([[0.01 * k, 1, 0.5] for k in range(297)], np.float32).reshape(-1), SA_P])
    N = np.concatenate([np.tile(np.array([0, 0, -1], np.float32), 297), SA_N])
    every = check_same(gpu, orc, plane(), P, N, HALF, 1, 2048)
    assert np.array_equal(gpu_rows(gpu, plane(), P, N, HALF, 1, 2048, i0=297), every[:, 297:])


@pytest.mark.parametrize("rg,rc,tau,N", [(0.5, 0.4, 0.5, 3), (0.9, 0.9, 0.1, 20)])
def test_sandwich(gpu, orc, rg, rc, tau, N):
    print("case: bounce sandwich N %d" % N, flush=True)
    t = plane(0.0, slot=0, half=1e5)
    t.quad((-1e5, -1e5, 2), (1e5, -1e5, 2), (1e5, 1e5, 2), (-1e5, 1e5, 2), 1)
    P = np.array([[0.1 * (k % 64), 0, 1] for k in range(128)], np.float32).reshape(-1)
    Nn = np.array([[0, 0, -1 if k < 64 else 1] for k in range(128)], np.float32).reshape(-1)
    check_same(gpu, orc, t, P, Nn, [(rg, 0.0, 0.0, 0), (rc, tau, 0.0, 0)], N, 2048)


def test_glass_thin_twin(gpu, orc):
    print("case: bounce glass sheet, thin box, twins", flush=True)
    tn = float(np.float32(orc.lib().orc_tn_from_Tn(0.64)))
    mT, mR = glass_means(orc, tn, 20000)
    a0, a1, z0, z1 = -1000, 1000, 1, 1.01

    def base():
        t = Tris()
        t.quad((-1e4, -1e4, 0), (1e4, -1e4, 0), (1e4, 1e4, 0), (-1e4, 1e4, 0), 0)
        return t
    sheet, box, twin = base(), base(), base()
    for t in (sheet, box, twin):
        t.quad((a0, a0, z0), (a0, a1, z0), (a1, a1, z0), (a1, a0, z0), 1)
    twin.quad((a0, a0, z0), (a1, a0, z0), (a1, a1, z0), (a0, a1, z0), 1)
    box.quad((a0, a0, z1), (a1, a0, z1), (a1, a1, z1), (a0, a1, z1), 1)
    box.quad((a0, a0, z0), (a0, a0, z1), (a0, a1, z1), (a0, a1, z0), 1)
    box.quad((a1, a0, z0), (a1, a1, z0), (a1, a1, z1), (a1, a0, z1), 1)
    box.quad((a0, a0, z0), (a1, a0, z0), (a1, a0, z1), (a0, a0, z1), 1)
    box.quad((a0, a1, z0), (a0, a1, z1), (a1, a1, z1), (a1, a1, z0), 1)
    P = np.array([[0.3 * (k % 8) - 1, 0.3 * (k // 8) - 1, 0.1 + 0.0125 * k] for k in range(64)], np.float32).reshape(-1)
    Nn = np.tile(np.array([0, 0, 1], np.float32), 64)
    g = (0.5, 0.0, 0.0, 0)
    rows = check_same(gpu, orc, sheet, P, Nn, [g, (0.0, 0.64, tn, GLASS)], 2, 4096)
    esc = rows[1:].astype(np.float64).sum(axis=0)
    expect = KPI * mR * 0.5 * mT * (1 + 0.5 * mR)
    assert abs(esc.mean() - expect) <= 5 * esc.std(ddof=1) / 8 + 1e-4 * expect
    assert np.array_equal(gpu_rows(gpu, box, P, Nn, [g, (0.0, 0.64, tn, GLASS | THIN)], 2, 4096), rows)
    assert np.array_equal(gpu_rows(gpu, twin, P, Nn, [g, (0.0, 0.64, tn, GLASS | THIN | TWIN)], 2, 4096), rows)


def test_black_wall_north(gpu, orc):
    print("case: bounce black wall, north 90", flush=True)
    t = Tris()
    t.quad((-1000, -1000, 0), (1000, -1000, 0), (1000, 1000, 0), (-1000, 1000, 0), 0)
    t.quad((1, -1000, 0), (1, -1000, 1000), (1, 1000, 1000), (1, 1000, 0), 1)
    check_same(gpu, orc, t, np.array([0, 0, 0.01], np.float32), np.array([0, 0, -1], np.float32), [(0.5, 0.0, 0.0, 0), (0.0, 0.0, 0.0, 0)], 1,
               8192, 2, 90.0)


def test_city_rows(gpu, orc):
    """a city of mixed slots (rho 0.1-0.6, porous, glass, thin glass) with sensors on the ground, roofs and facades"""
    print("case: bounce city", flush=True)
    rng = np.random.default_rng(17)
    t = Tris()
    for _ in range(60):
        x0, y0 = rng.uniform(-60, 60, 2)
        w, d = rng.uniform(5, 20, 2)
        t.box(x0, y0, 0.0, x0 + w, y0 + d, rng.uniform(5, 30), slot=int(rng.integers(1, 5)))
    t.quad((-150, -150, 0), (150, -150, 0), (150, 150, 0), (-150, 150, 0), 0)
    tn = float(np.float32(orc.lib().orc_tn_from_Tn(0.7)))
    slots = [(0.2, 0.0, 0.0, 0), (0.1, 0.0, 0.0, 0), (0.6, 0.0, 0.0, 0), (0.3, 0.4, 0.0, 0), (0.0, 0.7, tn, GLASS)]
    nc = 400
    P = np.zeros((nc, 3), np.float32)
    P[:, :2] = rng.uniform(-70, 70, (nc, 2))
    P[:, 2] = np.where(rng.uniform(size=nc) < 0.5, 0.0, rng.uniform(0.5, 20, nc))
    N = np.tile(np.array([0, 0, 1], np.float32), (nc, 1))
    side = rng.uniform(size=nc) < 0.3
    ang = rng.uniform(0, 2 * np.pi, nc)
    N[side] = np.stack([np.cos(ang[side]), np.sin(ang[side]), np.zeros(side.sum())], 1).astype(np.float32)
    for cull in (False, True):
        dc = check_same(gpu, orc, t, P.reshape(-1), N.reshape(-1), slots, 3, 256, 2, 30.0, cull)
        assert (dc.sum(axis=0) > 0).sum() > nc // 2


def test_reduce_series_dc(gpu, orc):
    """check_gpu.cpp:542-556: reduce_dc = the double loop over the rows; the series' reflected part in FP32 as the
    reference's kernel (bit for bit) and on MFMA within its bar; summed over the steps it equals the reduce of the summed
    bin radiances"""
    print("case: bounce reduce and series", flush=True)
    rng = np.random.default_rng(4)
    t = city(rng, 40, 40.0)
    nc = 700
    P = np.zeros((nc, 3), np.float32)
    P[:, :2] = rng.uniform(-50, 50, (nc, 2))
    P[:, 2] = 0.3
    N = np.tile(np.array([0, 0, 1], np.float32), (nc, 1))
    N[::3] = np.array([0, 0, -1], np.float32)
    mf = 1
    pa = orc.patches(mf)
    Pp, nb = pa["P"], pa["P"] + 1
    dni, dhi = orc.clear_sky(52.52, 13.41, 1)
    steps, st = steps_of(orc, dni, dhi)
    sky = gpu.sky(steps, 52.52, 13.41, 1.0, False, 0.0, 0.2, mf, tcos_table(), wantRad=True, series=True, reflect=True)
    up = np.nonzero(sky["up"])[0]
    D = np.concatenate([pa["centre"], -pa["centre"], sky["vec"][up]]).astype(np.float32).reshape(-1)
    gpu.build(t.v, t.i)
    gpu.set_tri_slots(np.zeros(len(t.i) // 3, np.uint32))
    gpu.set_slots([(0.35, 0.0, 0.0, 0)])
    gpu.set_scene(P.reshape(-1), N.reshape(-1), D)
    gpu.trace(0, nc)
    dc = gpu.bounce(0, nc, 2, 512, mf)
    W = rng.uniform(0, 2, nb).astype(np.float32)
    got = gpu.reduce_dc(nc, W)
    want = np.zeros(nc)
    for b in range(nb):
        want += dc[b].astype(np.float64) * np.float64(W[b])
    assert np.array_equal(got, want)
    Tup = sky["Tup"]
    colOf = np.cumsum(dni + dhi > 1e-4) - 1
    colOf[~(dni + dhi > 1e-4)] = -1
    sunStep = np.array([max(0, colOf[t_]) for t_ in up], np.int32)
    sunW = np.array([0.0 if colOf[t_] < 0 else sky["sunRad"][t_] * 1000.0 for t_ in up], np.float32)
    wsky = pa["omega"].astype(np.float32)
    row = np.arange(Pp, dtype=np.int32)
    G = np.full(nc, 0.5, np.float32)
    bits = gpu.bits(nc)
    S, g = np.ascontiguousarray(sky["S"]), np.ascontiguousarray(sky["g"])
    # the bins' radiance as the sky made it (row 0 g, 1 + p S plus the sun binned by AddDirect), against the oracle's
    # gendaymtx -d rows at the GPU's sun angles (host vs device libm); the contraction is compared on the device's Sdc
    Sdc = np.ascontiguousarray(sky["Sdc"])
    assert np.array_equal(Sdc[0], g) and np.all(Sdc[1:] >= S)
    k = math.pi / 180.0
    for t_ in [t_ for t_ in range(len(st)) if colOf[t_] >= 0][::7]:
        r = np.asarray(orc.gendaymtx_step(mf, st[t_].jd, sky["alt"][t_] * k, sky["azi"][t_] * k, st[t_].dni, st[t_].dhi, 0.2, 1, 0, 0), np.float64)
        Ls = 0.265074126 * r[3::3] + 0.670114631 * r[4::3] + 0.064811243 * r[5::3]
        want = S[:, colOf[t_]] + Ls.astype(np.float32)
        assert np.all(np.abs(Sdc[1:, colOf[t_]] - want) <= 2.5e-7 * np.abs(want) + 1e-30), t_
    args = (bits, nc, N.reshape(-1), D, Pp, row, wsky, S.reshape(-1), Tup, g, G, 2 * Pp, len(up), sunStep, sky["vec"][up].astype(np.float32).reshape(-1), sunW)
    E32o = np.zeros((nc, Tup), np.float32)
    orc.lib().orc_series32(*args, E32o.reshape(-1))
    E64 = np.zeros((nc, Tup))
    orc.lib().orc_series64(*args, E64.reshape(-1), 8)
    orc.lib().orc_series_dc(np.ascontiguousarray(dc.reshape(-1)), nc, nb, np.ascontiguousarray(Sdc.reshape(-1)), Tup, E32o.ctypes.data, E64.ctypes.data)
    E32 = gpu.series(nc, Pp, row, wsky, Pp, Tup, G, 2 * Pp, sunStep, sunW, 0, reflect=True)
    mism = np.argwhere(E32 != E32o)
    assert len(mism) == 0, (len(mism), mism[:5])
    scale = np.maximum(np.abs(E64), 1e-3 * np.abs(E64).max())
    for method, bar in ((1, 2e-5), (2, 2e-6), (3, 2e-6)):
        Em = gpu.series(nc, Pp, row, wsky, Pp, Tup, G, 2 * Pp, sunStep, sunW, method, reflect=True)
        err = np.max(np.abs(Em - E64) / scale)
        print("series with reflected light, method %d: max rel error vs FP64 %.3e" % (method, err))
        assert err <= bar
    # the sky's binned sun: the cumulative per patch in step order, as the oracle's gendaymtx -d rows give it
    binR = np.zeros(Pp)
    for t_ in range(len(st)):
        k = math.pi / 180.0
        r = np.asarray(orc.gendaymtx_step(mf, st[t_].jd, sky["alt"][t_] * k, sky["azi"][t_] * k, st[t_].dni, st[t_].dhi, 0.2, 1, 0, 0), np.float64)
        binR += (0.265074126 * r[3::3] + 0.670114631 * r[4::3] + 0.064811243 * r[5::3]) * st[t_].dt / 1000.0
    assert np.allclose(sky["sunBinRad"], binR, rtol=1e-6, atol=1e-9 * binR.max())


CLEAR = ["--clear-sky", "--lat", "42.35", "--lon", "-71.07", "--tz", "-5", "--from", "06-01T00:00", "--to", "07-01T00:00"]


def sq(native, name, z, half, xs):
    pts, idx = [], []
    for x in xs:
        b = len(pts) // 3
        pts += [x - half, -half, z, x + half, -half, z, x + half, half, z, x - half, half, z]
        idx += [b, b + 1, b + 2, b + 3]
    return native.SimulationMesh(name, np.array(pts, np.float64), np.array([4] * len(xs), np.int32), np.array(idx, np.int32))


def test_run_bounces(native, tmp_path):
    """check_gpu.cpp:951-1019: a face looking down over a ground mesh with one bounce gets about what the virtual ground
    gives it (3 %); the face looking up and the direct / sun-hours columns are unchanged; the series sums to the
    cumulative value (W/m2, and lux with illuminance first); ground halves of rho 0.3 / 0.1 read their own slot (the face
    1 m east of the border sees the west half with F = (1 - 1/sqrt 2) / 2 of its view); the slot-per-triangle scene is a
    resident scene of its own."""
    print("case: run, bounces", flush=True)
    s2 = native.SimulationMesh("s2", np.array([0.5, -0.5, 1, 1.5, -0.5, 1, 1.5, 0.5, 1, 0.5, 0.5, 1, -0.5, 4.5, 1, 0.5, 4.5, 1, 0.5, 5.5, 1,
                                               -0.5, 5.5, 1], np.float64), np.array([4, 4], np.int32),
                               np.array([3, 2, 1, 0, 4, 5, 6, 7], np.int32))
    east, west = sq(native, "east", 0, 500, [500]), sq(native, "west", 0, 500, [-500])
    base = CLEAR + ["--quantity", "radiation,direct,sunhours", "--no-self-occlude", "--bounce-samples", "65536"]
    virt = native.run("daylight", base, s2).values
    series = str(tmp_path / "s.npy")
    mesh = native.run("daylight", base + ["--bounces", "1", "--timeseries", series], s2, [east, west]).values
    assert abs(mesh[0, 0] / virt[0, 0] - 1) <= 0.03 and mesh[0, 1] == 0 and mesh[0, 2] == 0
    assert np.array_equal(mesh[1], virt[1])
    print("face looking down, June: ground mesh + 1 bounce %.3f kWh/m2, virtual ground %.3f" % (mesh[0, 0], virt[0, 0]))
    E = np.load(series)
    for e in range(2):
        assert abs(E[e].astype(np.float64).sum() / 1000.0 / mesh[e, 0] - 1) <= 1e-5
    lux = native.run("daylight", CLEAR + ["--quantity", "illuminance", "--no-self-occlude", "--bounce-samples", "65536", "--bounces", "1",
                                          "--timeseries", series], s2, [east, west])
    E = np.load(series)
    hours = float(lux.manifest["duration_h"])
    for e in range(2):
        assert lux.values[e, 0] > 0 and abs(E[e].astype(np.float64).sum() / hours / lux.values[e, 0] - 1) <= 1e-5
    halves = lambda opt, n=1: native.run("daylight", base + ["--bounces", str(n), "--optics", opt], s2, [east, west])
    h1 = halves("east=0.3;west=0.1")
    h2 = halves("east=0.1;west=0.3")
    h3 = halves("east=0.1;west=0.3", 0)
    h4 = halves("*=0.2")
    F = (1 - math.sqrt(0.5)) / 2
    x1, x2 = (0.3 * (1 - F) + 0.1 * F) / 0.2, (0.1 * (1 - F) + 0.3 * F) / 0.2
    print("ground halves 0.3 / 0.1 and 0.1 / 0.3: %.4f, %.4f x all 0.2 (expected %.4f, %.4f)" % (h1.values[0, 0] / mesh[0, 0],
                                                                                            h2.values[0, 0] / mesh[0, 0], x1, x2))
    assert abs(h1.values[0, 0] / (x1 * mesh[0, 0]) - 1) <= 0.02
    assert h2.stats["scene_reused"] and abs(h2.values[0, 0] / (x2 * mesh[0, 0]) - 1) <= 0.02
    assert not h3.stats["scene_reused"] and h3.values[0, 0] == 0
    assert h4.stats["scene_reused"] and np.array_equal(h4.values, mesh)
