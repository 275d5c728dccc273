"""Reflected light (the bounce pass) on the CPU: the oracle's restatement of __raygen__bounce (reference
src/trace_dev.cu:260-311) pinned by the closed forms of the reference's own checks (tests/check_gpu.cpp:456-667): a
row sums to pi x the weight of the reflected paths, a diffuse plane's reflections spread over the patch rows as cos^2
says, a ceiling sends everything to the ground bin, the ground / porous-ceiling sandwich's geometric series for 1 to 20
bounces, glass's cosine-weighted <T> and <R> under a pane, the black wall's shadowed patches with --north.
"""
import math

import numpy as np
import pytest

from scenes import Tris
from test_transmission_cpu import GLASS, THIN, TWIN, slot_rows

KPI = math.pi


def bounce(orc, tr, P, N, slots, bounces, M, mf=1, north=0.0, cull=False, i0=0, threads=8):
    L = orc.lib()
    sc = L.orc_scene_new(tr.v, len(tr.v) // 3, tr.i, len(tr.i) // 3)
    nc = len(P) // 3
    nb = 144 * mf * mf + 2
    dc = np.zeros(nb * nc, np.float32)
    ts = np.ascontiguousarray(tr.s, np.uint32)
    L.orc_bounce(sc, np.ascontiguousarray(P, np.float32), np.ascontiguousarray(N, np.float32), nc, i0, ts.ctypes.data, slot_rows(slots),
                 len(slots), bounces, M, mf, north, int(cull), dc, threads)
    L.orc_scene_free(sc)
    return dc.reshape(nb, nc)


def plane(z=0.0, away=False, slot=0, half=1000.0):
    t = Tris()
    a = (-half, -half, z), (half, -half, z), (half, half, z), (-half, half, z)
    t.quad(*(a if not away else (a[0], a[3], a[2], a[1])), slot)
    return t


SA_P = np.array([0, 0, 1, 3, 0, 0.01, 0, 0, 1], np.float32)
SA_N = np.array([0, 0, -1, 0, 0, -1, 0, 0, 1], np.float32)
HALF = [(0.5, 0.0, 0.0, 0)]


@pytest.mark.parametrize("mf", [1, 2, 4])
def test_plane_rows(orc, mf):
    """check_gpu.cpp:466-497: a 2 km plane of rho 0.5 under sensors looking down 1 m and 1 cm above it reflects every path
    once: a row sums to 0.5 pi, the ground bin stays empty, patch row r gets 0.5 pi (sin^2 of its top - of its bottom)
    within 5 sigma (the zenith cap the last bin); the sensor looking up sees nothing reflected."""
    M = 65536
    dc = bounce(orc, plane(), SA_P, SA_N, HALF, 1, M, mf)
    pm = orc.patches(mf)
    for i in range(2):
        assert abs(dc[:, i].astype(np.float64).sum() / (0.5 * KPI) - 1) <= 1e-4 and dc[0, i] == 0
    assert pm["row"][pm["P"] - 1] == 7 * mf
    for r in range(7 * mf + 1):
        c = dc[1:, 0][pm["row"] == r].astype(np.float64).sum()
        s0, s1 = math.sin(pm["alpha"] * r), 1.0 if r == 7 * mf else math.sin(pm["alpha"] * (r + 1))
        q = s1 * s1 - s0 * s0
        assert abs(c - 0.5 * KPI * q) <= 5 * 0.5 * KPI * math.sqrt(q * (1 - q) / M), r
    assert dc[:, 2].sum() == 0


def test_plane_winding_and_culling(orc):
    """check_gpu.cpp:498-505: wound away, the same rows (the normal on the ray's side); culled, none; wound towards the
    sensors, culling changes nothing."""
    rows = bounce(orc, plane(), SA_P, SA_N, HALF, 1, 4096)
    assert np.array_equal(bounce(orc, plane(away=True), SA_P, SA_N, HALF, 1, 4096), rows)
    assert not bounce(orc, plane(away=True), SA_P, SA_N, HALF, 1, 4096, cull=True).any()
    assert np.array_equal(bounce(orc, plane(), SA_P, SA_N, HALF, 1, 4096, cull=True), rows)


def test_own_plane_and_chunk(orc):
    """check_gpu.cpp:506-541: sensors in the plane of a 10 km quad see nothing reflected (a first hit on that plane is
    their own); a chunk's rows are those of the whole run (seeds from the global index)."""
    so_p = np.array([0, 0, 0, 0.01, 0.02, 0], np.float32)
    so_n = np.array([0, 0, 1, 0, 0, -1], np.float32)
    assert not bounce(orc, plane(half=5000.0), so_p, so_n, HALF, 1, 4096).any()
    P = np.concatenate([np.array([[0.01 * k, 1, 0.5] for k in range(297)], np.float32).reshape(-1), SA_P])
    N = np.concatenate([np.tile(np.array([0, 0, -1], np.float32), 297), SA_N])
    every = bounce(orc, plane(), P, N, HALF, 1, 2048)
    part = bounce(orc, plane(), P[297 * 3:], N[297 * 3:], HALF, 1, 2048, i0=297)
    assert np.array_equal(part, every[:, 297:])


def test_ceiling_to_ground_bin(orc):
    """check_gpu.cpp:557-565: a 2 km ceiling of rho 0.4 over a sensor looking up sends everything to the ground bin."""
    dc = bounce(orc, plane(1.0, away=True), np.zeros(3, np.float32), np.array([0, 0, 1], np.float32), [(0.4, 0.0, 0.0, 0)], 1, 4096)
    assert abs(dc[0, 0] / (0.4 * KPI) - 1) <= 1e-4 and dc[1:, 0].sum() == 0


@pytest.mark.parametrize("rg,rc,tau,N", [(0.5, 0.4, 0.5, 1), (0.5, 0.4, 0.5, 2), (0.5, 0.4, 0.5, 3), (0.9, 0.9, 0.1, 19), (0.9, 0.9, 0.1, 20)])
def test_sandwich_series(orc, rg, rc, tau, N):
    """check_gpu.cpp:566-596: between a ground (rho_g) and a porous ceiling (tau, rho_c), looking down light leaves after
    an odd number of reflections, pi rho_g tau sum_k (rho_c rho_g)^k; looking up after an even one, the first off the
    ceiling, pi tau rho_c rho_g sum_k (rho_c rho_g)^k; the ground bin stays empty. 64 sensors each (independent seeds):
    the mean within 5 sigma of the spread (Russian roulette at high albedo)."""
    t = plane(0.0, slot=0, half=1e5)
    t.quad((-1e5, -1e5, 2), (1e5, -1e5, 2), (1e5, 1e5, 2), (-1e5, 1e5, 2), 1)
    P = np.array([[0.1 * (k % 64), 0, 1] for k in range(128)], np.float32).reshape(-1)
    Nn = np.array([[0, 0, -1 if k < 64 else 1] for k in range(128)], np.float32).reshape(-1)
    dc = bounce(orc, t, P, Nn, [(rg, 0.0, 0.0, 0), (rc, tau, 0.0, 0)], N, 4096)
    for up in (0, 1):
        expect, f, k = 0.0, 1.0, 0
        while 2 * k + 1 + up <= N:
            expect += f
            f *= rc * rg
            k += 1
        expect *= KPI * tau * (rc * rg if up else rg)
        esc = dc[1:, 64 * up:64 * up + 64].astype(np.float64).sum(axis=0)
        mean, sigma = esc.mean(), esc.std(ddof=1) / math.sqrt(64)
        assert dc[0, 64 * up:64 * up + 64].sum() == 0
        assert abs(mean - expect) <= 5 * sigma + 1e-4 * expect, (up, mean, expect, sigma)


def glass_means(orc, tn, n=200000):
    """<T>, <R>: glass's transmittance and reflectance averaged over a cosine-weighted hemisphere (midpoint rule in c^2)"""
    L = orc.lib()
    import ctypes as C
    T, R = C.c_float(), C.c_float()
    mT = mR = 0.0
    for k in range(n):
        c = math.sqrt((k + 0.5) / n)
        L.orc_glass_TR(c, tn, C.byref(T), C.byref(R))
        mT += T.value / n
        mR += R.value / n
    return mT, mR


def test_glass_under_pane(orc):
    """check_gpu.cpp:597-649: glass (tau 0.64) 1 m over a ground (rho 0.5), sensors under it looking up: light leaves
    through the glass after the glass mirrored it and the ground reflected it, pi <R> rho_g <T> (1 + [N 2] rho_g <R>);
    a 1 cm pane box (thin, entry the sheet) and the sheet doubled with its twin give the sheet's rows; --fresnel off
    (no glass flag): nothing counts; opaque glass mirrors every path to the ground bin, pi <R>."""
    tn = float(np.float32(orc.lib().orc_tn_from_Tn(0.64)))
    mT, mR = glass_means(orc, tn, 20000)
    a0, a1, z0, z1 = -1000, 1000, 1, 1.01

    def base():
        t = Tris()
        t.quad((-1e4, -1e4, 0), (1e4, -1e4, 0), (1e4, 1e4, 0), (-1e4, 1e4, 0), 0)
        return t

    def lower(t):
        t.quad((a0, a0, z0), (a0, a1, z0), (a1, a1, z0), (a1, a0, z0), 1)

    def upper(t, z):
        t.quad((a0, a0, z), (a1, a0, z), (a1, a1, z), (a0, a1, z), 1)

    sheet, box, twin, twin2 = base(), base(), base(), base()
    lower(sheet); lower(box); lower(twin); upper(twin, z0); upper(twin2, z0); lower(twin2)
    upper(box, z1)
    box.quad((a0, a0, z0), (a0, a0, z1), (a0, a1, z1), (a0, a1, z0), 1)
    box.quad((a1, a0, z0), (a1, a1, z0), (a1, a1, z1), (a1, a0, z1), 1)
    box.quad((a0, a0, z0), (a1, a0, z0), (a1, a0, z1), (a0, a0, z1), 1)
    box.quad((a0, a1, z0), (a0, a1, z1), (a1, a1, z1), (a1, a1, z0), 1)
    P = np.array([[0.3 * (k % 8) - 1, 0.3 * (k // 8) - 1, 0.1 + 0.0125 * k] for k in range(64)], np.float32).reshape(-1)
    Nn = np.tile(np.array([0, 0, 1], np.float32), 64)
    slots = [(0.5, 0.0, 0.0, 0), (0.0, 0.64, tn, GLASS)]
    rows = None
    for N in (1, 2):
        dc = bounce(orc, sheet, P, Nn, slots, N, 4096)
        esc = dc[1:].astype(np.float64).sum(axis=0)
        expect = KPI * mR * 0.5 * mT * (1 + 0.5 * mR if N == 2 else 1)
        assert abs(esc.mean() - expect) <= 5 * esc.std(ddof=1) / 8 + 1e-4 * expect and dc[0].sum() == 0, (N, esc.mean(), expect)
        rows = dc
    assert np.array_equal(bounce(orc, box, P, Nn, [slots[0], (0.0, 0.64, tn, GLASS | THIN)], 2, 4096), rows)
    for t in (twin, twin2):
        assert np.array_equal(bounce(orc, t, P, Nn, [slots[0], (0.0, 0.64, tn, GLASS | THIN | TWIN)], 2, 4096), rows)
    assert not bounce(orc, sheet, P, Nn, [slots[0], (0.0, 0.64, tn, 0)], 2, 4096).any()
    pane = Tris()
    lower(pane)
    dc = bounce(orc, pane, P, Nn, [(0.0, 0.0, tn, GLASS)], 1, 4096)
    g = dc[0].astype(np.float64)
    assert abs(g.mean() - KPI * mR) <= 5 * g.std(ddof=1) / 8 + 1e-4 * KPI * mR and dc[1:].sum() == 0


@pytest.mark.parametrize("mf", [1, 2])
def test_black_wall_and_north(orc, mf):
    """check_gpu.cpp:650-667: a black wall 1 m east of a sensor looking down 1 cm over a ground (rho 0.5): no reflected
    light reaches a patch east of it below 72 degrees, some reaches every patch west of it; --north 90 turns that east to
    true south."""
    t = Tris()
    t.quad((-1000, -1000, 0), (1000, -1000, 0), (1000, 1000, 0), (-1000, 1000, 0), 0)
    t.quad((1, -1000, 0), (1, -1000, 1000), (1, 1000, 1000), (1, 1000, 0), 1)
    pm = orc.patches(mf)
    for north in (0.0, 90.0):
        dc = bounce(orc, t, np.array([0, 0, 0.01], np.float32), np.array([0, 0, -1], np.float32), [(0.5, 0.0, 0.0, 0), (0.0, 0.0, 0.0, 0)], 1,
                    16384 if mf == 1 else 65536, mf, north)
        dark = lit = wrong = 0
        for p in range(pm["P"]):
            if pm["alt"][p] > 72 * KPI / 180:
                continue
            az = math.fmod(pm["azi"][p] * 180 / KPI - north + 720, 360)
            if 30 <= az <= 150:
                dark += 1
                wrong += dc[p + 1, 0] != 0
            if 210 <= az <= 330:
                lit += 1
                wrong += not (dc[p + 1, 0] > 0)
        assert dark > 30 * mf and lit > 30 * mf and wrong == 0, (north, dark, lit, wrong)
