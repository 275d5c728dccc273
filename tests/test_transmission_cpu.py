"""Transmission (byte mode) on the CPU. The oracle's Radiance glass against the reference's rtrace figures
(tests/check_gpu.cpp:55-70) and the closed forms that fix it; its byte trace against the reference's hand-worked scenes
(check_gpu.cpp:251-339) and against brute force over the hits of small random scenes; the host's thin-solid and
triangle-slot rules (through the C ABI's host checks, no GPU) against the reference's own cases (check_mesh.cpp:266-310).
"""
import ctypes as C
import math

import numpy as np
import pytest

from scenes import Tris, dirs_check_gpu

KGLASS_DEG = [0, 30, 60, 75, 85]
KGLASS_T = [6.400000e-01, 6.248004e-01, 5.397539e-01, 3.696623e-01, 1.153073e-01]   # rtrace -ab 0, tn_from_Tn(0.64)
GLASS, THIN, INWARD, TWIN = 1, 2, 4, 8


def b255(b):
    return float(np.float32(b) / np.float32(255))


def slot_rows(slots):
    """[n][4] float32 rows rho, tau, tn, flags (flags as bits)"""
    a = np.zeros((len(slots), 4), np.float32)
    for k, (rho, tau, tn, flags) in enumerate(slots):
        a[k, :3] = (rho, tau, tn)
        a[k, 3] = np.array([flags], np.uint32).view(np.float32)[0]
    return a.reshape(-1)


def trace_bytes(orc, tr, P, N, D, slots, cull=False, tri_slots=None):
    L = orc.lib()
    sc = L.orc_scene_new(tr.v, len(tr.v) // 3, tr.i, len(tr.i) // 3)
    nc, nDir = len(P) // 3, len(D) // 3
    out = np.zeros(nDir * nc, np.uint8)
    ts = np.ascontiguousarray(tr.s if tri_slots is None else tri_slots, np.uint32)
    L.orc_trace_bytes(sc, np.ascontiguousarray(P, np.float32), np.ascontiguousarray(N, np.float32), nc, np.ascontiguousarray(D, np.float32),
                      nDir, int(cull), ts.ctypes.data, slot_rows(slots), len(slots), out, 2)
    L.orc_scene_free(sc)
    return out.reshape(nDir, nc)


# ---------------------------------------------------------------------------------------------------- glass


def test_glass_against_rtrace(orc):
    """check_gpu.cpp:61-67: the pane at 0 / 30 / 60 / 75 / 85 degrees within 1e-5 of Radiance's rtrace."""
    L = orc.lib()
    tn = L.orc_tn_from_Tn(0.64)
    assert abs(tn - 0.6975761815) < 1e-9   # the rtrace run's material
    for a, want in zip(KGLASS_DEG, KGLASS_T):
        assert abs(L.orc_glass_T(math.cos(math.radians(a)), float(np.float32(tn))) - want) <= 1e-5, a


def fresnel_slab(theta, n=1.52):
    """A lossless pane from the angles (Snell's law, Fresnel's sin / tan forms), the two faces' series summed:
    T = (1 - r) / (1 + r) per polarisation, averaged. An independent form of what glass_T computes with tn = 1."""
    if theta == 0:
        r = ((n - 1) / (n + 1)) ** 2
        return (1 - r) / (1 + r)
    tt = math.asin(math.sin(theta) / n)
    rs = (math.sin(theta - tt) / math.sin(theta + tt)) ** 2
    rp = (math.tan(theta - tt) / math.tan(theta + tt)) ** 2
    return 0.5 * (1 - rs) / (1 + rs) + 0.5 * (1 - rp) / (1 + rp)


def test_glass_closed_forms(orc):
    L = orc.lib()
    for Tn in (0.05, 0.2, 0.5, 0.64, 0.8, 0.9):   # tn_from_Tn inverts the normal-incidence transmittance (10-digit constants)
        assert abs(L.orc_glass_T(1.0, L.orc_tn_from_Tn(Tn)) - Tn) <= 1e-6, Tn
    for a in (0, 10, 30, 45, 60, 75, 89):          # no absorption: the lossless slab of the angle form
        assert abs(L.orc_glass_T(math.cos(math.radians(a)), 1.0) - fresnel_slab(math.radians(a))) <= 2e-7, a
    assert L.orc_glass_T(0.0, 0.7) == 0.0          # grazing passes nothing
    ts = [L.orc_glass_T(math.cos(math.radians(a)), 0.7) for a in range(0, 90, 5)]
    assert all(x > y for x, y in zip(ts, ts[1:]))  # less at every steeper angle
    assert L.orc_tn_from_Tn(1e-5) == 0.0           # below ~1.1e-4 the constants go negative: clamped


# ---------------------------------------------------------------------------------------------------- byte trace


def porous_scene():
    """check_gpu.cpp:251-255: porous panes (slot 1 at z 1 over x in [-50, 50], slot 2 at z 2 over x in [0, 50]), an opaque
    lid (slot 0) at z 0.5 over (30, 30)"""
    t = Tris()
    t.quad((-50, -50, 1), (50, -50, 1), (50, 50, 1), (-50, 50, 1), 1)
    t.quad((0, -50, 2), (50, -50, 2), (50, 50, 2), (0, 50, 2), 2)
    t.quad((25, 25, 0.5), (35, 25, 0.5), (35, 35, 0.5), (25, 35, 0.5), 0)
    P = np.array([-10, 0, 0, 10, 0, 0, 30, 30, 0, -10, 0, 0], np.float32)
    N = np.array([0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, -1], np.float32)
    return t, P, N


def test_porous_panes_bytes(orc):
    """check_gpu.cpp:262-265: s0 under one 0.5 pane, s1 under both, s2 under the lid, s3 looking down."""
    t, P, N = porous_scene()
    D = dirs_check_gpu()
    Z, NA = 145, 147
    by = trace_bytes(orc, t, P, N, D, [(0.2, 0, 0, 0), (0.2, 0.5, 0, 0), (0.2, 0.5, 0, 0)])
    assert by[Z, 0] == 128 and by[Z, 1] == 64 and by[Z, 2] == 0 and by[Z, 3] == 0 and by[NA, 3] == 255 and by[NA, 0] == 0
    # every slot opaque: the bits of the bits pass, as bytes of 255
    opaque = trace_bytes(orc, t, P, N, D, [(0.2, 0, 0, 0)] * 3)
    L = orc.lib()
    sc = L.orc_scene_new(t.v, len(t.v) // 3, t.i, len(t.i) // 3)
    nW = (len(D) // 3 + 31) // 32
    bits = np.zeros(nW * 4, np.uint32)
    L.orc_trace_bits(sc, P, N, 4, D, len(D) // 3, 0, bits, 1)
    L.orc_scene_free(sc)
    for d in range(len(D) // 3):
        want = ((bits[(d // 32) * 4:(d // 32) * 4 + 4] >> (d % 32)) & 1) * 255
        assert np.array_equal(opaque[d], want), d


def test_glass_pane_bytes(orc):
    """check_gpu.cpp:323-338: a glass pane at z 1 seen at the five rtrace angles within half a byte; --fresnel off
    (no glass flag) passes tau: 0.64 x 255 = 163.2 -> 163 at every angle."""
    t = Tris()
    t.quad((-100, -100, 1), (100, -100, 1), (100, 100, 1), (-100, 100, 1))
    tn = float(np.float32(orc.lib().orc_tn_from_Tn(0.64)))
    D = np.array([[math.sin(math.radians(a)), 0, math.cos(math.radians(a))] for a in KGLASS_DEG], np.float32).reshape(-1)
    P, N = np.zeros(3, np.float32), np.array([0, 0, 1], np.float32)
    gb = trace_bytes(orc, t, P, N, D, [(0, 0.64, tn, GLASS)])[:, 0]
    for k in range(5):
        assert abs(int(gb[k]) - 255 * KGLASS_T[k]) <= 0.5, (k, gb)
    assert list(trace_bytes(orc, t, P, N, D, [(0, 0.64, tn, 0)])[:, 0]) == [163] * 5


def test_thin_solid_rule(orc):
    """A 20 x 20 x 0.01 m pane box of tau 0.5 passes tau once as a thin solid (one of its two crossings passes 1) and
    twice without the flag, wound either way; with culling only front faces exist (one crossing). Which face passes:
    the back face (the exit of an outward box), or the front face with the inward flag: a ray up through the box's top
    face alone (normal +z, met from behind) passes all of it as a thin exit, tau as an inward one."""
    P, N = np.array([0, 0, 0], np.float32), np.array([0, 0, 1], np.float32)
    D = np.array([0, 0, 1], np.float32)
    out_ = Tris()
    out_.box(-10, -10, 1, 10, 10, 1.01, slot=0)
    inw = Tris()
    for k in range(0, len(out_._i), 3):   # every triangle wound the other way
        inw._i += [out_._i[k], out_._i[k + 2], out_._i[k + 1]]
    inw._v, inw._s = list(out_._v), list(out_._s)
    for box in (out_, inw):
        assert trace_bytes(orc, box, P, N, D, [(0.2, 0.5, 0, 0)])[0, 0] == 64
        assert trace_bytes(orc, box, P, N, D, [(0.2, 0.5, 0, THIN)])[0, 0] == 128
        assert trace_bytes(orc, box, P, N, D, [(0.2, 0.5, 0, THIN | INWARD)])[0, 0] == 128
    assert trace_bytes(orc, out_, P, N, D, [(0.2, 0.5, 0, 0)], cull=True)[0, 0] == 128
    top = Tris()
    top.quad((-10, -10, 1.01), (10, -10, 1.01), (10, 10, 1.01), (-10, 10, 1.01))
    assert trace_bytes(orc, top, P, N, D, [(0.2, 0.5, 0, THIN)])[0, 0] == 255
    assert trace_bytes(orc, top, P, N, D, [(0.2, 0.5, 0, THIN | INWARD)])[0, 0] == 128
    assert trace_bytes(orc, top, P, N, D, [(0.2, 0.5, 0, 0)])[0, 0] == 128


def test_own_plane_walk(orc):
    """A sensor in an opaque ground (its own plane, met only through float error) under a 0.5 pane: the ray walks past
    the own plane and passes the pane; a second, lower pane counts too; an opaque lid blocks."""
    t = Tris()
    t.quad((-1000, -1000, 0), (1000, -1000, 0), (1000, 1000, 0), (-1000, 1000, 0), 0)   # sensors lie in it
    t.quad((-50, -50, 3), (50, -50, 3), (50, 50, 3), (-50, 50, 3), 1)
    t.quad((0, -50, 2), (50, -50, 2), (50, 50, 2), (0, 50, 2), 1)
    t.quad((-40, 20, 1), (-20, 20, 1), (-20, 40, 1), (-40, 40, 1), 0)
    rng = np.random.default_rng(3)
    n = 64
    P = np.zeros((n, 3), np.float32)
    P[:, 0] = rng.uniform(-30, 30, n)
    P[:, 1] = rng.uniform(-30, 30, n)
    N = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    D = np.array([0, 0, 1], np.float32)
    by = trace_bytes(orc, t, P.reshape(-1), N.reshape(-1), D, [(0.2, 0, 0, 0), (0.2, 0.5, 0, 0)])[0]
    for i in range(n):
        x, y = P[i, 0], P[i, 1]
        lid = -40 <= x <= -20 and 20 <= y <= 40
        want = 0 if lid else (64 if x >= 0 else 128)
        assert by[i] == want, (i, x, y, by[i])


def brute_bytes(orc, tr, P, N, D, slots, cull):
    """The reference's byte semantics over every hit, from the oracle's sorted hit set (no BVH): the same byte for a
    scene without own planes is the product of the crossed factors, 0 when an opaque one or the product under 1/510."""
    V = tr.v.reshape(-1, 3).astype(np.float64)
    I = tr.i.reshape(-1, 3)
    nc, nDir = len(P) // 3, len(D) // 3
    out = np.zeros((nDir, nc), np.uint8)
    L = orc.lib()
    for i in range(nc):
        p, nrm = P[3 * i:3 * i + 3], N[3 * i:3 * i + 3]
        o = np.zeros(3, np.float32)
        L.orc_offset_ray(np.ascontiguousarray(p), np.ascontiguousarray(nrm), o)
        for d in range(nDir):
            dv = D[3 * d:3 * d + 3]
            if float(np.float32(nrm[0] * dv[0]) + np.float32(nrm[1] * dv[1]) + np.float32(nrm[2] * dv[2])) <= 0:
                continue
            prod, blocked = 1.0, False
            for k, (a, b, c) in enumerate(I):
                t, front = moller(o.astype(np.float64), dv.astype(np.float64), V[a], V[b], V[c])
                if t is None or t < 1e-4 or (cull and not front):
                    continue
                rho, tau, tn, flags = slots[tr.s[k]]
                if tau <= 0:
                    blocked = True
                    break
                prod *= tau
            if not blocked and prod >= 1 / 510:
                out[d, i] = int(np.rint(np.float32(prod) * np.float32(255)))
    return out


def moller(o, d, a, b, c):
    """Moller-Trumbore in double: (t, front) or (None, None)"""
    e1, e2 = b - a, c - a
    pv = np.cross(d, e2)
    det = float(np.dot(e1, pv))
    if abs(det) < 1e-12:
        return None, None
    tv = o - a
    u = float(np.dot(tv, pv)) / det
    qv = np.cross(tv, e1)
    v = float(np.dot(d, qv)) / det
    if u < 0 or v < 0 or u + v > 1:
        return None, None
    return float(np.dot(e2, qv)) / det, det > 0


def test_random_porous_vs_brute(orc):
    """Random porous and opaque quads floating over the sensors (no own planes, no grazing edges in play): the oracle's
    bytes equal the product of the crossed factors found by a double-precision test of every triangle."""
    rng = np.random.default_rng(11)
    t = Tris()
    taus = [0.0, 0.5, 0.25, 0.8, 0.9, 0.3]
    for k in range(40):
        x, y, z = rng.uniform(-20, 20), rng.uniform(-20, 20), rng.uniform(1, 10)
        w, h = rng.uniform(2, 12), rng.uniform(2, 12)
        t.quad((x, y, z), (x + w, y, z), (x + w, y + h, z), (x, y + h, z), int(rng.integers(0, len(taus))))
    slots = [(0.2, tau, 0, 0) for tau in taus]
    n = 24
    P = np.zeros((n, 3), np.float32)
    P[:, :2] = rng.uniform(-15, 15, (n, 2))
    N = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    dirs = rng.normal(size=(40, 3))
    dirs[:, 2] = np.abs(dirs[:, 2]) + 0.3
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    D = dirs.astype(np.float32).reshape(-1)
    got = trace_bytes(orc, t, P.reshape(-1), N.reshape(-1), D, slots)
    want = brute_bytes(orc, t, P.reshape(-1), N.reshape(-1), D, slots, False)
    # products of up to a few float factors: the oracle multiplies in double and rounds once; the brute force in
    # double too, so the bytes agree except where an edge is met within float error (none here by construction)
    assert (got != want).sum() <= 2, np.argwhere(got != want)[:5]
    assert (got > 0).sum() > 50 and ((got > 0) & (got < 255)).sum() > 20


def test_reduce_bytes_loop(orc):
    """k_reduce<Bytes> (gpu.cu:83-114) as the reference's own check (check_gpu.cpp:268-274): double(float(cos x byte /
    255)) x W over the non-zero bytes; sun-hours form without the cosine."""
    t, P, N = porous_scene()
    D = dirs_check_gpu()
    nDir = len(D) // 3
    by = trace_bytes(orc, t, P, N, D, [(0.2, 0, 0, 0), (0.2, 0.5, 0, 0), (0.2, 0.5, 0, 0)])
    W = (np.float32(0.04) * (np.float32(1) + np.float32(0.01) * np.arange(nDir, dtype=np.float32))).astype(np.float32)
    Dv, Nv = D.reshape(-1, 3), N.reshape(-1, 3)
    for useCos in (1, 0):
        out = np.zeros(4)
        orc.lib().orc_reduce_bytes(np.ascontiguousarray(by.reshape(-1)), 4, nDir, N, D, W, useCos, out)
        for i in range(4):
            acc = 0.0
            for d in range(nDir):
                if by[d, i]:
                    c = np.float32(max(np.float32(np.float32(Nv[i, 0] * Dv[d, 0]) + np.float32(Nv[i, 1] * Dv[d, 1])) +
                                       np.float32(Nv[i, 2] * Dv[d, 2]), np.float32(0))) if useCos else np.float32(1)
                    acc += float(np.float32(c * np.float32(np.float32(by[d, i]) / np.float32(255)))) * float(W[d])
            assert out[i] == acc, (useCos, i)


# ---------------------------------------------------------------------------------------------------- host rules


def box_mesh(native, name, x0, y0, z0, x1, y1, z1, soup, slots=None):
    """check_mesh.cpp:16-27: an axis box, faces wound outward (corner k: bit 0 x, bit 1 y, bit 2 z), indexed or a soup"""
    X, Y, Z = (x0, x1), (y0, y1), (z0, z1)
    q = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (1, 3, 7, 5), (3, 2, 6, 7), (2, 0, 4, 6)]
    corner = lambda k: [X[k & 1], Y[(k >> 1) & 1], Z[k >> 2]]
    pts, idx = [], []
    if not soup:
        for k in range(8):
            pts += corner(k)
    for f in q:
        for j in f:
            if soup:
                pts += corner(j)
                idx.append(len(pts) // 3 - 1)
            else:
                idx.append(j)
    return dict(points=pts, counts=[4] * 6, indices=idx, slots=slots)


def merged(a, b):
    base = len(a["points"]) // 3
    return dict(points=a["points"] + b["points"], counts=a["counts"] + b["counts"], indices=a["indices"] + [base + i for i in b["indices"]],
                slots=None)


def flipped(m):
    idx = list(m["indices"])
    for f in range(len(m["counts"])):
        idx[4 * f + 1], idx[4 * f + 3] = idx[4 * f + 3], idx[4 * f + 1]
    return dict(m, indices=idx)


def thin(native, m, slot=0, centre=None):
    import daylight_amd as da
    L = da.lib()
    h = native.SimulationMesh("m", np.array(m["points"], np.float64), np.array(m["counts"], np.int32), np.array(m["indices"], np.int32),
                              slots=None if m.get("slots") is None else np.array(m["slots"], np.uint16))
    try:
        out = C.c_uint32()
        c = None if centre is None else (C.c_double * 3)(*centre)
        assert L.dlx_thin_solid(h._h, slot, c, C.byref(out)) == 0
        return out.value
    finally:
        h.close()


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


def test_thin_solids_of_the_reference(native):
    """check_mesh.cpp:279-310, case by case."""
    pane = box_mesh(native, "pane", 0, 0, 1, 20, 20, 1.01, False)
    soup = box_mesh(native, "pane", 0, 0, 1, 20, 20, 1.01, True)
    far = (1e5, -3e5, 50)
    assert thin(native, pane) == THIN and thin(native, soup) == THIN
    assert thin(native, soup, centre=far) == THIN
    assert thin(native, flipped(soup)) == THIN | INWARD
    op = dict(pane, counts=pane["counts"][:-1], indices=pane["indices"][:20])
    assert thin(native, op) == 0 and thin(native, pane, slot=1) == 0   # open; a slot without faces
    bent = dict(pane, indices=[pane["indices"][0], pane["indices"][3], pane["indices"][2], pane["indices"][1]] + pane["indices"][4:])
    assert thin(native, bent, centre=(10, 10, 1.005)) == 0   # the bottom wound against the others
    assert thin(native, box_mesh(native, "cube", -1.5, -1.5, 1, 1.5, 1.5, 4, False)) == 0   # closed, 2V/S = 1 m
    assert thin(native, merged(box_mesh(native, "a", 0, 0, 0, 1, 1, 0.01, True), box_mesh(native, "b", 1, 0, 0, 2, 1, 0.01, True))) == THIN
    win = box_mesh(native, "window", 0, 0, 0, 1, 1, 0.01, False)
    win["points"] = win["points"] + [-0.1, -0.1, 0, 1.1, -0.1, 0, 1.1, 1.1, 0, -0.1, 1.1, 0]
    win["counts"] = win["counts"] + [4, 4, 4, 4]
    win["indices"] = win["indices"] + [8, 9, 1, 0, 9, 10, 3, 1, 10, 11, 2, 3, 11, 8, 0, 2]
    win["slots"] = [1, 1, 1, 1, 1, 1, 0, 0, 0, 0]
    assert thin(native, win, slot=1) == THIN and thin(native, win, slot=0) == 0
    twin = dict(points=[0, 0, 1, 2, 0, 1, 2, 2, 1, 0, 2, 1], counts=[4, 4], indices=[0, 1, 2, 3, 3, 2, 1, 0], slots=None)
    twin_first = dict(twin, indices=[3, 2, 1, 0, 0, 1, 2, 3])
    twin_soup = dict(twin, points=twin["points"] + [0, 2, 1, 2, 2, 1, 2, 0, 1, 0, 0, 1], indices=[0, 1, 2, 3, 4, 5, 6, 7])
    for m in (twin, twin_first, twin_soup):
        assert thin(native, m) == THIN | TWIN and thin(native, m, centre=far) == THIN | TWIN
    assert thin(native, merged(pane, flipped(soup))) == THIN | TWIN


def test_triangle_slots(native):
    """check_mesh.cpp:266-276: a triangle and a quad of face slots 2, 0 from mesh base 5 -> 7, 5, 5; no slots -> 0."""
    import daylight_amd as da
    L = da.lib()
    for slots, base, want in (([2, 0], 5, [7, 5, 5]), (None, 0, [0, 0, 0])):
        h = native.SimulationMesh("m", np.array([0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0], np.float64), np.array([3, 4], np.int32),
                                  np.array([0, 1, 2, 0, 1, 2, 3], np.int32), slots=None if slots is None else np.array(slots, np.uint16))
        try:
            out = np.zeros(8, np.uint32)
            n = C.c_int64()
            assert L.dlx_triangle_slots(h._h, base, out.ctypes.data, 8, C.byref(n)) == 0
            assert list(out[: n.value]) == want
        finally:
            h.close()
