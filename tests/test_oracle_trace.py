"""The oracle's ray semantics (reference src/trace_dev.cu __raygen__bits) pinned by tests/check_gpu.cpp's scenes: a CPU slab
test for every bit of a ground + box scene, the km-sized own planes, the 1 cm curb and the L foot; and its LBVH
traversal against brute force over every triangle."""
import math

import numpy as np
import pytest

from scenes import Tris, box_scene_check_gpu, dirs_check_gpu, quad_dirs, slab_hit


def bits_of(orc, tr, P, N, D, cull=False, threads=2):
    sc = orc.lib().orc_scene_new(tr.v, len(tr.v) // 3, tr.i, len(tr.i) // 3)
    nc, nDir = len(P) // 3, len(D) // 3
    nW = (nDir + 31) // 32
    bits = np.zeros(nW * nc, np.uint32)
    orc.lib().orc_trace_bits(sc, P, N, nc, D, nDir, int(cull), bits, threads)
    orc.lib().orc_scene_free(sc)
    b = np.zeros((nc, nDir), np.uint8)
    for d in range(nDir):
        b[:, d] = (bits[(d // 32) * nc:(d // 32 + 1) * nc] >> (d % 32)) & 1
    return b, bits


def test_box_scene_every_bit_vs_slab(orc):
    tr = box_scene_check_gpu()
    assert len(tr.i) // 3 == 14
    P = np.array([1, 1, 0, -1, -1, 0, -1, -1, 0, 1, 1, 2, 1, 1, 1], np.float32)
    N = np.array([0, 0, 1, 0, 0, 1, 0, 0, -1, 0, 0, -1, 0, 0, 1], np.float32)
    D = dirs_check_gpu()
    Z, H, NA = 145, 146, 147
    b, bits = bits_of(orc, tr, P, N, D)
    assert b[0, Z] == 0 and b[0, H] == 1 and b[1, Z] == 1 and b[1, NA] == 0
    assert b[2, NA] == 1 and b[2, Z] == 0 and b[3, NA] == 0 and b[4, Z] == 0
    Dv = D.reshape(-1, 3)
    for i in (0, 1):   # check_gpu.cpp:155-158
        for d in range(len(Dv)):
            up = float(np.dot(N[3 * i:3 * i + 3], Dv[d])) > 0
            want = up and not slab_hit(P[3 * i:3 * i + 3], Dv[d], 0.5, 1.5)
            assert b[i, d] == want, (i, d)
    out = np.zeros(5)
    orc.lib().orc_reduce(bits, 5, len(Dv), N, D, np.ones(len(Dv), np.float32), 0, out)
    assert out[2] == 1 and out[3] == 0 and out[4] == 0
    assert out[1] > out[0] and out[1] < 147
    cul, _ = bits_of(orc, tr, P, N, D, cull=True)   # s4 inside the box sees the zenith through culled back faces
    assert cul[4, Z] == 1 and cul[0, Z] == 0 and cul[3, NA] == 0


def test_reduce_against_double_loop(orc):
    tr = box_scene_check_gpu()
    P = np.array([1, 1, 0, -1, -1, 0, -1, -1, 0, 1, 1, 2, 1, 1, 1], np.float32)
    N = np.array([0, 0, 1, 0, 0, 1, 0, 0, -1, 0, 0, -1, 0, 0, 1], np.float32)
    D = dirs_check_gpu()
    nDir = len(D) // 3
    b, bits = bits_of(orc, tr, P, N, D)
    W = (np.float32(0.04) * (np.float32(1) + np.float32(0.01) * np.arange(nDir, dtype=np.float32))).astype(np.float32)
    out = np.zeros(5)
    orc.lib().orc_reduce(bits, 5, nDir, N, D, W, 1, out)
    Dv = D.reshape(-1, 3)
    for i in range(5):
        n = N[3 * i:3 * i + 3]
        acc = sum(float(max(np.float32(n[0] * Dv[d, 0] + n[1] * Dv[d, 1]) + np.float32(n[2] * Dv[d, 2]), np.float32(0))) * float(W[d])
                  for d in range(nDir) if b[i, d])
        assert abs(out[i] - acc) <= 1e-9


@pytest.mark.parametrize("L", [1000.0, 2000.0, 10000.0])
def test_km_own_planes(orc, L):
    """check_gpu.cpp:340-390: sensors in the plane of a km-sized quad or wall see their whole hemisphere."""
    Dq = quad_dirs()
    h = L / 2
    tq, tw = Tris(), Tris()
    tq.quad((-h, -h, 0), (h, -h, 0), (h, h, 0), (-h, h, 0))
    tw.quad((-h, 0, -h), (h, 0, -h), (h, 0, h), (-h, 0, h))
    so_p, so_n = [], []
    for p in ((0, 0, 0), (0.37, -0.21, 0), (0.3 * h, -0.4 * h, 0)):
        for nz in (1, -1):
            so_p += p; so_n += (0, 0, nz)
    so_p += (0.1, 0.1, 0.01); so_n += (0, 0, -1)   # looking down 1 cm over the quad: all blocked
    sw_p, sw_n = [], []
    for p in ((0, 0, 0), (0.3 * h, 0, -0.2 * h)):
        for ny in (1, -1):
            sw_p += p; sw_n += (0, ny, 0)
    for wall in (0, 1):
        if wall and L == 1000:
            continue
        P = np.array(sw_p if wall else so_p, np.float32)
        N = np.array(sw_n if wall else so_n, np.float32)
        b, _ = bits_of(orc, tw if wall else tq, P, N, Dq)
        Dv = Dq.reshape(-1, 3)
        want = (N.reshape(-1, 3) @ Dv.T > 0).astype(np.uint8)
        if not wall:
            want[6, :] = 0
        assert np.array_equal(b, want)
        assert int(want.sum()) == (4 if wall else 6) * len(Dv) // 2


def test_curb_and_l_foot(orc):
    Dq = quad_dirs()
    Dv = Dq.reshape(-1, 3)
    nq = len(Dv)
    curb = Tris()
    curb.quad((0.2, -1, 0), (0.2, 1, 0), (0.2, 1, 0.01), (0.2, -1, 0.01))
    curb.quad((0.2, -1, 0.01), (0.2, 1, 0.01), (0.3, 1, 0.01), (0.3, -1, 0.01))
    both = curb.copy()
    both.quad((-5000, -5000, 0), (5000, -5000, 0), (5000, 5000, 0), (-5000, 5000, 0))
    P, N = np.array([0, 0, 0], np.float32), np.array([0, 0, 1], np.float32)
    a, _ = bits_of(orc, curb, P, N, Dq)
    b, _ = bits_of(orc, both, P, N, Dq)
    assert np.array_equal(a, b) and int((a[0, :nq // 2] == 0).sum()) >= 64   # check_gpu.cpp:393-406
    s = 1.0 / math.sqrt(2)
    P = np.array([0.37, 0, -10], np.float32)
    N = np.array([0, -s, s], np.float32)
    for tiles in (False, True):   # check_gpu.cpp:413-440
        t = Tris()
        st = 100 if tiles else 2000
        for x in np.arange(-1000, 1000, st):
            for y in np.arange(-1000, 1000, st):
                t.quad((x, y, -10), (x + st, y, -10), (x + st, y + st, -10), (x, y + st, -10))
        t.quad((-10, 0, -10), (10, 0, -10), (10, 0, 10), (-10, 0, 10))
        b, _ = bits_of(orc, t, P, N, Dq)
        want = ((Dv @ N > 0) & (Dv[:, 2] > 0) & (Dv[:, 1] < 0)).astype(np.uint8)
        assert np.array_equal(b[0], want) and int(want.sum()) == nq // 4


def test_empty_scene(orc):
    D = dirs_check_gpu()
    P = np.array([1, 1, 0, -1, -1, 0, -1, -1, 0], np.float32)
    N = np.array([0, 0, 1, 0, 0, 1, 0, 0, -1], np.float32)
    b, _ = bits_of(orc, Tris(), P, N, D)
    assert b[0].sum() == 147 and b[2].sum() == 1


def random_scene(rng, nbox=60, extent=50.0):
    t = Tris()
    for _ in range(nbox):
        x0, y0 = rng.uniform(-extent, extent, 2)
        w, d, h = rng.uniform(1, 8, 3)
        t.box(x0, y0, 0.0, x0 + w, y0 + d, h)
    t.quad((-2 * extent, -2 * extent, 0), (2 * extent, -2 * extent, 0), (2 * extent, 2 * extent, 0), (-2 * extent, 2 * extent, 0))
    return t


def test_bvh_traversal_equals_brute_force(orc):
    rng = np.random.default_rng(3)
    t = random_scene(rng)
    sc = orc.lib().orc_scene_new(t.v, len(t.v) // 3, t.i, len(t.i) // 3)
    nt = len(t.i) // 3
    mism = 0
    for _ in range(4000):
        p = np.array([rng.uniform(-60, 60), rng.uniform(-60, 60), rng.choice([0.0, rng.uniform(0, 10)])], np.float32)
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        if rng.uniform() < 0.5:
            n = np.array([0, 0, 1], np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        for cull in (0, 1):
            a = orc.lib().orc_visible(sc, p, n, d, cull)
            b = orc.lib().orc_visible_brute(t.v, t.i, nt, p, n, d, cull)
            mism += a != b
    orc.lib().orc_scene_free(sc)
    assert mism == 0


def test_lbvh_structure(orc):
    """Karras 2012 invariants: every leaf once, parents' boxes contain their children's, 2^31-flagged leaves."""
    import ctypes as C
    rng = np.random.default_rng(5)
    t = random_scene(rng, nbox=200)
    nt = len(t.i) // 3
    sc = orc.lib().orc_scene_new(t.v, len(t.v) // 3, t.i, nt)
    node, tri, prim = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nn = orc.lib().orc_scene_nodes(sc, C.byref(node), C.byref(tri), C.byref(prim))
    nodes = np.ctypeslib.as_array(C.cast(node, C.POINTER(C.c_float)), shape=(nn * 16,)).reshape(nn, 16).copy()
    prims = np.ctypeslib.as_array(C.cast(prim, C.POINTER(C.c_uint32)), shape=(nt,)).copy()
    orc.lib().orc_scene_free(sc)
    assert nn == nt - 1 and sorted(prims.tolist()) == list(range(nt))
    kids = nodes[:, 12:14].copy().view(np.uint32).reshape(-1)
    leaves = kids[kids >= 0x80000000] & 0x7fffffff
    inner = kids[kids < 0x80000000]
    assert sorted(leaves.tolist()) == list(range(nt)) and sorted(inner.tolist()) == list(range(1, nn))
