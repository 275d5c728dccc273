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


def scene_arrays(orc, sc, nt):
    import ctypes as C
    node, tri, prim = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nn = orc.lib().orc_scene_nodes(sc, C.byref(node), C.byref(tri), C.byref(prim))
    nodes = np.ctypeslib.as_array(C.cast(node, C.POINTER(C.c_float)), shape=(nn * 16,)).reshape(nn, 16).copy()
    tris = np.ctypeslib.as_array(C.cast(tri, C.POINTER(C.c_float)), shape=(nt * 12,)).reshape(nt, 12).copy()
    prims = np.ctypeslib.as_array(C.cast(prim, C.POINTER(C.c_uint32)), shape=(nt,)).copy()
    return nodes, tris, prims


def sah_scene(orc, t):
    nt = len(t.i) // 3
    sc = orc.lib().orc_scene_new(t.v, len(t.v) // 3, t.i, nt)
    h = orc.lib().orc_scene_sah(sc)
    return sc, nt, h


def test_sah_traversal_equals_brute_force(orc):
    """Any tree gives the same ray results: the SAH tree's traversal against every triangle, both cull modes."""
    rng = np.random.default_rng(3)
    t = random_scene(rng)
    sc, nt, _ = sah_scene(orc, t)
    mism = 0
    for _ in range(4000):
        p = np.array([rng.uniform(-60, 60), rng.uniform(-60, 60), rng.choice([0.0, rng.uniform(0, 10)])], np.float32)
        n = np.array([0, 0, 1], np.float32) if rng.uniform() < 0.5 else rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        for cull in (0, 1):
            mism += orc.lib().orc_visible(sc, p, n, d, cull) != orc.lib().orc_visible_brute(t.v, t.i, nt, p, n, d, cull)
    orc.lib().orc_scene_free(sc)
    assert mism == 0


def subtree_leaves(nodes, root=0):
    """{slot: sorted leaf positions} for every inner node, by a walk from the root."""
    out, order = {}, []
    st = [root]
    while st:
        x = st.pop()
        order.append(x)
        for c in nodes[x, 12:14].copy().view(np.uint32):
            if c < 0x80000000:
                st.append(int(c))
    for x in reversed(order):
        s = []
        for c in nodes[x, 12:14].copy().view(np.uint32):
            s += [int(c & 0x7fffffff)] if c >= 0x80000000 else out[int(c)]
        out[x] = sorted(s)
    return out


def test_sah_structure(orc):
    """Every leaf once; a node's two child boxes are exactly its children's boxes (a leaf's: its triangle's min / max);
    every subtree holds a contiguous run of leaf positions, left before right; the node of a run split at m sits at slot
    m - 1, the root and slot 0 swapped; the height the build reports is the tree's."""
    rng = np.random.default_rng(5)
    t = random_scene(rng, nbox=200)
    sc, nt, h = sah_scene(orc, t)
    nodes, tris, prims = scene_arrays(orc, sc, nt)
    assert h == orc.lib().orc_scene_height(sc)
    orc.lib().orc_scene_free(sc)
    nn = nt - 1
    assert len(nodes) == nn and sorted(prims.tolist()) == list(range(nt))
    kids = nodes[:, 12:14].copy().view(np.uint32).reshape(-1)
    assert sorted((kids[kids >= 0x80000000] & 0x7fffffff).tolist()) == list(range(nt))
    assert sorted(kids[kids < 0x80000000].tolist()) == list(range(1, nn))
    runs = subtree_leaves(nodes)
    assert len(runs) == nn and runs[0] == list(range(nt))
    splits = {}
    for x, s in runs.items():
        assert s == list(range(s[0], s[-1] + 1))   # contiguous
        left = nodes[x, 12:13].copy().view(np.uint32)[0]
        nl = 1 if left >= 0x80000000 else len(runs[int(left)])
        splits[x] = s[0] + nl   # the right run starts here
    root_id = splits[0] - 1
    for x, m in splits.items():
        want = 0 if m - 1 == root_id else (root_id if m - 1 == 0 else m - 1)
        assert x == want
    tv = tris[:, :9].reshape(nt, 3, 3)
    for x in range(nn):
        for side in (0, 1):
            c = nodes[x, 12 + side:13 + side].copy().view(np.uint32)[0]
            box = nodes[x, 6 * side:6 * side + 6]
            if c >= 0x80000000:
                k = int(c & 0x7fffffff)
                assert np.array_equal(box[:3], tv[k].min(0)) and np.array_equal(box[3:], tv[k].max(0))
            else:
                cb = nodes[int(c)]
                assert np.array_equal(box[:3], np.minimum(cb[0:3], cb[6:9])) and np.array_equal(box[3:], np.maximum(cb[3:6], cb[9:12]))
    # the leaves carry their triangles: leaf k = input triangle prims[k]
    V = t.v.reshape(-1, 3)
    I = t.i.reshape(-1, 3)
    assert np.array_equal(tv, V[I[prims]])


def test_sah_known_cases(orc):
    """Two clusters 1 km apart: the root separates them. Identical triangles: no plane splits them, so every range halves
    (height ceil(log2 n)). Random boxes over a ground grid: less summed node area (the SAH's measure) than the LBVH."""
    t = Tris()
    rng = np.random.default_rng(1)
    for off in (0.0, 1000.0):
        for _ in range(40):
            x, y = rng.uniform(0, 5, 2)
            t.box(off + x, y, 0, off + x + 1, y + 1, 1)
    sc, nt, _ = sah_scene(orc, t)
    nodes, tris, _ = scene_arrays(orc, sc, nt)
    orc.lib().orc_scene_free(sc)
    runs = subtree_leaves(nodes)
    left = nodes[0, 12:13].copy().view(np.uint32)[0]
    lk = runs[int(left)]
    xs = tris[:, [0, 3, 6]].mean(1)
    assert len(lk) == nt // 2 and len(set((xs[lk] > 500).tolist())) == 1
    t = Tris()
    for _ in range(150):
        t.quad((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
    sc, nt, h = sah_scene(orc, t)
    orc.lib().orc_scene_free(sc)
    assert nt == 300 and h == 9
    t = random_scene(np.random.default_rng(7), nbox=300)

    def area(nodes):
        b = np.concatenate([nodes[:, 0:6], nodes[:, 6:12]]).astype(np.float64)
        d = b[:, 3:] - b[:, :3]
        return float((d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0]).sum())

    nt = len(t.i) // 3
    sc = orc.lib().orc_scene_new(t.v, len(t.v) // 3, t.i, nt)
    a_lbvh = area(scene_arrays(orc, sc, nt)[0])
    orc.lib().orc_scene_sah(sc)
    a_sah = area(scene_arrays(orc, sc, nt)[0])
    orc.lib().orc_scene_free(sc)
    assert a_sah < 0.8 * a_lbvh
