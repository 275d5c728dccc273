"""Mesh ingest on the GPU (dlg_mesh_*): the scene's points and faces triangulated where the scene is built, against the
oracle's restatement of the host triangulation (triangulate_into, reference mesh.cpp:137-168: points recentred in
double then rounded to float; a triangle (a, b, c), a quad (a, b, c) (a, c, d), faces in order, meshes in order at
their vertex bases), and the tree built from the staged arrays against the tree built from the same arrays uploaded.

Bars: vertices, triangles and per-triangle slots bit for bit; the built tree's nodes and leaves bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import daylight_amd
    g = daylight_amd.Gpu()
    yield g
    g.close()


def quad_grid(rng, nx, ny, origin, cell, jitter):
    """An nx x ny grid of quads (jittered heights) far from the origin, plus its diagonal triangles every fifth cell."""
    xs, ys = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    P = np.stack([xs.ravel() * cell, ys.ravel() * cell, rng.uniform(0, jitter, xs.size)], 1) + origin
    counts, idx = [], []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            q = [a, a + 1, a + nx + 2, a + nx + 1]
            if (i + j) % 5 == 0:
                counts += [3, 3]
                idx += [q[0], q[1], q[2], q[0], q[2], q[3]]
            else:
                counts.append(4)
                idx += q
    return P, np.array(counts, np.int32), np.array(idx, np.int32)


def random_tris(rng, n, origin):
    P = rng.uniform(-30, 30, (3 * n, 3)) + origin
    return P, np.full(n, 3, np.int32), rng.permutation(3 * n).astype(np.int32)


def oracle_triangulate(orc, meshes):
    """The oracle's triangulation of the meshes in order, with its bounding-box centre: V [nv, 3], I [nt, 3]."""
    L = orc.lib()
    arr = (orc.Mesh * len(meshes))(*[m.c() for m in meshes])
    c = np.zeros(3)
    L.orc_bbox_centre(C.cast(arr, C.c_void_p), len(meshes), c)
    Vs, Is, vb = [], [], 0
    skipped = C.c_int(0)
    for k, m in enumerate(meshes):
        Vk = np.zeros(len(m.points), np.float32)
        Ik = np.zeros(3 * max(1, int(np.maximum(m.counts - 2, 0).sum())), np.uint32)
        n = L.orc_triangulate(C.byref(arr[k]), c, Vk, Ik, vb, C.byref(skipped))
        Vs.append(Vk.reshape(-1, 3))
        Is.append(Ik[:3 * n].reshape(-1, 3))
        vb += len(m.points) // 3
    return np.concatenate(Vs), np.concatenate(Is), c


def scene_of(orc, seed):
    rng = np.random.default_rng(seed)
    A = orc.MeshArrays("grid", *quad_grid(rng, 40, 30, np.array([3.1e5, 5.8e6, 41.0]), 0.75, 2.0))
    B = orc.MeshArrays("tris", *random_tris(rng, 500, np.array([3.1e5 + 10, 5.8e6 + 5, 60.0])))
    return rng, [A, B]


@pytest.mark.parametrize("seed", [1, 2])
def test_staged_vs_oracle(gpu, orc, seed):
    rng, meshes = scene_of(orc, seed)
    V0, I0, c = oracle_triangulate(orc, meshes)
    fsA = rng.integers(0, 5, len(meshes[0].counts)).astype(np.uint16)
    gpu.mesh_begin(len(V0), int(sum(int((m.counts - 2).sum()) for m in meshes)), True, c)
    made = [gpu.mesh_add(meshes[0].points, meshes[0].counts, meshes[0].indices, fsA, 0),
            gpu.mesh_add(meshes[1].points, meshes[1].counts, meshes[1].indices, None, 5)]
    V, I, S = gpu.mesh_staged(slotted=True)
    gpu.mesh_build()
    assert made == [int((m.counts - 2).sum()) for m in meshes]
    assert V.view(np.uint32).tolist() == V0.view(np.uint32).tolist()
    assert np.array_equal(I, I0)
    # a face's slot on each of its triangles: base + the face's slot (none: 0)
    want = np.concatenate([np.repeat(fsA.astype(np.uint32), meshes[0].counts - 2), np.full(int((meshes[1].counts - 2).sum()), 5, np.uint32)])
    assert np.array_equal(S, want)


def test_build_from_staged_equals_upload(gpu, orc):
    rng, meshes = scene_of(orc, 3)
    V0, I0, c = oracle_triangulate(orc, meshes)
    gpu.build(V0, I0)
    nodes0, leaves0, h0 = gpu.scene()[:3]
    gpu.mesh_begin(len(V0), len(I0), False, c)
    for m in meshes:
        gpu.mesh_add(m.points, m.counts, m.indices)
    gpu.mesh_build()
    nodes, leaves, h = gpu.scene()[:3]
    assert h == h0
    assert nodes.view(np.uint32).tolist() == nodes0.view(np.uint32).tolist()
    assert leaves.view(np.uint32).tolist() == leaves0.view(np.uint32).tolist()


def test_ear_clipped_faces_placed_in_face_order(gpu, orc):
    """Faces of 5+ corners come from the host's list and land at their face's place; a face the list says made 0
    triangles (skipped) leaves no gap."""
    rng = np.random.default_rng(9)
    counts = np.array([3, 5, 4, 6, 5, 3, 7], np.int32)
    P = rng.uniform(-5, 5, (60, 3))
    idx = rng.integers(0, 60, int(counts.sum())).astype(np.int32)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    ng_face, ng_made, ng_tri = [1, 3, 4, 6], [3, 4, 0, 2], []
    for f, k in zip(ng_face, ng_made):
        corners = idx[starts[f]:starts[f] + counts[f]]
        for j in range(k):   # any corners will do: the device places what the list says
            ng_tri += [corners[0], corners[j + 1], corners[j + 2]]
    want, t = [], iter(np.array(ng_tri, np.uint32).reshape(-1, 3))
    for f, n in enumerate(counts):
        q = idx[starts[f]:starts[f] + n]
        if n == 3:
            want.append(q[[0, 1, 2]])
        elif n == 4:
            want += [q[[0, 1, 2]], q[[0, 2, 3]]]
        else:
            want += [next(t) for _ in range(ng_made[ng_face.index(f)])]
    want = np.array(want, np.uint32) + 7   # behind a first mesh of 7 points
    first = orc.MeshArrays("first", rng.uniform(-1, 1, (7, 3)), [3], [0, 1, 2])
    gpu.mesh_begin(67, 1 + int((counts - 2).sum()), False, np.zeros(3))
    assert gpu.mesh_add(first.points, first.counts, first.indices) == 1
    made = gpu.mesh_add(P, counts, idx, ngon=(np.array(ng_face), np.array(ng_made), np.array(ng_tri)))
    V, I, _ = gpu.mesh_staged()
    gpu.mesh_build()
    assert made == len(want)
    assert np.array_equal(I[1:], want)
    assert np.array_equal(V[7:], P.astype(np.float32))


def test_bad_faces_fail_without_reading_out_of_range(gpu):
    import daylight_amd
    P = np.zeros((4, 3))
    gpu.mesh_begin(8, 8, False, np.zeros(3))
    with pytest.raises(daylight_amd.GpuError, match="vertex index out of range"):
        gpu.mesh_add(P, [3, 3], [0, 1, 2, 0, 1, 4])
    gpu.mesh_begin(8, 8, False, np.zeros(3))
    with pytest.raises(daylight_amd.GpuError, match="corners past the index array"):
        gpu.mesh_add(P, [3, 4], [0, 1, 2, 0, 1, 3])
    gpu.mesh_begin(8, 8, False, np.zeros(3))
    with pytest.raises(daylight_amd.GpuError, match="fewer than 3 corners"):
        gpu.mesh_add(P, [3, 2], [0, 1, 2, 0, 1])
    # a clean mesh after the failures
    gpu.mesh_begin(4, 2, False, np.zeros(3))
    assert gpu.mesh_add(P + 1, [4], [0, 1, 2, 3]) == 2
    gpu.mesh_build()


def test_ngon_context_run_equals_its_triangles(orc):
    """A run whose context has faces of 5 to 12 corners (ear-clipped on the host, placed by the GPU ingest) gives the
    values of the same run over those faces' triangles given as triangles (dlx_triangulate): the same scene, bit for
    bit."""
    import os
    import native
    from test_ngon_cpu import triangulate
    rng = np.random.default_rng(11)
    pts, counts, idx = [], [], []
    for k in range(60):   # star-shaped roofs over the box, 5 to 12 corners
        n = int(rng.integers(5, 13))
        ang = (np.arange(n) + rng.uniform(0, 0.8, n)) * (2 * np.pi / n)   # gaps under 180 degrees: star-shaped, simple
        r = rng.uniform(0.2, 0.8, n)
        c = rng.uniform(-4, 4, 2)
        z = rng.uniform(2.5, 6.0)
        base = len(pts)
        pts += [[c[0] + r[j] * np.cos(ang[j]), c[1] + r[j] * np.sin(ang[j]), z] for j in range(n)]
        counts.append(n)
        idx += list(range(base, base + n))
    V, T, sk = triangulate(pts, counts, idx)
    assert sk == 0 and len(T) == sum(c - 2 for c in counts)
    box = native.SimulationMesh.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "box.obj"))
    args = ["--clear-sky", "--lat", "52.52", "--lon", "13.41", "--tz", "1", "--quantity", "radiation,sunhours"]
    with native.SimulationMesh("roofs", pts, counts, idx) as ng, native.SimulationMesh("roofs", pts, [3] * len(T), T.ravel()) as tr:
        a = native.run("daylight", args, box, context=[ng])
        b = native.run("daylight", args, box, context=[tr])
    box.close()
    assert a.values.shape == b.values.shape
    assert np.array_equal(a.values, b.values)
    assert a.stats["rays"] == b.stats["rays"]
