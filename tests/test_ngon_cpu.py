"""Faces of 5 or more corners (reference mesh.cpp:137-168, earcut there; this library's own ear clipper here), through the
C ABI's dlx_triangulate (the host triangulation the scene ingest places on the GPU), against the reference's own checks
(tests/check_mesh.cpp:103-133): a concave L hexagon after a triangle draws 4 triangles of total area 3, every one wound
as the polygon (+Z); a collinear pentagon is skipped and counted; and what any triangulation of a simple polygon must
satisfy (n - 2 triangles, total area = the shoelace area, corners from the face)."""
import ctypes as C
import math

import numpy as np
import pytest


def triangulate(points, counts, indices, centre=None):
    import daylight_amd
    import native
    L = daylight_amd.lib()
    with native.SimulationMesh("m", points, counts, indices) as m:
        n, sk = C.c_int64(), C.c_int32()
        cap = int(sum(max(c - 2, 0) for c in counts))
        v = np.zeros(3 * len(np.asarray(points).reshape(-1, 3)), np.float32)
        t = np.zeros(3 * max(cap, 1), np.uint32)
        c = None if centre is None else np.ascontiguousarray(centre, np.float64)
        assert L.dlx_triangulate(m._h, None if c is None else c.ctypes.data, v.ctypes.data, t.ctypes.data, cap, C.byref(n), C.byref(sk)) == 0
        return v.reshape(-1, 3), t[:3 * n.value].reshape(-1, 3), sk.value


def area_z(V, t):
    a, b, c = V[t[0]].astype(np.float64), V[t[1]].astype(np.float64), V[t[2]].astype(np.float64)
    return 0.5 * ((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))


def test_concave_L_after_a_triangle():
    P = [[0, 0, 0], [2, 0, 0], [2, 1, 0], [1, 1, 0], [1, 2, 0], [0, 2, 0]]
    V, T, sk = triangulate(P, [3, 6], [0, 1, 2, 0, 1, 2, 3, 4, 5])
    assert len(T) == 5 and sk == 0
    assert T[0].tolist() == [0, 1, 2]
    areas = [area_z(V, t) for t in T[1:]]
    assert all(a > 0 for a in areas) and abs(sum(areas) - 3) < 1e-12


def test_collinear_pentagon_skipped():
    V, T, sk = triangulate([[k, k, k] for k in range(5)], [5], [0, 1, 2, 3, 4])
    assert len(T) == 0 and sk == 1


@pytest.mark.parametrize("n,seed", [(5, 1), (8, 2), (13, 3), (40, 4)])
def test_star_polygons_cover_their_area(n, seed):
    """Random star-shaped (simple, often concave) polygons in a tilted plane: n - 2 triangles, their areas summing to the
    polygon's, each wound as the polygon, corners the face's own."""
    rng = np.random.default_rng(seed)
    ang = (np.arange(n) + rng.uniform(0, 0.8, n)) * (2 * math.pi / n)   # gaps under 180 degrees: star-shaped about 0, so simple
    r = rng.uniform(0.3, 1.0, n)
    q = np.stack([r * np.cos(ang), r * np.sin(ang)], 1)
    shoelace = 0.5 * np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])
    u, v = np.array([1.0, 0.0, 0.5]) / math.sqrt(1.25), np.array([0.0, 1.0, 0.0])
    P = q[:, :1] * u + q[:, 1:] * v + np.array([10.0, -4.0, 2.0])
    V, T, sk = triangulate(P, [n], list(range(n)), centre=[10.0, -4.0, 2.0])
    assert sk == 0 and len(T) == n - 2
    N = np.cross(u, v)
    tot = 0.0
    for t in T:
        a, b, c = (V[k].astype(np.float64) for k in t)
        s = 0.5 * np.dot(np.cross(b - a, c - a), N)
        assert s > -1e-9
        tot += s
    assert abs(tot - shoelace) < 1e-5 * abs(shoelace)
