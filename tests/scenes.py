"""Scenes of the reference's GPU tests (tests/check_gpu.cpp) and synthetic workloads, as float32 triangle soups and meshes."""
import math

import numpy as np


class Tris:
    """A triangle soup: v float32 [3 nv], i uint32 [3 nt]."""

    def __init__(self):
        self._v, self._i, self._s = [], [], []

    def quad(self, a, b, c, d, slot=0):   # CCW seen from outside: (0, 1, 2), (0, 2, 3); slot: both triangles' global slot
        k = len(self._v) // 3
        for p in (a, b, c, d):
            self._v += [float(x) for x in p]
        self._i += [k, k + 1, k + 2, k, k + 2, k + 3]
        self._s += [slot, slot]

    def box(self, x0, y0, z0, x1, y1, z1, slot=0):   # faces wound outward
        lo, hi = (x0, y0, z0), (x1, y1, z1)
        c = [(hi[0] if k & 1 else lo[0], hi[1] if k & 2 else lo[1], hi[2] if k & 4 else lo[2]) for k in range(8)]
        for q in ((0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (1, 3, 7, 5), (3, 2, 6, 7), (2, 0, 4, 6)):
            self.quad(*[c[j] for j in q], slot=slot)

    def copy(self):
        t = Tris()
        t._v, t._i, t._s = list(self._v), list(self._i), list(self._s)
        return t

    @property
    def s(self):
        """the global slot of every triangle"""
        return np.array(self._s, np.uint32)

    @property
    def v(self):
        return np.array(self._v if self._v else [0.0], np.float32)

    @property
    def i(self):
        return np.array(self._i if self._i else [0], np.uint32)[: len(self._i)] if self._i else np.zeros(0, np.uint32)


def box_scene_check_gpu():
    """check_gpu.cpp:114-125: ground [-2,2]^2 at z 0 and the box [0.5,1.5]^3, faces outward."""
    t = Tris()
    lo, hi = 0.5, 1.5
    t.quad((-2, -2, 0), (2, -2, 0), (2, 2, 0), (-2, 2, 0))
    t.quad((lo, lo, lo), (lo, hi, lo), (hi, hi, lo), (hi, lo, lo))
    t.quad((lo, lo, hi), (hi, lo, hi), (hi, hi, hi), (lo, hi, hi))
    t.quad((lo, lo, lo), (lo, lo, hi), (lo, hi, hi), (lo, hi, lo))
    t.quad((hi, lo, lo), (hi, hi, lo), (hi, hi, hi), (hi, lo, hi))
    t.quad((lo, lo, lo), (hi, lo, lo), (hi, lo, hi), (lo, lo, hi))
    t.quad((lo, hi, lo), (lo, hi, hi), (hi, hi, hi), (hi, hi, lo))
    return t


def f32v(x, y, z):
    return [np.float32(x), np.float32(y), np.float32(z)]


def dirs_check_gpu():
    """check_gpu.cpp:135-145: 145 patch-like directions, zenith, near-horizontal, nadir (148)."""
    D = []
    naz = [30, 30, 24, 24, 18, 12, 6]
    for r in range(7):
        alt = (6 + 12 * r) * math.pi / 180
        for j in range(naz[r]):
            az = 2 * math.pi * j / naz[r]
            D += f32v(math.cos(alt) * math.sin(az), math.cos(alt) * math.cos(az), math.sin(alt))
    D += f32v(0, 0, 1) + f32v(0, 0, 1)
    l = math.sqrt(0.999 ** 2 + 0.0447 ** 2)
    D += f32v(0.999 / l, 0, 0.0447 / l) + f32v(0, 0, -1)
    return np.array(D, np.float32)


def quad_dirs():
    """check_gpu.cpp:347-353: 12 altitudes x 64 azimuths up, mirrored down (1536)."""
    D = []
    for alt in (0.25, 0.5, 1., 2., 4., 8., 15., 30., 45., 60., 75., 89.):
        for j in range(64):
            a, az = alt * math.pi / 180, 2 * math.pi * (j + 0.5) / 64
            D.append(f32v(math.cos(a) * math.sin(az), math.cos(a) * math.cos(az), math.sin(a)))
    D += [[d[0], d[1], -d[2]] for d in D]
    return np.array(D, np.float32).reshape(-1)


def slab_hit(o, d, lo, hi):
    """check_gpu.cpp:82-90: ray vs the box [lo, hi]^3 in double, t > 0."""
    t0, t1 = 0.0, 1e30
    for a in range(3):
        oo, dd = float(o[a]), float(d[a])
        if dd == 0:
            if oo < lo or oo > hi:
                return False
            continue
        ta, tb = (lo - oo) / dd, (hi - oo) / dd
        if ta > tb:
            ta, tb = tb, ta
        t0, t1 = max(t0, ta), min(t1, tb)
    return t0 <= t1
