"""dlg_chunk_noalloc - the chunk without a series that allocates nothing - against dlg_chunk, on the GPU.

Bit for bit, through the C ABI: two calls in a row over the same directions, the second over fewer sensors (so that
the fixed device array holds the first call's larger visibility when the second reads it), in bit mode and in byte
mode (translucent slots); and on a comb of 2^20 triangles whose walks go past the stacks' LDS part into the array's
stack area (where deep trees get a smaller persistent grid). Sensors, weights and sums live in device memory that the test allocates through the HIP
runtime the library links (one runtime in the process). A call over more sensors than the array holds fails with the
limit in its message, before any kernel runs.
"""
import ctypes as C
import math

import numpy as np
import pytest

from scenes import Tris

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import daylight_amd
    g = daylight_amd.Gpu()
    yield g
    g.close()


def scene(rng):
    """boxes with opaque, porous and translucent slots over an opaque ground"""
    t = Tris()
    for _ in range(150):
        x0, y0 = rng.uniform(-90, 90, 2)
        w, d = rng.uniform(5, 30, 2)
        t.box(x0, y0, 0.0, x0 + w, y0 + d, rng.lognormal(math.log(15), 0.5), slot=int(rng.integers(0, 3)))
    t.quad((-140, -140, 0), (140, -140, 0), (140, 140, 0), (-140, 140, 0), 0)
    return t


def sensors(rng, n):
    P = np.zeros((n, 3), np.float32)
    P[:, :2] = rng.uniform(-120, 120, (n, 2))
    P[:, 2] = np.where(rng.uniform(size=n) < 0.5, 0.0, rng.uniform(0, 30, n))
    N = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    side = rng.uniform(size=n) < 0.3
    ang = rng.uniform(0, 2 * np.pi, n)
    N[side] = np.stack([np.cos(ang[side]), np.sin(ang[side]), np.zeros(side.sum())], 1).astype(np.float32)
    return P, N


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime already loaded with the library"""
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]), ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int])):
            getattr(self.L, name).argtypes = args
            getattr(self.L, name).restype = C.c_int
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.held.append(p.value)
        assert self.L.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0     # device to host
        return out

    def free(self):
        for p in self.held:
            self.L.hipFree(p)
        self.held = []


def test_noalloc_equals_chunk(gpu):
    rng = np.random.default_rng(5)
    t = scene(rng)
    gpu.build(t.v, t.i)
    gpu.set_tri_slots(t.s)
    gpu.set_slots([(0.2, 0.0, 0.0, 0), (0.2, 0.5, 0.0, 0), (0.1, 0.3, 0.0, 0)])
    D = rng.normal(size=(700, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    D = D.astype(np.float32)
    n = 4000
    P, N = sensors(rng, n)
    gpu.set_scene(P.reshape(-1), N.reshape(-1), D.reshape(-1))
    useCos = [1, 0, 1, 1]
    W = rng.uniform(0, 2, (len(useCos), len(D))).astype(np.float32)
    hip = Hip()
    try:
        dW = hip.put(W.reshape(-1))
        limit = gpu.noalloc_limit()
        assert limit == 64 << 20
        for translucent in (False, True):
            for i0, nc in ((0, 4000), (1000, 700)):   # the second call smaller than the first
                print("case: no-alloc chunk, translucent %s, sensors [%d, %d)" % (translucent, i0, i0 + nc), flush=True)
                want = gpu.chunk(i0, nc, W, useCos, translucent=translucent)
                dP = hip.put(P[i0:i0 + nc].reshape(-1))
                dN = hip.put(N[i0:i0 + nc].reshape(-1))
                blank = np.full(len(useCos) * nc, np.nan)
                dS = hip.put(blank)
                gpu.chunk_noalloc(dP, dN, nc, dW, useCos, dS, translucent=translucent)
                got = hip.get(dS, blank).reshape(len(useCos), nc)
                assert np.array_equal(got, want), (translucent, i0, nc)
                assert np.count_nonzero(got) > 0
        # beyond the array: refused on the host with the limit named, nothing launched
        per = ((len(D) + 31) // 32) * 4
        big = limit // per + 1
        with pytest.raises(Exception, match="device array"):
            gpu.chunk_noalloc(dP, dN, big, dW, useCos, dS)
    finally:
        hip.free()


def test_noalloc_equals_chunk_deep_stacks(gpu):
    from test_gpu_configs import _comb, _comb_sensors
    rng = np.random.default_rng(9)
    V, I = _comb()
    nt = len(I) // 3
    gpu.build(V, I)
    gpu.set_tri_slots((np.arange(nt, dtype=np.uint32) // 7) % 3)
    gpu.set_slots([(0.2, 0.0, 0.0, 0), (0.2, 0.5, 0.0, 0), (0.1, 0.3, 0.0, 0)])
    n = 256
    P, N = _comb_sensors(rng, n)
    D = rng.normal(size=(48, 3))
    D[:32] = D[:32] * 0.2 + [1.0, 0.0, 0.0]   # along the comb
    D = (D / np.linalg.norm(D, axis=1, keepdims=True)).astype(np.float32)
    gpu.set_scene(P.reshape(-1), N.reshape(-1), D.reshape(-1))
    useCos = [1, 0]
    W = rng.uniform(0, 2, (len(useCos), len(D))).astype(np.float32)
    hip = Hip()
    try:
        dW = hip.put(W.reshape(-1))
        for translucent in (False, True):
            for i0, nc in ((0, n), (150, 100)):   # the second call smaller, and with open sensors (those past 192)
                print("case: no-alloc chunk on the comb, translucent %s, sensors [%d, %d)" % (translucent, i0, i0 + nc), flush=True)
                want = gpu.chunk(i0, nc, W, useCos, translucent=translucent)
                dP = hip.put(P[i0:i0 + nc].reshape(-1))
                dN = hip.put(N[i0:i0 + nc].reshape(-1))
                blank = np.full(len(useCos) * nc, np.nan)
                dS = hip.put(blank)
                gpu.chunk_noalloc(dP, dN, nc, dW, useCos, dS, translucent=translucent)
                got = hip.get(dS, blank).reshape(len(useCos), nc)
                assert np.array_equal(got, want), (translucent, i0, nc)
                assert np.count_nonzero(got) > 0
    finally:
        hip.free()
