"""Ladybug mode on the CPU: the oracle's restatement pinned, and the product's host emulation of gendaymtx -A against it.

Pins of the oracle: the reference's own ladybug sensor checks (tests/check_mesh.cpp:89-150: a non-planar quad's normal
normalize(n1 + n2) and its corners' mean, the n-gon error, a cube's vertex normals, area weighting); the sky emulation
(reference run.cpp:362-470) against (a) a single step, where the averages are the step's own gendaymtx rows printed
%.3g, (b) the same rows summed in float32 step order by numpy, (c) what gendaymtx conserves: the -d rows carry the
direct horizontal-normal energy (sum_p avg dir_p omega_p = mean DNI), the -s green channel the diffuse horizontal
(sum_p avg green_p tsin_p omega_p = 1.004 mean DHI), the ground row (DHI + DNI sin alt) / pi x ground, all to the
%.3g rounding. gendaymtx_step and sun_radiance themselves are pinned to the reference's sky.cpp (test_oracle_sky.py).

The product: dlx_ladybug_sky (the host code the run uses) bit for bit against the oracle; the mode's flag rules
through the CLI before any GPU work.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kit-sim_daylight_amd", "lib")
BOX = os.path.join(ROOT, "tests", "golden", "box.obj")


def g3(x):
    return float("%.3g" % x)


def poly(orc, pts, counts=None, idx=None):
    pts = np.array(pts, np.float64).reshape(-1, 3)
    n = len(pts)
    return orc.MeshArrays("p", pts, counts or [n], idx or list(range(n)))


def near3(a, b, tol=1e-6):
    return np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))) <= tol


def test_ladybug_sensor_rules_match_reference_checks(orc):
    q = poly(orc, [0, 0, 0, 1, 0, 0, 1, 1, 0.2, 0, 1, 0])
    P, N = orc.sensors(q, ladybug=True)
    s = np.array([-0.2, -0.4, 2.0])
    assert near3(N[0], s / np.linalg.norm(s)) and near3(P[0], [0.5, 0.5, 0.05])
    P2, N2 = orc.sensors(q, ladybug=False)   # physical: Newell
    n = np.array([-0.2, -0.2, 2.0])
    assert near3(N2[0], n / np.linalg.norm(n)) and abs(N[0][1] - N2[0][1]) > 1e-6
    L = orc.MeshArrays("L", np.array([0, 0, 0, 2, 0, 0, 2, 1, 0, 1, 1, 0, 1, 2, 0, 0, 2, 0], np.float64), [3, 6], [0, 1, 2, 0, 1, 2, 3, 4, 5])
    assert orc.sensors(L, ladybug=True) is None
    cube = orc.MeshArrays("cube", np.array([0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 0, 0, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1], np.float64), [4] * 6,
                          [0, 3, 2, 1, 4, 5, 6, 7, 0, 1, 5, 4, 1, 2, 6, 5, 2, 3, 7, 6, 3, 0, 4, 7])
    k = 1 / math.sqrt(3)
    Pc, Nc = orc.sensors(cube, ladybug=True, onVertices=True)
    assert near3(Nc[6], [k, k, k]) and near3(Nc[0], [-k, -k, -k]) and near3(Pc[6], [1, 1, 1])
    two = orc.MeshArrays("two", np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 2, 2, 0, 0], np.float64), [3, 3], [0, 1, 2, 0, 3, 4])
    w = math.sqrt(4.25)
    for lb in (True, False):
        assert near3(orc.sensors(two, ladybug=lb, onVertices=True)[1][0], [0, 2 / w, 0.5 / w])


def boston_steps(orc, every=1):
    dni, dhi = orc.clear_sky(42.35, -71.07, -5)
    st = orc.make_steps(dni=dni, dhi=dhi)[::every]
    return (np.array([s.jd for s in st], np.int32), np.array([s.hour for s in st]), np.array([s.dni for s in st], np.float32),
            np.array([s.dhi for s in st], np.float32), float(sum(s.dt for s in st)))


def rows_of(orc, mf, jd, hour, dni, dhi, lat, lon, tz, grefl):
    """Each step's gendaymtx -d and -s rows at the .wea-rounded site, float32 [T, P + 1, 3]."""
    L = orc.lib()
    latW, lonW, merW = g3w(lat), g3w(-lon), float(int(-tz * 15.0))
    D, S, alts = [], [], []
    for t in range(len(jd)):
        a, z = C.c_double(), C.c_double()
        L.orc_sun_radiance(latW, lonW, merW, int(jd[t]), float(hour[t]), C.byref(a), C.byref(z))
        dn, dh = float(int(dni[t])), float(int(dhi[t]))
        D.append(orc.gendaymtx_step(mf, int(jd[t]), a.value, z.value, dn, dh, grefl, 1, 0, 0).reshape(-1, 3))
        S.append(orc.gendaymtx_step(mf, int(jd[t]), a.value, z.value, dn, dh, grefl, 0, 1, 0).reshape(-1, 3))
        alts.append(a.value)
    return np.array(D), np.array(S), np.array(alts)


def g3w(x):
    return float("%.2f" % x)


@pytest.mark.parametrize("mf", [1, 2])
def test_oracle_single_step_is_its_rows_printed(orc, mf):
    jd, hour, dni, dhi, _ = boston_steps(orc)
    t = int(np.argmax(dni))   # a clear noon
    sl = slice(t, t + 1)
    out = orc.ladybug_sky(mf, jd[sl], hour[sl], dni[sl], dhi[sl], 42.35, -71.07, -5.0, 0.2, 1.0)
    D, S, _ = rows_of(orc, mf, jd[sl], hour[sl], dni[sl], dhi[sl], 42.35, -71.07, -5.0, 0.2)
    want_s = np.vectorize(g3)(S[0].astype(np.float64)).astype(np.float32)
    want_d = np.vectorize(g3)(D[0].astype(np.float64)).astype(np.float32)
    assert np.array_equal(out["dump"][0], want_s)
    assert np.array_equal(out["dump"][1][1:], np.repeat(want_d[1:, :1], 3, axis=1))


def test_oracle_year_sums_in_float_step_order_and_conserves(orc):
    mf = 1
    jd, hour, dni, dhi, dur = boston_steps(orc, every=3)
    grefl = 0.2
    out = orc.ladybug_sky(mf, jd, hour, dni, dhi, 42.35, -71.07, -5.0, grefl, dur)
    D, S, alts = rows_of(orc, mf, jd, hour, dni, dhi, 42.35, -71.07, -5.0, grefl)
    T = len(jd)
    inv = np.float32(1.0 / T)
    sumD = np.cumsum(D, axis=0, dtype=np.float32)[-1]   # sequential, as a float accumulator in step order
    sumS = np.cumsum(S, axis=0, dtype=np.float32)[-1]
    avgD, avgS = sumD * inv, sumS * inv
    pr = np.vectorize(g3)
    assert np.array_equal(out["dump"][0], pr(avgS.astype(np.float64)).astype(np.float32))
    assert np.array_equal(out["dump"][1][1:, 0], pr(avgD[1:, 0].astype(np.float64)).astype(np.float32))
    pa = orc.patches(mf)
    om = pa["omega"].astype(np.float64)
    rd = pr(avgD[1:, 0].astype(np.float64))   # the printed average as read back (double), before its float copy in the dump
    assert np.array_equal(out["dir"], rd * om * dur / 1000.0)
    # energy: AddDirect spreads dni x 179 / 179 over omega; the -s green channel integrates to 1.004 DHI
    dniI, dhiI = np.floor(dni).astype(np.float64), np.floor(dhi).astype(np.float64)
    up = (dniI + dhiI) > 1e-4
    assert abs((rd * om).sum() / np.mean(np.where(up & (dniI * 179 > 1e-4), dniI, 0.0)) - 1) < 5e-3
    tsin = np.sin(pa["alt"].astype(np.float64))
    green = out["dump"][0][1:, 1].astype(np.float64)
    assert abs((green * tsin * om).sum() / (1.004 * np.mean(np.where(up, dhiI, 0.0))) - 1) < 5e-3
    gnd = np.mean(np.where(up, (dhiI + np.where(alts > 0, dniI * np.sin(alts), 0.0)) / math.pi * grefl, 0.0))
    assert abs(out["dump"][0][0, 0] / gnd - 1) < 5e-3
    assert abs(out["gnd"] - grefl * out["rad"].mean()) <= 1e-15 * out["gnd"]


def product_sky(mf, jd, hour, dni, dhi, lat, lon, tz, grefl, dur):
    import daylight_amd
    L = daylight_amd.lib()
    P = 145 if mf == 1 else 577
    d, f, r, dump = np.zeros(P), np.zeros(P), np.zeros(P), np.zeros(6 * (P + 1), np.float32)
    g = C.c_double()
    jd, hour = np.ascontiguousarray(jd, np.int32), np.ascontiguousarray(hour, np.float64)
    dni, dhi = np.ascontiguousarray(dni, np.float32), np.ascontiguousarray(dhi, np.float32)
    rc = L.dlx_ladybug_sky(mf, len(jd), jd.ctypes.data, hour.ctypes.data, dni.ctypes.data, dhi.ctypes.data, lat, lon, tz, grefl, dur,
                           d.ctypes.data, f.ctypes.data, r.ctypes.data, C.byref(g), dump.ctypes.data)
    assert rc == 0
    return dict(dir=d, dif=f, rad=r, gnd=g.value, dump=dump.reshape(2, P + 1, 3))


@pytest.mark.parametrize("mf,lat,lon,tz", [(1, 42.35, -71.07, -5.0), (2, 52.52, 13.41, 1.0)])
def test_product_host_sky_bit_exact_vs_oracle(orc, mf, lat, lon, tz):
    dni, dhi = orc.clear_sky(lat, lon, tz)
    rng = np.random.default_rng(mf)
    dni = (dni * rng.uniform(0, 1.2, dni.size)).astype(np.float32)   # clouds, and values past the clear sky
    st = orc.make_steps(dni=dni, dhi=dhi)
    jd = np.array([s.jd for s in st], np.int32)
    hour = np.array([s.hour for s in st])
    dn = np.array([s.dni for s in st], np.float32)
    dh = np.array([s.dhi for s in st], np.float32)
    dur = float(sum(s.dt for s in st))
    o = orc.ladybug_sky(mf, jd, hour, dn, dh, lat, lon, tz, 0.2, dur)
    p = product_sky(mf, jd, hour, dn, dh, lat, lon, tz, 0.2, dur)
    for k in ("dir", "dif", "rad"):
        assert np.array_equal(p[k], o[k]), k
    assert p["gnd"] == o["gnd"]
    assert np.array_equal(p["dump"], o["dump"])


def tool(*args):
    exe = os.path.join(LIB, "daylight")
    return subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=60)


def test_mode_rules_before_any_gpu_work(tmp_path):
    base = [BOX, "--clear-sky", "--lat", "52.52", "--lon", "13.41", "--tz", "1", "--mode", "ladybug"]
    r = tool(*base, "--quantity", "illuminance")
    assert r.returncode == 2 and "illuminance is not available in ladybug mode" in r.stderr
    r = tool(*base, "--sky-density", "4")
    assert r.returncode == 2 and "--sky-density 1 or 2 only" in r.stderr
    r = tool(BOX, "--clear-sky", "--lat", "52.52", "--lon", "13.41", "--tz", "1", "--dump-sky", str(tmp_path / "s.txt"))
    assert r.returncode == 2 and "--dump-sky is ladybug-mode only" in r.stderr
    hexagon = tmp_path / "hex.obj"
    hexagon.write_text("".join("v %g %g 0\n" % (math.cos(a), math.sin(a)) for a in np.arange(6) * math.pi / 3) + "f 1 2 3 4 5 6\n")
    r = tool(str(hexagon), "--clear-sky", "--lat", "52.52", "--lon", "13.41", "--tz", "1", "--mode", "ladybug")
    assert r.returncode == 2 and "ladybug mode supports tri/quad faces only" in r.stderr
