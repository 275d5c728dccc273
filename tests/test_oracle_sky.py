"""The oracle's sky model (reference src/sky.cpp) pinned by the reference's own known answers (tests/check_sky.cpp), by
closed forms, and bit for bit by the reference's sky.cpp itself where it was built (oracle/_ref)."""
import math

import numpy as np
import pytest

PI = math.pi


def test_patch_counts_and_solid_angles(orc):
    for mf, P in ((1, 145), (2, 577), (4, 2305)):   # check_sky.cpp:23-31
        p = orc.patches(mf)
        assert p["P"] == P
        assert abs(float(np.sum(p["omega"].astype(np.float64))) - 2 * PI) <= 1e-6
        assert p["row"][-1] == 7 * mf and p["ninrow"][-1] == 1 and p["row"][0] == 0 and p["ninrow"][0] == 30 * mf


def test_patch_altitudes_and_vectors(orc, known):
    t = orc.patches(1)
    alt_c = [6, 18, 30, 42, 54, 66, 78]
    p0 = 0
    for i in range(7):   # check_sky.cpp:33-42: centres every 12 degrees, Ladybug trapezoid altitudes
        assert abs(math.degrees(t["alt"][p0]) - alt_c[i]) <= 1e-4
        assert abs(math.degrees(math.asin(t["lb"][p0][2])) - known["alt_ladybug_mf1"][i]) <= 1e-5
        assert abs(math.atan2(t["lb"][p0][0], t["lb"][p0][1]) - t["azi"][p0]) <= 1e-6
        p0 += t["ninrow"][p0]
    assert p0 == 144
    c = t["centre"]
    assert c[0][0] == 0 and abs(c[0][1] - math.cos(math.radians(6))) <= 1e-6 and c[1][0] > 0 and c[1][1] > 0
    assert abs(c[144][2] - 1) <= 1e-6 and list(t["lb"][144]) == [0, 0, 1]


def test_patch_samples_stay_inside(orc):
    t = orc.patches(1)
    for p in range(t["P"]):
        c = np.zeros(3, np.float32)
        orc.lib().orc_patch_sample(1, p, 0, 1, c)
        assert np.array_equal(c, t["centre"][p])
        for k in range(4):
            s = np.zeros(3, np.float32)
            orc.lib().orc_patch_sample(1, p, k, 4, s)
            alt = math.asin(float(s[2]))
            assert abs(float(np.dot(s.astype(np.float64), s)) - 1) <= 1e-6
            a = float(t["alpha"])
            if p == t["P"] - 1:
                assert alt >= PI / 2 - a / 2 - 1e-6
                continue
            assert a * t["row"][p] - 1e-6 <= alt <= a * (t["row"][p] + 1) + 1e-6
            daz = math.fmod(math.atan2(s[0], s[1]) - t["azi"][p] + 3 * PI, 2 * PI) - PI
            assert abs(daz) <= PI / t["ninrow"][p] + 1e-6


def test_night_and_normalisation(orc):
    t = orc.patches(1)
    out = orc.gendaymtx_step(1, 172, PI / 4, PI, 5e-5, 5e-5, 0.2, True, True, False)
    assert not out.any()
    out = orc.gendaymtx_step(1, 172, PI / 4, PI, 0, 100, 0.2, True, True, False)
    expect = 100.0 * float(np.float32(1.004))
    s_tab = sum(out[3 * (p + 1) + 1] * orc.lib().orc_tsin(float(t["alt"][p])) * float(t["omega"][p]) for p in range(t["P"]))
    s_true = sum(out[3 * (p + 1) + 1] * math.sin(float(t["alt"][p])) * float(t["omega"][p]) for p in range(t["P"]))
    assert abs(s_tab - expect) <= 1e-6 * expect   # check_sky.cpp:83-92: table sine identity
    assert abs(s_true - expect) <= 1e-3 * expect
    assert abs(out[4] / out[3] - 1.004 / 0.96) <= 1e-6 * 1.004 / 0.96 and abs(out[5] / out[3] - 1.118 / 0.96) <= 1e-6 * 1.118 / 0.96
    assert abs(out[0] - 0.2 * 100 / PI) <= 1e-5 * 0.2 * 100 / PI and out[1] == out[0] == out[2]
    assert (out[3::3][:t["P"]] > 0).all()


def perez_L(c, z, t, p, south):
    zeta = PI / 2 - float(t["alt"][p])
    g = math.acos(math.cos(z) * math.cos(zeta) + math.sin(z) * math.sin(zeta) * math.cos(abs(float(t["azi"][p]) - south)))
    return (1 + c[0] * math.exp(c[1] / math.cos(zeta))) * (1 + c[2] * math.exp(c[3] * g) + c[4] * math.cos(g) ** 2)


def test_perez_ratios_categories_0_and_7(orc):
    """Hand-evaluated Perez 1993 Table 1 rows 1 and 8 (check_sky.cpp:97-118)."""
    t = orc.patches(1)
    z, da, south = PI / 4, 171 * 2 * PI / 365, PI
    E0 = 1.00011 + 0.034221 * math.cos(da) + 0.00128 * math.sin(da) + 0.000719 * math.cos(2 * da) + 0.000077 * math.sin(2 * da)
    D = 100 / (math.cos(z) + 0.15 * (93.885 - 45) ** -1.253) / (1367 * E0)
    c0 = [1.3525 - 0.2576 * z + D * (-0.2690 - 1.4366 * z), -0.7670 + 0.0007 * z + D * (1.2734 - 0.1233 * z),
          math.exp((D * (2.8 + 0.6004 * z)) ** 1.2375) - 1.0, -math.exp(D * (1.8734 + 0.6297 * z)) + 0.9738 + D * 0.2809,
          0.0356 - 0.1246 * z + D * (-0.5718 + 0.9938 * z)]
    c7 = [-1.0500 + 0.0289 * z + D * (0.4260 + 0.3590 * z), -0.3250 + 0.1156 * z + D * (0.7781 + 0.0025 * z),
          31.0625 - 14.5 * z + D * (-46.1148 + 55.375 * z), -7.2312 + 0.4050 * z + D * (13.35 + 0.6234 * z),
          1.5 - 0.6426 * z + D * (1.8564 + 0.5636 * z)]
    P = t["P"]
    for dni, c in ((0, c0), (800, c7)):
        out = orc.gendaymtx_step(1, 172, PI / 4, south, dni, 100, 0.2, False, True, False)
        r1 = perez_L(c, z, t, P - 1, south) / perez_L(c, z, t, 0, south)
        r2 = perez_L(c, z, t, 15, south) / perez_L(c, z, t, 0, south)
        assert abs(out[3 * P + 1] / out[4] - r1) <= 1e-5 * abs(r1)
        assert abs(out[3 * 16 + 1] / out[4] - r2) <= 1e-5 * abs(r2)


def test_ground_row_and_add_direct(orc):
    t = orc.patches(1)
    out = orc.gendaymtx_step(1, 172, PI / 4, PI, 800, 100, 0.2, True, False, False)
    assert out[0] == out[1] == out[2] == 0   # -d zeroes the ground reflectance
    out = orc.gendaymtx_step(1, 172, PI / 4, PI, 800, 100, 0.2, False, True, False)
    e = 0.2 * (100 + 800 * math.sin(PI / 4)) / PI
    assert abs(out[0] - e) <= 1e-5 * e
    alt, azi = math.radians(35), math.radians(200)
    out = orc.gendaymtx_step(1, 172, alt, azi, 800, 100, 0.2, True, False, False)
    g = out[4::3][:t["P"]]
    assert abs(float(np.sum(g.astype(np.float64) * t["omega"])) - 800) <= 800e-5   # check_sky.cpp:124-139
    assert np.count_nonzero(g) == 4 and np.array_equal(out[3::3][:t["P"]], g) and np.array_equal(out[5::3][:t["P"]], g)
    sv = np.array([math.cos(alt) * math.sin(azi), math.cos(alt) * math.cos(azi), math.sin(alt)], np.float32).astype(np.float64)
    dots = t["centre"].astype(np.float64) @ sv
    assert set(np.nonzero(g)[0]) == set(np.argsort(-dots)[:4])
    for vis in (0, 1):   # DHI 0: clearness inf -> 11.9, nothing NaN
        out = orc.gendaymtx_step(1, 172, PI / 4, PI, 800, 0, 0.2, True, True, bool(vis))
        assert np.isfinite(out).all() and out[0] > 0


def test_luminous_efficacy(orc):
    """-O0 path, category 7: Perez 1990 Eqns 7/8 by hand (check_sky.cpp:151-167)."""
    t = orc.patches(1)
    z, W, da = PI / 4, math.exp(0.07 * 11 - 0.075), 171 * 2 * PI / 365
    E0 = 1.00011 + 0.034221 * math.cos(da) + 0.00128 * math.sin(da) + 0.000719 * math.cos(2 * da) + 0.000077 * math.sin(2 * da)
    m = 1 / (math.cos(z) + 0.15 * (93.885 - 45) ** -1.253)
    delta = 100 * m / (1367 * E0)
    dif = 100 * (152.23 + 0.35 * W - 45.27 * math.cos(z) - 7.98 * math.log(delta))
    dr = 800 * (101.18 + 1.58 * W - 1.10 * math.exp(5.73 * z - 5) - 8.29 * delta)
    out = orc.gendaymtx_step(1, 172, PI / 4, PI, 800, 100, 0.2, True, True, True)
    e = 0.2 * (dif + dr * math.sin(z)) / PI / 179
    assert abs(out[0] - e) <= 1e-5 * e
    out = orc.gendaymtx_step(1, 172, PI / 4, PI, 800, 100, 0.2, False, True, True)
    s = sum(out[3 * (p + 1) + 1] * orc.lib().orc_tsin(float(t["alt"][p])) * float(t["omega"][p]) for p in range(t["P"]))
    assert abs(s - dif / 179 * float(np.float32(1.004))) <= 1e-5 * dif / 179
    out = orc.gendaymtx_step(1, 172, PI / 4, PI, 800, 100, 0.2, True, False, True)
    s = float(np.sum(out[4::3][:t["P"]].astype(np.float64) * t["omega"]))
    assert abs(s - dr / 179) <= 1e-5 * dr / 179


def test_round_fmt_and_sun_c(orc):
    L = orc.lib()
    assert L.orc_round_fmt(0.123456, b"%.3g") == 0.123 and L.orc_round_fmt(1234.5, b"%.3g") == 1230
    assert L.orc_round_fmt(42.346, b"%.2f") == 42.35
    import ctypes as C
    a, z = C.c_double(), C.c_double()
    L.orc_sun_radiance(90, 0, 0, 172, 0, C.byref(a), C.byref(z))
    assert abs(a.value - 0.409296209774023) <= 1e-12   # check_sky.cpp:180-182
    L.orc_sun_radiance(0, 0, 0, 81, 0, C.byref(a), C.byref(z))
    assert abs(a.value + PI / 2 - 0.118294008300896 * PI / 12) <= 1e-12
    L.orc_sun_radiance(42.35, 71.07, 75, 172, 12.5, C.byref(a), C.byref(z))
    assert abs(math.degrees(a.value) - 68.97980) <= 1e-3 and abs(math.degrees(z.value) - 209.4386) <= 1e-3


# ---------------------------------------------------------------------------------------------- the reference itself
@pytest.fixture(scope="module")
def ref(orc):
    r = orc.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (no reference sources here)")
    return r


def test_patches_bit_exact_vs_reference(orc, ref):
    import ctypes as C
    for mf in (1, 2, 4):
        P = orc.lib().orc_patch_count(mf)
        a = [np.zeros(P, np.float32) for _ in range(3)] + [np.zeros(P, np.int32) for _ in range(2)] + [np.zeros(3 * P, np.float32) for _ in range(2)]
        alpha = C.c_float()
        assert ref.ref_patches(mf, *a, C.byref(alpha)) == P
        mine = orc.patches(mf)
        for k, key in enumerate(("alt", "azi", "omega", "row", "ninrow")):
            assert np.array_equal(mine[key], a[k]), key
        assert np.array_equal(mine["centre"].reshape(-1), a[5]) and np.array_equal(mine["lb"].reshape(-1), a[6])
        for p in range(0, P, 7):
            for K in (1, 4, 9):
                x, y = np.zeros(3, np.float32), np.zeros(3, np.float32)
                orc.lib().orc_patch_sample(mf, p, K - 1, K, x)
                ref.ref_patch_sample(mf, p, K - 1, K, y)
                assert np.array_equal(x, y)


def test_gendaymtx_rows_bit_exact_vs_reference(orc, ref):
    rng = np.random.default_rng(7)
    cases = [(172, PI / 4, PI, 0, 100), (172, PI / 4, PI, 800, 100), (172, PI / 4, PI, 800, 0), (1, -0.1, 1.0, 5e-5, 5e-5),
             (355, 0.05, 2.5, 300, 20), (80, 1.55, 0.3, 900, 50)]
    for _ in range(40):
        cases.append((int(rng.integers(1, 366)), float(rng.uniform(-0.2, 1.6)), float(rng.uniform(0, 2 * PI)),
                      float(rng.uniform(0, 1000)), float(rng.uniform(0, 400))))
    for mf in (1, 2, 4):
        P = orc.lib().orc_patch_count(mf)
        for jd, alt, azi, dni, dhi in cases:
            for sun, sky, vis in ((0, 1, 0), (1, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0)):
                mine = orc.gendaymtx_step(mf, jd, alt, azi, dni, dhi, 0.2, sun, sky, vis)
                theirs = np.zeros(3 * (P + 1), np.float32)
                ref.ref_gendaymtx_step(mf, jd, alt, azi, dni, dhi, 0.2, sun, sky, vis, theirs)
                assert np.array_equal(mine, theirs), (mf, jd, alt, azi, dni, dhi, sun, sky, vis)
