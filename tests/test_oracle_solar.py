"""The oracle's sun / clear sky / steps (reference src/solar.cpp) pinned by tests/check_solar.cpp's known answers and,
where it was built, bit for bit by the reference's solar.cpp (oracle/_ref)."""
import ctypes as C
import math

import numpy as np
import pytest


def test_ladybug_sun_goldens(orc, known):
    for lat, lon, tz, hoy, leap, alt, azi, up in known["ladybug_sun"]:   # check_solar.cpp:18-56, 1e-5 deg
        m, d, h = orc.hoy_to_date(hoy, bool(leap))
        a, z, u, _ = orc.sun_noaa(lat, lon, tz, m, d, h, bool(leap))
        assert abs(a - alt) < 1e-5 and abs(z - azi) < 1e-5 and u == bool(up), (lat, hoy, a, z)


def test_sun_vector(orc, known):
    _, _, _, v = orc.sun_noaa(42.35, -71.07, -5, 6, 16, 16.5, False)
    assert np.allclose(v, known["sun_vec_4000_5"], atol=1e-6, rtol=0)
    n = np.zeros(3, np.float32)
    orc.lib().orc_rotate_z(v, 30, n)
    assert np.allclose(n[:2], known["sun_vec_4000_5_north30"], atol=1e-6, rtol=0) and abs(n[2] - v[2]) < 1e-7


def test_ineichen_and_spencer(orc, known):
    g, b, d = C.c_double(), C.c_double(), C.c_double()
    for z, TL, alt, ghi, dni, dhi in known["ineichen"]:   # pvlib 0.16.1 goldens, 1e-6 relative
        orc.lib().orc_ineichen(z, TL, alt, 1366.1, C.byref(g), C.byref(b), C.byref(d))
        for got, want in ((g.value, ghi), (b.value, dni), (d.value, dhi)):
            assert abs(got - want) / want < 1e-6
    for doy, want in known["spencer"].items():
        assert abs(orc.lib().orc_extraterrestrial(int(doy)) - want) / want < 1e-6


def test_boston_clear_sky_sums(orc, known):
    dni, dhi = orc.clear_sky(42.35, -71.07, -5, 3, 0)
    sdni, sdhi = float(np.sum(dni.astype(np.float64))) / 1000, float(np.sum(dhi.astype(np.float64))) / 1000
    want_dni, want_dhi = known["boston_clear_sums"]
    assert abs(sdni - want_dni) < 0.005 * want_dni and abs(sdhi - want_dhi) < 0.01 * want_dhi
    assert dni[11] > 0 and dhi[11] > 0 and dni[0] == 0 and dni[23] == 0
    with pytest.raises(ValueError):
        orc.clear_sky(42.35, -71.07, -5, 0.5, 0)


def test_annual_steps(orc):
    dni = np.arange(8760, dtype=np.float32) % 1000
    dhi = np.arange(8760, dtype=np.float32) % 500
    st = orc.make_steps(dni=dni, dhi=dhi)
    assert len(st) == 8760
    assert (st[0].hoy, st[0].hour, st[0].month, st[0].day, st[0].jd, st[0].dt) == (0.5, 0.5, 1, 1, 1, 1)
    assert (st[8759].hoy, st[8759].hour, st[8759].month, st[8759].day, st[8759].jd) == (8759.5, 23.5, 12, 31, 365)
    assert (st[1415].month, st[1415].day, st[1415].jd, st[1416].month, st[1416].day, st[1416].jd) == (2, 28, 59, 3, 1, 60)
    leap = orc.make_steps(dni=np.zeros(8784, np.float32), dhi=np.zeros(8784, np.float32), leap=True)
    assert (leap[1428].month, leap[1428].day, leap[1428].jd, leap[1428].hour) == (2, 29, 60, 12.5)
    assert (leap[8783].jd, leap[8783].hoy) == (366, 8783.5)


def test_range_steps(orc):
    dni = np.arange(8760, dtype=np.float32) % 1000
    dhi = np.arange(8760, dtype=np.float32) % 500
    st = orc.make_steps(t0=4110, t1=4113, stepMin=30, dni=dni, dhi=dhi)   # check_solar.cpp:159-167
    assert len(st) == 6
    for k, s in enumerate(st):
        assert s.hour == 6.25 + 0.5 * k and s.hoy == 4110.25 + 0.5 * k and s.dt == 0.5 and (s.month, s.day, s.jd) == (6, 21, 172)
    assert abs(st[0].dni - (dni[4109] * 0.25 + dni[4110] * 0.75)) < 1e-3
    st = orc.make_steps(t0=4110, t1=4111, stepMin=10, dni=dni, dhi=dhi)
    assert len(st) == 6 and abs(st[5].hour - (6 + 55 / 60.0)) < 1e-9
    st = orc.make_steps(t0=4110, dni=dni, dhi=dhi)
    assert len(st) == 1 and st[0].hour == 6 and st[0].dt == 1 and abs(st[0].dni - (dni[4109] + dni[4110]) / 2) < 1e-3
    st = orc.make_steps(t0=24 * 151 + 8, t1=24 * 153 + 18, dayFrom=480, dayTo=1080, dni=dni, dhi=dhi)
    assert len(st) == 30 and st[9].hour == 17.5 and st[10].day == 2 and st[10].hoy == st[0].hoy + 24
    st = orc.make_steps(t0=24 * 151 + 8, t1=24 * 153 + 18, dayFrom=480, dayTo=1050, dni=dni, dhi=dhi)
    assert len(st) == 30 and st[9].dt == 0.5 and st[9].hour == 17.25
    st = orc.make_steps(t0=0.5 / 24 * 24 * 0 + 20 / 60.0, t1=24 * 4 + 20 / 60.0, stepMin=10, dni=dni, dhi=dhi)
    assert len(st) == 576
    st = orc.make_steps(t0=20 / 60.0, t1=24 * 4 + 20 / 60.0, stepMin=10, dayTo=60, dni=dni, dhi=dhi)
    assert len(st) == 24
    a = orc.make_steps(t0=0, t1=8760, dni=dni, dhi=dhi)
    b = orc.make_steps(dni=dni, dhi=dhi)
    assert all((x.month, x.day, x.hour, x.dni, x.dhi) == (y.month, y.day, y.hour, y.dni, y.dhi) for x, y in zip(a, b))
    jan = orc.make_steps(t0=0.5, dni=dni, dhi=dhi, shift=1)
    assert (jan[0].month, jan[0].day, jan[0].hour, jan[0].dni) == (12, 31, 23.5, dni[8759])
    with pytest.raises(ValueError):
        orc.make_steps(t0=10, t1=5, dni=dni, dhi=dhi)


def test_parse_hoy(orc):
    L = orc.lib()
    h = C.c_double()
    assert L.orc_parse_hoy(b"2017-06-21T12:30", 0, C.byref(h)) == 0 and h.value == 4116.5
    assert L.orc_parse_hoy(b"02-29T00:00", 1, C.byref(h)) == 0 and h.value == 1416
    assert L.orc_parse_hoy(b"02-29T00:00", 0, C.byref(h)) != 0
    assert L.orc_parse_hoy(b"06-21T24:00", 0, C.byref(h)) != 0 and L.orc_parse_hoy(b"garbage", 0, C.byref(h)) != 0
    for r in range(0, 8784, 97):
        m, d, hr = orc.hoy_to_date(r + 0.25, True)
        s = ("%02d-%02dT%02d:%02d" % (m, d, int(hr), round((hr - int(hr)) * 60))).encode()
        assert L.orc_parse_hoy(s, 1, C.byref(h)) == 0 and h.value == r + 0.25


@pytest.fixture(scope="module")
def ref(orc):
    r = orc.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (no reference sources here)")
    return r


def test_sun_and_clear_sky_bit_exact_vs_reference(orc, ref):
    for lat, lon, tz in ((52.52, 13.41, 1), (42.35, -71.07, -5), (-33.87, 151.21, 10), (69.65, 18.96, 1)):
        a, b = orc.clear_sky(lat, lon, tz), (np.zeros(8760, np.float32), np.zeros(8760, np.float32))
        assert ref.ref_clear_sky(lat, lon, tz, 3.0, 0.0, *b) == 0
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        for hoy in np.linspace(0.25, 8783.75, 301):
            for leap in (0, 1):
                m, d, h = orc.hoy_to_date(float(hoy), bool(leap))
                x = orc.sun_noaa(lat, lon, tz, m, d, h, bool(leap))
                al, az, up, v = C.c_double(), C.c_double(), C.c_int(), np.zeros(3, np.float32)
                ref.ref_sun_noaa(lat, lon, tz, m, d, h, leap, C.byref(al), C.byref(az), C.byref(up), v)
                assert (x[0], x[1], x[2]) == (al.value, az.value, bool(up.value)) and np.array_equal(x[3], v)


def test_steps_bit_exact_vs_reference(orc, ref):
    dni = (np.arange(8760, dtype=np.float32) * 7) % 1000
    dhi = (np.arange(8760, dtype=np.float32) * 3) % 500
    for t0, t1, step, df, dt_, shift in ((math.nan, math.nan, 60, 0, 1440, 0), (math.nan, math.nan, 60, 0, 1440, 1.5),
                                         (4110, 4113, 30, 0, 1440, 0), (100.3, 400.7, 7, 480, 1050, -2), (8759.5, math.nan, 60, 0, 1440, 0)):
        mine = orc.make_steps(t0=t0, t1=t1, stepMin=step, dayFrom=df, dayTo=dt_, dni=dni, dhi=dhi, shift=shift)
        out = np.zeros(8 * 20000)
        n = ref.ref_make_steps(t0, t1, step, df, dt_, dni.ctypes.data, dhi.ctypes.data, 8760, 0, shift, out, 20000)
        assert n == len(mine)
        rows = out[: 8 * n].reshape(n, 8)
        got = np.array([[s.month, s.day, s.jd, s.hour, s.hoy, s.dt, s.dni, s.dhi] for s in mine])
        assert np.array_equal(rows, got)
