"""--mode ladybug end to end on the GPU (reference run.cpp:221-227, 304-330, 362-470, 476-505, 639-679, 725-736): the
daylight tool's columns, its --dump-sky blocks and its series against the oracle's ladybug run of the same flags.

Bars: sun hours exact; the energy columns within 2e-6 relative to the oracle (the reduce's bar); the dumped blocks as
the same text; the FP32 series within 2e-6 of the oracle's FP64 series.
"""
import os

import numpy as np
import pytest

from test_gpu_e2e import BERLIN, clear_weather, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOX = os.path.join(ROOT, "tests", "golden", "box.obj")
QS = ("radiation", "direct", "diffuse", "irradiance", "sunhours")


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


@pytest.fixture(scope="module")
def box(native):
    m = native.SimulationMesh.load(BOX)
    yield m
    m.close()


CASES = [
    ("faces, 145 patches", [], dict()),
    ("577 patches, north 30", ["--sky-density", "2", "--north", "30"], dict(mf=2, north=30.0)),
    ("vertices, both sides, offset", ["--on", "vertices", "--sides", "both", "--offset", "0.1"], dict(onVertices=True, sides="both", offset=0.1)),
]


@pytest.mark.parametrize("name,flags,kw", CASES, ids=[c[0] for c in CASES])
def test_ladybug_columns_vs_oracle(native, box, orc, box_meshes, name, flags, kw):
    print("case: box, --mode ladybug, %s" % name, flush=True)
    r = native.run("daylight", ["--clear-sky"] + BERLIN + ["--tz", "1", "--mode", "ladybug", "--quantity", ",".join(QS)] + flags, box)
    o = orc.daylight(box_meshes, clear_weather(orc), quantities=QS, ladybug=True, **kw)["values"]
    assert r.values.shape == o.shape
    nS = 2 if kw.get("sides") == "both" else 1
    for qi, q in enumerate(QS):
        for s in range(nS):
            c = qi * nS + s
            if q == "sunhours":
                assert np.array_equal(r.values[:, c], o[:, c]), (q, s)
            else:
                assert rel_err(r.values[:, c], o[:, c]) <= 2e-6, (q, s, rel_err(r.values[:, c], o[:, c]))
    # the sun is inside the patches: radiation = direct + diffuse, and physical mode differs
    rad, dr, df = r.values[:, 0:nS], r.values[:, nS:2 * nS], r.values[:, 2 * nS:3 * nS]
    assert np.allclose(rad, dr + df, rtol=1e-6, atol=1e-9)


def test_dump_sky_vs_oracle(native, box, orc, tmp_path):
    print("case: box, --mode ladybug --dump-sky", flush=True)
    path = tmp_path / "sky.txt"
    native.run("daylight", ["--clear-sky"] + BERLIN + ["--tz", "1", "--mode", "ladybug", "--dump-sky", str(path)], box)
    w = clear_weather(orc)
    st = orc.make_steps(dni=w["dni"], dhi=w["dhi"])
    jd = np.array([s.jd for s in st], np.int32)
    hour = np.array([s.hour for s in st])
    dn = np.array([s.dni for s in st], np.float32)
    dh = np.array([s.dhi for s in st], np.float32)
    o = orc.ladybug_sky(1, jd, hour, dn, dh, 52.52, 13.41, 1.0, 0.2, float(sum(s.dt for s in st)))
    lines = path.read_text().splitlines()
    want = []
    for b, head in enumerate(("-s", "-d")):
        want.append("# gendaymtx %s -A ascii rows (ground row first), %%.3g" % head)
        want += ["%.3g %.3g %.3g" % tuple(float(x) for x in row) for row in o["dump"][b]]
    assert lines == want


def test_dump_sun_radiance_columns(native, box, orc, tmp_path):
    print("case: box, --dump-sun", flush=True)
    path = tmp_path / "sun.txt"
    native.run("daylight", ["--clear-sky"] + BERLIN + ["--tz", "1", "--from", "06-21T00:00", "--to", "06-22T00:00", "--dump-sun", str(path)], box)
    rows = np.loadtxt(path)
    assert rows.shape == (24, 12)
    import ctypes as C
    L = orc.lib()
    for r in rows:
        a, z = C.c_double(), C.c_double()
        L.orc_sun_radiance(52.52, -13.41, -15.0, 172, r[3], C.byref(a), C.byref(z))
        assert abs(r[7] - np.degrees(a.value)) < 1e-5 and abs(r[8] - np.degrees(z.value)) < 1e-5
        assert int(r[9]) == int(a.value > -0.2665 * np.pi / 180)
    w = clear_weather(orc)   # the weather columns stay last, as the reference's dump has them
    assert np.allclose(rows[:, 10], w["dni"][24 * 171:24 * 172], atol=5e-5) and np.allclose(rows[:, 11], w["dhi"][24 * 171:24 * 172], atol=5e-5)


def test_ladybug_series_vs_oracle_fp64(native, box, orc, box_meshes, tmp_path, monkeypatch):
    print("case: box, --mode ladybug --timeseries fp32", flush=True)
    monkeypatch.setenv("DAYLIGHT_SERIES", "fp32")
    path = str(tmp_path / "series.npy")
    r = native.run("daylight", ["--clear-sky"] + BERLIN + ["--tz", "1", "--mode", "ladybug", "--quantity", "radiation", "--from", "06-21T00:00",
                                                         "--to", "06-23T00:00", "--timeseries", path], box)
    E = np.load(path)
    t0 = 24 * 171
    o = orc.daylight(box_meshes, clear_weather(orc), quantities=("radiation",), t0=t0, t1=t0 + 48, series=True, ladybug=True)
    assert E.shape == o["series"].shape
    assert rel_err(E.astype(np.float64), o["series"]) <= 2e-6
    assert np.all(np.isfinite(r.values))
