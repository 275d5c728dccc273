"""The daylight tool end to end through the C ABI (include/daylight_c.h, the reference's daylight_c.h contract) on the
GPU: against the outputs the reference prints in its docs, and against the oracle's run of the same flags.

Bars: sun hours exact; radiation and the other energy columns within 2e-6 relative to the oracle (scale max(|x|,
1e-3 max|x|), the validation scale rule of the reference's validate/README.md:160); the series per contraction method
within its stated bar of the oracle's FP64 series.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOX = os.path.join(ROOT, "tests", "golden", "box.obj")
BERLIN = ["--lat", "52.52", "--lon", "13.41"]


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


@pytest.fixture(scope="module")
def box(native):
    m = native.SimulationMesh.load(BOX)
    yield m
    m.close()


def rel_err(got, want):
    scale = np.maximum(np.abs(want), 1e-3 * np.abs(want).max() if want.size else 1.0)
    scale = np.where(scale > 0, scale, 1.0)
    return float(np.max(np.abs(got - want) / scale)) if want.size else 0.0


def clear_weather(orc, tz=1.0):
    dni, dhi = orc.clear_sky(52.52, 13.41, tz)
    return dict(lat=52.52, lon=13.41, tz=tz, leap=False, dni=dni, dhi=dhi)


def test_box_clear_sky_year_golden(native, box, known, orc, box_meshes):
    """docs/daylight.md:133-150: radiation to the printed digits, sun hours exactly; and the oracle's run."""
    print("case: box, --clear-sky year, radiation,sunhours", flush=True)
    r = native.run("daylight", ["--clear-sky"] + BERLIN + ["--quantity", "radiation,sunhours"], box)
    assert r.manifest["columns"] == ["radiation_front", "sunhours_front"]
    gold = np.array(known["box_berlin_radiation_sunhours"])
    v = r.values
    assert np.array_equal(v[:, 1], gold[:, 1])
    assert np.all(np.abs(v[:, 0] - gold[:, 0]) <= 5e-10 * np.abs(gold[:, 0]) + 1e-9)
    o = orc.daylight(box_meshes, clear_weather(orc), quantities=("radiation", "sunhours"))["values"]
    assert np.array_equal(v[:, 1], o[:, 1])
    assert rel_err(v[:, 0], o[:, 0]) <= 2e-6
    assert r.stats["rays"] == 16 * 4746


def test_box_sunhours_june21_golden(native, box, known):
    """docs/daylight.md:163-180: 21 June in 10-minute steps, no weather; the printed values exactly."""
    print("case: box, sunhours 21 June, 10 min steps", flush=True)
    r = native.run("daylight", ["--quantity", "sunhours"] + BERLIN + ["--tz", "1", "--from", "06-21T00:00", "--to", "06-22T00:00",
                                                                     "--step", "10"], box)
    assert ["%.10g" % x for x in r.values[:, 0]] == ["%.10g" % x[0] for x in known["box_sunhours_june21"]]


OPTION_SETS = [
    ("sides both", ["--sides", "both"], dict(sides="both")),
    ("on vertices", ["--on", "vertices"], dict(onVertices=True)),
    ("cull backfaces", ["--cull-backfaces"], dict(cullBackfaces=True)),
    ("north 30", ["--north", "30"], dict(north=30.0)),
    ("offset 0.05", ["--offset", "0.05"], dict(offset=0.05)),
    ("no self occlusion", ["--no-self-occlude"], dict(selfOcclude=False)),
    ("577 patches", ["--sky-density", "2"], dict(mf=2)),
    ("4 samples per patch", ["--patch-samples", "4"], dict(patchSamples=4)),
    ("ground 0.5, no virtual ground", ["--ground", "0.5", "--virtual-ground", "off"], dict(ground=0.5, virtualGround=False)),
]


@pytest.mark.parametrize("name,flags,kw", OPTION_SETS, ids=[o[0] for o in OPTION_SETS])
def test_options_vs_oracle(native, box, orc, box_meshes, name, flags, kw):
    """Every quantity under one option set against the oracle's run of the same options."""
    print("case: box, %s" % name, flush=True)
    qs = ("radiation", "direct", "diffuse", "irradiance", "illuminance", "sunhours")
    r = native.run("daylight", ["--clear-sky"] + BERLIN + ["--tz", "1", "--quantity", ",".join(qs)] + flags, box)
    o = orc.daylight(box_meshes, clear_weather(orc), quantities=qs, **kw)["values"]
    assert r.values.shape == o.shape
    nS = 2 if kw.get("sides") == "both" else 1
    for qi, q in enumerate(qs):
        for s in range(nS):
            c = qi * nS + s
            if q == "sunhours":
                assert np.array_equal(r.values[:, c], o[:, c]), (q, s)
            else:
                assert rel_err(r.values[:, c], o[:, c]) <= 2e-6, (q, s, rel_err(r.values[:, c], o[:, c]))


SERIES_BARS = {"fp32": 2e-6, "bf16x3": 2e-5, "bf16x6": 2e-6, "fp16x3": 2e-6}


@pytest.mark.parametrize("method", sorted(SERIES_BARS))
def test_series_file_vs_oracle_fp64(native, box, orc, box_meshes, tmp_path, method, monkeypatch):
    """--timeseries (two days, both sides) per contraction method against the oracle's FP64 series; the series'
    step sums are the cumulative radiation (check_gpu.cpp:981-995)."""
    print("case: box, --timeseries, %s" % method, flush=True)
    monkeypatch.setenv("DAYLIGHT_SERIES", method)
    path = str(tmp_path / "series.npy")
    r = native.run("daylight", ["--clear-sky"] + BERLIN + ["--tz", "1", "--quantity", "radiation", "--from", "06-21T00:00",
                                                         "--to", "06-23T00:00", "--sides", "both", "--timeseries", path], box)
    E = np.load(path)
    t0 = 24 * 171
    o = orc.daylight(box_meshes, clear_weather(orc), quantities=("radiation",), t0=t0, t1=t0 + 48, series=True, sides="both")
    assert E.shape == o["series"].shape
    err = rel_err(E.astype(np.float64), o["series"])
    print("series %s: max rel error vs FP64 %.3e" % (method, err))
    assert err <= SERIES_BARS[method]
    hoys = np.loadtxt(path + ".hoys.txt")
    assert len(hoys) == E.shape[1]
    assert np.allclose(E.astype(np.float64).sum(axis=1) / 1000.0, r.values.reshape(-1), rtol=1e-5)


def test_series_device_sink_matches_file(native, box, tmp_path):
    """dlx_run_series_device hands the same series chunk by chunk in device memory."""
    print("case: box, device series sink", flush=True)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    path = str(tmp_path / "series.npy")
    flags = ["--clear-sky"] + BERLIN + ["--tz", "1", "--quantity", "radiation", "--from", "03-01T00:00", "--to", "03-08T00:00",
                                        "--timeseries", path]
    native.run("daylight", flags, box)
    E = np.load(path)
    got = np.full(E.shape, np.nan, np.float32)
    calls = []

    def sink(row0, rows, cols, ptr):
        calls.append((row0, rows, cols))
        assert ptr and cols == E.shape[1]
        buf = np.empty((rows, cols), np.float32)
        assert hip.hipMemcpy(buf.ctypes.data, C.c_void_p(ptr), buf.nbytes, 2) == 0
        got[row0:row0 + rows] = buf
        return 0

    r = native.run_series_device(flags, box, sink=sink)
    assert calls and sum(c[1] for c in calls) == E.shape[0]
    assert np.array_equal(got, E)
    assert r.values.shape == (16, 1)


def test_scene_resident_across_runs(native, box):
    """A second run over the same mesh handles reuses the resident scene and gives the same values."""
    print("case: box, resident scene", flush=True)
    flags = ["--quantity", "sunhours"] + BERLIN + ["--tz", "1", "--from", "06-21T00:00", "--to", "06-22T00:00"]
    a = native.run("daylight", flags, box)
    b = native.run("daylight", flags, box)
    assert b.stats["scene_reused"]
    assert np.array_equal(a.values, b.values)


def test_progress_and_cancel(native, box):
    print("case: box, control", flush=True)
    flags = ["--quantity", "sunhours"] + BERLIN + ["--tz", "1", "--from", "06-21T00:00", "--to", "06-22T00:00"]
    c = native.Control()
    native.run("daylight", flags, box, control=c)
    assert c.progress == 1.0
    c2 = native.Control()
    c2.cancel()
    with pytest.raises(native.DaylightError) as e:
        native.run("daylight", flags, box, control=c2)
    assert e.value.status == 2 and "cancelled" in str(e.value)


def write_epw(path, dni, dhi, lat=52.52, lon=13.41, tz=1.0):
    rows = ["LOCATION,Test,-,DEU,test,000000,%.2f,%.2f,%.1f,0.0" % (lat, lon, tz), "DESIGN CONDITIONS,0",
            "TYPICAL/EXTREME PERIODS,0", "GROUND TEMPERATURES,0", "HOLIDAYS/DAYLIGHT SAVINGS,No,0,0,0", "COMMENTS 1,",
            "COMMENTS 2,", "DATA PERIODS,1,1,Data,Sunday, 1/ 1,12/31"]
    r = 0
    for m, nd in enumerate([31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31], 1):
        for d in range(1, nd + 1):
            for h in range(1, 25):
                f = [1999, m, d, h, 0, "?", 10.0, 5.0, 70, 101325, 0, 0, 300, "%.1f" % (dni[r] + dhi[r]), "%.1f" % dni[r],
                     "%.1f" % dhi[r]] + [0] * 19
                rows.append(",".join(str(x) for x in f))
                r += 1
    with open(path, "w") as fh:
        fh.write("\n".join(rows) + "\r\n")


def test_epw_year_vs_oracle(native, box, orc, box_meshes, tmp_path):
    """--epw: a cloudy year written as an EPW, read by the library's reader, against the oracle given the same numbers
    (parsed here from the same text)."""
    print("case: box, --epw cloudy year", flush=True)
    dni, dhi = orc.clear_sky(52.52, 13.41, 1.0)
    rng = np.random.default_rng(3)
    c = rng.beta(2, 3, 8760).astype(np.float32)
    dni, dhi = dni * (1 - c), dhi + 0.5 * c * dni
    path = str(tmp_path / "cloudy.epw")
    write_epw(path, dni, dhi)
    back = np.array([[float(x) for x in l.split(",")[14:16]] for l in open(path).read().splitlines()[8:]], np.float32)
    w = dict(lat=52.52, lon=13.41, tz=1.0, leap=False, dni=back[:, 0].copy(), dhi=back[:, 1].copy())
    qs = ("radiation", "illuminance", "sunhours")
    r = native.run("daylight", ["--epw", path, "--quantity", ",".join(qs)], box)
    o = orc.daylight(box_meshes, w, quantities=qs)["values"]
    assert np.array_equal(r.values[:, 2], o[:, 2])
    for c_ in (0, 1):
        assert rel_err(r.values[:, c_], o[:, c_]) <= 2e-6


def test_large_grid_vs_oracle(native, box, orc, box_meshes):
    """A sensor mesh large enough to be made on several host threads (9,600 quads; both sides, offset) over the box
    scene as context: every value against the oracle, sun hours exactly."""
    print("case: 120 x 80 grid over the box scene, both sides", flush=True)
    nx, ny = 120, 80
    xs, ys = np.meshgrid(np.linspace(-3, 5, nx + 1), np.linspace(-3, 5, ny + 1), indexing="ij")
    pts = np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, 0.25)], 1)
    i = np.arange(nx)[:, None] * (ny + 1) + np.arange(ny)[None, :]
    quads = np.stack([i, i + ny + 1, i + ny + 2, i + 1], -1).reshape(-1, 4).astype(np.int32)
    counts = np.full(len(quads), 4, np.int32)
    grid = native.SimulationMesh("grid", pts, counts, quads.ravel())
    qs = ("radiation", "sunhours")
    r = native.run("daylight", ["--clear-sky"] + BERLIN + ["--tz", "1", "--quantity", ",".join(qs), "--sides", "both", "--offset",
                                                         "0.01"], grid, context=box)
    g = orc.MeshArrays("grid", pts, counts, quads.ravel(), sensor=True)
    ctx = [orc.MeshArrays(m.name, m.points, m.counts, m.indices, sensor=False) for m in box_meshes]
    o = orc.daylight([g] + ctx, clear_weather(orc), quantities=qs, sides="both", offset=0.01)["values"]
    assert r.values.shape == o.shape == (nx * ny, 4)
    assert np.array_equal(r.values[:, 2:], o[:, 2:])
    assert rel_err(r.values[:, :2], o[:, :2]) <= 2e-6


def test_chunking_invariance(native, box, tmp_path):
    """check_gpu.cpp:778-797: the chunk size changes no bit. The 9,600-quad grid (both sides, 19,200 sensors) in ~11
    chunks against one; the box's series handed to the device sink two sensors at a time against the default."""
    print("case: chunking invariance (grid, series sink)", flush=True)
    nx, ny = 120, 80
    xs, ys = np.meshgrid(np.linspace(-3, 5, nx + 1), np.linspace(-3, 5, ny + 1), indexing="ij")
    pts = np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, 0.25)], 1)
    i = np.arange(nx)[:, None] * (ny + 1) + np.arange(ny)[None, :]
    quads = np.stack([i, i + ny + 1, i + ny + 2, i + 1], -1).reshape(-1, 4).astype(np.int32)
    grid = native.SimulationMesh("grid", pts, np.full(len(quads), 4, np.int32), quads.ravel())
    flags = ["--clear-sky"] + BERLIN + ["--tz", "1", "--quantity", "radiation,irradiance,sunhours", "--sides", "both"]
    one = native.run("daylight", flags, grid, context=box)
    sflags = ["--clear-sky"] + BERLIN + ["--tz", "1", "--quantity", "radiation", "--from", "03-01T00:00", "--to", "03-08T00:00",
                                         "--timeseries", str(tmp_path / "s.npy")]
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def series(calls):
        out = {}

        def sink(row0, rows, cols, ptr):
            calls.append(rows)
            buf = np.empty((rows, cols), np.float32)
            assert hip.hipMemcpy(buf.ctypes.data, C.c_void_p(ptr), buf.nbytes, 2) == 0
            out[row0] = buf
            return 0

        native.run_series_device(sflags, box, sink=sink)
        return np.concatenate([out[k] for k in sorted(out)])

    calls1, callsN = [], []
    e1 = series(calls1)
    try:
        native.chunk_bytes(1 << 20)
        many = native.run("daylight", flags, grid, context=box)
        native.chunk_bytes(e1.shape[1] * 4 * 2)   # two sensors of series per chunk
        eN = series(callsN)
    finally:
        native.chunk_bytes(0)
    assert np.array_equal(one.values, many.values)
    assert len(calls1) == 1 and len(callsN) == 8 and set(callsN) == {2}
    assert np.array_equal(e1, eN)


def test_runs_in_turn_keep_their_own_sensors(native, box):
    """Sensors kept between runs over the same mesh handles follow each run's own sensor options (offset, sides,
    vertices), also when views runs sit in between: every run gives the values it gives alone."""
    print("case: sensors kept between runs", flush=True)
    base = ["--quantity", "sunhours"] + BERLIN + ["--tz", "1", "--from", "06-21T00:00", "--to", "06-22T00:00"]
    variants = [[], ["--offset", "0.5"], ["--sides", "both"], ["--on", "vertices"], []]
    alone = {}
    for v in variants[:4]:
        alone[tuple(v)] = native.run("daylight", base + v, box).values
    native.run("views", [], box)
    for v in variants:
        got = native.run("daylight", base + v, box).values
        assert np.array_equal(got, alone[tuple(v)]), v
        native.run("views", ["--offset", "0.3"], box)
    # the variants differ, so a run handed another run's sensors would show
    assert not np.array_equal(alone[()], alone[("--offset", "0.5")])
    assert alone[()].shape != alone[("--sides", "both")].shape and alone[()].shape != alone[("--on", "vertices")].shape
