"""The oracle's daylight run end to end against the outputs the reference prints in its docs (known_answers.json)."""
import math

import numpy as np


def fmt10(x):
    return "%.10g" % x


def test_box_clear_sky_berlin_radiation_sunhours(orc, known, box_meshes):
    """docs/daylight.md:133-150 (the small-scene workload): radiation to the printed 10 digits, sun hours exact."""
    tz = round(13.41 / 15)
    dni, dhi = orc.clear_sky(52.52, 13.41, tz)
    r = orc.daylight(box_meshes, dict(lat=52.52, lon=13.41, tz=tz, leap=False, dni=dni, dhi=dhi), quantities=("radiation", "sunhours"))
    gold = np.array(known["box_berlin_radiation_sunhours"])
    assert r["nSun"] == 4456 and r["nDir"] == 4746 and r["durationHours"] == 8760
    v = r["values"]
    assert np.array_equal(v[:, 1], gold[:, 1])
    assert np.all(np.abs(v[:, 0] - gold[:, 0]) <= 1e-9 * np.abs(gold[:, 0]) + 1e-12)


def test_box_sunhours_one_day(orc, known, box_meshes):
    """docs/daylight.md:163-180: 21 June in 10-minute steps, no weather; printed values exactly."""
    t0 = 24 * (31 + 28 + 31 + 30 + 31 + 20)
    r = orc.daylight(box_meshes, None, quantities=("sunhours",), lat=52.52, lon=13.41, tz=1, t0=t0, t1=t0 + 24, stepMin=10)
    assert [fmt10(x) for x in r["values"][:, 0]] == [fmt10(x[0]) for x in known["box_sunhours_june21"]]


def test_series_sums_equal_cumulative(orc, box_meshes):
    """check_gpu.cpp:981-995: the FP64 series summed over its steps is the cumulative radiation (W h -> kWh)."""
    tz = 1
    dni, dhi = orc.clear_sky(52.52, 13.41, tz)
    w = dict(lat=52.52, lon=13.41, tz=tz, leap=False, dni=dni, dhi=dhi)
    t0 = 24 * 171
    r = orc.daylight(box_meshes, w, quantities=("radiation",), t0=t0, t1=t0 + 48, series=True, sides="both")
    total = r["series"].sum(axis=1) / 1000.0
    cum = r["values"].reshape(-1)
    assert np.allclose(total, cum, rtol=1e-5, atol=1e-9)
