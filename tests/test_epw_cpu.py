"""The EPW reader (what a --epw run reads) against an independent restatement of its rule: columns 14 and 15 (DNI, DHI)
of every data row read as C's atof reads a field, 9999 and above a missing value (0); the LOCATION line's latitude,
longitude and time zone (reference src/solar.cpp read_epw, sky.cpp's atof use). No GPU involved."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_NUM = re.compile(r"[ \t]*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")


def atof(field):
    """C atof: the longest leading decimal of the field after blanks, else 0 (a double, then the reader's float)."""
    m = _NUM.match(field)
    if not m:
        return 0.0
    s = m.group(1)
    if s.startswith("+") and len(s) > 1 and s[1] in "+-":
        return 0.0
    return float(s)


FORMATS = ["%.1f", "%.3f", "%d", "+%.2f", " %.2f", "%.6e", "%.17g", "%.20f", "%.2f ", "%.0f."]


def write(path, rows, crlf=False, trailing_comma=False):
    head = ["LOCATION,Test,-,DEU,test,000000,48.15,11.58,1.0,520.0", "DESIGN CONDITIONS,0", "TYPICAL/EXTREME PERIODS,0",
            "GROUND TEMPERATURES,0", "HOLIDAYS/DAYLIGHT SAVINGS,No,0,0,0", "COMMENTS 1,", "COMMENTS 2,",
            "DATA PERIODS,1,1,Data,Sunday, 1/ 1,12/31"]
    nl = "\r\n" if crlf else "\n"
    with open(path, "w", newline="") as f:
        f.write(nl.join(head + [",".join(r) + ("," if trailing_comma else "") for r in rows]) + nl)


def rows_of(rng, n=8760):
    rows, want = [], []
    for k in range(n):
        vals = []
        for c in range(2):
            fmt = FORMATS[rng.integers(len(FORMATS))]
            x = float(rng.uniform(0, 1000)) * (10.0 ** rng.integers(-3, 2))
            if rng.uniform() < 0.01:
                x = 9999.0
            s = (fmt % x) if "d" not in fmt else (fmt % int(x))
            if rng.uniform() < 0.005:
                s = "12abc"
            vals.append(s)
        r = ["1999", "1", "1", "1", "0", "?", "10.0", "5.0", "70", "101325", "0", "0", "300", "0", vals[0], vals[1]] + ["0"] * 19
        rows.append(r)
        w = [atof(v) for v in vals]
        w = [0.0 if x >= 9999 else x for x in w]
        want.append(w)
    return rows, np.array(want, np.float64).astype(np.float32)


@pytest.mark.parametrize("crlf,trailing", [(False, False), (True, False), (False, True)])
def test_values_bit_exact(tmp_path, crlf, trailing):
    import native
    rng = np.random.default_rng(7 + crlf + 2 * trailing)
    rows, want = rows_of(rng)
    p = tmp_path / "w.epw"
    write(p, rows, crlf, trailing)
    dni, dhi, loc = native.read_epw(p)
    assert loc == (48.15, 11.58, 1.0)
    assert np.array_equal(dni.view(np.uint32), want[:, 0].view(np.uint32))
    assert np.array_equal(dhi.view(np.uint32), want[:, 1].view(np.uint32))


def test_short_rows_and_counts(tmp_path):
    import native
    rng = np.random.default_rng(1)
    rows, _ = rows_of(rng)
    p = tmp_path / "short.epw"
    write(p, rows[:100])
    with pytest.raises(native.DaylightError, match="100 data rows, expected 8760"):
        native.read_epw(p)
    bad = [r[:15] for r in rows]   # 15 fields: no DHI column
    write(p, bad)
    with pytest.raises(native.DaylightError, match="short data row 1"):
        native.read_epw(p)
    with pytest.raises(native.DaylightError, match="cannot open EPW"):
        native.read_epw(tmp_path / "none.epw")
