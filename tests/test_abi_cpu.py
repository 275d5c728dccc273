"""The built library loads and exports every function include/*.h declares (no GPU call is made)."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kit-sim_daylight_amd")


def declared():
    names = []
    for h in ("daylight_c.h", "dl_device.h"):
        p = os.path.join(ROOT, "include", h)
        if not os.path.exists(p):
            continue
        txt = re.sub(r"/\*.*?\*/", "", open(p).read(), flags=re.S)
        names += re.findall(r"\b(dl[gx]?_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(n for n in names if not n.endswith("_t")))


def test_library_exports_every_declared_symbol():
    sys.path.insert(0, PKG)
    import daylight_amd
    if not os.path.exists(daylight_amd.LIB_PATH):
        pytest.fail("libdaylight_amd.so is not built (python kit-sim_daylight_amd/build.py)")
    names = declared()
    assert len(names) > 20
    missing = sorted(set(names) - set(daylight_amd.exported(names)))
    assert not missing, missing


def test_binding_declares_every_symbol():
    """Every entry point of include/*.h has a ctypes signature in the binding (a call never goes through untyped)."""
    sys.path.insert(0, PKG)
    import daylight_amd
    typed = {n for n, _, _ in daylight_amd.C_ABI + daylight_amd.SEAM}
    missing = sorted(set(declared()) - typed)
    assert not missing, missing


def _native():
    sys.path.insert(0, PKG)
    import native
    return native


def test_abi_version_usage_and_usage_errors():
    """dl_version / dl_usage, and flag errors that the CLI's parser reports before any GPU work (status 1, the CLI's
    message)."""
    native = _native()
    import daylight_amd
    assert daylight_amd.lib().dl_version() == 1
    assert "--quantity" in native.usage("daylight") and "--fov" in native.usage("views")
    box = native.SimulationMesh.load(os.path.join(ROOT, "tests", "golden", "box.obj"))
    assert box.names == ("ground", "box")
    for tool, flags, msg in (("daylight", ["--bogus"], "unknown flag --bogus"),
                             ("daylight", ["--step", "0"], "--step must be positive"),
                             ("daylight", ["--out", "x.txt"], "--out"),
                             ("daylight", ["--clear-sky", "--epw", "x.epw"], "exclusive"),
                             ("views", ["--fov", "200"], "--fov H must be"),
                             ("views", ["--distance", "bogus"], "--distance"),
                             ("views", ["--cell", "0.001"], "exceed 2^20")):
        with pytest.raises(native.DaylightError) as e:
            native.run(tool, flags, box)
        assert e.value.status == 1 and msg in str(e.value), (flags, str(e.value))


def test_mesh_create_rejects_bad_indices():
    native = _native()
    with pytest.raises(native.DaylightError) as e:
        native.SimulationMesh("bad", [[0, 0, 0], [1, 0, 0], [0, 1, 0]], [3], [0, 1, 7])
    assert e.value.status == 1


def test_epw_reader_errors(tmp_path):
    """read_epw's failures (reference solar.cpp / cli.cpp): a missing file, a short data row, a wrong row count; all
    reported before the run reaches the GPU."""
    native = _native()
    box = native.SimulationMesh.load(os.path.join(ROOT, "tests", "golden", "box.obj"))
    with pytest.raises(native.DaylightError) as e:
        native.run("daylight", ["--epw", str(tmp_path / "none.epw")], box)
    assert e.value.status == 2 and "cannot open EPW" in str(e.value)
    head = ["LOCATION,X,-,DEU,x,0,52.52,13.41,1.0,0"] + ["X,0"] * 7
    row = ",".join(["1999", "1", "1", "1", "0", "?"] + ["0"] * 29)
    p = tmp_path / "short.epw"
    p.write_text("\n".join(head + [row] * 10 + ["1,2,3"]) + "\n")
    with pytest.raises(native.DaylightError) as e:
        native.run("daylight", ["--epw", str(p)], box)
    assert e.value.status == 2 and "EPW: short data row 11" in str(e.value)
    p.write_text("\n".join(head + [row] * 100) + "\n")
    with pytest.raises(native.DaylightError) as e:
        native.run("daylight", ["--epw", str(p)], box)
    assert e.value.status == 2 and "EPW: 100 data rows, expected 8760" in str(e.value)
