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
