"""The command-line tools' failures that need no GPU: usage errors (exit 1 with the usage text) and unreadable inputs
(exit 2), as the reference's main.cpp reports them."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kit-sim_daylight_amd", "lib")
BOX = os.path.join(ROOT, "tests", "golden", "box.obj")


def tool(name, *args):
    exe = os.path.join(LIB, name)
    if not os.path.exists(exe):
        pytest.fail("%s is not built (python kit-sim_daylight_amd/build.py)" % exe)
    return subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("name", ["daylight", "views"])
def test_help_and_usage_errors(name):
    r = tool(name, "-h")
    assert r.returncode == 0 and r.stdout.startswith("usage: " + name)
    r = tool(name)
    assert r.returncode == 1 and "error: no mesh given" in r.stderr and "usage: " + name in r.stderr
    r = tool(name, BOX, "--bogus")
    assert r.returncode == 1 and "unknown flag --bogus" in r.stderr


def test_unreadable_mesh_and_unknown_extension(tmp_path):
    r = tool("daylight", str(tmp_path / "missing.obj"), "--quantity", "sunhours", "--lat", "52.5", "--lon", "13.4")
    assert r.returncode == 2 and "error:" in r.stderr
    p = tmp_path / "mesh.xyz"
    p.write_text("")
    r = tool("views", str(p))
    assert r.returncode == 2 and "unknown extension" in r.stderr


def test_flag_rules_before_the_run():
    r = tool("daylight", BOX, "--clear-sky")
    assert r.returncode == 1 and "--clear-sky needs --lat and --lon" in r.stderr
    r = tool("views", BOX, "--fov", "360", "--cell", "0.01")
    assert r.returncode == 1 and "exceed 2^20" in r.stderr
