"""The command-line tools on the GPU: the reference docs' printed outputs (docs/daylight.md:133-150, docs/views.md:333-350),
the text and NumPy writers and the manifest beside them."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kit-sim_daylight_amd", "lib")
BOX = os.path.join(ROOT, "tests", "golden", "box.obj")


def tool(name, *args):
    print("case: %s %s" % (name, " ".join(args)), flush=True)
    return subprocess.run([os.path.join(LIB, name)] + list(args), capture_output=True, text=True, timeout=300)


def test_daylight_stdout_golden(known):
    r = tool("daylight", BOX, "--clear-sky", "--lat", "52.52", "--lon", "13.41", "--quantity", "radiation,sunhours")
    assert r.returncode == 0, r.stderr
    rows = [[float(x) for x in l.split()] for l in r.stdout.splitlines()]
    gold = np.array(known["box_berlin_radiation_sunhours"])
    v = np.array(rows)
    assert v.shape == gold.shape
    assert np.array_equal(v[:, 1], gold[:, 1])
    assert np.all(np.abs(v[:, 0] - gold[:, 0]) <= 5e-10 * gold[:, 0] + 1e-9)
    # "%.10g" text: every number printed with at most 10 significant digits
    assert all(len(x.replace(".", "").replace("-", "").lstrip("0")) <= 10 for l in r.stdout.splitlines() for x in l.split())


def test_daylight_npy_and_manifest(tmp_path):
    out = str(tmp_path / "r.npy")
    args = ["--clear-sky", "--lat", "52.52", "--lon", "13.41", "--quantity", "radiation,sunhours", "--sides", "both"]
    r = tool("daylight", BOX, *args, "--out", out)
    assert r.returncode == 0, r.stderr
    v = np.load(out)
    m = json.load(open(out + ".json"))
    assert v.dtype == np.float64 and v.shape == (16, 4)
    assert m["columns"] == ["radiation_front", "radiation_back", "sunhours_front", "sunhours_back"]
    txt = str(tmp_path / "r.txt")
    assert tool("daylight", BOX, *args, "--out", txt).returncode == 0
    t = np.loadtxt(txt)
    assert np.allclose(t, v, rtol=1e-9, atol=0)
    assert json.load(open(txt + ".json"))["columns"] == m["columns"]


def test_views_stdout_golden(known):
    r = tool("views", BOX)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines()]
    gold = known["box_views_defaults"]
    assert len(rows) == len(gold)
    for got, want in zip(rows, gold):
        assert got[:3] == ["%.10g" % x for x in want[:3]]
        assert all(abs(float(g) - w) <= 2e-7 * abs(w) for g, w in zip(got[3:], want[3:]))


def test_views_context_file_and_targets(tmp_path):
    """--context FILE meshes block but have no rows; --target names a context mesh."""
    obj = tmp_path / "ctx.obj"
    obj.write_text("o wall\nv 3 -2 0\nv 3 2 0\nv 3 2 3\nv 3 -2 3\nf 1 2 3 4\n")
    r = tool("views", BOX, "--context", str(obj), "--target", "wall=wall", "--out", str(tmp_path / "v.npy"))
    assert r.returncode == 0, r.stderr
    v = np.load(str(tmp_path / "v.npy"))
    m = json.load(open(str(tmp_path / "v.npy") + ".json"))
    cols = [k for k, c in enumerate(m["columns"]) if c.startswith("wall")]
    assert v.shape[0] == 16 and cols, m["columns"]
    assert v[:, cols].max() > 0
