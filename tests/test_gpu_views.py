"""The views tool end to end through the C ABI on the GPU: against the reference docs' printed box output and against
the oracle's run of the same flags, with the per-ray classes and distances (--dump-rays) bit for bit.

Bars: ray classes and distances identical to the oracle's; shares and reach shares exact; log-mean distances within
1e-12 relative (device log vs host log); isovist sums exact (the same float terms in the same wave64 tree).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOX = os.path.join(ROOT, "tests", "golden", "box.obj")


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


@pytest.fixture(scope="module")
def box(native):
    m = native.SimulationMesh.load(BOX)
    yield m
    m.close()


def compare(got, want, columns):
    assert got.shape == want.shape
    for c, name in enumerate(columns):
        g, w = got[:, c], want[:, c]
        assert np.array_equal(np.isnan(g), np.isnan(w)), name
        g, w = g[~np.isnan(g)], w[~np.isnan(w)]
        if "distance" in name:
            assert np.all(np.abs(g - w) <= 1e-12 * np.abs(w)), (name, np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-300)))
        else:
            assert np.array_equal(g, w), (name, np.max(np.abs(g - w)))


def test_box_defaults_golden_and_rays(native, box, orc, box_meshes, known, tmp_path):
    """docs/views.md:333-350 and the oracle: the values, then every ray's class and distance."""
    print("case: views box defaults", flush=True)
    dump = str(tmp_path / "rays.npy")
    r = native.run("views", ["--dump-rays", dump], box)
    gold = np.array(known["box_views_defaults"])
    v = r.values
    assert r.manifest["columns"] == ["sky_front", "ground_front", "obstruction_front", "distance_front", "distance_min_front"]
    assert [["%.10g" % x for x in row[:3]] for row in v] == [["%.10g" % x for x in row[:3]] for row in gold]
    assert np.all(np.abs(v[:, 3:] - gold[:, 3:]) <= 2e-7 * np.abs(gold[:, 3:]))
    o = orc.views(box_meshes, raw=True)
    compare(v, o["values"], o["columns"])
    cls = np.load(dump)
    dist = np.load(dump + ".depth.npy")
    assert cls.shape == o["cls"].shape == (16, 3600)
    assert np.array_equal(cls, o["cls"])
    assert np.array_equal(dist.view(np.uint32), o["dist"].view(np.uint32))


CASES = [
    ("360 isovist reach", ["--fov", "360", "--offset", "0.3", "--cell", "4", "--distance", "isovist,mean,min", "--reach", "en17037"],
     dict(fov=(360, -70, 50), cell=4.0, offset=0.3, distance=("isovist", "mean", "min"), reach=(6.0, 20.0, 50.0))),
    ("360 cull own surface", ["--fov", "360", "--cell", "5", "--distance", "isovist"],
     dict(fov=(360, -70, 50), cell=5.0, distance=("isovist",))),
    ("area plan", ["--fov", "360,0,0", "--cell", "3", "--offset", "0.2", "--distance", "isovist,min"],
     dict(fov=(360, 0, 0), cell=3.0, offset=0.2, distance=("isovist", "min"))),
    ("no jitter both sides", ["--jitter", "off", "--sides", "both", "--cell", "3"], dict(jitter=False, sides="both", cell=3.0)),
    ("target group", ["--target", "box=boxes", "--distance", "mean,min,group", "--cell", "3"],
     dict(targets={"box": "boxes"}, distance=("mean", "min", "group"), cell=3.0)),
    ("north 30 far 50", ["--north", "30", "--fov", "360", "--offset", "0.1", "--far", "50", "--cell", "4"],
     dict(north=30.0, fov=(360, -70, 50), offset=0.1, far=50.0, cell=4.0)),
    ("no virtual ground", ["--virtual-ground", "off", "--cell", "3"], dict(virtualGround=False, cell=3.0)),
    ("ground z -1", ["--ground-z", "-1", "--cell", "3"], dict(groundZ=-1.0, cell=3.0)),
    ("cull backfaces", ["--cull-backfaces", "--cell", "3"], dict(cullBackfaces=True, cell=3.0)),
    ("no self occlusion", ["--no-self-occlude", "--cell", "3"], dict(selfOcclude=False, cell=3.0)),
    ("on vertices", ["--on", "vertices", "--cell", "3", "--offset", "0.05"], dict(onVertices=True, cell=3.0, offset=0.05)),
    ("narrow fov back", ["--fov", "60,-10,30", "--cell", "1", "--sides", "back"], dict(fov=(60, -10, 30), cell=1.0, sides="back")),
]


@pytest.mark.parametrize("name,flags,kw", CASES, ids=[c[0] for c in CASES])
def test_views_vs_oracle(native, box, orc, box_meshes, tmp_path, name, flags, kw):
    print("case: views box, %s" % name, flush=True)
    dump = str(tmp_path / "rays.npy")
    r = native.run("views", flags + ["--dump-rays", dump], box)
    o = orc.views(box_meshes, raw=True, **kw)
    assert r.manifest["columns"] == o["columns"]
    assert np.array_equal(np.load(dump), o["cls"])
    assert np.array_equal(np.load(dump + ".depth.npy").view(np.uint32), o["dist"].view(np.uint32))
    compare(r.values, o["values"], o["columns"])


def test_views_chunking_invariance(native, box, tmp_path):
    """The chunk size changes no bit (views.cpp chunks by the same rule as daylight): the box's default views and the
    360-degree isovist case in chunks of about two sensors against one chunk, values and every ray."""
    print("case: views chunking invariance", flush=True)
    for flags in ([], ["--fov", "360", "--offset", "0.3", "--cell", "4", "--distance", "isovist,mean,min", "--reach", "en17037"]):
        a = native.run("views", flags + ["--dump-rays", str(tmp_path / "a.npy")], box)
        try:
            native.chunk_bytes(40000)
            b = native.run("views", flags + ["--dump-rays", str(tmp_path / "b.npy")], box)
        finally:
            native.chunk_bytes(0)
        assert np.array_equal(a.values, b.values, equal_nan=True)
        assert np.array_equal(np.load(tmp_path / "a.npy"), np.load(tmp_path / "b.npy"))
        assert np.array_equal(np.load(str(tmp_path / "a.npy") + ".depth.npy").view(np.uint32),
                              np.load(str(tmp_path / "b.npy") + ".depth.npy").view(np.uint32))


def test_views_sensor_blocks(native, box, tmp_path):
    """The rays' work items come in blocks of 512 sensors (tile-group major within a block, the last block partial):
    a 1,300-sensor grid around the box in one chunk (blocks 512 + 512 + 276) against chunks of about 100 and 700
    sensors (one partial block; a full and a partial one), values and every ray."""
    print("case: views sensor blocks (1,300-sensor grid, 360 degrees in 4-degree cells)", flush=True)
    nx, ny = 26, 50
    xs, ys = np.meshgrid(np.linspace(-3, 5, nx + 1), np.linspace(-3, 5, ny + 1), indexing="ij")
    pts = np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, 0.25)], 1)
    i = np.arange(nx)[:, None] * (ny + 1) + np.arange(ny)[None, :]
    quads = np.stack([i, i + ny + 1, i + ny + 2, i + 1], -1).reshape(-1, 4).astype(np.int32)
    grid = native.SimulationMesh("grid", pts, np.full(len(quads), 4, np.int32), quads.ravel())
    try:
        flags = ["--fov", "360", "--cell", "4", "--distance", "isovist,mean,min"]
        a = native.run("views", flags + ["--dump-rays", str(tmp_path / "a.npy")], grid, context=box)
        rays = np.load(tmp_path / "a.npy")
        assert a.values.shape[0] == nx * ny and rays.shape[0] == nx * ny
        per_sensor = rays[0].size * 5   # a class byte and a float distance per ray
        for n in (100, 700):
            try:
                native.chunk_bytes(per_sensor * n)
                b = native.run("views", flags + ["--dump-rays", str(tmp_path / "b.npy")], grid, context=box)
            finally:
                native.chunk_bytes(0)
            assert np.array_equal(a.values, b.values, equal_nan=True), n
            assert np.array_equal(rays, np.load(tmp_path / "b.npy")), n
            assert np.array_equal(np.load(str(tmp_path / "a.npy") + ".depth.npy").view(np.uint32),
                                  np.load(str(tmp_path / "b.npy") + ".depth.npy").view(np.uint32)), n
    finally:
        grid.close()
