"""The serving loop (dlx_server_*) on the GPU: every request's result is the one it gets alone through dl_run.

- A mix of daylight requests over one context (merged into shared runs), self-occluding daylight requests (served
  alone), views requests and a second resident context, submitted at once to a server of three lanes: values and
  manifests equal the alone runs', bit for bit; merged runs happened.
- A budget of one byte admits one run at a time (the device otherwise idle) with the same results.
- Flag errors surface at submit; a request waited for twice is refused.
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


@pytest.fixture(scope="module")
def world(native, tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("server"))
    epw, _, _ = W.weather_epw(native, tmp)
    boxes, central = W.city()
    full = W.sensor_mesh(boxes[central])
    cuts = [0, 37, 400, 1500, 1540, 3000, 4100]
    parts = [native.SimulationMesh("part%d" % k, *W.submesh(full, cuts[k], cuts[k + 1])) for k in range(len(cuts) - 1)]
    ctx = native.SimulationMesh("city", *W.box_mesh(boxes))
    far = native.SimulationMesh("far", *W.box_mesh(boxes[:50]))   # a second context scene
    grid = native.SimulationMesh("grid", *W.submesh(full, 5000, 5200))
    return epw, parts, ctx, far, grid


def requests(world):
    epw, parts, ctx, far, grid = world
    day = ["--epw", epw, "--sky-density", "1", "--quantity", "radiation,sunhours", "--no-self-occlude"]
    self_occ = ["--epw", epw, "--sky-density", "1", "--quantity", "radiation"]
    views = ["--fov", "360", "--cell", "6", "--far", "500", "--jitter", "on", "--distance", "isovist,mean"]
    reqs = [("daylight", day, [p], [ctx]) for p in parts]
    reqs += [("daylight", self_occ, [parts[1]], [ctx]), ("views", views, [grid], [ctx]), ("daylight", day, [parts[2]], [far]),
             ("views", views, [parts[0]], [far]), ("daylight", day, [parts[4], parts[3]], [ctx])]
    return reqs


def alone(native, reqs):
    return [native.run(t, a, s, c) for t, a, s, c in reqs]


def check_same(got, want):
    for k, ((r, times), w) in enumerate(zip(got, want)):
        assert np.array_equal(r.values, w.values, equal_nan=True), k
        assert r.manifest == w.manifest, k


def test_mix_equals_alone(native, world):
    reqs = requests(world)
    print("case: %d requests alone through dl_run" % len(reqs), flush=True)
    want = alone(native, reqs)
    srv = native.Server(lanes=3)
    try:
        print("case: the same requests submitted at once to 3 lanes", flush=True)
        ids = [srv.submit(t, a, s, c) for t, a, s, c in reqs]
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    print("server stats", st, "batches", [t["batch"] for _, t in got])
    check_same(got, want)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["merged"] >= 2 and st["runs"] < len(reqs)
    assert max(t["batch"] for _, t in got) >= 2


def test_budget_one_run_at_a_time(native, world):
    reqs = requests(world)[:5]
    want = alone(native, reqs)
    srv = native.Server(lanes=3, hbm_budget=1)
    try:
        print("case: budget 1 byte, 3 lanes", flush=True)
        ids = [srv.submit(t, a, s, c) for t, a, s, c in reqs]
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    check_same(got, want)
    assert st["peak_running"] == 1


def test_errors(native, world):
    epw, parts, ctx, far, grid = world
    srv = native.Server(lanes=1)
    try:
        with pytest.raises(native.DaylightError, match="bogus"):
            srv.submit("daylight", ["--epw", epw, "--bogus"], [parts[0]], [ctx])
        rid = srv.submit("daylight", ["--epw", epw, "--quantity", "radiation", "--no-self-occlude"], [parts[0]], [ctx])
        r, t = srv.wait(rid)
        assert r.values.shape == (37, 1) and t["batch"] == 1
        with pytest.raises(native.DaylightError, match="already waited"):
            srv.wait(rid)
    finally:
        srv.close()
