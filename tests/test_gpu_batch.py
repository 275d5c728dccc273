"""One launch per kernel over a run's sensor meshes, as a replayed HIP graph (dlg_chunk), on the GPU.

- Every replay of a chunk's graph gives the eager path's values and series bit for bit (DAYLIGHT_GRAPH=off runs the
  same kernels one by one), with the program's own work and waits between the replays: three runs of one graph, then a
  run in many chunks (one capture for the full chunks, one for the tail) against the one-chunk run.
- Meshes of uneven size in one run give the rows each gives alone.
- Every tile shape of the MFMA contraction (64 WM sensors x 128 steps per block, tile-group order) gives the same bits,
  at every split method: each element's products are summed in the same order whatever the shape; the tuned run too.
"""
import os

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    import native as n
    return n


@pytest.fixture(scope="module")
def block(native, tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("block"))
    epw, _, _ = W.weather_epw(native, tmp)
    boxes, central = W.city()
    full = W.sensor_mesh(boxes[central])
    meshes = [native.SimulationMesh("a", *W.submesh(full, 0, 37)), native.SimulationMesh("b", *W.submesh(full, 37, 1500)),
              native.SimulationMesh("c", *W.submesh(full, 1500, 1700))]
    ctx = native.SimulationMesh("city", *W.box_mesh(boxes))
    return epw, meshes, ctx, tmp


def run_pair(native, args, meshes, ctx, tmp, tag):
    path = os.path.join(tmp, tag + ".npy")
    r = native.run("daylight", args + ["--timeseries", path], meshes, context=ctx)
    return r.values, np.load(path)


def test_graph_replays_equal_eager(native, block, monkeypatch):
    epw, meshes, ctx, tmp = block
    args = ["--epw", epw, "--sky-density", "2", "--quantity", "radiation,sunhours,illuminance", "--no-self-occlude"]
    monkeypatch.setenv("DAYLIGHT_GRAPH", "off")
    print("case: eager chunk, 1700 sensors", flush=True)
    v0, e0 = run_pair(native, args, meshes, ctx, tmp, "eager")
    monkeypatch.delenv("DAYLIGHT_GRAPH")
    c0 = native.gpu_counters()
    for k in range(3):
        print("case: graph run %d" % k, flush=True)
        v, e = run_pair(native, args, meshes, ctx, tmp, "graph%d" % k)
        assert np.array_equal(v, v0) and np.array_equal(e, e0), k
    c1 = native.gpu_counters()
    assert c1["graph_launches"] - c0["graph_launches"] == 3
    assert c1["graph_captures"] - c0["graph_captures"] <= 1
    # many chunks: a full chunk's graph replayed, the tail captured once
    native.chunk_bytes(400 * 1024)
    try:
        print("case: graph runs in chunks of 400 KiB", flush=True)
        for k in range(2):
            v, e = run_pair(native, args, meshes, ctx, tmp, "chunks%d" % k)
            assert np.array_equal(v, v0) and np.array_equal(e, e0), k
    finally:
        native.chunk_bytes(0)
    c2 = native.gpu_counters()
    print("graph launches %d, captures %d" % (c2["graph_launches"] - c1["graph_launches"], c2["graph_captures"] - c1["graph_captures"]))
    assert c2["graph_launches"] - c1["graph_launches"] >= 6


def test_uneven_meshes_in_one_run(native, block):
    epw, meshes, ctx, tmp = block
    args = ["--epw", epw, "--sky-density", "1", "--quantity", "radiation,direct", "--no-self-occlude"]
    together, et = run_pair(native, args, meshes, ctx, tmp, "together")
    at = 0
    for m in meshes:
        alone, ea = run_pair(native, args, [m], ctx, tmp, "alone")
        n = alone.shape[0]
        assert np.array_equal(together[at:at + n], alone) and np.array_equal(et[at:at + n], ea), m.names
        at += n
    assert at == together.shape[0]


SHAPES = ["2x4", "2x8", "2x16", "4x2", "4x4", "4x8"]


@pytest.mark.parametrize("method", ["bf16x3", "bf16x6", "fp16x3"])
def test_tile_shapes_bit_identical(native, block, monkeypatch, method):
    epw, meshes, ctx, tmp = block
    args = ["--epw", epw, "--sky-density", "1", "--quantity", "radiation", "--no-self-occlude"]
    monkeypatch.setenv("DAYLIGHT_SERIES", method)
    ref = None
    for shape in SHAPES:
        monkeypatch.setenv("DAYLIGHT_SERIES_TILE", shape)
        print("case: series %s, tile shape %s, 1700 sensors" % (method, shape), flush=True)
        v, e = run_pair(native, args, meshes, ctx, tmp, "shape")
        if ref is None:
            ref = (v, e)
        assert np.array_equal(v, ref[0]) and np.array_equal(e, ref[1]), shape
    monkeypatch.delenv("DAYLIGHT_SERIES_TILE")
    print("case: series %s, tuned shape" % method, flush=True)
    v, e = run_pair(native, args, meshes, ctx, tmp, "tuned")
    assert np.array_equal(e, ref[1])
