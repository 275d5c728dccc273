"""Sensor preparation on the GPU (dlg_sensors) against the oracle's make_sensors_into, bit for bit, on the GPU.

Cases: the urban block's roof and facade quads; a mesh of triangles, quads, pentagons, a concave hexagon, a degenerate
(zero-area) quad and a point no face uses; per face and per point; ladybug's face rules; front, back and both sides;
an offset along the normal; a recentring centre; several meshes in one reservation (each from its own first sensor).
Also the counts the host reports: degenerate faces and zero vertex normals.
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import daylight_amd
    g = daylight_amd.Gpu()
    yield g
    g.close()


def mixed_mesh(rng):
    """triangles, quads, pentagons, a concave hexagon, a zero-area quad and an unused point, jittered"""
    pts, counts, idx = [], [], []

    def face(corners):
        base = len(pts)
        pts.extend(corners)
        counts.append(len(corners))
        idx.extend(range(base, base + len(corners)))

    for k in range(200):
        o = rng.uniform(-50, 50, 3)
        n = [3, 4, 5][k % 3]
        ang = np.sort(rng.uniform(0, 2 * np.pi, n))
        r = rng.uniform(0.5, 3)
        tilt = rng.normal(size=3)
        u = np.cross(tilt, [0.3, 0.5, 0.8])
        u /= np.linalg.norm(u)
        v = np.cross(tilt / np.linalg.norm(tilt), u)
        face([o + r * (np.cos(a) * u + np.sin(a) * v) + rng.normal(scale=0.01, size=3) for a in ang])
    face([[0, 0, 0], [4, 0, 0], [4, 3, 0], [2, 1, 0], [0, 3, 0], [1, 1.5, 0]])   # concave
    face([[5, 5, 5], [5, 5, 5], [5, 5, 5], [5, 5, 5]])                           # zero area
    pts.append([9, 9, 9])                                                        # used by no face
    # shared corners for the vertex mode: the first 40 faces' corners welded onto a few points
    P = np.array(pts, np.float64)
    I = np.array(idx, np.int32)
    I[:40] = I[:40] % 7
    return P, np.array(counts, np.int32), I


@pytest.mark.parametrize("mode", ["faces", "points"])
@pytest.mark.parametrize("ladybug", [False, True])
def test_sensors_bit_exact(gpu, orc, mode, ladybug):
    rng = np.random.default_rng(3)
    boxes, central = W.city()
    block = W.sensor_mesh(boxes[central[:6]])
    mixed = mixed_mesh(rng)
    meshes = [block, mixed]
    if ladybug:   # ladybug mode takes tri / quad faces only
        keep = mixed[1] <= 4
        starts = np.concatenate([[0], np.cumsum(mixed[1])])
        idx = np.concatenate([mixed[2][starts[f]:starts[f + 1]] for f in np.nonzero(keep)[0]])
        meshes = [block, (mixed[0], mixed[1][keep], idx)]
    centre = (12.5, -3.25, 1.0)
    for sides in ("front", "back", "both"):
        for offset in (0.0, 0.1):
            print("case: sensors, %s, ladybug %s, sides %s, offset %g" % (mode, ladybug, sides, offset), flush=True)
            P, N, made, deg, zero = gpu.sensors(meshes, onVertices=mode == "points", sides=sides, offset=offset, centre=centre, ladybug=ladybug)
            at = 0
            for k, (p, c, i) in enumerate(meshes):
                m = orc.MeshArrays("m%d" % k, p, c, i)
                want = orc.sensors(m, ladybug=ladybug, onVertices=mode == "points", sides=sides, offset=offset, centre=centre)
                n = len(want[0])
                assert made[k] * (2 if sides == "both" else 1) == n
                assert np.array_equal(P[at:at + n], want[0]), (k, sides, offset)
                assert np.array_equal(N[at:at + n], want[1]), (k, sides, offset)
                at += n
            assert at == len(P)
            assert deg[0] == 0 and deg[1] >= 1      # the zero-area quad (and faces whose welded corners coincide)
            if mode == "points":
                assert zero[1] >= 1                  # the unused point (and the zero-area quad's)
            else:
                assert zero == [0, 0]


def test_run_uses_device_sensors(gpu, orc):
    """An end-to-end run's sensors are the device's: a run, and a second one over the same handles that keeps them."""
    import native
    boxes, central = W.city()
    s = native.SimulationMesh("s", *W.sensor_mesh(boxes[central[:3]]))
    ctx = native.SimulationMesh("c", *W.box_mesh(boxes))
    args = ["--clear-sky", "--lat", "52.52", "--lon", "13.41", "--tz", "1", "--quantity", "radiation", "--no-self-occlude", "--sides", "both",
            "--offset", "0.05"]
    a = native.run("daylight", args, s, ctx)
    b = native.run("daylight", args, s, ctx)
    assert np.array_equal(a.values, b.values) and b.stats["scene_reused"]
