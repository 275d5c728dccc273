"""The oracle's views run against the outputs the reference prints in its docs (docs/views.md:333-350) and against
what the geometry fixes."""
import math

import numpy as np


def test_box_views_defaults_golden(orc, known, box_meshes):
    """Shares (the ray classes, jitter and frames) to all 10 printed digits; distances within one float ulp of the
    reference's (its hardware triangle test and the watertight test may round a hit distance differently)."""
    r = orc.views(box_meshes)
    assert r["columns"] == ["sky_front", "ground_front", "obstruction_front", "distance_front", "distance_min_front"]
    gold = np.array(known["box_views_defaults"])
    v = r["values"]
    assert [["%.10g" % x for x in row[:3]] for row in v] == [["%.10g" % x for x in row[:3]] for row in gold]
    assert np.all(np.abs(v[:, 3:] - gold[:, 3:]) <= 2e-7 * np.abs(gold[:, 3:]))


def test_box_sides_sky_share_closed_form(orc, box_meshes):
    """A vertical face with nothing above the horizon sees sky over sin 50 / (sin 50 + sin 70) of its field
    (docs/views.md:355-357), whatever the jitter."""
    for jitter in (True, False):
        v = orc.views(box_meshes, jitter=jitter)["values"]
        want = math.sin(math.radians(50)) / (math.sin(math.radians(50)) + math.sin(math.radians(70)))
        assert np.allclose(v[8:, 0], want, rtol=1e-12, atol=0)
        assert np.allclose(v[:, :3].sum(axis=1), 1.0, rtol=1e-12)


def test_open_top_sees_far(orc, box_meshes):
    """The box's top looks at open sky only: share 1 and every distance at --far."""
    for far in (1000.0, 250.0):
        v = orc.views(box_meshes, far=far)["values"]
        assert np.array_equal(v[6:8], [[1, 0, 0, far, far]] * 2)


def test_isovist_volume_of_closed_room(orc):
    """Inside a closed cube of side s, seen all around from near its centre, the isovist volume is the cube's, s^3, up
    to the spherical grid's error. With --offset 0 the directions behind the sensor (below its upward normal) are its
    own surface (views.cpp:272, cull in 360 mode): then only the upper half, s^3 / 2, counts."""
    s = 4.0
    pts = np.array([[0, 0, 0], [s, 0, 0], [s, s, 0], [0, s, 0], [0, 0, s], [s, 0, s], [s, s, s], [0, s, s]], float)
    q = np.array([[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]])   # facing in
    room = orc.MeshArrays("room", pts, np.full(6, 4, np.int32), q.ravel().astype(np.int32), sensor=False)
    eye = orc.MeshArrays("eye", np.array([[1.9, 1.9, 2.0], [2.1, 1.9, 2.0], [2.1, 2.1, 2.0], [1.9, 2.1, 2.0]]), [4],
                         np.array([0, 1, 2, 3], np.int32))
    for offset, want in ((0.0, s ** 3 / 2), (1e-3, s ** 3)):
        r = orc.views([eye, room], fov=(360, -90, 90), cell=1.0, selfOcclude=False, jitter=False, virtualGround=False,
                      distance=("isovist", "mean", "min"), offset=offset)
        vol = r["values"][0, r["columns"].index("volume_front")]
        assert abs(vol - want) <= 0.01 * want, (offset, vol)
        dmin = r["values"][0, r["columns"].index("distance_min_front")]
        assert abs(dmin - (2.0 - offset)) <= 2e-3


def test_reach_shares_in_closed_room(orc):
    """Seen all around from the centre of a closed cube of side s, every ray ends between s/2 and s sqrt(3)/2 away:
    the share of rays reaching D is 1 below s/2 and 0 beyond s sqrt(3)/2, and falls in between."""
    s = 4.0
    pts = np.array([[0, 0, 0], [s, 0, 0], [s, s, 0], [0, s, 0], [0, 0, s], [s, 0, s], [s, s, s], [0, s, s]], float)
    q = np.array([[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]])
    room = orc.MeshArrays("room", pts, np.full(6, 4, np.int32), q.ravel().astype(np.int32), sensor=False)
    eye = orc.MeshArrays("eye", np.array([[1.9, 1.9, 2.0], [2.1, 1.9, 2.0], [2.1, 2.1, 2.0], [1.9, 2.1, 2.0]]), [4],
                         np.array([0, 1, 2, 3], np.int32))
    reach = (1.9, 2.5, 3.3, 3.5)
    r = orc.views([eye, room], fov=(360, -90, 90), cell=2.0, selfOcclude=False, virtualGround=False, offset=1e-3,
                  distance=("mean",), reach=reach)
    v = dict(zip(r["columns"], r["values"][0]))
    assert v["reach_1.9_front"] == 1.0
    assert 0.0 < v["reach_3.3_front"] < v["reach_2.5_front"] < 1.0
    assert v["reach_3.5_front"] == 0.0
    assert v["obstruction_front"] == 1.0 and v["sky_front"] == 0.0


def test_target_group_splits_obstruction(orc, box_meshes, known):
    """Naming the box a target moves exactly its rays out of obstruction: on the ground quads (which see only the box
    as an obstruction) the target share is the golden obstruction share and obstruction is 0."""
    r = orc.views(box_meshes, targets={"box": "boxes"}, distance=("mean", "min", "group"))
    assert r["columns"][:4] == ["sky_front", "ground_front", "obstruction_front", "boxes_front"]
    gold = np.array(known["box_views_defaults"])
    v = r["values"]
    assert ["%.10g" % x for x in v[:4, 3]] == ["%.10g" % x for x in gold[:4, 2]]
    assert np.all(v[:4, 2] == 0.0)
    assert ["%.10g" % x for x in v[:4, 0]] == ["%.10g" % x for x in gold[:4, 0]]
