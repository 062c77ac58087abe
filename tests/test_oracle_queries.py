"""Pins the oracle's brute-force ray and point queries to closed forms."""
import numpy as np
import pytest

from oracle import native as on


def unit_cube():
    v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    f = []
    for a, b, c, d in quads:
        f += [(a, b, c), (a, c, d)]
    return v, np.array(f, np.int32)


def test_closest_hit_unit_cube_closed_forms():
    v, f = unit_cube()
    o = np.array([[0.3, 0.4, 5.0], [0.3, 0.4, 0.5], [-2, 0.5, 0.5], [3, 3, 3], [0.25, 0.25, 0.5]],
                 np.float32)
    d = np.array([[0, 0, -1], [0, 0, 1], [1, 0, 0], [1, 0, 0], [0.6, 0, 0.8]], np.float32)
    face, t = on.closest_hit_brute(v, f, o, d)
    assert t[0] == pytest.approx(4.0) and face[0] in (2, 3)          # top z=1 quad
    assert t[1] == pytest.approx(0.5) and face[1] in (2, 3)          # from inside, two-sided
    assert t[2] == pytest.approx(2.0) and face[2] in (8, 9)          # x=0 side
    assert face[3] == -1 and t[3] == -1
    assert t[4] == pytest.approx(0.5 / 0.8)                          # exits the top first
    face, t = on.closest_hit_brute(v, f, o[:1], d[:1], tmax=3.9)
    assert face[0] == -1                                             # beyond tmax


def test_closest_point_plane_and_corner():
    v, f = unit_cube()
    p = np.array([[0.5, 0.5, 3.0], [2.0, 2.0, 2.0], [0.5, 0.5, 0.6]], np.float32)
    face, d = on.closest_point_brute(v, f, p)
    assert d[0] == pytest.approx(2.0)
    assert d[1] == pytest.approx(np.sqrt(3.0))
    assert d[2] == pytest.approx(0.4)                                 # inside: nearest wall


def test_closest_point_ties_go_to_lowest_face_id():
    """A point nearest to a shared vertex or edge is equally near several faces; the query
    returns the lowest face id among them (closed forms on the cube's faces)."""
    v, f = unit_cube()
    # outside a corner vertex (1, 1, 1): faces of the top, y=1 and x=1 quads touch it
    p = np.array([[1.5, 1.5, 1.5], [0.5, 1.5, 1.5], [0.5, 0.5, 1.5], [-0.5, -0.5, -0.5]],
                 np.float32)
    face, d = on.closest_point_brute(v, f, p)
    touching = lambda vid: min(k for k in range(len(f)) if vid in f[k])
    assert face[0] == touching(7) and d[0] == pytest.approx(np.sqrt(0.75))
    # nearest to the edge (x, 1, 1): the top quad (faces 2, 3) and the y=1 quad (6, 7) share it
    edge = [k for k in range(len(f)) if 6 in f[k] and 7 in f[k]]
    assert face[1] == min(edge) and d[1] == pytest.approx(np.sqrt(0.5))
    # a face interior: no tie
    assert face[2] in (2, 3) and d[2] == pytest.approx(0.5)
    assert face[3] == touching(0)


def test_oracle_bvh_closest_point_ties_match_brute_force():
    """On a marching-cubes trench (grid points, many nearest to shared edges and vertices) the
    oracle's BVH and brute force pick the same face everywhere."""
    from oracle import levelset as ol
    phi = ol.init_sdf(3.0, 0.9, 4.0, 0.15, 0.9, 0.3, 3.0, False).astype(np.float32)
    v, f = on.marching_cubes(phi, 0.15)
    ii, jj, kk = np.where(np.abs(phi) < 0.6)
    pts = np.stack([ii * 0.15, jj * 0.15, kk * 0.15], 1).astype(np.float32)
    fb, db = on.closest_point_brute(v, f, pts)
    fo, do = on.OracleBvh(v, f).closest_point(pts)
    assert len(pts) > 500
    np.testing.assert_array_equal(fo, fb)
    np.testing.assert_allclose(do, db)
