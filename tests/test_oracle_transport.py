"""Pins of the oracle's deterministic neutral transports (oracle/transport.py) by closed forms
and independent solvers: the tridiagonal solve against scipy's banded solver, a uniform duct's
concentration profile against its closed form, box slice geometry, the nearest-slice rule, the
march on a closed cavity (0), on a flat open field (the reach condition of its step count), at
the floor centre of a long slot (the opening's view factor sin(theta0)), full sticking, a hand
example of the re-emission field, and the Aitken step on a geometric sequence."""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from oracle import transport as ot


def test_tridiagonal_matches_banded_solver():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 60):
        lo, up = -rng.uniform(0.1, 1.0, n - 1), -rng.uniform(0.1, 1.0, n - 1)
        d = rng.uniform(2.5, 4.0, n)
        b = rng.normal(size=n)
        ab = np.zeros((3, n))
        ab[0, 1:], ab[1], ab[2, :-1] = up, d, lo
        np.testing.assert_allclose(ot.solve_tridiagonal(lo, d, up, b), solve_banded((1, 1), ab, b),
                                   rtol=1e-12, atol=1e-14)


def _duct(m, a, p, nz_extra=3):
    """area/perimeter of a uniform duct of m slices below sub_top, nz_extra slices above."""
    nz = m + nz_extra
    zs = np.arange(nz) * 0.5
    sub_top = zs[m]                     # slices 0..m-1 are below
    area, perim = np.zeros(nz), np.zeros(nz)
    area[:m], perim[:m] = a, p
    return zs, sub_top, area, perim


@pytest.mark.parametrize("s_b", [0.0, 0.3, 5.0])
def test_uniform_duct_profile_closed_form(s_b):
    """Chain of m equal conductances C (links C too), unit concentration above, sink s_b at the
    floor only: the flux J = s_b / (1 + m s_b / C) crosses every link, so from the top slice
    down x_i = 1 - (i + 1) J / C."""
    dx, m, a, p = 0.5, 9, 3.0, 7.0
    zs, sub_top, area, perim = _duct(m, a, p)
    loss = np.zeros(len(zs))
    loss[0] = s_b                       # slice 0 is the deepest
    prof = ot.conductance_profile(area, perim, zs, dx, sub_top, loss, 1.0)
    C = (2.0 / 3.0) * max(2 * a / p, dx) * a / dx
    J = s_b / (1 + m * s_b / C)
    top_down = 1.0 - (np.arange(m) + 1) * J / C
    np.testing.assert_allclose(prof[:m][::-1], top_down, rtol=1e-12, atol=1e-14)
    assert (prof[m:] == 1.0).all()


def test_broken_chain_starves_below_the_gap():
    dx = 0.5
    zs, sub_top, area, perim = _duct(8, 2.0, 5.0)
    area[3] = 0.0                       # slices 0-2 are cut off from the opening
    loss = np.zeros(len(zs))
    loss[0] = 0.2
    prof = ot.conductance_profile(area, perim, zs, dx, sub_top, loss, 1.0)
    assert (prof[:3] == 0.0).all() and prof[3] == 1.0
    np.testing.assert_allclose(prof[4:8], 1.0, rtol=0, atol=1e-14)


def test_slice_geometry_of_boxes():
    dx = 0.25
    nx, ny, nz = 20, 18, 12
    zs = np.arange(nz) * dx
    sub_top = zs[8]
    phi = np.ones((nx, ny, nz))
    phi[5:12, 4:9, 2:] = -1.0           # 7 x 5 hole from slice 2 up
    a, p = ot.slice_geometry(phi, zs, dx, sub_top, "hole")
    np.testing.assert_allclose(a[2:8], 35 * dx * dx)
    np.testing.assert_allclose(p[2:8], 24 * dx)
    assert (a[:2] == 0).all() and (a[8:] == 0).all()
    phi = np.ones((nx, ny, nz))
    phi[6:10, :, 3:] = -1.0             # trench 4 cells wide along y
    phi[6:8, 0, 5] = 1.0                # one row narrower in slice 5
    a, p = ot.slice_geometry(phi, zs, dx, sub_top, "trench")
    np.testing.assert_allclose(a[3:8], [4 * dx] * 2 + [(17 * 4 + 2) * dx / 18] + [4 * dx] * 2)
    assert (p[3:8] == 2.0).all() and p[2] == 0.0


def test_nearest_valid_slice_ties_to_the_lower():
    v = np.array([2, 3, 7, 12])
    got = ot.nearest_valid_slice(v, np.array([0, 2, 4, 5, 6, 9, 10, 11, 15]))
    np.testing.assert_array_equal(got, [2, 2, 3, 3, 7, 7, 12, 12, 12])
    assert (ot.nearest_valid_slice(np.array([], int), np.array([1, 2])) == -1).all()


def test_knudsen_faces_above_the_substrate_read_one():
    dx = 0.25
    phi = np.ones((12, 12, 16))
    phi[4:8, 4:8, 4:] = -1.0
    phi[:, :, 12:] = -1.0
    zs = np.arange(16) * dx
    sub_top = zs[12]
    c = np.array([[1.5, 1.5, sub_top + 0.1], [1.5, 1.5, 1.5], [1.5, 1.5, 1.0]])
    m = ot.knudsen_face_flux(phi, zs, dx, sub_top, "hole", c, np.array([1.0, 1.0, 1.0]),
                             np.ones(3) * dx * dx, 3.0, np.ones(3) * 0.5,
                             center=(1.5, 1.5), half_width=0.5)
    assert m[0] == 1.0 and 0.0 < m[2] <= m[1] < 1.0


def test_fib_sphere_directions():
    d, w = ot.fib_sphere(512)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert abs(d.mean(axis=0)).max() < 5e-3 and (w == 1).all()
    np.testing.assert_allclose(np.sort(d[:, 2]), np.sort(1 - (np.arange(512) + 0.5) / 256))


def _march(phi, dx, o, n, dirs, periodic_y=0):
    nz = phi.shape[2]
    zs = np.arange(nz) * dx
    z_top = zs[-1] - 0.5 * dx
    ms = (int((z_top - zs[0]) / dx) + 4) * (3 if periodic_y else 1)
    return ot.dda_gather(phi, dx, np.atleast_2d(o), np.atleast_2d(n), dirs.astype(np.float32),
                         np.zeros(phi.size, np.float32), z_top, 1.0, ms, periodic_y), z_top, ms


def test_march_closed_cavity_sees_nothing():
    phi = np.ones((10, 10, 12), np.float32)
    phi[3:7, 3:7, 3:8] = -1.0
    d, _ = ot.fib_sphere(128)
    v, _, _ = _march(phi, 1.0, [5.0, 5.0, 3.0], [0, 0, 1], d)
    assert v[0] == 0.0


def test_march_flat_field_reach_condition():
    """Open floor at z = h: a direction reaches z_top iff its last checked point
    o + (1.01 + max_steps - 1) dx d is at or above z_top, so the direct flux is the
    cosine-weighted share of those directions (none of them lies near the threshold here)."""
    dx, nz, h = 1.0, 30, 9.0
    z = np.arange(nz) * dx
    phi = np.broadcast_to((h - z)[None, None, :], (12, 12, nz)).astype(np.float32).copy()
    d, _ = ot.fib_sphere(256)
    v, z_top, ms = _march(phi, dx, [6.0, 6.0, h], [0, 0, 1], d)
    up = d[:, 2] > 1e-6
    reach = up & ((1.01 + ms - 1) * dx * d[:, 2] >= z_top - h)
    margin = np.abs((1.01 + ms - 1) * dx * d[up, 2] - (z_top - h)).min()
    assert margin > 1e-3
    assert 0.3 < reach.sum() / up.sum() < 0.9
    np.testing.assert_allclose(v[0], d[reach, 2].sum() / d[up, 2].sum(), rtol=1e-6)


def test_march_slot_floor_view_factor():
    """Floor centre of a long slot (half width b, walls H high, periodic along y): the
    cosine-weighted share of the sky is the view factor of the opening, sin(theta0) with
    tan(theta0) = b / H. The march samples walls every dx and the trilinear level set rounds
    the wall tops: at dx = H / 17 the share is 8 % high, at H / 34 and H / 68 it is within 0.9 %
    (what remains is the quadrature of 1024 upward directions)."""
    dx, nx, ny, nz = 0.5, 80, 6, 88
    cx, b, floor, top = 19.5, 4.0, 20.0, 37.0
    x = np.arange(nx)[:, None, None] * dx
    z = np.arange(nz)[None, None, :] * dx
    walls = np.minimum(np.abs(x - cx) - b, top - z)
    phi = np.maximum(floor - z, walls) * np.ones((1, ny, 1))
    d, _ = ot.fib_sphere(2048)
    v, _, _ = _march(phi.astype(np.float32), dx, [cx, 1.25, floor], [0, 0, 1], d, periodic_y=1)
    theta0 = np.arctan(b / (top - floor))
    assert v[0] == pytest.approx(np.sin(theta0), rel=0.02)


def test_remit_hand_example():
    """Two faces in one solid cell (areas 1 and 3, (1-s) vf = 0.2 and 0.6) give it their
    area-weighted mean 0.5; its unseeded solid neighbours (six, wrapping) take 0.5; a gas cell
    and a cell two steps away stay 0."""
    phi = np.ones((5, 5, 5))
    phi[2, 2, 3] = -1.0
    cell = np.array([(2 * 5 + 2) * 5 + 2, (2 * 5 + 2) * 5 + 2])
    areas = np.array([1.0, 3.0])
    carea = np.zeros(125)
    np.add.at(carea, cell, areas)
    seeded = np.zeros(125, bool)
    seeded[cell] = True
    R = ot.dda_remit(phi, cell, carea, seeded.reshape(5, 5, 5), areas, np.array([0.5, 1.0]),
                     np.array([0.4, 0.6])).reshape(5, 5, 5)
    assert R[2, 2, 2] == pytest.approx(0.5)
    for q in [(1, 2, 2), (3, 2, 2), (2, 1, 2), (2, 3, 2), (2, 2, 1)]:
        assert R[q] == pytest.approx(0.5)
    assert R[2, 2, 3] == 0.0 and R[0, 2, 2] == 0.0
    # wrap-around neighbours of a seeded cell on the boundary
    cell2 = np.array([(0 * 5 + 2) * 5 + 2])
    s2 = np.zeros(125, bool)
    s2[cell2] = True
    R2 = ot.dda_remit(phi, cell2, np.where(s2, 1.0, 0.0), s2.reshape(5, 5, 5), np.ones(1),
                      np.ones(1), np.ones(1)).reshape(5, 5, 5)
    assert R2[4, 2, 2] == pytest.approx(1.0) and R2[1, 2, 2] == pytest.approx(1.0)


def test_aitken_is_exact_on_a_geometric_sequence():
    r, a, lim = 0.8, np.array([2.0, -1.0]), np.array([5.0, 1.0])
    it = [lim - a * r ** k for k in range(3)]
    np.testing.assert_allclose(ot.aitken(it[2], it[1], it[0]), lim, rtol=1e-12)
    flat = np.array([1.0])
    assert ot.aitken(flat, flat, flat)[0] == 1.0


def test_full_sticking_has_no_reemission():
    dx = 1.0
    nx, ny, nz = 16, 6, 24
    x = np.arange(nx)[:, None, None] * dx
    z = np.arange(nz)[None, None, :] * dx
    phi = (np.maximum(8.0 - z, np.minimum(np.abs(x - 7.5) - 3.0, 18.0 - z)) *
           np.ones((1, ny, 1))).astype(np.float32)
    c = np.array([[7.5, 2.5, 8.0], [4.5, 2.5, 12.0]])
    n = np.array([[0, 0, 1.0], [1.0, 0, 0]])
    zs = np.arange(nz) * dx
    one = ot.dda_neutral_flux(phi, dx, zs, c, n, np.ones(2), np.ones(2), n_dir=32, n_reemit=2,
                              periodic_y=1)
    dirs = ot.fib_sphere(64)[0].astype(np.float32)
    direct = _march(phi, dx, c, n, dirs, periodic_y=1)[0].astype(np.float32)
    np.testing.assert_allclose(one, direct, rtol=1e-6)
    half = ot.dda_neutral_flux(phi, dx, zs, c, n, np.full(2, 0.5), np.ones(2), n_dir=32,
                               n_reemit=2, periodic_y=1)
    assert (half >= one - 1e-9).all() and half[1] > one[1]
