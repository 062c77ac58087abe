"""Pins the level-set oracle (oracle/levelset.py) to hand-computed values, closed forms and
brute force, so that the GPU parity tests compare against something trustworthy."""
import numpy as np
import pytest

from oracle import levelset as ol


# ------------------------------------------------------------------ advection (threed.py:1248)
def _col(vals):
    return np.asarray(vals, np.float64).reshape(1, 1, -1)


def test_advect_hand_values_peak_valley_boundary():
    # peak [0,1,0], dx=1, F=1, dt=0.1: centre g = max(1,0)^2 + min(-1,0)^2 = 2
    out = ol.advect_3d(_col([0, 1, 0]), np.ones((1, 1, 3)), 1.0, 0.1).ravel()
    assert out[1] == pytest.approx(1 - 0.1 * np.sqrt(2.0), abs=1e-15)
    # ends: k=0 has no D-, D+ = 1 -> min(1,0)=0 -> g=0 ; k=2: D- = -1 -> 0 ; no D+ -> 0
    assert out[0] == 0.0 and out[2] == 0.0
    # valley [1,0,1]: centre D- = -1, D+ = 1 -> g = 0 (a minimum of phi does not move)
    out = ol.advect_3d(_col([1, 0, 1]), np.ones((1, 1, 3)), 1.0, 0.1).ravel()
    assert out[1] == 0.0
    # k=0 of [1,0,1]: D+ = -1 -> g = 1 -> 1 - 0.1
    assert out[0] == pytest.approx(0.9, abs=1e-15)
    # two-node domain [1,0]: node 0 moves by F*dt, node 1 (no D+, D- = -1) is fixed
    out = ol.advect_3d(_col([1, 0]), np.ones((1, 1, 2)), 1.0, 0.1).ravel()
    assert out[0] == pytest.approx(0.9, abs=1e-15) and out[1] == 0.0


def test_advect_plane_moves_F_dt_and_F0_identity():
    dx, dt = 0.25, 0.05
    z = np.arange(40) * dx
    phi = np.broadcast_to(3.0 - z, (6, 5, 40)).copy()         # solid below z=3
    F = np.full(phi.shape, 0.7)
    out = ol.advect_3d(phi, F, dx, dt)
    inner = out[:, :, :-1]                                     # top face has no D-, stays put
    np.testing.assert_allclose(inner, (phi - 0.7 * dt)[:, :, :-1], atol=1e-14)
    np.testing.assert_array_equal(out[:, :, -1], phi[:, :, -1])
    np.testing.assert_array_equal(ol.advect_3d(phi, np.zeros_like(phi), dx, dt), phi)


def test_advect_pinned_mask_is_held():
    rng = np.random.default_rng(1)
    phi = rng.normal(size=(5, 6, 7))
    mask = rng.random(phi.shape) < 0.3
    mp = rng.normal(size=phi.shape)
    out = ol.advect_pinned(phi, np.ones_like(phi), mask, mp, 0.5, 0.01, 3)
    np.testing.assert_array_equal(out[mask], mp[mask])


# ------------------------------------------------------------------ Eikonal (threed.py:1421)
def test_eik_solve_satisfies_godunov_equation():
    rng = np.random.default_rng(2)
    dx = 0.3
    a = rng.uniform(0, 1.0, size=(3, 20000))
    a[:, :3000] = a[0, :3000]                                  # exact ties
    x = ol.eik_solve(a[0], a[1], a[2], dx)
    resid = np.sum(np.maximum(x[None] - a, 0.0) ** 2, axis=0) - dx * dx
    np.testing.assert_allclose(resid, 0.0, atol=1e-12)
    assert np.all(x > a.min(axis=0))                           # strictly above the smallest
    for p in [(1, 0, 2), (2, 1, 0), (0, 2, 1)]:                # symmetric in its arguments
        np.testing.assert_array_equal(ol.eik_solve(a[p[0]], a[p[1]], a[p[2]], dx), x)


def test_eik_solve_term_counts():
    dx = 1.0
    assert ol.eik_solve(0.0, 5.0, 5.0, dx) == 1.0                        # one term
    assert ol.eik_solve(0.0, 0.0, 5.0, dx) == pytest.approx(np.sqrt(0.5))  # two equal terms
    assert ol.eik_solve(0.0, 0.0, 0.0, dx) == pytest.approx(1 / np.sqrt(3))  # three


@pytest.mark.parametrize("slope", [2.0, 0.7])
def test_reinit_axis_plane_exact(slope):
    dx = 0.25
    z = np.arange(48) * dx
    phi = np.broadcast_to(slope * (5.3 - z), (8, 7, 48)).copy()
    band = 6 * dx
    out = ol.reinit_fsm(phi, dx, band)
    exact = 5.3 - z
    m = np.abs(exact) < band
    np.testing.assert_allclose(out[:, :, m], np.broadcast_to(exact[m], (8, 7, m.sum())), atol=1e-12)
    far = np.abs(exact) > band + 2 * dx
    np.testing.assert_allclose(np.abs(out[:, :, far]), band + dx)


def test_reinit_diagonal_plane_exact_in_band():
    dx = 0.2
    n = 60
    x = np.arange(n) * dx
    X, Y, Z = np.meshgrid(x, x, np.arange(6) * dx, indexing='ij')
    exact = (X + Y - 11.9) / np.sqrt(2.0)
    out = ol.reinit_fsm(1.7 * exact, dx, 6 * dx)
    # the clamped central gradient is one-sided on the domain faces, so keep to nodes whose
    # 10-sweep domain of dependence stays off them
    lo, hi = 10 * dx, x[-1] - 10 * dx
    m = (np.abs(exact) < 4 * dx) & (X > lo) & (Y > lo) & (X < hi) & (Y < hi)
    assert m.sum() > 100
    np.testing.assert_allclose(out[m], exact[m], atol=1e-9)


def test_reinit_perturbed_sphere_gates():
    """Gates of the reference's own reinit checks (scripts/reinit_correctness_3d.py:57-59,
    scripts/check_reinit_fidelity.py:11-25): |grad| ~ 1 in the band, near the exact distance."""
    N, dx = 48, 0.25
    xs = (np.arange(N) - N / 2) * dx
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing='ij')
    exact = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 3.0
    pert = exact * (1.0 + 0.4 * np.sin(X) * np.cos(Y))
    band = 6 * dx
    out = ol.reinit_fsm(pert, dx, band)
    inner = np.abs(out) < band - dx
    gx, gy, gz = np.gradient(out, dx)
    g = np.sqrt(gx ** 2 + gy ** 2 + gz ** 2)[inner]
    assert abs(g.mean() - 1.0) < 0.05 and np.percentile(g, 95) < 1.15
    near = np.abs(exact) < 4 * dx
    err = np.abs(out - exact)[near]
    # the frozen seed |phi0|/|grad phi0| is first order: O(dx) worst case, small on average
    assert err.mean() < 0.1 * dx and err.max() < 1.5 * dx
    nzero = near & (exact != 0)
    assert np.all(np.sign(out[nzero]) == np.sign(exact[nzero]))     # contour never moves


# ------------------------------------------------------------------ initial SDF (threed.py:32)
def _brute_distance(points_rz, hw, s, t, rmax, n=4000):
    """Distance to a dense sampling of the boundary polyline (floor, wall, mask top)."""
    floor = np.stack([np.linspace(0, hw, n), np.full(n, s)], 1)
    wall = np.stack([np.full(n, hw), np.linspace(s, t, n)], 1)
    top = np.stack([np.linspace(hw, rmax, 4 * n), np.full(4 * n, t)], 1)
    B = np.concatenate([floor, wall, top])
    d = np.full(len(points_rz), np.inf)
    for c in range(0, len(B), 2000):
        blk = B[c:c + 2000]
        d = np.minimum(d, np.sqrt(((points_rz[:, None, :] - blk[None]) ** 2).sum(-1)).min(1))
    return d


@pytest.mark.parametrize("hole", [True, False])
def test_init_sdf_matches_brute_force(hole):
    Lx, Ly, Lz, dx, w, m, s = 6.0, 5.0, 9.0, 0.25, 2.5, 1.0, 6.0
    phi = ol.init_sdf(Lx, Ly, Lz, dx, w, m, s, hole)
    xs, ys, zs = ol.grid_axes(Lx, Ly, Lz, dx)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
    r = np.sqrt((X - Lx / 2) ** 2 + (Y - Ly / 2) ** 2) if hole else np.abs(X - Lx / 2)
    rng = np.random.default_rng(0)
    idx = rng.choice(phi.size, 3000, replace=False)
    pts = np.stack([r.ravel()[idx], Z.ravel()[idx]], 1)
    bd = _brute_distance(pts, w / 2, s, s + m, 20.0)
    np.testing.assert_allclose(np.abs(phi.ravel()[idx]), bd, atol=3e-3)
    # sign: the analytic CSG field of threed.py:51-55 is positive in solid
    d_sub = s - Z
    d_slab = np.minimum(Z - s, s + m - Z)
    ana = np.maximum(d_sub, np.minimum(d_slab, r - w / 2))
    nz = np.abs(ana) > 1e-9
    assert np.all(np.sign(phi[nz]) == np.sign(ana[nz]))
    # mask definition threed.py:36-43
    mk = ol.trench_mask(Lx, Ly, Lz, dx, w, m, s, hole)
    assert mk.sum() > 0 and np.all(phi[mk] >= 0)


# ------------------------------------------------------------------ depth metrics (threed.py:1798)
def test_depth_metrics_on_known_floor():
    Lx, Ly, Lz, dx = 8.0, 8.0, 12.0, 0.25
    xs, ys, zs = ol.grid_axes(Lx, Ly, Lz, dx)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing='ij')
    sub_top, w = 9.0, 3.0
    r = np.sqrt((X - 4) ** 2 + (Y - 4) ** 2)
    floor = 5.0                                             # etched to z=5 inside r<1.5
    phi = np.where(Z < sub_top, 1.0, -1.0)
    phi = np.where((r < w / 2) & (Z >= floor), -1.0, phi)
    geo = dict(phi=phi, xs=xs, ys=ys, zs=zs, dx=dx, Lx=Lx, Ly=Ly, sub_top=sub_top,
               trench_width=w, hole=True)
    assert ol.depth3d(geo) == pytest.approx(sub_top - floor)
    # footprint r <= w/2+dx includes unetched columns at the rim: 5th percentile stays the floor
    assert ol.max_depth_3d(geo) == pytest.approx(sub_top - floor)
    # an enclosed gas pocket below the floor is not connected to the top: ignored
    phi2 = phi.copy()
    phi2[16, 16, 4] = -1.0
    geo['phi'] = phi2
    assert ol.depth3d(geo) == pytest.approx(sub_top - floor)


def _adjoint_case(seed):
    rng = np.random.default_rng(seed)
    shape, dx, dt = (7, 6, 9), 0.1, 0.02
    phi = rng.normal(size=shape)
    F = rng.uniform(0.2, 2.0, shape)
    mask = rng.uniform(size=shape) < 0.2
    mp = rng.normal(size=shape)
    go = rng.normal(size=shape)
    return phi, F, mask, mp, go, dx, dt


@pytest.mark.parametrize("seed", [0, 1])
def test_advect_adjoint_matches_central_differences(seed):
    """The oracle's adjoint of a pinned advection substep against float64 central differences
    of advect_3d (directional derivatives along random directions in phi and in F)."""
    phi, F, mask, mp, go, dx, dt = _adjoint_case(seed)

    def loss(p, f):
        return float((go * np.where(mask, mp, ol.advect_3d(p, f, dx, dt))).sum())

    gphi, gF = ol.advect_3d_adjoint(phi, F, mask, dx, dt, go)
    rng = np.random.default_rng(seed + 10)
    h = 1e-6
    for _ in range(3):
        v = rng.normal(size=phi.shape)
        fd = (loss(phi + h * v, F) - loss(phi - h * v, F)) / (2 * h)
        assert abs(fd - (gphi * v).sum()) <= 1e-6 * max(1.0, abs(fd))
        u = rng.normal(size=phi.shape)
        fd = (loss(phi, F + h * u) - loss(phi, F - h * u)) / (2 * h)
        assert abs(fd - (gF * u).sum()) <= 1e-6 * max(1.0, abs(fd))
    # mask nodes pass nothing through: their F gradient is exactly 0
    assert np.all(gF[mask] == 0.0)


def test_profile_rz_of_a_cylinder_and_a_plane():
    """The smoothed slice area of an exact signed distance: a vertical cylinder of radius R
    (gas inside) gives r(z) = R up to the smoothing's curvature term (relative O(eps^2/R^2));
    a horizontal interface at z0 gives A(z) = Lx Ly H_eps(z - z0) exactly, slice by slice."""
    dx, R = 0.1, 2.0
    n, nz = 64, 8
    x = (np.arange(n) - (n - 1) / 2) * dx
    X, Y = np.meshgrid(x, x, indexing='ij')
    phi = np.repeat((np.hypot(X, Y) - R)[:, :, None], nz, axis=2)
    for eps in (1.0 * dx, 1.5 * dx):
        r = ol.profile_rz(phi, dx, eps)
        assert np.allclose(r, r[0]) and abs(r[0] - R) < 2e-3 * R, (eps, r[0])
    z = np.arange(40) * dx
    z0 = 1.234
    phi = np.broadcast_to(z0 - z, (10, 12, 40))                     # gas above z0
    A = ol.slice_gas_area(phi, dx, 0.15)
    np.testing.assert_allclose(A, 10 * 12 * dx * dx * ol.smooth_heaviside(z - z0, 0.15),
                               rtol=1e-12, atol=1e-15)
    assert A[0] == 0.0 and A[-1] == pytest.approx(10 * 12 * dx * dx)


def test_profile_area_adjoint_matches_central_differences():
    rng = np.random.default_rng(4)
    dx, eps = 0.1, 0.15
    phi = rng.normal(scale=0.2, size=(6, 5, 7))
    garea = rng.normal(size=7)
    g = ol.slice_gas_area_adjoint(phi, dx, eps, garea)
    h = 1e-6
    for _ in range(4):
        u = rng.normal(size=phi.shape)
        fd = (garea @ ol.slice_gas_area(phi + h * u, dx, eps) -
              garea @ ol.slice_gas_area(phi - h * u, dx, eps)) / (2 * h)
        assert abs(fd - (g * u).sum()) < 1e-6 * max(1.0, abs(fd))


def _reinit_adjoint_case(n=20, dx=0.25):
    x = (np.arange(n) - n / 2 + 0.3) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.sqrt(X ** 2 + 1.3 * Y ** 2 + 0.8 * Z ** 2) - 1.6 + 0.1 * np.sin(2 * X) * np.cos(Y)
    return 1.2 * phi + 0.05 * Z, X, Y, Z


def test_reinit_adjoint_matches_central_differences():
    """The oracle's reinit adjoint against float64 central differences of reinit_fsm along a
    random and a smooth direction (|grad phi| != 1, a curved front, every case of the sweeps);
    the map is piecewise smooth, and at these steps no choice flips."""
    dx = 0.25
    phi, X, Y, Z = _reinit_adjoint_case(dx=dx)
    band = 6 * dx
    rng = np.random.default_rng(0)
    g = rng.standard_normal(phi.shape) * (np.abs(phi) < band)
    gc = ol.reinit_fsm_adjoint(phi, dx, band, g)
    for v in (rng.standard_normal(phi.shape), np.cos(X) * np.sin(0.7 * Y + Z)):
        h = 1e-6
        fd = ((ol.reinit_fsm(phi + h * v, dx, band) - ol.reinit_fsm(phi - h * v, dx, band))
              * g).sum() / (2 * h)
        assert abs(fd - (gc * v).sum()) < 1e-6 * abs(fd), (fd, (gc * v).sum())


def test_reinit_adjoint_of_a_shifted_plane():
    """An axis plane's exact distance is reinit's fixed point and moves with a uniform shift, so
    the Jacobian maps the constant 1 to 1 on every unclamped node: sum(J^T g) = sum of g over
    the nodes below the clamp. Holds whatever the internal choices, which pins the adjoint's
    bookkeeping (a dropped neighbour, seed or edge term breaks it). Planes normal to each axis,
    both orientations, off the grid nodes."""
    dx = 0.2
    shape = (14, 12, 16)
    band = 5 * dx
    rng = np.random.default_rng(1)
    for ax in range(3):
        for sgn, off in ((1.0, 0.537), (-1.0, 0.811)):
            c = np.arange(shape[ax]) * dx - (shape[ax] // 2) * dx - off * dx
            sh = [1, 1, 1]
            sh[ax] = shape[ax]
            phi = np.broadcast_to(sgn * c.reshape(sh), shape).copy()
            g = rng.standard_normal(shape)
            r = ol.reinit_fsm(phi, dx, band)
            assert np.abs(r - np.clip(phi, -(band + dx), band + dx)).max() < 1e-12
            live = np.abs(r) < band + dx
            for dt in (np.float64, np.float32):
                gc = ol.reinit_fsm_adjoint(phi, dx, band, g, state_dtype=dt)
                assert abs(gc.sum() - g[live].sum()) < 1e-9 * np.abs(g).sum(), (ax, sgn, dt)


@pytest.mark.parametrize("sigma", [0.5, 0.8, 1.3, 2.6])
def test_gaussian_smooth_equals_scipy(sigma):
    """The surf_smooth restatement against scipy.ndimage.gaussian_filter itself (the reference
    calls it at threed.py:1781): float32 grids, small axes where the reflection wraps more
    than once, bit for bit."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(int(sigma * 10))
    for shape in ((9, 7, 11), (3, 2, 20), (16, 12, 5)):
        a = rng.normal(size=shape).astype(np.float32)
        got = ol.gaussian_smooth(a, sigma)
        ref = gaussian_filter(a, sigma)
        assert got.dtype == ref.dtype == np.float32
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_reinit_pde_keeps_an_axis_plane_distance(axis):
    """reinit_gpu's PDE step (threed.py:1349-1398) leaves an exact plane distance unchanged,
    boundaries included (the clamped one-sided differences of a plane along an axis stay exact),
    and drives the interface nodes of a scaled plane to the true sub-cell distance."""
    dx = 0.25
    shape = (10, 8, 12)
    x = np.indices(shape)[axis] * dx
    d = (x - 1.13).astype(np.float32)
    np.testing.assert_array_equal(ol.reinit_pde(d, dx, 24), d)
    r = ol.reinit_pde(3.0 * d, dx, 300)
    iface = np.abs(d) < dx
    np.testing.assert_allclose(r[iface], d[iface], atol=2e-6)
    assert np.abs(r - d)[np.abs(d) < 4 * dx].max() < 1e-4         # the band relaxes too


def test_reinit_pde_signs_and_clip():
    """One step at an interface node by hand: D = phi0 / |grad phi0|_central (clipped to
    1.8 dx) and phi <- phi - 0.3 (sgn(phi0) |phi| - D)."""
    dx = 1.0
    p = np.full((3, 3, 4), 5.0, np.float32)
    p[:, :, :2] = -5.0
    p[1, 1, 1], p[1, 1, 2] = -0.1, 0.1
    r = ol.reinit_pde(p, dx, 1)
    # node (1,1,1): s0 = -0.1, o_zp = 0.1 -> interface; grad0 = max(|(0.1 - (-5)) / 2|, ...) = 2.55
    D = max(min(-0.1 / 2.55, 1.8), -1.8)
    np.testing.assert_allclose(r[1, 1, 1], -0.1 - 0.3 * (-1 * 0.1 - D), rtol=1e-6)
