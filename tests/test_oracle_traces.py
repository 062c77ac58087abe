"""Pins the oracle's sources and traces (oracle/petch_oracle.c) to closed forms: source moments,
the flat open field (every primary ray deposits exactly once, m = 1), the normal-incidence ion
yields integrated over the Gaussian energy distribution, and its BVH against brute force."""
import numpy as np
import pytest
from scipy import integrate, stats

from oracle import chem as oc
from oracle import native as on
from petch.params import PAR


def flat_mesh(Lx, Ly, z0, nx=8, ny=6, margin=0.0):
    xs = np.linspace(-margin, Lx + margin, nx + 1)
    ys = np.linspace(-margin, Ly + margin, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing='ij')
    v = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z0)], 1).astype(np.float32)
    f = []
    for i in range(nx):
        for j in range(ny):
            a = i * (ny + 1) + j
            b, c, d = a + (ny + 1), a + (ny + 1) + 1, a + 1
            f += [(a, b, c), (a, c, d)]
    return v, np.array(f, np.int32)


def test_source_moments():
    box = on.Box(5.0, 3.0, 9.0, 8.75, 0)
    n = 200000
    o, d = on.gen_source(box, on.Source(0, 0.0, 0.0), n, 11)
    assert np.all(o[:, 2] == np.float32(8.75))
    assert abs(o[:, 0].mean() - 2.5) < 0.02 and abs(o[:, 0].var() - 25 / 12) < 0.03
    assert stats.kstest(o[:, 1] / 3.0, 'uniform').pvalue > 1e-3
    ct = -d[:, 2]
    assert abs(ct.mean() - 2 / 3) < 3e-3 and abs((ct ** 2).mean() - 0.5) < 3e-3  # cosine law
    sig = np.deg2rad(2.5)
    o, d = on.gen_source(box, on.Source(1, sig, 0.0), n, 12)
    assert abs(np.arcsin(d[:, 0]).std() / sig - 1) < 0.02                      # Gaussian tilt
    for expo in (100.0, 1000.0):
        o, d = on.gen_source(box, on.Source(2, 0.0, expo), n, 13)
        assert abs((-d[:, 2]).mean() - (expo + 1) / (expo + 2)) < 1e-4          # power cosine
    nrm = np.linalg.norm(d, axis=1)
    np.testing.assert_allclose(nrm, 1.0, atol=1e-6)


@pytest.mark.parametrize("periodic", [0, 1])
def test_rr_neutral_flat_open_field_is_unity(periodic):
    Lx, Ly = 4.0, 3.0
    v, f = flat_mesh(Lx, Ly, 1.0)
    bvh = on.OracleBvh(v, f)
    box = on.Box(Lx, Ly, 6.0, 5.75, periodic)
    n = 60000
    for beta in (0.1, 0.7, 1.0):
        fl = bvh.trace_neutral_rr(box, on.Source(0, 0, 0), np.ones(len(f)), beta, n, 5)
        if periodic:
            # one deposit of weight 1 per ray; the wrap of threed.py:327-346 loses the rare ray
            # that leaves exactly through a lateral corner (tx == ty)
            assert n - 5 <= fl.sum() <= n
        else:
            assert fl.sum() <= n
        A = 0.5 * Lx / 8 * Ly / 6
        m = (fl / A) * (Lx * Ly / n)
        if periodic:
            assert abs(m.mean() - 1.0) < 1e-4
            assert m.std() < 5 * np.sqrt(len(f) / n) + 0.02


def test_faithful_ion_normal_incidence_closed_form():
    """sigma = 0: every ion hits the flat floor at normal incidence (theta < thetaRMin), so it is
    absorbed at its first hit and deposits A*max(sqrt(E)-sqrt(Eth),0) with f_sp = f_ie = 1."""
    Lx, Ly = 4.0, 3.0
    v, f = flat_mesh(Lx, Ly, 1.0)
    bvh = on.OracleBvh(v, f)
    box = on.Box(Lx, Ly, 6.0, 5.75, 1)
    p = dict(PAR)
    ion = on.Ion(**oc.ion_params(p))
    n = 200000
    out = bvh.trace_ion_yield(box, on.Source(1, 0.0, 0.0), ion, n, 9)
    assert out[3].sum() == pytest.approx(n) and out[4].sum() == pytest.approx(n)

    def expect(A, Eth):
        g = lambda E: A * max(np.sqrt(max(E, 0)) - np.sqrt(Eth), 0) * stats.norm.pdf(E, 100, 10)
        return integrate.quad(g, 0, 200, limit=200)[0]
    for ch, (A, Eth) in enumerate([(p['A_sp'], p['Eth_sp']), (p['A_ie'], p['Eth_ie']),
                                   (p['A_p'], p['Eth_p'])]):
        mean = out[ch].sum() / n
        assert mean == pytest.approx(expect(A, Eth), rel=2e-3)


def test_legacy_first_hit_count():
    Lx, Ly = 4.0, 3.0
    v, f = flat_mesh(Lx, Ly, 1.0, margin=2.0)               # floor wider than the source plane
    bvh = on.OracleBvh(v, f)
    box = on.Box(Lx, Ly, 6.0, 5.75, 0)
    n = 20000
    out = bvh.trace_legacy(box, on.Source(1, np.deg2rad(2.5), 0), 1.0, 0, 0, 0.34, 0.8, n, 3)
    assert out[0].sum() == pytest.approx(n, abs=0.5)       # near-vertical ions all hit the floor
    assert out[1].sum() / n == pytest.approx(1.0, abs=3e-3)  # at cos ~ 1
    out = bvh.trace_legacy(box, on.Source(0, 0, 0), 0.3, 12, 0, 0.34, 0.8, n, 4)
    assert out[0].sum() <= n                                # re-emitted neutrals may escape


def test_oracle_bvh_matches_brute_force():
    x = np.arange(30) * 0.1
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    phi = (1.1 - np.sqrt((X - 1.45) ** 2 + (Y - 1.45) ** 2 + (Z - 1.45) ** 2)).astype(np.float32)
    v, f = on.marching_cubes(phi, 0.1)
    rng = np.random.default_rng(0)
    o = rng.uniform(0, 2.9, (3000, 3)).astype(np.float32)
    d = rng.normal(size=(3000, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    fb, tb = on.closest_hit_brute(v, f, o, d)
    fo, to = on.OracleBvh(v, f).closest_hit(o, d)
    np.testing.assert_array_equal(fo, fb)
    np.testing.assert_allclose(to, tb)


@pytest.mark.parametrize("beta", [0.7, 1.0])
def test_rr_path_adjoint_matches_frozen_decision_differences(beta):
    """The oracle's path adjoint of the RR neutral trace (stored paths, backward sums) against
    central differences of the same trace with its paths and roulette / termination decisions
    frozen and survivors reweighted by their frozen survival probability
    (oracle_trace_neutral_rr_frozen, which equals the trace where the two bare arrays agree).
    Roulette is frequent at these stickings, so the survivors' reweighting is exercised."""
    from oracle import levelset as ol
    phi = ol.init_sdf(6.0, 1.5, 12.0, 0.15, 1.5, 7.0, 3.0, False).astype(np.float32)
    v, f = on.marching_cubes(phi, 0.15)
    bvh = on.OracleBvh(v, f)
    box = on.Box(6.0, 1.5, 12.0, 12.0 - 0.15, 1)
    src = on.Source(0, 0, 0)
    F = len(f)
    rng = np.random.default_rng(int(beta * 10))
    g = rng.normal(size=F)
    bare = rng.uniform(0.2, 1.0, F).astype(np.float32)
    n = 20000
    plain = bvh.trace_neutral_rr(box, src, bare, beta, n, 7)
    frozen = bvh.trace_neutral_rr_frozen(box, src, bare, bare, beta, n, 7)
    np.testing.assert_allclose(frozen, plain, rtol=0, atol=1e-6 * np.abs(plain).max())
    gb = bvh.trace_neutral_rr_adjoint(box, src, bare, beta, n, 7, g)
    h = 2e-3    # the weights are products over whole paths: keep the second-order term small
    for _ in range(3):
        u = rng.normal(size=F)
        lp = g @ bvh.trace_neutral_rr_frozen(box, src, bare, bare + h * u, beta, n, 7)
        lm = g @ bvh.trace_neutral_rr_frozen(box, src, bare, bare - h * u, beta, n, 7)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gb @ u) <= 1e-3 * abs(fd) + 1e-3, (fd, gb @ u)


def test_coupled_coverage_adjoint_matches_frozen_decision_differences():
    """The oracle's unrolled adjoint of the MC coverage loop (two rounds of F and O traces,
    clip-normalisation, smoothing, Belen update) in m_i against central differences of the same
    loop with every round's paths and decisions frozen. The smoothing matrix equals the
    smoothing functions it transposes."""
    from oracle import flux as ofl
    from oracle import levelset as ol
    from oracle import mesh as om
    from petch.params import Flags
    phi = ol.init_sdf(6.0, 1.5, 12.0, 0.15, 1.5, 7.0, 3.0, False).astype(np.float32)
    v, f = on.marching_cubes(phi, 0.15)
    M = ofl.OracleMesh(v, f)
    geo = dict(Lx=6.0, Ly=1.5, Lz=12.0, dx=0.15)
    par = dict(PAR)
    par['periodic_y'] = 1
    fl = Flags(chemistry="belen", ion_reflection=True)
    rng = np.random.default_rng(8)
    x = rng.uniform(0.0, 2.0, M.nf)
    for dev, n_s, a in ((True, 1, 1.0), (True, 2, 0.6), (False, 2, 0.6)):
        T = ofl._smooth_matrix(M, n_s, a, dev)
        ref = x.copy()
        if dev:
            for _ in range(n_s):
                ref = om.smooth_flux_dev(ref, M.normals, M.pairs, a)
        else:
            ref = om.smooth_flux(x, M.normals, M.pairs, n_s, a)
        np.testing.assert_allclose(T.dot(x), ref, rtol=1e-12, atol=1e-12)
    m_i = rng.uniform(0.05, 1.2, M.nf)
    cos_i = rng.uniform(0.0, 1.0, M.nf)
    w = rng.normal(size=M.nf)
    n = 6000
    bare, _, _, bares, grads = ofl.coupled_coverage(M, geo, par, fl, m_i, cos_i, n, 3, 2, w=w)
    h = 1e-4
    for _ in range(2):
        u = rng.normal(size=M.nf)
        lp = w @ ofl.coupled_coverage(M, geo, par, fl, m_i + h * u, cos_i, n, 3, 2,
                                      frozen=bares)[0]
        lm = w @ ofl.coupled_coverage(M, geo, par, fl, m_i - h * u, cos_i, n, 3, 2,
                                      frozen=bares)[0]
        fd = (lp - lm) / (2 * h)
        an = grads["m_i"] @ u
        assert abs(fd - an) <= 2e-3 * abs(fd) + 1e-4, (fd, an)


def test_rr_path_adjoint_matches_common_random_number_differences():
    """The gradient the path adjoint estimates is that of the expected flux: against central
    differences of the plain trace with common random numbers (decisions free to flip) it agrees
    within 2 % for a positive loss along positive directions, where the flips' noise is small
    next to the signal (1M rays). Cutting the derivative at the roulette instead (survivors'
    weight 0.3 taken as a constant) misses about a fifth of it here."""
    from oracle import levelset as ol
    phi = ol.init_sdf(6.0, 1.5, 12.0, 0.15, 1.5, 7.0, 3.0, False).astype(np.float32)
    v, f = on.marching_cubes(phi, 0.15)
    bvh = on.OracleBvh(v, f)
    box = on.Box(6.0, 1.5, 12.0, 12.0 - 0.15, 1)
    src = on.Source(0, 0, 0)
    F = len(f)
    rng = np.random.default_rng(1)
    bare = rng.uniform(0.2, 1.0, F).astype(np.float32)
    g = rng.uniform(0.5, 1.5, F)
    n, beta, h = 1000000, 0.7, 0.05
    gb = bvh.trace_neutral_rr_adjoint(box, src, bare, beta, n, 7, g)
    for u in (np.ones(F), rng.uniform(0.0, 1.0, F)):
        lp = g @ bvh.trace_neutral_rr(box, src, (bare + h * u).astype(np.float32), beta, n, 7)
        lm = g @ bvh.trace_neutral_rr(box, src, (bare - h * u).astype(np.float32), beta, n, 7)
        fd = (lp - lm) / (2 * h)
        assert abs(gb @ u - fd) <= 0.02 * abs(fd), (fd, gb @ u)


def _cube_cavity(R=1.5, dx=0.25):
    """MC mesh of a gas-filled cube of half-width R in solid, shifted so that its x, y range is
    [-0.25, 2R - 0.25]: a source plane [0, 2R - 0.5]^2 at its mid height lies inside it."""
    n = int(np.ceil(2 * R / dx)) + 5
    c = 0.5 * (n - 1) * dx
    x = np.arange(n) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    phi = (np.maximum(np.maximum(np.abs(X - c), np.abs(Y - c)), np.abs(Z - c)) - R)
    v, f = on.marching_cubes(phi.astype(np.float32), dx)
    v = v.copy()
    v[:, :2] -= c - R + 0.25
    return v, f, c


def test_capped_trace_closed_forms():
    """_trace3d_cov (threed.py:511-546): in a closed cavity every ray deposits 1 per hit until it
    sticks or reaches the cap, so with S = 0 exactly n_reemit + 1 per ray, with S = 1 exactly 1,
    and with S = 1/2 on average (1 - 2^-(n_reemit + 1)) / (1/2); on a flat field wide enough to
    catch every source ray each ray deposits exactly once (the re-emission escapes). The box's
    periodic flag changes nothing (the reference kernel applies no boundary condition)."""
    v, f, c = _cube_cavity()
    bvh = on.OracleBvh(v, f)
    box, src = on.Box(2.5, 2.5, 3.0, float(c), 0), on.Source(0, 0, 0)
    n, nre = 20000, 5
    ones = np.ones(len(f), np.float32)
    assert bvh.trace_neutral_cap(box, src, np.zeros(len(f), np.float32), 0.7, nre, n, 3).sum() \
        == n * (nre + 1)
    assert bvh.trace_neutral_cap(box, src, ones, 1.0, nre, n, 3).sum() == n
    half = bvh.trace_neutral_cap(box, src, ones, 0.5, nre, n, 3).sum() / n
    mean = (1.0 - 0.5 ** (nre + 1)) / 0.5
    assert abs(half - mean) < 5 * 1.0 / np.sqrt(n)                 # per-ray sd below 1
    wrap = bvh.trace_neutral_cap(on.Box(2.5, 2.5, 3.0, float(c), 1), src, ones, 0.5, nre, n, 3)
    np.testing.assert_array_equal(wrap, bvh.trace_neutral_cap(box, src, ones, 0.5, nre, n, 3))
    m = 20.0
    vf, ff = flat_mesh(4.0, 3.0, 1.0, margin=m)
    fb = on.OracleBvh(vf, ff)
    fbox = on.Box(4.0, 3.0, 6.0, 5.75, 0)
    o, d = on.gen_source(fbox, src, 5000, 9)
    t = (1.0 - o[:, 2]) / d[:, 2]                    # every source ray that lands on the plane
    px, py = o[:, 0] + t * d[:, 0], o[:, 1] + t * d[:, 1]
    lands = ((px > -m + 1e-3) & (px < 4.0 + m - 1e-3) & (py > -m + 1e-3) &
             (py < 3.0 + m - 1e-3)).sum()
    assert 4500 < lands < 5000
    for beta in (0.0, 0.5):
        fl = fb.trace_neutral_cap(fbox, src, np.ones(len(ff), np.float32), beta, 24, 5000, 9)
        assert fl.sum() == lands
