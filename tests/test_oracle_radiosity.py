"""The oracle's radiosity adjoint (oracle/flux.py radiosity_adjoint) pinned by what the
mathematics fixes: central differences of the direct solve in float64, the linear-system identity
lambda . D = gG . G, the clip masks, and convergence of the transposed Jacobi sweeps to the
direct adjoint."""
import numpy as np
import pytest

from oracle import flux as ofl


def _case(seed, nf=48, n_ff=16, p_esc=0.3, escape_each=False):
    rng = np.random.default_rng(seed)
    hit = rng.integers(0, nf, nf * n_ff).astype(np.int64)
    hit[rng.random(nf * n_ff) < p_esc] = -1
    if escape_each:
        hit[::n_ff] = -1               # D >= 1/n_ff, so G > 0 stays off the clip at 0
    D, Aff = ofl.radiosity_matrix(hit, nf, n_ff)
    return rng, hit, D, Aff, nf, n_ff


def _loss(D, Aff, bare, beta, w):
    G = ofl.radiosity_direct(D, Aff, bare, beta)
    return float(w @ np.clip(G, 0.0, 8.0))


@pytest.mark.parametrize("seed", [0, 1])
def test_radiosity_adjoint_matches_central_differences(seed):
    rng, hit, D, Aff, nf, n_ff = _case(seed)
    beta = 0.7
    bare = rng.uniform(0.1, 1.0, nf)
    bare[:4] = 1.6                     # s = clip(bare beta) saturates at 1: no gradient there
    w = rng.normal(size=nf)
    G = ofl.radiosity_direct(D, Aff, bare, beta)
    gD, gbare = ofl.radiosity_adjoint(D, Aff, G, bare, beta, w)
    assert np.all(gbare[:4] == 0.0)
    h = 1e-6
    for _ in range(3):
        v = rng.normal(size=nf)
        fd = (_loss(D + h * v, Aff, bare, beta, w) - _loss(D - h * v, Aff, bare, beta, w)) / (2 * h)
        assert abs(fd - gD @ v) <= 1e-6 * max(1.0, abs(fd)), (fd, gD @ v)
        u = rng.normal(size=nf)
        u[:4] = 0.0
        fd = (_loss(D, Aff, bare + h * u, beta, w) - _loss(D, Aff, bare - h * u, beta, w)) / (2 * h)
        assert abs(fd - gbare @ u) <= 1e-6 * max(1.0, abs(fd)), (fd, gbare @ u)


def test_radiosity_adjoint_identity_and_transpose():
    """G is linear in D, so lambda . D = gG . G; the transpose is the rays of each source face
    (row j of Aff^T gathers lambda at the hit faces of face j's n_ff rays)."""
    rng, hit, D, Aff, nf, n_ff = _case(5)
    bare = rng.uniform(0.0, 1.0, nf)
    w = rng.normal(size=nf)
    for beta in (0.7, 1.0):
        G = ofl.radiosity_direct(D, Aff, bare, beta)
        gD, gbare = ofl.radiosity_adjoint(D, Aff, G, bare, beta, w)
        assert abs(gD @ D - w @ G) <= 1e-10 * np.abs(w @ G) + 1e-12
        lam = rng.normal(size=nf)
        gather = np.array([sum(lam[h] for h in hit[j * n_ff:(j + 1) * n_ff] if h >= 0)
                           for j in range(nf)]) / n_ff
        np.testing.assert_allclose(Aff.T.dot(lam), gather, rtol=1e-12, atol=1e-14)


def test_radiosity_adjoint_clip_and_jacobi():
    """Faces whose G leaves [0, 8] pass no gradient; the transposed Jacobi sweeps converge to the
    direct adjoint."""
    rng, hit, D, Aff, nf, n_ff = _case(9)
    bare = rng.uniform(0.2, 1.0, nf)
    w = rng.normal(size=nf)
    G = ofl.radiosity_direct(D, Aff, bare, 0.7)
    Gc = G.copy()
    Gc[:3] = 9.0
    gD, _ = ofl.radiosity_adjoint(D, Aff, Gc, bare, 0.7, np.where(np.arange(nf) < 3, w, 0.0))
    assert np.all(gD == 0.0)
    d_direct = ofl.radiosity_adjoint(D, Aff, G, bare, 0.7, w)
    d_jac = ofl.radiosity_adjoint(D, Aff, G, bare, 0.7, w, n_iter=400)
    for a, b in zip(d_direct, d_jac):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)


def _chem_case(seed, F=40):
    from petch.params import PAR, Flags
    rng = np.random.default_rng(seed)
    par = dict(PAR)
    fl = Flags(chemistry="belen", ion_reflection=True)
    m_i = rng.uniform(0.1, 1.2, F)
    cos_i = rng.uniform(0.0, 1.0, F)
    iy = tuple(rng.uniform(0.0, 0.5, F) for _ in range(3))
    return rng, par, fl, m_i, cos_i, iy


@pytest.mark.parametrize("yields", [False, True])
def test_belen_adjoint_matches_central_differences(yields):
    """The complex-step derivative of the Belen bare update against float64 central differences,
    for the m_F, m_O and ion inputs (m_i, or the ion yields)."""
    from oracle import chem as oc
    rng, par, fl, m_i, cos_i, iy = _chem_case(3)
    F = len(m_i)
    m_F, m_O = rng.uniform(0.01, 1.0, F), rng.uniform(0.01, 1.0, F)
    iy = iy if yields else None
    g = rng.normal(size=F)
    out = oc.belen_bare_adjoint(m_i, m_F, m_O, cos_i, par, fl, g, ion_yield=iy)

    def loss(**kw):
        a = dict(m_i=m_i, m_F=m_F, m_O=m_O)
        y = list(iy) if iy is not None else None
        for k, v in kw.items():
            if k in a:
                a[k] = v
            else:
                y[{"ie": 1, "ip": 2}[k]] = v
        thF, thO = oc.belen_coverages(a["m_i"], a["m_F"], a["m_O"], cos_i, par, fl,
                                      ion_yield=tuple(y) if y is not None else None)
        return float(g @ np.clip(1.0 - thF - thO, 0.0, 1.0))

    base = dict(m_i=m_i, m_F=m_F, m_O=m_O)
    if iy is not None:
        base.update(ie=iy[1], ip=iy[2])
    for k, grad in out.items():
        v = rng.normal(size=F)
        h = 1e-7 * max(1.0, float(np.abs(base[k]).max()))
        fd = (loss(**{k: base[k] + h * v}) - loss(**{k: base[k] - h * v})) / (2 * h)
        assert abs(fd - grad @ v) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad @ v)
    assert set(out) == ({"m_F", "m_O", "ie", "ip"} if yields else {"m_F", "m_O", "m_i"})


def test_coverage_fixed_point_adjoint_matches_central_differences():
    """The unrolled adjoint of the radiosity coverage loop (n_fp rounds of two solves and a Belen
    update) against float64 central differences in m_i and D. Every face has an escaping ray:
    a face with D = 0 and q = 0 upstream has G = 0 on the clip's kink, where central differences
    average the two one-sided slopes."""
    rng, hit, D, Aff, nf, n_ff = _case(2, nf=40, escape_each=True)
    _, par, fl, m_i, cos_i, _ = _chem_case(4, F=nf)
    w = rng.normal(size=nf)
    _, _, _, grads = ofl.radiosity_coverage(D, Aff, m_i, cos_i, par, fl, n_fp=3, w=w)

    def loss(mi, DD):
        return float(w @ ofl.radiosity_coverage(DD, Aff, mi, cos_i, par, fl, n_fp=3)[0])

    for _ in range(2):
        v = rng.normal(size=nf)
        h = 1e-6
        fd = (loss(m_i + h * v, D) - loss(m_i - h * v, D)) / (2 * h)
        assert abs(fd - grads["m_i"] @ v) <= 1e-5 * max(1.0, abs(fd)), (fd, grads["m_i"] @ v)
        fd = (loss(m_i, D + h * v) - loss(m_i, D - h * v)) / (2 * h)
        assert abs(fd - grads["D"] @ v) <= 1e-5 * max(1.0, abs(fd)), (fd, grads["D"] @ v)
