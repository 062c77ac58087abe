"""The oracle's radiosity adjoint (oracle/flux.py radiosity_adjoint) pinned by what the
mathematics fixes: central differences of the direct solve in float64, the linear-system identity
lambda . D = gG . G, the clip masks, and convergence of the transposed Jacobi sweeps to the
direct adjoint."""
import numpy as np
import pytest

from oracle import flux as ofl


def _case(seed, nf=48, n_ff=16, p_esc=0.3):
    rng = np.random.default_rng(seed)
    hit = rng.integers(0, nf, nf * n_ff).astype(np.int64)
    hit[rng.random(nf * n_ff) < p_esc] = -1
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
