"""Host-side checks of the differentiable pieces without a GPU: the adjoint wrappers reject
inputs of the wrong size before anything reaches the library, the autograd module exposes its
pieces, and the oracle's radiosity adjoint is linear."""
import numpy as np
import pytest
import torch

from oracle import flux as ofl
from petch import adjoint, ops


def test_adjoint_module_exposes_the_differentiable_pieces():
    for name in ("advect", "radiosity", "belen", "radiosity_coverage", "neutral_rr",
                 "neutral_rr_pair", "finish", "coupled_coverage", "surface_rate", "extend",
                 "advance", "ion_yield", "ion_terms"):
        assert callable(getattr(adjoint, name)), name


@pytest.mark.parametrize("which", ["hitface", "G", "bare", "gm"])
def test_radiosity_adjoint_rejects_short_inputs(which):
    nf, n_ff = 10, 4
    args = dict(hitface=torch.zeros(nf * n_ff, dtype=torch.int32), G=torch.zeros(nf, dtype=torch.float64),
                bare=torch.zeros(nf), gm=torch.zeros(nf))
    args[which] = args[which][:-1]
    with pytest.raises(ValueError, match=which):
        ops.radiosity_adjoint(args["hitface"], nf, n_ff, args["G"], args["bare"], 0.7, args["gm"])


def test_smooth_and_normalize_adjoints_reject_short_inputs():
    g = torch.zeros(6)
    with pytest.raises(ValueError, match="3 values per face"):
        ops.smooth_adjoint(torch.zeros(17, dtype=torch.int32), torch.zeros(18), 1, 1.0, True, g)
    with pytest.raises(ValueError, match="areas"):
        ops.normalize_adjoint(torch.zeros(6), torch.zeros(5), 1.0, 0.0, 8.0, g)


def test_extend_adjoint_rejects_a_gradient_of_another_shape():
    phi = torch.zeros((4, 5, 6))
    with pytest.raises(ValueError, match="gFs shape"):
        ops.extend_adjoint(None, phi, 0.1, 0.3, torch.zeros((4, 5, 7)))


def test_oracle_radiosity_adjoint_is_linear_in_the_output_gradient():
    """The oracle's radiosity adjoint (a transposed linear solve) scales with the output
    gradient."""
    rng = np.random.default_rng(3)
    nf, n_ff = 30, 8
    hit = rng.integers(0, nf, nf * n_ff)
    hit[::n_ff] = -1
    D, Aff = ofl.radiosity_matrix(hit, nf, n_ff)
    bare = rng.uniform(0.2, 1.0, nf)
    G = ofl.radiosity_direct(D, Aff, bare, 0.7)
    gm = rng.normal(size=nf)
    gD, gb = ofl.radiosity_adjoint(D, Aff, G, bare, 0.7, gm)
    gD2, gb2 = ofl.radiosity_adjoint(D, Aff, G, bare, 0.7, 2.0 * gm)
    np.testing.assert_allclose(gD2, 2.0 * gD, rtol=1e-12)
    np.testing.assert_allclose(gb2, 2.0 * gb, rtol=1e-12)


@pytest.mark.parametrize("model", ["langmuir", "belen"])
def test_surface_rate_adjoint_with_yields_follows_the_chemistry(monkeypatch, model):
    """The Langmuir rate reads m_i even when ion yields are passed (only the Belen rate reads the
    yields, belen.py:34), so its adjoint must hand the library an m_i gradient buffer and return
    zero yield gradients; with Belen the yields get the gradient and m_i none. The library call
    is intercepted: this checks which buffers the wrapper passes and returns."""
    from petch.chemistry import chem_params
    from petch.params import PAR, Flags
    calls = []

    class FakeCtx:
        def call(self, name, *args):
            calls.append((name, args))

    monkeypatch.setattr(ops, "_ctx", lambda t: FakeCtx())
    monkeypatch.setattr(ops, "ptr", lambda t, *a, **k: t)
    F = 7
    z = torch.zeros(F)
    chem = chem_params(dict(PAR), Flags(chemistry=model))
    iy = (torch.ones(F), torch.ones(F), torch.ones(F))
    gm_F, gm_O, gm_i, g_sp, g_ie, g_ip = ops.surface_rate_adjoint(
        z, z, z, z, iy, torch.zeros(F, 3), ops.MaskGeom(), chem, torch.ones(F))
    (name, args), = calls
    assert name == "petch_surface_rate_adjoint"
    passed_gm_i, passed_g = args[15], args[16:19]
    if model == "langmuir":
        assert gm_i is not None and passed_gm_i is gm_i
        for g in (g_sp, g_ie, g_ip):
            assert g is not None and float(g.abs().sum()) == 0.0
        assert all(p is None or float(p.abs().sum()) == 0.0 for p in passed_g)
    else:
        assert gm_i is None and passed_gm_i is None
        assert all(p is g for p, g in zip(passed_g, (g_sp, g_ie, g_ip)))
