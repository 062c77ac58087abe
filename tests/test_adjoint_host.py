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
