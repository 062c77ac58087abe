"""The differentiable etch run (petch.diffetch) in the deterministic mode (fixed-point flux sums,
so a run is a function of its inputs to the bit): its forward is run_etch_3d's loop, and the
gradient of the profile loss after the whole run with respect to (Emean, Fflux, Oflux,
rate_scale) against central differences with common random numbers (the same Philox streams at
theta +- h) of the same run with each step's surface held (the record of the forward: meshes,
BVHs, warm-start face maps, substep counts), which is the function the adjoint differentiates.

Workload: the small hole at dx 0.2 um (70 x 70 x 120), 20 steps over 3.0 s, 200k ions and 200k
F + 200k O neutrals per fixed-point round, SF6O2 defaults; the loss is the sum over the hole's
slices (8 um below the substrate top to the mask top) of (r(z) - r*(z))^2 with r* the profile
of a run with seed 1234 at perturbed parameters (Emean +10 %, Fflux -10 %, Oflux +10 %,
rate_scale +10 %)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASE = dict(Lx=14.0, Ly=14.0, Lz=24.0, dx=0.2, trench_width=6.0, mask_th=2.0, sub_top=18.0,
            hole=True, t_end=3.0, n_steps=20)
PERTURB = np.array([1.10, 0.90, 1.10, 1.10])


@pytest.fixture
def det():
    from petch import ops
    ops.set_deterministic(True)
    yield
    ops.set_deterministic(False)


def _setup(n_steps=20, seed_offset=0, n_rays=200_000):
    from petch.api import SF6O2
    from petch import diffetch as de
    from petch.params import Flags
    m = SF6O2()
    case = dict(CASE, n_steps=n_steps, t_end=CASE['t_end'] * n_steps / CASE['n_steps'])
    return de.DiffEtch(**case, par=dict(m.par), flags=Flags(**m.flags), n_ion=n_rays,
                       n_neu=n_rays, seed_offset=seed_offset)


def test_forward_is_the_etch_loop(det):
    """DiffEtch.run at the recipe's own parameters is run_etch_3d (same seeds, warm start,
    rounds, substeps, reinit) to the bit."""
    import torch
    from petch import threed as t3
    from petch.api import SF6O2
    from petch.params import Flags
    dd = _setup(n_steps=6, n_rays=100_000)
    with torch.no_grad():
        phi = dd.run(dd.theta0).cpu().numpy()
    m = SF6O2()
    case = dict(CASE, n_steps=6, t_end=CASE['t_end'] * 6 / 20)
    g = t3.run_etch_3d(**case, par=dict(m.par), flags=Flags(**m.flags), n_ion=100_000,
                       n_neu=100_000, reinit_method="fsm", verbose=False)
    print(f"max |phi - run_etch_3d phi| = {np.abs(phi - g['phi']).max():.2e} um", flush=True)
    assert np.array_equal(phi, g['phi'])


def test_profile_gradient_matches_held_surface_differences(det):
    import torch
    from petch import diffetch as de
    dd = _setup()
    zsel = de.zsel_feature(dd, below=8.0)
    tgt_run = _setup(seed_offset=1234)
    target = de.profile_target(tgt_run, tgt_run.theta0 * torch.from_numpy(PERTURB))
    theta0 = dd.theta0.clone()
    loss, grad = de.loss_and_gradient(dd, theta0, target, zsel, record=True)
    print(f"loss {loss:.6e}; substeps {sorted(dd.record['nsub'].items())}", flush=True)
    assert np.isfinite(loss) and loss > 0
    rel = []
    for k, name in enumerate(de.PARAMS):
        h = 1e-3 * abs(float(theta0[k]))
        tp, tm = theta0.clone(), theta0.clone()
        tp[k] += h
        tm[k] -= h
        fd = (de.loss_only(dd, tp, target, zsel, frozen=dd.record) -
              de.loss_only(dd, tm, target, zsel, frozen=dd.record)) / (2 * h)
        an = float(grad[k])
        r = abs(an - fd) / max(abs(fd), 1e-300)
        rel.append(r)
        print(f"d loss / d {name}: adjoint {an:+.6e} CRN fd (surface held) {fd:+.6e} rel {r:.4f}",
              flush=True)
    assert max(rel) < 0.05, rel


def test_checkpointed_gradient_equals_the_full_graph(det):
    """Segments of 5 steps recomputed in the backward pass give the gradient of the full graph:
    the recompute replays the forward's surfaces and redraws the same streams."""
    import torch
    from petch import diffetch as de
    dd = _setup(n_steps=10, n_rays=100_000)
    zsel = de.zsel_feature(dd, below=8.0)
    with torch.no_grad():
        target = de.profile_target(dd, dd.theta0 * torch.from_numpy(PERTURB))
    l1, g1 = de.loss_and_gradient(dd, dd.theta0, target, zsel)
    l2, g2 = de.loss_and_gradient(dd, dd.theta0, target, zsel, checkpoint_every=5)
    print(f"full graph {g1.tolist()} checkpointed {g2.tolist()}", flush=True)
    assert l2 == l1
    np.testing.assert_allclose(g2.numpy(), g1.numpy(), rtol=1e-6)
