"""The differentiable etch run (petch.diffetch) in the deterministic mode (fixed-point flux sums,
so a run is a function of its inputs to the bit): its forward is run_etch_3d's loop, and the
gradient of the profile loss after the whole run with respect to (Emean, Fflux, Oflux,
rate_scale) against differences with common random numbers (the same Philox streams at
theta +- h) of the same run with each step's surface held (the record of the forward: meshes,
BVHs, warm-start face maps, substep counts), which is the function the adjoint differentiates.

Workload: the small hole at dx 0.2 um (70 x 70 x 120), 3.0 s over 20 steps (the configuration;
one step of 0.15 s for the one-step check), SF6O2 defaults; the loss is the sum over the hole's
slices (8 um below the substrate top to the mask top) of (r(z) - r*(z))^2 with r* the profile
of a run with seed 1234 at perturbed parameters (Emean +10 %, Fflux -10 %, Oflux +10 %,
rate_scale +10 %). Central differences of the run with its surface rebuilt every step (the
plain forward) jump with every change of a marching-cubes cell or a nearest-face choice and
are not derivatives at any step size tried (1e-3 to 3e-2)."""
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


def _held_slopes(dd, theta0, target, zsel, record, span, n_pts, params):
    """Least-squares slope of the loss over theta_k * (1 + t), t in [-span, span], with every
    step's surface held at the record's: the secant of the function the adjoint differentiates."""
    from petch import diffetch as de
    out = {}
    for k in params:
        ts = np.linspace(-span, span, n_pts)
        Ls = []
        for t in ts:
            tp = theta0.clone()
            tp[k] *= 1.0 + t
            Ls.append(de.loss_only(dd, tp, target, zsel, frozen=record))
        out[k] = np.polyfit(ts * abs(float(theta0[k])), Ls, 1)[0]
    return out


def _gradient_case(n_steps, n_rays):
    import torch
    from petch import diffetch as de
    dd = _setup(n_steps=n_steps, n_rays=n_rays)
    zsel = de.zsel_feature(dd, below=8.0)
    tgt_run = _setup(n_steps=n_steps, seed_offset=1234, n_rays=n_rays)
    target = de.profile_target(tgt_run, tgt_run.theta0 * torch.from_numpy(PERTURB))
    theta0 = dd.theta0.clone()
    loss, grad = de.loss_and_gradient(dd, theta0, target, zsel, record=True)
    assert np.isfinite(loss) and loss > 0
    return dd, zsel, target, theta0, grad


def test_one_step_gradient_matches_crn_differences(det):
    """One step, 1M rays: the loss with the step's surface held is smooth at these scales, and
    the adjoint equals its central differences at h = 1e-3 relative (common random numbers)."""
    from petch import diffetch as de
    dd, zsel, target, theta0, grad = _gradient_case(1, 1_000_000)
    rel = []
    for k, name in enumerate(de.PARAMS):
        h = 1e-3 * abs(float(theta0[k]))
        tp, tm = theta0.clone(), theta0.clone()
        tp[k] += h
        tm[k] -= h
        fd = (de.loss_only(dd, tp, target, zsel, frozen=dd.record) -
              de.loss_only(dd, tm, target, zsel, frozen=dd.record)) / (2 * h)
        r = abs(float(grad[k]) - fd) / abs(fd)
        rel.append(r)
        print(f"1 step d loss / d {name}: adjoint {float(grad[k]):+.6e} CRN fd {fd:+.6e} "
              f"rel {r:.5f}", flush=True)
    assert max(rel) < 0.02, rel


def test_twenty_step_gradient_against_the_held_surface_secant(det):
    """The configuration's 20 steps with 4M rays: the adjoint against the secant of the loss
    (21 points over +-1e-4 relative, surfaces held). Between the jumps of the discretisation
    (a node entering the reinit's interface set re-seeds its distance) the increments equal the
    adjoint; the jumps lower the secant by a few percent after 20 steps."""
    from petch import diffetch as de
    dd, zsel, target, theta0, grad = _gradient_case(20, 4_000_000)
    sl = _held_slopes(dd, theta0, target, zsel, dd.record, 1e-4, 21, range(4))
    rel = []
    for k, name in enumerate(de.PARAMS):
        r = abs(float(grad[k]) - sl[k]) / abs(sl[k])
        rel.append(r)
        print(f"20 steps d loss / d {name}: adjoint {float(grad[k]):+.6e} held-surface secant "
              f"{sl[k]:+.6e} rel {r:.4f}", flush=True)
    assert max(rel) < 0.10, rel


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


def test_trench_variant_gradient_in_five_parameters(det):
    """The batched differentiable configuration's trench (60 x 10 x 240, periodic), one step of
    its 30 (0.2 s), 1M rays: the five-parameter gradient (with the F sticking factor betaE, whose
    derivative comes from the species' own path-replay adjoints) against CRN central differences
    at h = 1e-3 relative with the step's surface held."""
    import torch
    from petch import diffetch as de
    from petch.params import Flags
    case = dict(de.CONFIG4, n_steps=1, t_end=de.CONFIG4['t_end'] / de.CONFIG4['n_steps'])
    (par, flags), = de.variant_recipes(1)
    dd = de.DiffEtch(**case, par=par, flags=Flags(**flags), n_ion=1_000_000, n_neu=1_000_000,
                     params=de.PARAMS5)
    tg = de.DiffEtch(**case, par=par, flags=Flags(**flags), n_ion=1_000_000, n_neu=1_000_000,
                     seed_offset=1234, params=de.PARAMS5)
    target = de.profile_target(tg, tg.theta0 * torch.tensor(de.PERTURB5, dtype=torch.float64))
    zsel = de.zsel_feature(dd)
    theta0 = dd.theta0.clone()
    loss, grad = de.loss_and_gradient(dd, theta0, target, zsel, record=True)
    assert np.isfinite(loss) and loss > 0
    rel = []
    for k, name in enumerate(de.PARAMS5):
        h = 1e-3 * abs(float(theta0[k]))
        tp, tm = theta0.clone(), theta0.clone()
        tp[k] += h
        tm[k] -= h
        fd = (de.loss_only(dd, tp, target, zsel, frozen=dd.record) -
              de.loss_only(dd, tm, target, zsel, frozen=dd.record)) / (2 * h)
        r = abs(float(grad[k]) - fd) / abs(fd)
        rel.append(r)
        print(f"trench d loss / d {name}: adjoint {float(grad[k]):+.6e} CRN fd {fd:+.6e} "
              f"rel {r:.5f}", flush=True)
    assert max(rel) < 0.02, rel


def test_variant_driver_runs_checkpointed_gradients(det):
    """variant_gradients: two variants of the trench, 6 steps with checkpoints every 3, 200k
    rays: a finite positive loss and a finite five-parameter gradient per variant."""
    from petch import diffetch as de
    case = dict(de.CONFIG4, n_steps=6, t_end=de.CONFIG4['t_end'] * 6 / de.CONFIG4['n_steps'])
    res, secs = de.variant_gradients(2, n_rays=200_000, checkpoint_every=3, case=case)
    print(f"2 variants x 6 steps: {secs:.2f} s; " + "; ".join(
        f"loss {l:.4e} grad {[round(float(x), 6) for x in g]}" for l, g in res), flush=True)
    assert len(res) == 2
    for loss, g in res:
        assert np.isfinite(loss) and loss > 0
        assert g.shape == (5,) and np.isfinite(g.numpy()).all()
