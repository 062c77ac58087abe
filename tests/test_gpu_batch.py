"""Batched features: one launch per trace for many features must give each feature what its own
single-feature launch gives (same Philox streams; only the order of float atomic adds differs),
and the lockstep loop must follow the single-feature runs."""
import numpy as np
import pytest

from oracle import levelset as ol
from oracle import native as on
from petch.params import PAR, Flags

pytestmark = pytest.mark.gpu

FEATS = [
    dict(Lx=8.0, Ly=8.0, Lz=14.0, dx=0.25, w=3.0, mask=8.0, sub=4.0, hole=True, periodic=0),
    dict(Lx=6.0, Ly=1.5, Lz=12.0, dx=0.15, w=1.5, mask=7.0, sub=3.0, hole=False, periodic=1),
    dict(Lx=5.0, Ly=5.0, Lz=9.0, dx=0.2, w=2.0, mask=4.0, sub=3.0, hole=True, periodic=0),
]


def _surface(g):
    import torch
    from petch import ops
    phi = ol.init_sdf(g['Lx'], g['Ly'], g['Lz'], g['dx'], g['w'], g['mask'], g['sub'], g['hole'])
    v, f = on.marching_cubes(phi.astype(np.float32), g['dx'])
    dev = torch.device("cuda", 0)
    bvh = ops.Bvh(dev).build(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev))
    box = ops.make_box(g['Lx'], g['Ly'], g['Lz'], g['Lz'] - g['dx'], g['periodic'])
    return bvh, box, len(f)


def test_neutral_batch_equals_single_launches():
    import torch
    from petch import ops
    src = ops.make_source(0)
    items, singles = [], []
    for k, g in enumerate(FEATS):
        bvh, box, nf = _surface(g)
        bare = torch.from_numpy(np.random.default_rng(k).uniform(0.2, 1, nf).astype(np.float32)).cuda()
        it = dict(bvh=bvh, box=box, bare=bare, beta_a=0.7, n_a=20000 + 5000 * k, seed_a=100 + k,
                  beta_b=1.0, n_b=15000, seed_b=200 + k)
        items.append(it)
        singles.append(ops.trace_neutral_rr_pair(bvh, box, src, bare, 0.7, it['n_a'], it['seed_a'],
                                                 1.0, it['n_b'], it['seed_b']))
    # an empty surface takes part without rays and gets zero outputs
    empty = ops.Bvh(torch.device("cuda", 0)).build(torch.zeros((0, 3), device="cuda"),
                                                   torch.zeros((0, 3), dtype=torch.int32, device="cuda"))
    items.append(dict(bvh=empty, box=items[0]['box'], bare=torch.ones(1, device="cuda"), beta_a=0.7,
                      n_a=1000, seed_a=1, beta_b=1.0, n_b=1000, seed_b=2))
    outs = ops.trace_neutral_rr_batch(src, items)
    for (fa, fb), (sa, sb) in zip(outs, singles):
        for x, y in ((fa, sa), (fb, sb)):
            x, y = x.cpu().numpy(), y.cpu().numpy()
            assert y.sum() > 0
            np.testing.assert_allclose(x, y, rtol=2e-5, atol=1e-3)
    assert outs[-1][0].numel() == 0


def test_ion_batch_equals_single_launches():
    import torch
    from petch import flux, ops
    items, singles = [], []
    for k, g in enumerate(FEATS):
        bvh, box, nf = _surface(g)
        par = dict(PAR, Emean=60.0 + 40.0 * k)
        src = ops.make_source(2, 0.0, 500.0 + 250 * k) if k != 1 else ops.make_source(1, 0.04, 0.0)
        it = dict(bvh=bvh, box=box, src=src, ion=flux.ion_params(par), n=20000 + 7000 * k,
                  seed=300 + k)
        items.append(it)
        singles.append(ops.trace_ion_yield(bvh, box, src, it['ion'], it['n'], it['seed']))
    outs = ops.trace_ion_yield_batch(items)
    for a, b in zip(outs, singles):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert b[3].sum() > 0
        np.testing.assert_allclose(a, b, rtol=2e-5, atol=1e-3)


def test_batch_rejects_bad_arguments():
    import torch
    from petch import ops
    bvh, box, nf = _surface(FEATS[0])
    bare = torch.ones(nf, device="cuda")
    it = dict(bvh=bvh, box=box, bare=bare, beta_a=0.7, n_a=-5, seed_a=1, beta_b=1.0, n_b=10, seed_b=2)
    with pytest.raises(Exception):
        ops.trace_neutral_rr_batch(ops.make_source(0), [it])


def _specs():
    fl = Flags(chemistry="belen", yield_angular="viennaps", coverage_sticking=True,
               warm_start_coverage=True, ion_reflection=True)
    specs = []
    for k, (g, rs) in enumerate(zip(FEATS, (0.05, 0.1, 0.08))):
        par = dict(PAR, rate_scale=rs, Emean=80.0 + 30 * k, periodic_y=g['periodic'])
        specs.append(dict(Lx=g['Lx'], Ly=g['Ly'], Lz=g['Lz'], dx=g['dx'], trench_width=g['w'],
                          mask_th=2.0, sub_top=g['Lz'] - 4.0, hole=g['hole'], par=par, flags=fl,
                          n_ion=20000, n_neu=20000, seed_offset=10 * k))
    return specs


def test_run_etch_3d_batch_first_step_equals_single_runs():
    """One step of three features of different grids and recipes in lockstep against their own
    run_etch_3d. The fluxes agree up to the order of float atomic adds (a Russian-roulette draw
    within a few ulp of its threshold may flip, moving one ray), so phi agrees to ~1e-3 dx;
    a wrong stream or feature mapping would move it by the Monte-Carlo noise of V dt, ~0.1 dx."""
    from petch import batch, threed as t3
    specs = _specs()
    geos = batch.run_etch_3d_batch(specs, t_end=0.1, n_steps=1)
    for sp, gb in zip(specs, geos):
        gs = t3.run_etch_3d(**sp, t_end=0.1, n_steps=1, reinit_method="fsm", verbose=False)
        gi = t3.run_etch_3d(**{**sp, "seed_offset": sp["seed_offset"] + 1}, t_end=0.1,
                            n_steps=1, reinit_method="fsm", verbose=False)
        band = np.abs(gs['phi']) < 3 * sp['dx']
        d = np.abs(gb['phi'] - gs['phi'])[band].max()
        di = np.abs(gi['phi'] - gs['phi'])[band].max()
        print(f"feature dx={sp['dx']}: one step, batch-vs-single max |dphi| {d:.3e}; "
              f"another seed {di:.3e}", flush=True)
        assert d <= 0.01 * sp['dx']
        assert di > 10 * d                    # the check can tell a different stream apart


def test_run_etch_3d_batch_depths_follow_single_runs():
    """Four steps: the depths agree to one cell (after the first step float-order differences
    are amplified like any two runs with the same seeds)."""
    from petch import batch, threed as t3
    specs = _specs()
    geos = batch.run_etch_3d_batch(specs, t_end=0.4, n_steps=4)
    for sp, gb in zip(specs, geos):
        gs = t3.run_etch_3d(**sp, t_end=0.4, n_steps=4, reinit_method="fsm", verbose=False)
        assert abs(t3.max_depth_3d(gb) - t3.max_depth_3d(gs)) <= sp['dx'] + 1e-9
        assert abs(t3.center_depth_3d(gb) - t3.center_depth_3d(gs)) <= sp['dx'] + 1e-9
