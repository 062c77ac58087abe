"""Etch-product redeposition (mc_redep_3d, threed.py:1582-1605) on the GPU through the C ABI
(petch_trace_redeposition) against the oracle's C restatement, and the etch loops with
flags.redeposition.

Both sides draw every ray's face, lobe and bounces from the same Philox streams, so with equal
seeds the per-face results agree to float rounding (same-seed L1 bar as the RR neutral trace);
with independent seeds the per-face z-scores look like N(0, 1). The closed forms: a closed
cavity with s_redep = 1 returns the emitted total; zero, negative or non-finite rates emit
nothing."""
import numpy as np
import pytest

from oracle import levelset as ol
from oracle import native as on
from petch.params import PAR, Flags
from test_gpu_traces import GEOMS, N_BATCH, SAME_SEED_BAR, _check

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda", 0)


def _surface(phi, dx):
    """(bvh, centroids, areas, gas normals) on the device and the oracle's BVH of one surface."""
    import torch
    from petch import ops
    tphi = torch.from_numpy(np.ascontiguousarray(phi, np.float32)).to(_dev())
    v, f = ops.marching_cubes(tphi, dx)
    cen, area, nrm = ops.mesh_attrs(v, f)
    gn = ops.gas_normals(tphi, dx, cen, nrm)
    bvh = ops.Bvh(_dev()).build(v, f)
    return bvh, cen, area, gn, on.OracleBvh(v.cpu().numpy(), f.cpu().numpy())


def _hole():
    g = GEOMS["hole"]
    phi = ol.init_sdf(g['Lx'], g['Ly'], g['Lz'], g['dx'], g['w'], g['mask'], g['sub'], g['hole'])
    return _surface(phi, g['dx'])


def test_redeposition_matches_oracle():
    import torch
    from petch import ops
    bvh, cen, area, gn, obvh = _hole()
    F = bvh.nf
    rng = np.random.default_rng(8)
    rate = rng.uniform(0.0, 1.0, F).astype(np.float32)
    rate[rng.random(F) < 0.3] = 0.0
    tr = torch.from_numpy(rate).to(_dev())
    c, g, a = cen.cpu().numpy(), gn.cpu().numpy(), area.cpu().numpy()
    n, s = 40000, 0.5
    print(f"redeposition hole nf={F}", flush=True)
    gs = ops.trace_redeposition(bvh, cen, gn, area, tr, s, n, 77).cpu().numpy()
    os_ = obvh.trace_redeposition(c, g, a, rate, s, n, 77)
    gb = [ops.trace_redeposition(bvh, cen, gn, area, tr, s, n, 500 + k).cpu().numpy()
          for k in range(N_BATCH)]
    ob = [obvh.trace_redeposition(c, g, a, rate, s, n, 900 + k) for k in range(N_BATCH)]
    assert bvh.status() == 0
    assert os_.sum() > 0
    _check(gs, os_, gb, ob, "redeposition/hole", bar=SAME_SEED_BAR)


def test_closed_cavity_conserves_and_empty_rates_emit_nothing():
    import torch
    from petch import ops
    R, dx = 2.0, 0.25
    m = int(np.ceil(2 * R / dx)) + 5
    ctr = 0.5 * (m - 1) * dx
    x = np.arange(m) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    phi = np.sqrt((X - ctr) ** 2 + (Y - ctr) ** 2 + (Z - ctr) ** 2) - R
    bvh, cen, area, gn, _ = _surface(phi, dx)
    F = bvh.nf
    rate = torch.rand(F, device=_dev(), generator=torch.Generator(_dev()).manual_seed(3))
    out = ops.trace_redeposition(bvh, cen, gn, area, rate, 1.0, 30000, 5)
    a = area.double().clamp(min=1e-9)
    tot = float((rate.double() * area.double()).sum())
    got = float((out.double() * a).sum())
    print(f"closed cavity nf={F}: returned / emitted {got / tot:.7f}", flush=True)
    assert abs(got / tot - 1.0) < 1e-5
    bad = torch.zeros(F, device=_dev())
    bad[0], bad[1], bad[2] = float("nan"), float("inf"), -1.0
    z = ops.trace_redeposition(bvh, cen, gn, area, bad, 0.5, 1000, 5)
    assert int((z != 0).sum()) == 0


def test_deterministic_mode_repeats_bit_for_bit():
    import torch
    from petch import ops
    bvh, cen, area, gn, _ = _hole()
    rate = torch.rand(bvh.nf, device=_dev(), generator=torch.Generator(_dev()).manual_seed(1))
    ops.set_deterministic(True)
    try:
        a = ops.trace_redeposition(bvh, cen, gn, area, rate, 0.5, 30000, 11)
        b = ops.trace_redeposition(bvh, cen, gn, area, rate, 0.5, 30000, 11)
    finally:
        ops.set_deterministic(False)
    c = ops.trace_redeposition(bvh, cen, gn, area, rate, 0.5, 30000, 11)
    assert torch.equal(a, b)
    np.testing.assert_allclose(a.cpu().numpy(), c.cpu().numpy(), rtol=1e-5, atol=1e-6)


FL = Flags(chemistry="belen", yield_angular="viennaps", coverage_sticking=True,
           ion_reflection=True, redeposition=True)
SPEC = dict(Lx=3.0, Ly=0.9, Lz=6.5, dx=0.15, trench_width=0.9, mask_th=0.3, sub_top=6.0,
            n_ion=4000, n_neu=2000)


def test_loop_runs_with_redeposition_and_it_slows_the_etch():
    """run_etch_3d with flags.redeposition: finite, etches, and a strong k_redep etches less deep
    than none; the reference-API mc_redep_3d returns a non-negative (F,) flux on the result."""
    from petch import threed as t3
    print("loop redeposition", flush=True)
    kw = dict(SPEC, t_end=0.3, n_steps=4, verbose=False, reinit_method="fsm")
    g1 = t3.run_etch_3d(par=dict(PAR, k_redep=1.0), flags=FL, **kw)
    g5 = t3.run_etch_3d(par=dict(PAR, k_redep=20.0), flags=FL, **kw)
    g0 = t3.run_etch_3d(par=dict(PAR), flags=Flags(**{**FL.to_dict(), "redeposition": False}),
                        **kw)
    d1, d5, d0 = (t3.max_depth_3d(g) for g in (g1, g5, g0))
    print(f"depths: no redep {d0:.3f}, k_redep 1 {d1:.3f}, k_redep 20 {d5:.3f}", flush=True)
    assert np.isfinite(g1['phi']).all() and d1 > 0.0
    assert d5 < d0
    verts, faces, cen, areas = t3.extract_mesh_3d(g1['phi'], g1['dx'])
    mesh = t3.Mesh(verts, faces)
    gn = t3._gas_normals(verts, faces, cen, g1)
    Rf = t3.mc_redep_3d(mesh, cen, areas, gn, np.ones(len(faces)), n_redep=5000, s_redep=0.5,
                        seed=2)
    assert Rf.shape == (len(faces),) and (Rf >= 0).all() and Rf.sum() > 0


def test_batch_first_step_with_redeposition_equals_single_run():
    """One batched step of a feature with redeposition beside one without, each against its
    own run_etch_3d (phi within 1e-2 dx, as the batch tests bound float-order differences)."""
    from petch import batch, threed as t3
    specs = [dict(SPEC, par=dict(PAR, k_redep=5.0), flags=FL),
             dict(SPEC, par=dict(PAR), flags=Flags(**{**FL.to_dict(), "redeposition": False}),
                  seed_offset=3)]
    geos = batch.run_etch_3d_batch(specs, t_end=0.1, n_steps=1)
    for sp, gb in zip(specs, geos):
        gs = t3.run_etch_3d(**sp, t_end=0.1, n_steps=1, reinit_method="fsm", verbose=False)
        band = np.abs(gs['phi']) < 3 * sp['dx']
        d = np.abs(gb['phi'] - gs['phi'])[band].max()
        print(f"batch vs single (redeposition={sp['flags'].redeposition}): {d:.3e}", flush=True)
        assert d <= 0.01 * sp['dx']
