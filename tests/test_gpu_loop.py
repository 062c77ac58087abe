"""Flux orchestration and the whole etch loop against the oracle's CPU restatement.

The GPU and the oracle draw the same Philox streams, so the coupled flux of one step agrees to
a few parts in 1e4 (only transcendental rounding separates the paths); over a whole run the
final depths agree within the run-to-run spread of independent seeds."""
import numpy as np
import pytest

from oracle import flux as ofl
from oracle import levelset as ol
from oracle import native as on
from petch.params import PAR, Flags

pytestmark = pytest.mark.gpu

FAITHFUL = Flags(chemistry="belen", yield_angular="viennaps", coverage_sticking=True,
                 warm_start_coverage=True, sampling="sobol", ion_reflection=True)


def _dev():
    import torch
    return torch.device("cuda", 0)


def _geo(name):
    if name == "hole":
        g = dict(Lx=8.0, Ly=8.0, Lz=14.0, dx=0.25, w=3.0, mask=8.0, sub=4.0, hole=True)
    else:
        g = dict(Lx=6.0, Ly=1.5, Lz=12.0, dx=0.15, w=1.5, mask=7.0, sub=3.0, hole=False)
    phi = ol.init_sdf(g['Lx'], g['Ly'], g['Lz'], g['dx'], g['w'], g['mask'], g['sub'], g['hole'])
    geo = dict(Lx=g['Lx'], Ly=g['Ly'], Lz=g['Lz'], dx=g['dx'])
    return phi.astype(np.float32), geo


def _rel_l1(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30)


@pytest.mark.parametrize("name,flags_kw", [
    ("hole", dict()),
    ("trench", dict()),
    ("hole", dict(ion_reflection=False)),
])
def test_coupled_flux_matches_oracle(name, flags_kw):
    import torch
    from petch import flux
    phi, geo = _geo(name)
    par = dict(PAR)
    par['periodic_y'] = 1 if name == "trench" else 0
    fl = Flags(**{**FAITHFUL.to_dict(), **flags_kw})
    mesh = flux.SurfaceMesh(_dev()).build(torch.from_numpy(phi).to(_dev()), geo['dx'])
    v, f = mesh.verts.cpu().numpy(), mesh.faces.cpu().numpy()
    M = ofl.OracleMesh(v, f)
    n = 30000
    for bare_init in (None, np.random.default_rng(0).uniform(0.2, 1, len(f))):
        bi = None if bare_init is None else torch.as_tensor(bare_init, dtype=torch.float32, device=_dev())
        g = flux.coupled_flux(mesh, geo, par, n, n, 3, fl, bare_init=bi)
        o = ofl.coupled_flux(M, geo, par, n, n, 3, fl,
                             bare_init=None if bare_init is None else bare_init.astype(np.float32))
        for k in ("m_i", "m_F", "m_O", "cos_i", "bare"):
            r = _rel_l1(g[k].cpu().numpy(), o[k])
            print(f"{name} {flags_kw} warm={bare_init is not None} {k}: rel L1 {r:.2e}", flush=True)
            assert r < 5e-4, k
        if o['ion_yield'] is not None:
            for a, b in zip(g['ion_yield'], o['ion_yield']):
                assert _rel_l1(a.cpu().numpy(), b) < 5e-4


def test_radiosity_flux_matches_oracle():
    import torch
    from petch import flux
    phi, geo = _geo("hole")
    par = dict(PAR)
    mesh = flux.SurfaceMesh(_dev()).build(torch.from_numpy(phi).to(_dev()), geo['dx'])
    M = ofl.OracleMesh(mesh.verts.cpu().numpy(), mesh.faces.cpu().numpy())
    g = flux.radiosity_flux(mesh, torch.from_numpy(phi).to(_dev()), geo, par, 30000, 2, FAITHFUL)
    o = ofl.radiosity_flux(M, phi, geo, par, 30000, 2, FAITHFUL)
    for k in ("m_i", "m_F", "m_O", "cos_i", "bare"):
        r = _rel_l1(g[k].cpu().numpy(), o[k])
        print(f"radiosity {k}: rel L1 {r:.2e}", flush=True)
        assert r < 1e-5, k


def test_uncoupled_flux_matches_oracle():
    import torch
    from petch import flux
    phi, geo = _geo("hole")
    par = dict(PAR)
    mesh = flux.SurfaceMesh(_dev()).build(torch.from_numpy(phi).to(_dev()), geo['dx'])
    M = ofl.OracleMesh(mesh.verts.cpu().numpy(), mesh.faces.cpu().numpy())
    for refl in (False, True):
        g = flux.uncoupled_flux(mesh, geo, par, 30000, 30000, 4, refl)
        o = ofl.uncoupled_flux(M, geo, par, 30000, 30000, 4, refl)
        for k in ("m_i", "m_F", "m_O", "cos_i"):
            r = _rel_l1(g[k].cpu().numpy(), o[k])
            print(f"uncoupled reflection={refl} {k}: rel L1 {r:.2e}", flush=True)
            assert r < 5e-4, k


SMOKE = dict(Lx=3.0, Ly=0.9, Lz=6.5, dx=0.15, trench_width=0.9, mask_th=0.3, sub_top=6.0,
             t_end=1.5, n_steps=20)


def test_reference_smoke_loop():
    """tests/test_smoke.py:37-50 of the reference (periodic trench, 20 steps, 8k rays, fsm):
    finite phi, max_depth > 0.3, center_depth > 0.1."""
    from petch import threed as t3
    par = dict(PAR)
    par['rate_scale'] = 0.1
    par['periodic_y'] = 1
    geo = t3.run_etch_3d(**SMOKE, hole=False, par=dict(par), flags=FAITHFUL, n_ion=8000,
                         n_neu=8000, reinit_method="fsm", verbose=False)
    md, cd = t3.max_depth_3d(geo), t3.center_depth_3d(geo)
    assert np.isfinite(geo['phi']).all()
    assert md > 0.3 and cd > 0.1
    assert geo['timings']['nsub_max'] >= 1


def _floor_depth(phi, dx, sub_top):
    """Sub-cell depth of the trench floor along the centre plane x = Lx/2: in each column the
    first gas-to-solid crossing below the top, linearly interpolated, averaged over y."""
    phi = np.asarray(phi, np.float64)
    col = phi[phi.shape[0] // 2]                       # (ny, nz)
    zs = []
    for p in col:
        k = len(p) - 1
        while k > 0 and p[k] < 0.0:
            k -= 1
        if p[k] < 0.0 or k == len(p) - 1:
            continue
        zs.append((k + p[k] / (p[k] - p[k + 1])) * dx)
    if not zs:                                         # etched through to the domain bottom
        return float("nan")
    return sub_top - float(np.mean(zs))


# The smoke trench of tests/test_smoke.py:37-50 with a 12 um substrate under the mask instead of
# 6 um, so that even at rate_scale 0.1 the front ends 40 cells above the domain bottom: the
# comparison stays clear of the clamped boundary stencils and the mesh opening at the bottom.
SMOKE_DEEP = dict(SMOKE, Lz=12.5, sub_top=12.0)


@pytest.mark.parametrize("rate_scale", [0.1, 0.01])
def test_smoke_loop_oracle_parity(rate_scale):
    """The smoke trench's depth distribution against the oracle's: eight seeds on each side.

    Both sides draw the same Philox streams and per step agree closely (same fluxes to float
    rounding, identical nearest-face extension), but 20 coupled steps amplify the float-atomic
    and float32-vs-float64 differences, so single runs of one seed end a cell or so apart; the
    test compares the seed distributions instead. The interpolated floor depth (not quantised)
    must agree in mean within 3 standard errors of the difference (each side's spread floored
    at 0.02 dx); the max and center depths are whole grid cells, so their means must agree
    within one cell."""
    from petch import threed as t3
    par = dict(PAR)
    par['rate_scale'] = rate_scale
    par['periodic_y'] = 1
    dx = SMOKE_DEEP['dx']
    seeds = tuple(100 * k for k in range(8))
    g_md, g_cd, g_fd, o_md, o_cd, o_fd = [], [], [], [], [], []
    for s in seeds:
        gg = t3.run_etch_3d(**SMOKE_DEEP, hole=False, par=dict(par), flags=FAITHFUL,
                            n_ion=8000, n_neu=8000, reinit_method="fsm", verbose=False,
                            seed_offset=s)
        g_md.append(t3.max_depth_3d(gg))
        g_cd.append(t3.center_depth_3d(gg))
        g_fd.append(_floor_depth(gg['phi'], dx, SMOKE_DEEP['sub_top']))
        og = ofl.run_etch_3d(**SMOKE_DEEP, hole=False, par=dict(par), flags=FAITHFUL,
                             n_ion=8000, n_neu=8000, seed_offset=s)
        o_md.append(ol.max_depth_3d(og))
        o_cd.append(ol.depth3d(og))
        o_fd.append(_floor_depth(og['phi'], dx, SMOKE_DEEP['sub_top']))
    g_fd, o_fd = np.array(g_fd), np.array(o_fd)
    assert np.all(np.isfinite(g_fd)) and np.all(np.isfinite(o_fd))
    # the front stays well above the domain bottom
    assert max(np.max(g_fd), np.max(o_fd)) < SMOKE_DEEP['sub_top'] - 5 * dx
    sg = max(np.std(g_fd, ddof=1), 0.02 * dx)
    so = max(np.std(o_fd, ddof=1), 0.02 * dx)
    se = np.sqrt(sg ** 2 / len(seeds) + so ** 2 / len(seeds))
    print(f"rate_scale {rate_scale}: gpu max {np.mean(g_md):.3f} center {np.mean(g_cd):.3f} "
          f"floor {np.mean(g_fd):.4f}+-{np.std(g_fd):.4f} | oracle max {np.mean(o_md):.3f} "
          f"center {np.mean(o_cd):.3f} floor {np.mean(o_fd):.4f}+-{np.std(o_fd):.4f} | "
          f"diff {np.mean(g_fd) - np.mean(o_fd):+.4f} = {(np.mean(g_fd) - np.mean(o_fd)) / se:+.2f} se",
          flush=True)
    assert abs(np.mean(g_fd) - np.mean(o_fd)) <= 3 * se
    assert abs(np.mean(g_md) - np.mean(o_md)) <= dx
    assert abs(np.mean(g_cd) - np.mean(o_cd)) <= dx


def test_api_hole_runs_and_saves(tmp_path):
    """tests/test_smoke.py:55-63 of the reference through Domain / SF6O2 / Process / Result."""
    import petch
    dom = petch.Domain.trench(extent=3.0, dx=0.15, width=0.9, mask=0.3, depth=6.0)
    res = petch.Process(dom, petch.SF6O2(rate_scale=0.1), duration=1.5).run(steps=20)
    assert res.max_depth > 0.3 and res.aspect_ratio > 0.3 and res.wall_time > 0.0
    v, f = res.mesh
    assert len(v) > 0 and len(f) > 0
    p = res.save(str(tmp_path / "s.vtk"))
    txt = open(p).read()
    assert txt.startswith("# vtk DataFile Version 3.0") and f"POLYGONS {len(f)} {4 * len(f)}" in txt
    p = res.save(str(tmp_path / "s.obj"))
    assert sum(1 for ln in open(p) if ln.startswith("f ")) == len(f)


def test_capped_coverage_trace_runs_in_both_loops():
    """par['cov_transport'] = 'cap' (threed.py:881-885): the coupled flux's neutrals come from
    the capped trace; the loop and the batched loop run, etch, and agree after one step."""
    from petch import batch, threed as t3
    par = dict(PAR, cov_transport="cap", n_reemit_cov=30)
    kw = dict(SMOKE, n_steps=1, t_end=SMOKE['t_end'] / SMOKE['n_steps'])
    print("loop cov_transport=cap", flush=True)
    g = t3.run_etch_3d(**kw, par=par, flags=FAITHFUL, n_ion=4000, n_neu=4000, verbose=False,
                       reinit_method="fsm")
    assert np.isfinite(g['phi']).all() and t3.max_depth_3d(g) > 0.0
    spec = {k: v for k, v in kw.items() if k not in ("t_end", "n_steps")}
    gb, = batch.run_etch_3d_batch([dict(spec, par=par, flags=FAITHFUL, n_ion=4000, n_neu=4000)],
                                  t_end=kw['t_end'], n_steps=1)
    band = np.abs(g['phi']) < 3 * kw['dx']
    assert np.abs(gb['phi'] - g['phi'])[band].max() <= 0.01 * kw['dx']


def test_differentiable_step_matches_finite_differences():
    """A loss on the level set after a differentiable step on a fixed trench surface: the
    radiosity coverage loop (two rounds), the surface rate, the extension and two advection
    substeps (adjoint.radiosity_coverage + adjoint.advance). The whole forward is deterministic
    here, so its gradient in m_i must match float64 central differences of the float32 loss
    along random directions (5 %, for float32 rounding of phi)."""
    import torch
    from petch import adjoint, flux, ops
    from petch.chemistry import chem_params
    phi, geo = _geo("trench")
    dx = geo['dx']
    dev = _dev()
    tphi = torch.from_numpy(phi).to(dev)
    mesh = flux.SurfaceMesh(dev).build(tphi, dx)
    gn = ops.gas_normals(tphi, dx, mesh.centroids, mesh.normals)
    box = ops.make_box(geo['Lx'], geo['Ly'], geo['Lz'], geo['Lz'] - dx, 1)
    n_ff = 64
    hit = ops.trace_ff(mesh.bvh, box, mesh.centroids, gn, n_ff, 21)
    row_ptr, col, D = ops.radiosity_assemble(hit, mesh.nf, n_ff)
    par = dict(PAR)
    fl = Flags(chemistry="belen", ion_reflection=True)
    chem = chem_params(par, fl)
    rng = np.random.default_rng(17)
    m_i = torch.from_numpy(rng.uniform(0.2, 1.2, mesh.nf).astype(np.float32)).to(dev)
    cos_i = torch.from_numpy(rng.uniform(0.3, 1.0, mesh.nf).astype(np.float32)).to(dev)
    mask = torch.zeros_like(tphi, dtype=torch.uint8)
    mg = ops.MaskGeom()
    band = 3 * dx
    w = torch.from_numpy(rng.normal(size=phi.shape)).to(dev)

    def loss(x):
        _, m_F, m_O = adjoint.radiosity_coverage(x, cos_i, None, D, hit, row_ptr, col, n_ff,
                                                 chem, par.get('betaE', 0.7),
                                                 par.get('betaO', 1.0), n_fp=2, tol=1e-10)
        phi1 = adjoint.advance(mesh, tphi, mask, tphi, dx, dt, 2, band, x, m_F, m_O, cos_i,
                               None, mg, chem)
        return (phi1.double() * w).sum()

    with torch.no_grad():
        _, m_F0, m_O0 = adjoint.radiosity_coverage(m_i, cos_i, None, D, hit, row_ptr, col, n_ff,
                                                   chem, 0.7, 1.0, n_fp=2)
        V0, _ = ops.surface_rate(m_i, m_F0, m_O0, cos_i, None, mesh.centroids, mg, chem)
    dt = 0.4 * dx / float(V0.max())
    x = m_i.clone().requires_grad_(True)
    loss(x).backward()
    g = x.grad.double()
    assert float(g.abs().max()) > 0
    h = 3e-3
    for k in range(3):
        u = torch.from_numpy(rng.normal(size=mesh.nf).astype(np.float32)).to(dev)
        with torch.no_grad():
            fd = float(loss(m_i + h * u) - loss(m_i - h * u)) / (2 * h)
        an = float((g * u.double()).sum())
        print(f"direction {k}: fd {fd:.6e} adjoint {an:.6e}", flush=True)
        assert abs(fd - an) <= 0.05 * abs(fd) + 1e-6 * float((g.abs() * u.abs()).sum())


def test_ion_energy_gradient_through_a_step():
    """d loss / d (mean ion energy) through a whole step on a fixed trench surface with the
    deterministic radiosity transport: the faithful ion trace (forward mode in the energy),
    its normalisation and smoothing, the radiosity coverage loop with the ion yields, the rate,
    the extension and two advection substeps. The ions follow the same paths at E +- h, so the
    forward is smooth in E and float64 central differences must match within 5 %."""
    import torch
    from petch import adjoint, flux, ops
    from petch.chemistry import chem_params
    phi, geo = _geo("trench")
    dx = geo['dx']
    dev = _dev()
    tphi = torch.from_numpy(phi).to(dev)
    mesh = flux.SurfaceMesh(dev).build(tphi, dx)
    par = dict(PAR)
    par['periodic_y'] = 1
    fl = Flags(chemistry="belen", ion_reflection=True)
    gn = ops.gas_normals(tphi, dx, mesh.centroids, mesh.normals)
    n_ff = 64
    hit = ops.trace_ff(mesh.bvh, flux.box_of(geo, par), mesh.centroids, gn, n_ff, 21)
    row_ptr, col, D = ops.radiosity_assemble(hit, mesh.nf, n_ff)
    chem = chem_params(par, fl)
    mask = torch.zeros_like(tphi, dtype=torch.uint8)
    mg = ops.MaskGeom()
    band = 3 * dx
    w = torch.from_numpy(np.random.default_rng(23).normal(size=phi.shape)).to(dev)
    n_ion = 30000

    def loss(E):
        m_i, cos_i, iy = adjoint.ion_terms(mesh, geo, par, n_ion, 5, E)
        _, m_F, m_O = adjoint.radiosity_coverage(m_i, cos_i, iy, D, hit, row_ptr, col, n_ff, chem,
                                                 0.7, 1.0, n_fp=2, tol=1e-10)
        phi1 = adjoint.advance(mesh, tphi, mask, tphi, dx, dt, 2, band, m_i, m_F, m_O, cos_i, iy,
                               mg, chem)
        return (phi1.double() * w).sum()

    E0 = float(par['Emean'])
    with torch.no_grad():
        m_i, cos_i, iy = adjoint.ion_terms(mesh, geo, par, n_ion, 5, torch.tensor(E0))
        _, m_F, m_O = adjoint.radiosity_coverage(m_i, cos_i, iy, D, hit, row_ptr, col, n_ff, chem,
                                                 0.7, 1.0, n_fp=2)
        V0, _ = ops.surface_rate(m_i, m_F, m_O, cos_i, iy, mesh.centroids, mg, chem)
    dt = 0.4 * dx / float(V0.max())
    E = torch.tensor(E0, dtype=torch.float64, requires_grad=True)
    loss(E).backward()
    h = 0.5
    with torch.no_grad():
        fd = float(loss(torch.tensor(E0 + h)) - loss(torch.tensor(E0 - h))) / (2 * h)
    print(f"d loss / dE: fd {fd:.6e} adjoint {float(E.grad):.6e}", flush=True)
    assert abs(fd) > 0
    assert abs(float(E.grad) - fd) <= 0.05 * abs(fd)


def test_mc_step_energy_gradient_matches_common_random_number_differences():
    """The gradient of a whole Monte-Carlo step in the mean ion energy (forward-mode ion trace,
    paired RR neutral traces with their path-replay adjoints over two coverage rounds, finish,
    Belen, rate, extension, two advection substeps) against central differences of the same step
    with common random numbers (the same seeds at E +- h; ion and roulette decisions free to
    flip), with a positive loss on the level set near the front: within 2 %."""
    import torch
    from petch import adjoint, flux, ops
    from petch.chemistry import chem_params
    phi, geo = _geo("trench")
    dx = geo['dx']
    dev = _dev()
    tphi = torch.from_numpy(phi).to(dev)
    mesh = flux.SurfaceMesh(dev).build(tphi, dx)
    par = dict(PAR)
    par['periodic_y'] = 1
    fl = Flags(chemistry="belen", ion_reflection=True)
    chem = chem_params(par, fl)
    mask = torch.zeros_like(tphi, dtype=torch.uint8)
    mg = ops.MaskGeom()
    band = 4 * dx
    w = (tphi.abs() < 2 * dx).double()
    n_ion, n_neu = 400000, 400000

    def loss(E, dt):
        m_i, cos_i, iy = adjoint.ion_terms(mesh, geo, par, n_ion, 3, E)
        _, m_F, m_O = adjoint.coupled_coverage(mesh, geo, par, fl, m_i, cos_i, iy, n_neu, 3,
                                               n_fp=2)
        phi1 = adjoint.advance(mesh, tphi, mask, tphi, dx, dt, 2, band, m_i, m_F, m_O, cos_i,
                               iy, mg, chem)
        return (phi1.double() * w).sum()

    E0 = float(par['Emean'])
    dt = 0.05
    E = torch.tensor(E0, dtype=torch.float64, requires_grad=True)
    loss(E, dt).backward()
    h = 2.0
    with torch.no_grad():
        fd = float(loss(torch.tensor(E0 + h), dt) - loss(torch.tensor(E0 - h), dt)) / (2 * h)
    an = float(E.grad)
    print(f"MC step d loss / dE: CRN fd {fd:.6e} adjoint {an:.6e} rel {abs(an - fd) / abs(fd):.4f}",
          flush=True)
    assert abs(an - fd) <= 0.02 * abs(fd)


def test_pde_reinit_runs_in_both_loops():
    """reinit_method='gpu' (the PDE reinit of threed.py:1399-1418) in run_etch_3d and in the
    batched driver: both finish with a finite level set and etch."""
    from petch import batch as B
    from petch import threed as t3
    kw = dict(SMOKE, n_steps=3, t_end=SMOKE['t_end'] * 3 / SMOKE['n_steps'])
    print("pde reinit loop", flush=True)
    a = t3.run_etch_3d(**kw, par=dict(PAR), flags=FAITHFUL, n_ion=3000, n_neu=3000,
                       verbose=False, reinit_method="gpu")
    assert np.isfinite(a['phi']).all() and t3.max_depth_3d(a) > 0.0
    geo_keys = {k: kw[k] for k in ("Lx", "Ly", "Lz", "dx", "trench_width", "mask_th", "sub_top")}
    outs = B.run_etch_3d_batch([dict(geo_keys, par=dict(PAR), flags=FAITHFUL, n_ion=3000,
                                     n_neu=3000)], t_end=kw['t_end'], n_steps=3,
                               reinit_method="gpu")
    assert np.isfinite(outs[0]['phi']).all()


def test_surf_smooth_runs_and_keeps_the_mask_pinned():
    """run_etch_3d(surf_smooth=...) smooths phi after every reinit (threed.py:1779-1782): the
    run completes, phi differs from the unsmoothed run, and the mask nodes hold their initial
    values; the batched driver runs the same option."""
    from petch import batch as B
    from petch import threed as t3
    kw = dict(SMOKE, n_steps=3, t_end=SMOKE['t_end'] * 3 / SMOKE['n_steps'])  # the smoke's dt
    print("surf_smooth loop", flush=True)
    a = t3.run_etch_3d(**kw, par=dict(PAR), flags=FAITHFUL, n_ion=3000, n_neu=3000,
                       verbose=False, reinit_method="fsm", surf_smooth=0.6)
    b = t3.run_etch_3d(**kw, par=dict(PAR), flags=FAITHFUL, n_ion=3000, n_neu=3000,
                       verbose=False, reinit_method="fsm")
    init = t3.make_trench_3d(kw['Lx'], kw['Ly'], kw['Lz'], kw['dx'], kw['trench_width'],
                             kw['mask_th'], kw['sub_top'])
    m = init['mask']
    assert np.isfinite(a['phi']).all() and np.isfinite(b['phi']).all()
    assert np.abs(a['phi'] - b['phi']).max() > 1e-3
    np.testing.assert_array_equal(a['phi'][m], init['phi'][m])
    geo_keys = {k: kw[k] for k in ("Lx", "Ly", "Lz", "dx", "trench_width", "mask_th", "sub_top")}
    outs = B.run_etch_3d_batch([dict(geo_keys, par=dict(PAR), flags=FAITHFUL, n_ion=3000,
                                     n_neu=3000, surf_smooth=0.6)],
                               t_end=kw['t_end'], n_steps=3)
    assert np.isfinite(outs[0]['phi']).all()
    np.testing.assert_array_equal(outs[0]['phi'][m], init['phi'][m])
