"""The surface-charging hooks on the device (petch/charging.py) against the oracle
(oracle/charging.py): the alpha factor on given fluxes, the floor throttle and the
Hwang-Giapis closure on a carved trench's faces, and the flux paths that apply them. Bars:
float32 rounding of float64 arithmetic (rtol 1e-6)."""
import numpy as np
import pytest

from oracle import charging as och
from petch.params import PAR, Flags

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda", 0)


def _t(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, np.float32), device=_dev())


def test_alpha_factor_matches_oracle():
    from petch import charging
    rng = np.random.default_rng(5)
    m_i = rng.uniform(0.0, 2.0, 3000).astype(np.float32)
    m_e = (m_i * rng.uniform(0.2, 1.3, 3000)).astype(np.float32)
    m_i[:50] = 0.0
    for alpha in (0.3, 1.0):
        got = charging.charge_factor(_t(m_i), _t(m_e), alpha).cpu().numpy()
        np.testing.assert_allclose(got, och.charge_factor(m_i, m_e, alpha), rtol=1e-6, atol=1e-7)


def _trench_surface():
    import torch
    from petch import ops
    from test_gpu_transport import _carved
    dx, nx, ny, nz = 0.25, 40, 12, 64
    phi, _ = _carved("trench", nx, ny, nz, dx, 12.0, 6.0, 1.0)
    tphi = torch.from_numpy(phi).to(_dev())
    v, f = ops.marching_cubes(tphi, dx)
    cen, area, nrm = ops.mesh_attrs(v, f)
    gn = ops.gas_normals(tphi, dx, cen, nrm)

    class M:
        pass
    mesh = M()
    mesh.centroids, mesh.areas, mesh.nf = cen, area, int(f.shape[0])
    geo = dict(Lx=nx * dx, Ly=ny * dx, dx=dx, trench_width=2.0, sub_top=12.0, hole=False)
    return mesh, gn, geo


def test_floor_throttle_and_hg_closure_match_oracle():
    from petch import charging
    mesh, gn, geo = _trench_surface()
    F = mesh.nf
    rng = np.random.default_rng(9)
    m = rng.uniform(0.1, 1.0, F).astype(np.float32)
    iy = tuple(rng.uniform(0.1, 1.0, F).astype(np.float32) for _ in range(3))
    c, g = mesh.centroids.cpu().numpy(), gn.cpu().numpy()
    a = np.maximum(mesh.areas.cpu().numpy().astype(float), 1e-9)
    print(f"charging hooks F={F}", flush=True)
    m2, iy2 = charging.floor_throttle(mesh, gn, geo, _t(m), tuple(_t(x) for x in iy))
    rm, riy = och.floor_throttle(c, g[:, 2], geo, m, iy)
    np.testing.assert_allclose(m2.cpu().numpy(), rm, rtol=1e-6)
    for x, y in zip(iy2, riy):
        np.testing.assert_allclose(x.cpu().numpy(), y, rtol=1e-6)
    assert (rm < m).any()
    for mode in ("hg", "petch"):
        par = dict(PAR, Emean=120.0)
        fl = Flags(surface_charging=mode)
        m3, iy3 = charging.hg_closure(mesh, gn, geo, par, fl, _t(m), tuple(_t(x) for x in iy))
        rm3, riy3 = och.hg_closure(c, g[:, 2], a, geo, par, m, iy, table=mode)
        np.testing.assert_allclose(m3.cpu().numpy(), rm3, rtol=1e-6)
        for x, y in zip(iy3, riy3):
            np.testing.assert_allclose(x.cpu().numpy(), y, rtol=1e-6, atol=1e-7)
        assert (riy3[1] > iy[1] + 1e-6).any()              # the foot gained yield


@pytest.mark.parametrize("flag_kw,par_kw", [
    (dict(surface_charging=True), dict(charge_alpha=0.6)),
    (dict(floor_charge_throttle=True), {}),
    (dict(neutral_transport="knudsen", surface_charging="hg"), {}),
])
def test_loop_runs_with_charging(flag_kw, par_kw):
    """run_etch_3d with each hook: the loop runs on the device and etches."""
    from petch import threed as t3
    fl = Flags(chemistry="belen", yield_angular="viennaps", coverage_sticking=True,
               ion_reflection=True, **flag_kw)
    print(f"loop charging {flag_kw} {par_kw}", flush=True)
    geo = t3.run_etch_3d(Lx=3.0, Ly=0.9, Lz=6.5, dx=0.15, trench_width=0.9, mask_th=0.3,
                         sub_top=6.0, t_end=0.3, n_steps=4, par=dict(PAR, **par_kw), flags=fl,
                         n_ion=4000, n_neu=2000, verbose=False, reinit_method="fsm")
    assert np.isfinite(geo['phi']).all() and t3.max_depth_3d(geo) > 0.0


def test_alpha_model_only_lowers_the_ion_arrival():
    """The coupled flux with charge_alpha > 0 on the same streams: every face's ion arrival is
    the uncharged one times a factor in [1 - alpha, 1]."""
    from petch import flux as fx
    from petch import threed as t3
    geo = t3.make_trench_3d(Lx=3.0, Ly=0.9, Lz=6.5, dx=0.15, trench_width=0.9, mask_th=0.3,
                            sub_top=6.0)
    v, f, _, _ = t3.extract_mesh_3d(geo['phi'], geo['dx'])
    mesh = t3.Mesh(v, f)
    fl = Flags(chemistry="belen", yield_angular="viennaps", coverage_sticking=True,
               ion_reflection=True, surface_charging=True)
    base = fx.coupled_flux(mesh, geo, dict(PAR), 4000, 2000, 3, fl)
    ch = fx.coupled_flux(mesh, geo, dict(PAR, charge_alpha=0.6), 4000, 2000, 3, fl)
    a, b = base['m_i'].cpu().numpy(), ch['m_i'].cpu().numpy()
    ok = a > 1e-6
    r = b[ok] / a[ok]
    assert (r <= 1.0 + 1e-6).all() and (r >= 0.4 - 1e-6).all() and (r < 0.999).any()
