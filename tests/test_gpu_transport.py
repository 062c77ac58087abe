"""The deterministic neutral transports (Knudsen conductance tail, DDA march) on the GPU against
the oracle (oracle/transport.py), on carved holes and trenches through the C ABI, and the
etch loop's neutral_transport='knudsen' / 'dda' paths end to end.

Bars: Knudsen float64 on both sides (rtol 1e-9: the slice means and the sink sums are added in
another order). DDA: the march takes the same float32 steps on both sides, so the wall
decisions agree and the gathered values differ only by the order of their float64 sums; without
the Aitken step rtol 1e-6, with it 1e-4 (the step amplifies last-bit differences of the
iterates by up to r / (1 - r) = 199)."""
import numpy as np
import pytest

from oracle import transport as ot
from petch.params import PAR, Flags

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda", 0)


def _carved(shape, nx, ny, nz, dx, sub_top, depth, half):
    """SDF-like phi (> 0 solid) of a substrate below sub_top with a hole (radius half) or a
    trench along y (half width half) carved depth deep, centred in x (and y)."""
    x = np.arange(nx)[:, None, None] * dx
    y = np.arange(ny)[None, :, None] * dx
    z = np.arange(nz)[None, None, :] * dx
    cx, cy = 0.5 * nx * dx, 0.5 * ny * dx
    r = np.hypot(x - cx, y - cy) if shape == "hole" else np.abs(x - cx) + 0 * y
    carve = np.maximum(r - half, (sub_top - depth) - z)
    return np.minimum(sub_top - z, carve).astype(np.float32), (cx, cy)


def _surface(phi, dx):
    import torch
    from petch import ops
    tphi = torch.from_numpy(phi).to(_dev())
    v, f = ops.marching_cubes(tphi, dx)
    cen, area, nrm = ops.mesh_attrs(v, f)
    gn = ops.gas_normals(tphi, dx, cen, nrm)
    return tphi, cen, area, gn


CASES = [("hole", False, False), ("hole", True, True), ("trench", False, True),
         ("trench", True, False)]


@pytest.mark.parametrize("shape,field,front", CASES)
def test_knudsen_face_flux_matches_oracle(shape, field, front):
    import torch
    from petch import ops
    dx, nx, ny, nz = 0.25, 40, 36 if shape == "hole" else 12, 64
    sub_top, depth, half = 12.0, 8.0, 2.0
    phi, (cx, cy) = _carved(shape, nx, ny, nz, dx, sub_top, depth, half)
    tphi, cen, area, gn = _surface(phi, dx)
    F = int(cen.shape[0])
    rng = np.random.default_rng(7)
    bare = rng.uniform(0.05, 1.0, F).astype(np.float32)
    betas = (0.7, 1.0)
    Ly = ny * dx
    print(f"knudsen {shape} field={field} front={front} F={F}", flush=True)
    got = ops.knudsen_face_flux(tphi, dx, sub_top, shape == "trench", cen, gn, area,
                                torch.from_numpy(bare).to(_dev()), betas, Ly,
                                wall_loss_scale=2.9, front_loss=front, center=(cx, cy),
                                half_width=half, sink_field=field)
    c, g, a = cen.cpu().numpy(), gn.cpu().numpy(), np.maximum(area.cpu().numpy().astype(float), 1e-9)
    sb = bare.astype(float)
    if field:
        fld = c[:, 2] > sub_top - 0.5 * dx
        assert fld.any()
        sb = np.minimum(sb, sb[fld].mean())
    zs = np.arange(nz) * dx
    for beta, m in zip(betas, got):
        ref = ot.knudsen_face_flux(phi, zs, dx, sub_top, "trench" if shape == "trench" else "via",
                                   c, g[:, 2], a, Ly, np.clip(sb * beta, 0.0, 1.0), wls=2.9,
                                   center=(cx, cy), half_width=half, front_loss=front)
        np.testing.assert_allclose(m.cpu().numpy(), ref.astype(np.float32), rtol=1e-6, atol=1e-7)
        deep = (c[:, 2] < sub_top - depth + 0.5) & (g[:, 2] > 0.7)
        assert deep.any() and ref[deep].max() < 1.0      # the floor is starved


@pytest.mark.parametrize("shape,n_re,periodic", [("hole", 1, 0), ("hole", 3, 0),
                                                 ("trench", 3, 1)])
def test_dda_neutral_flux_matches_oracle(shape, n_re, periodic):
    import torch
    from petch import ops
    dx, nx, ny, nz = 0.25, 24, 24 if shape == "hole" else 8, 36
    sub_top, depth, half = 7.0, 4.0, 1.25
    phi, _ = _carved(shape, nx, ny, nz, dx, sub_top, depth, half)
    tphi, cen, area, gn = _surface(phi, dx)
    F = int(cen.shape[0])
    rng = np.random.default_rng(3)
    bare = rng.uniform(0.05, 1.0, F).astype(np.float32)
    betas = (0.3, 0.9)
    print(f"dda {shape} n_reemit={n_re} periodic={periodic} F={F}", flush=True)
    got = ops.dda_neutral_flux(tphi, dx, cen, gn, area, torch.from_numpy(bare).to(_dev()), betas,
                               n_dir=16, n_reemit=n_re, periodic_y=periodic)
    c, g = cen.cpu().numpy(), gn.cpu().numpy()
    a = np.maximum(area.cpu().numpy().astype(float), 1e-9)
    zs = np.arange(nz) * dx
    rtol = 1e-6 if n_re < 2 else 1e-4
    for beta, m in zip(betas, got):
        s = np.clip(bare.astype(float) * beta, 0.0, 1.0)
        ref = ot.dda_neutral_flux(phi, dx, zs, c, g, s, a, n_dir=16, n_reemit=n_re,
                                  periodic_y=periodic)
        np.testing.assert_allclose(m.cpu().numpy(), ref, rtol=rtol, atol=1e-6)
        if n_re == 1:                                   # sky plus one re-emission: the floor
            floor = (c[:, 2] < sub_top - depth + 0.3) & (g[:, 2] > 0.7)   # sees less than 1
            assert floor.any() and (ref[floor] < 1.0).all()


@pytest.mark.parametrize("nt", ["knudsen", "dda"])
def test_etch_loop_runs_deterministic_transports(nt):
    """run_etch_3d with neutral_transport = 'knudsen' / 'dda': the loop runs its deterministic
    neutrals on the device and etches; the reference-API flux wrapper returns the same
    multipliers in [0, 8] on a fresh surface."""
    from petch import threed as t3
    par = dict(PAR, dda_n_dir=16, dda_n_reemit=3, periodic_y=1)
    fl = Flags(chemistry="belen", yield_angular="viennaps", coverage_sticking=True,
               ion_reflection=True, neutral_transport=nt)
    print(f"loop {nt}", flush=True)
    geo = t3.run_etch_3d(Lx=3.0, Ly=0.9, Lz=6.5, dx=0.15, trench_width=0.9, mask_th=0.3,
                         sub_top=6.0, t_end=0.3, n_steps=4, par=par, flags=fl, n_ion=4000,
                         n_neu=1000, verbose=False, reinit_method="fsm")
    d = t3.max_depth_3d(geo)
    assert np.isfinite(geo['phi']).all() and d > 0.0
    verts, faces, cen, areas = t3.extract_mesh_3d(geo['phi'], geo['dx'])
    mesh = t3.Mesh(verts, faces)
    fn = t3.mc_flux_3d_knudsen if nt == "knudsen" else t3.mc_flux_3d_dda
    m_i, m_F, m_O, cos_i = fn(mesh, verts, faces, cen, areas, geo, dict(par), n_ion=4000, seed=1,
                              flags=fl)
    for m in (m_F, m_O):
        assert m.shape == (len(faces),) and (m >= 0).all() and (m <= 8).all()
