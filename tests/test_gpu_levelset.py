"""GPU parity of the level-set kernels (initial SDF, upwind advection, FSM reinit) against the
FP64 oracle, through the C ABI. Bar: max-abs <= 1e-5*dx; mask bit-exact."""
import numpy as np
import pytest

from oracle import levelset as ol

pytestmark = pytest.mark.gpu

# (Lx, Ly, Lz, dx, width, mask, sub_top, hole): hole of the 40-step etch (56x56x96), the AR-10
# trench of the single-flux-step workload (160x40x260), and a small ragged case
GEOMS = [
    (14.0, 14.0, 24.0, 0.25, 6.0, 2.0, 18.0, True),
    (16.0, 4.0, 26.0, 0.1, 2.0, 2.0, 20.0, False),
    (3.3, 2.1, 4.7, 0.1, 0.9, 0.3, 3.0, True),
]


def _dev():
    import torch
    return torch.device("cuda", 0)


@pytest.mark.parametrize("g", GEOMS)
def test_init_sdf_parity(g):
    import torch
    from petch import ops
    Lx, Ly, Lz, dx, w, m, s, hole = g
    nx, ny, nz = int(round(Lx / dx)), int(round(Ly / dx)), int(round(Lz / dx))
    print(f"init_sdf {nx}x{ny}x{nz} hole={hole}", flush=True)
    phi, mask = ops.init_sdf(nx, ny, nz, dx, Lx, Ly, w, m, s, hole, device=_dev())
    torch.cuda.synchronize()
    ref = ol.init_sdf(Lx, Ly, Lz, dx, w, m, s, hole)
    refm = ol.trench_mask(Lx, Ly, Lz, dx, w, m, s, hole)
    np.testing.assert_array_equal(mask.cpu().numpy().astype(bool), refm)
    np.testing.assert_allclose(phi.cpu().numpy(), ref, rtol=0, atol=1e-5 * dx)


def _perturbed_state(g, seed):
    Lx, Ly, Lz, dx, w, m, s, hole = g
    phi = ol.init_sdf(Lx, Ly, Lz, dx, w, m, s, hole)
    rng = np.random.default_rng(seed)
    X = np.arange(phi.shape[0])[:, None, None] * dx
    Y = np.arange(phi.shape[1])[None, :, None] * dx
    phi = phi * (1.0 + 0.3 * np.sin(0.7 * X) * np.cos(0.9 * Y))      # |grad| != 1
    phi = phi.astype(np.float32)
    F = np.where(np.abs(phi) < 4 * dx, rng.uniform(0.0, 2.0, phi.shape), 0.0).astype(np.float32)
    mask = ol.trench_mask(Lx, Ly, Lz, dx, w, m, s, hole)
    return phi, F, mask


@pytest.mark.parametrize("g", GEOMS)
def test_advect_parity(g):
    import torch
    from petch import ops
    dx = g[3]
    phi, F, mask = _perturbed_state(g, 1)
    mask_phi = phi.copy()
    nsub, dt = 7, 0.3 * dx / 2.0
    print(f"advect shape={phi.shape} nsub={nsub}", flush=True)
    d = _dev()
    tp = torch.from_numpy(phi).to(d)
    out = ops.advect(tp, torch.from_numpy(F).to(d), torch.from_numpy(mask.astype(np.uint8)).to(d),
                     torch.from_numpy(mask_phi).to(d), dx, dt, nsub)
    torch.cuda.synchronize()
    ref = ol.advect_pinned(phi.astype(np.float64), F.astype(np.float64), mask,
                           mask_phi.astype(np.float64), dx, dt, nsub)
    got = out.cpu().numpy()
    assert np.isfinite(got).all()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * dx)
    assert np.abs(got - phi).max() > 10 * dx * 1e-5                 # it actually moved


@pytest.mark.parametrize("g", GEOMS)
def test_reinit_fsm_parity(g):
    import torch
    from petch import ops
    dx = g[3]
    phi, _, _ = _perturbed_state(g, 2)
    band = 4 * dx + 2.0 * dx            # the driver's call (threed.py:1771)
    print(f"reinit shape={phi.shape} band={band}", flush=True)
    t = torch.from_numpy(phi).to(_dev())
    ops.reinit_fsm(t, dx, band)
    torch.cuda.synchronize()
    ref = ol.reinit_fsm(phi.astype(np.float64), dx, band)
    np.testing.assert_allclose(t.cpu().numpy(), ref, rtol=0, atol=1e-5 * dx)


def test_reinit_then_advect_is_stable_and_moves_front():
    """Size-independent property at the 56x56x96 hole: a uniform etch speed in the band moves the
    floor down by ~F*dt and keeps |grad phi| ~ 1 after reinit."""
    import torch
    from petch import ops
    g = GEOMS[0]
    Lx, Ly, Lz, dx, w, m, s, hole = g
    nx, ny, nz = int(round(Lx / dx)), int(round(Ly / dx)), int(round(Lz / dx))
    phi, mask = ops.init_sdf(nx, ny, nz, dx, Lx, Ly, w, m, s, hole, device=_dev())
    mask_phi = phi.clone()
    F = torch.where(phi.abs() < 4 * dx, 1.0, 0.0).float()
    for _ in range(4):
        ops.advect(phi, F, mask, mask_phi, dx, 0.1 * dx, 4)
        ops.reinit_fsm(phi, dx, 6 * dx)
    p = phi.cpu().numpy()
    # centre column: floor (first solid node from the top) went from z=18 to about 18 - 1.6dx
    col = p[nx // 2, ny // 2, :]
    zc = np.where(np.diff(np.sign(col)) != 0)[0]
    k = zc[-1]
    z0 = (k + col[k] / (col[k] - col[k + 1])) * dx
    assert z0 == pytest.approx(s - 16 * 0.1 * dx, abs=0.3 * dx)
    gx, gy, gz = np.gradient(p, dx)
    gn = np.sqrt(gx ** 2 + gy ** 2 + gz ** 2)[np.abs(p) < 4 * dx]
    assert abs(np.median(gn) - 1.0) < 0.05


def test_advect_adjoint_matches_oracle():
    """One substep: the HIP adjoint equals the oracle's (pinned by central differences) on the
    same float32 state; nsub substeps through the autograd Function: the directional derivative
    of a loss equals float64 central differences of the oracle's pinned advection."""
    import torch
    from petch import adjoint, ops
    d = _dev()
    g = GEOMS[2]
    dx = g[3]
    phi, F, mask = _perturbed_state(g, 21)
    rng = np.random.default_rng(22)
    F = (F + rng.uniform(0.1, 1.0, F.shape)).astype(np.float32)
    mask_phi = phi.copy()
    dt = 0.3 * dx / 2.5
    go = rng.normal(size=phi.shape).astype(np.float32)
    tp, tF = torch.from_numpy(phi).to(d), torch.from_numpy(F).to(d)
    tm = torch.from_numpy(mask.astype(np.uint8)).to(d)
    gF = torch.zeros_like(tF)
    gphi = ops.advect_adjoint(tp, tF, tm, torch.from_numpy(go).to(d), dx, dt, gF)
    rphi, rF = ol.advect_3d_adjoint(phi, F, mask, dx, dt, go)
    scale = np.abs(rphi).max()
    np.testing.assert_allclose(gphi.cpu().numpy(), rphi, rtol=0, atol=2e-5 * scale)
    np.testing.assert_allclose(gF.cpu().numpy(), rF, rtol=0, atol=2e-5 * np.abs(rF).max())
    # three substeps, back-propagated through the autograd Function
    nsub = 3
    x = tp.clone().requires_grad_()
    f = tF.clone().requires_grad_()
    w = torch.from_numpy(go).to(d)
    out = adjoint.advect(x, f, tm, torch.from_numpy(mask_phi).to(d), dx, dt, nsub)
    np.testing.assert_allclose(out.detach().cpu().numpy(),
                               ops.advect(tp.clone(), tF, tm, torch.from_numpy(mask_phi).to(d),
                                          dx, dt, nsub).cpu().numpy(), rtol=0, atol=0)
    (out * w).sum().backward()
    # the same gradient when the substep states are rebuilt in the backward pass
    x2 = tp.clone().requires_grad_()
    f2 = tF.clone().requires_grad_()
    out2 = adjoint.advect(x2, f2, tm, torch.from_numpy(mask_phi).to(d), dx, dt, nsub,
                          recompute=True)
    (out2 * w).sum().backward()
    assert torch.equal(out2, out)
    assert torch.equal(x2.grad, x.grad)
    np.testing.assert_allclose(f2.grad.cpu().numpy(), f.grad.cpu().numpy(), rtol=0, atol=0)

    def loss(p, fs):
        return float((go * ol.advect_pinned(p, fs, mask, mask_phi.astype(np.float64), dx, dt,
                                            nsub)).sum())

    h = 1e-5
    for k in range(2):
        v = rng.normal(size=phi.shape)
        fd = (loss(phi + h * v, F) - loss(phi - h * v, F)) / (2 * h)
        an = float((x.grad.cpu().numpy().astype(np.float64) * v).sum())
        print(f"phi direction {k}: fd {fd:.6f} adjoint {an:.6f}", flush=True)
        assert abs(fd - an) <= 2e-3 * abs(fd) + 1e-3
        u = rng.normal(size=phi.shape)
        fd = (loss(phi, F + h * u) - loss(phi, F - h * u)) / (2 * h)
        an = float((f.grad.cpu().numpy().astype(np.float64) * u).sum())
        print(f"F direction {k}: fd {fd:.6f} adjoint {an:.6f}", flush=True)
        assert abs(fd - an) <= 2e-3 * abs(fd) + 1e-3


def test_profile_area_and_its_adjoint_match_the_oracle():
    """petch_profile_area (smoothed gas area per z slice, fixed-order float64 sums) and its
    adjoint against the oracle on the small hole's initial level set, and the autograd profile
    loss (sum over the hole's slices of (r(z) - r*(z))^2) against the oracle's chain."""
    import torch
    from petch import adjoint, ops
    dx = 0.25
    phi = ol.init_sdf(14.0, 14.0, 24.0, dx, 6.0, 2.0, 18.0, True).astype(np.float32)
    dev = torch.device("cuda", 0)
    tphi = torch.from_numpy(phi).to(dev)
    eps = 1.5 * dx
    A = ops.slice_gas_area(tphi, dx, eps).cpu().numpy()
    Ao = ol.slice_gas_area(phi.astype(np.float64), dx, eps)
    np.testing.assert_allclose(A, Ao, rtol=1e-12, atol=1e-12)
    rng = np.random.default_rng(8)
    garea = rng.normal(size=A.shape)
    g = ops.slice_gas_area_adjoint(tphi, dx, eps, torch.from_numpy(garea).to(dev)).cpu().numpy()
    go = ol.slice_gas_area_adjoint(phi.astype(np.float64), dx, eps, garea)
    np.testing.assert_allclose(g, go, rtol=1e-6, atol=1e-7 * np.abs(go).max())
    # the profile loss over the hole's slices against a target 0.3 um narrower
    nz = phi.shape[2]
    zsel = (np.arange(nz) * dx >= 10.0) & (np.arange(nz) * dx < 18.0)
    target = np.full(nz, 2.7)
    x = tphi.clone().requires_grad_(True)
    loss = adjoint.profile_loss(x, dx, torch.from_numpy(target), torch.from_numpy(zsel), eps)
    loss.backward()
    ro = ol.profile_rz(phi.astype(np.float64), dx, eps)
    lo = float(((ro - target)[zsel] ** 2).sum())
    assert float(loss.detach()) == pytest.approx(lo, rel=1e-10)
    with np.errstate(divide="ignore", invalid="ignore"):
        gr = np.where(zsel & (ro > 0), 2 * (ro - target) / (2 * np.pi * ro), 0.0)
    gref = ol.slice_gas_area_adjoint(phi.astype(np.float64), dx, eps, gr)
    got = x.grad.cpu().numpy()
    np.testing.assert_allclose(got, gref, rtol=1e-5, atol=1e-7 * np.abs(gref).max())
    assert np.abs(got).max() > 0


def test_batched_advection_and_reinit_equal_single_calls():
    """petch_advect_batch and petch_reinit_fsm_batch (one launch per substep / sweep for many
    grids, bricks that provably do not change skipped) give exactly what petch_advect and
    petch_reinit_fsm give grid by grid: grids of different shapes (one large enough that most
    bricks lie far from the interface), substep counts 0-3 (odd counts end in the work buffer and
    are copied back), sweep counts, mask nodes off and at their pinned values, and a work buffer
    full of NaN (a skipped brick's stale work values must never be read)."""
    import torch
    from petch import ops
    dev = _dev()
    rng = np.random.default_rng(9)
    shapes = [(30, 22, 41), (17, 33, 25), (40, 8, 12), (72, 56, 100)]
    grids = []
    for k, (nx, ny, nz) in enumerate(shapes):
        dx = 0.1 + 0.05 * (k % 3)
        phi = ol.init_sdf(nx * dx, ny * dx, nz * dx, dx, nx * dx / 3, 3 * dx, nz * dx / 2,
                          k % 2 == 0).astype(np.float32)
        phi += (0.05 * dx * rng.normal(size=phi.shape)).astype(np.float32)
        F = np.where(np.abs(phi) < 4 * dx, rng.uniform(0.5, 1.5, phi.shape), 0.0).astype(np.float32)
        mask = (rng.random(phi.shape) < 0.05).astype(np.uint8)
        mp = phi + np.where(mask & (rng.random(phi.shape) < 0.3), 0.2 * dx, 0.0).astype(np.float32)
        grids.append((dx, torch.from_numpy(phi).to(dev), torch.from_numpy(F).to(dev),
                      torch.from_numpy(mask).to(dev), torch.from_numpy(mp).to(dev)))
    for nsubs in ((1, 2, 3, 3), (0, 3, 2, 2)):
        ref, items = [], []
        for (dx, phi, F, mask, mp), nsub in zip(grids, nsubs):
            a = phi.clone()
            ops.advect(a, F, mask, mp, dx, 0.3 * dx, nsub)
            ref.append(a)
            b = phi.clone()
            items.append((b, torch.full_like(b, float("nan")), F, mask, mp, dx, 0.3 * dx, nsub))
        ops.advect_batch(items)
        for r, it in zip(ref, items):
            assert torch.equal(r, it[0])
    ref, items = [], []
    for (dx, phi, F, mask, mp), n_iter in zip(grids, (None, 7, 10, None)):
        a = phi.clone()
        ops.reinit_fsm(a, dx, 6 * dx, n_iter)
        ref.append(a)
        b = phi.clone()
        items.append((b, dx, 6 * dx, n_iter))
    ops.reinit_fsm_batch(items)
    for r, it in zip(ref, items):
        assert torch.equal(r, it[0])


@pytest.mark.parametrize("g", [GEOMS[0], GEOMS[2]])
def test_reinit_adjoint_matches_the_oracle(g):
    """petch_reinit_fsm_adjoint against the oracle's exact adjoint (pinned by central
    differences and the plane shift) on the perturbed hole and the ragged case, band of the
    etch loop, a random cotangent in the band. The device stores the sweep states in float32,
    so where a node converges within rounding its last lowering comes a sweep earlier than in
    float64, and the gradient reaches the same seeds by another path: against the float64
    oracle that moves 1e-4 to 8e-4 of the L1 mass. The oracle with float32 states takes the
    device's decisions and is the tight comparison."""
    import torch
    from petch import adjoint, ops
    dx = g[3]
    phi, _, _ = _perturbed_state(g, 3)
    band = 4 * dx + 2.0 * dx
    rng = np.random.default_rng(5)
    go = (rng.standard_normal(phi.shape) * (np.abs(phi) < band + dx)).astype(np.float32)
    print(f"reinit adjoint shape={phi.shape} band={band}", flush=True)
    d = _dev()
    tp = torch.from_numpy(phi).to(d)
    got = ops.reinit_fsm_adjoint(tp, dx, band, torch.from_numpy(go).to(d)).cpu().numpy()
    ref = ol.reinit_fsm_adjoint(phi.astype(np.float64), dx, band, go.astype(np.float64),
                                state_dtype=np.float32)
    ref64 = ol.reinit_fsm_adjoint(phi.astype(np.float64), dx, band, go.astype(np.float64))
    rel = np.abs(got - ref).sum() / np.abs(ref).sum()
    rel64 = np.abs(got - ref64).sum() / np.abs(ref64).sum()
    print(f"L1 rel {rel:.2e} (float64-state oracle {rel64:.2e}); max abs "
          f"{np.abs(got - ref).max():.2e} of {np.abs(ref).max():.2e}", flush=True)
    assert rel < 1e-5
    assert rel64 < 2e-3
    # the autograd wrapper: forward is ops.reinit_fsm out of place, backward the same adjoint
    a = tp.clone().requires_grad_(True)
    r = adjoint.reinit(a, dx, band)
    ref_fwd = tp.clone()
    ops.reinit_fsm(ref_fwd, dx, band)
    assert torch.equal(r.detach(), ref_fwd) and torch.equal(a.detach(), tp)
    r.backward(torch.from_numpy(go).to(d))
    assert torch.equal(a.grad, torch.from_numpy(got).to(d))


@pytest.mark.parametrize("sigma", [0.5, 1.3])
def test_gaussian_smooth_matches_scipy_restatement(sigma):
    """petch_gaussian_smooth (surf_smooth) against the oracle's restatement of
    scipy.ndimage.gaussian_filter (itself equal to scipy bit for bit), then the mask re-pinned:
    bit for bit, on a trench level set and a ragged random grid."""
    import torch
    from petch import ops
    d = torch.device("cuda", 0)
    cases = [ol.init_sdf(6.0, 3.0, 10.0, 0.2, 2.0, 1.0, 6.0, False).astype(np.float32),
             np.random.default_rng(3).normal(size=(5, 3, 17)).astype(np.float32)]
    for phi in cases:
        print(f"smooth sigma={sigma} shape={phi.shape}", flush=True)
        rng = np.random.default_rng(1)
        mask = (rng.random(phi.shape) < 0.2).astype(np.uint8)
        mphi = rng.normal(size=phi.shape).astype(np.float32)
        t = torch.from_numpy(phi.copy()).to(d)
        ops.gaussian_smooth(t, sigma, torch.from_numpy(mask).to(d), torch.from_numpy(mphi).to(d))
        ref = ol.gaussian_smooth(phi, sigma)
        ref[mask.astype(bool)] = mphi[mask.astype(bool)]
        np.testing.assert_array_equal(t.cpu().numpy(), ref)


def test_reinit_pde_matches_the_oracle():
    """petch_reinit_pde (reinit_method='gpu') against the oracle's float32 restatement of
    reinit_gpu, bit for bit (same operations in the same order, correctly rounded sqrt and
    division on both sides), on a perturbed trench and a ragged random grid."""
    import torch
    from petch import ops
    d = torch.device("cuda", 0)
    rng = np.random.default_rng(4)
    base = ol.init_sdf(6.0, 3.0, 10.0, 0.2, 2.0, 1.0, 6.0, False)
    cases = [(base * (1.0 + 0.3 * rng.random(base.shape))).astype(np.float32),
             rng.normal(size=(7, 5, 13)).astype(np.float32)]
    for phi in cases:
        print(f"reinit_pde {phi.shape}", flush=True)
        t = torch.from_numpy(phi.copy()).to(d)
        ops.reinit_pde(t, 0.2, 24)
        np.testing.assert_array_equal(t.cpu().numpy(), ol.reinit_pde(phi, 0.2, 24))
