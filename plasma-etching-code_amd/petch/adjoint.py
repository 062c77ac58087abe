"""Differentiable pieces of the etch step (torch.autograd.Function over hand-written HIP adjoints).

The reference differentiates nothing inside its etch loop (its only gradients are the demo tapes
scripts/inverse_design.py:95-101 and scripts/warp_3d_flux.py:106-113). Here the level-set
advection carries an adjoint: `advect(phi, F, mask, mask_phi, dx, dt, nsub)` runs the same
substeps as ops.advect and back-propagates through them with petch_advect_adjoint, the upwind
choices frozen at each substep's state (the states are kept from the forward pass, nsub x the
grid). Gradients flow to phi and to the speed F; the mask and its pinned values are constants.

    phi1 = adjoint.advect(phi0.requires_grad_(), F.requires_grad_(), mask, mask_phi, dx, dt, 3)
    loss = (phi1 * w).sum(); loss.backward()          # phi0.grad, F.grad

The radiosity solve of mc_flux_3d_radiosity (threed.py:965-985) carries its implicit adjoint:
`radiosity(bare, D, hitface, row_ptr, col, n_ff, beta, n_iter, tol)` returns
m = clip(G, 0, 8) of G = D + Aff diag(1 - clip(bare beta, 0, 1)) G and back-propagates through
the transposed solve (petch_radiosity_adjoint) to bare and to the escape fractions D. The form
factors (hitface, the CSR) are constants: they come from discrete first hits.

The Belen coverage update bare = clip(1 - thF - thO, 0, 1) (threed.py:715-737) carries its
closed-form derivative (petch_belen_coverage_adjoint): `belen(m_i, m_F, m_O, cos_i, ion_yield,
chem)`. `radiosity_coverage` chains both into the coverage fixed point of mc_flux_3d_radiosity
(threed.py:963-990): n_fp rounds of the F and O solves and the Belen update from bare = 1,
differentiable in the ion terms (m_i, or the ion yields) and the escape fractions D, the
adjoint unrolled over the n_fp rounds.

The Russian-roulette neutral trace (_trace3d_cov_rr, threed.py:549-596) carries a path-replay
adjoint (petch_trace_neutral_rr_adjoint): `neutral_rr(bvh, box, src, bare, beta, n, seed)`
returns the raw per-face accumulator and back-propagates to bare with every ray's path and its
roulette decisions held fixed and a roulette survivor's weight read as w / p (p, its survival
probability, held fixed), which makes the gradient's expectation the derivative of the expected
flux: it matches common-random-number differences to a fraction of a percent. The backward pass
redraws the same Philox streams, so nothing is stored per ray. `coupled_coverage` chains it with the clip-normalisation
and smoothing of the fluxes (petch_normalize_adjoint, petch_smooth_adjoint) and the Belen update
into the neutral rounds of the MC coupled flux (threed.py:855-904), differentiable in the ion
terms.

The etch rate and the velocity extension carry their adjoints too: `surface_rate(...)`
(petch_surface_rate_adjoint, closed form for the Belen and Langmuir chemistries) and
`extend(bvh, V, phi, dx, band)` (petch_extend_adjoint: each band node's gradient goes back to its
nearest face; the nearest-face assignment is the geometry and stays fixed). With advect they
take a loss on the advanced level set back to the fluxes of a step.

The FSM reinit (threed.py:1515-1537) carries the exact adjoint of the piecewise-smooth map its
sweeps compute (petch_reinit_fsm_adjoint): `reinit(a, dx, band)`. Every discrete choice
(interface flags, the smaller neighbour per axis, the active Godunov terms, which sweep lowers a
node, the clamp) is held at the forward's; a lowered node moves with its upwind neighbours by
the Godunov weights (x - a_d) / sum_e (x - a_e), and the interface seeds |phi| / |grad phi|
pass their derivative to the node and its neighbours.
"""
import math

import torch

from . import ops


class _Advect(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, F, mask, mask_phi, dx, dt, nsub, recompute):
        states = [phi.detach().contiguous()]
        Fd = F.detach().contiguous()
        for _ in range(int(nsub)):
            states.append(ops.advect_step(states[-1], Fd, mask, mask_phi, dx, dt))
        kept = states[:1] if recompute else states[:-1]
        ctx.save_for_backward(*kept, Fd, mask, mask_phi)
        ctx.args = (float(dx), float(dt), int(nsub), bool(recompute))
        return states[-1] if nsub > 0 else states[0].clone()

    @staticmethod
    def backward(ctx, gout):
        *states, F, mask, mask_phi = ctx.saved_tensors
        dx, dt, nsub, recompute = ctx.args
        if recompute:  # the substeps are deterministic: rebuild the states from the input
            for _ in range(nsub - 1):
                states.append(ops.advect_step(states[-1], F, mask, mask_phi, dx, dt))
        gF = torch.zeros_like(F)
        g = gout.contiguous()
        for s in reversed(states[:nsub]):
            g = ops.advect_adjoint(s, F, mask, g, dx, dt, gF)
        return g, gF, None, None, None, None, None, None


def advect(phi, F, mask, mask_phi, dx, dt, nsub, recompute=False):
    """nsub upwind substeps of size dt with the mask re-pin (ops.advect), differentiable in phi
    and F. recompute=True keeps only the input level set for the backward pass and rebuilds the
    substep states there (one grid instead of nsub; the substeps are deterministic, so the
    gradient is the same bit for bit)."""
    return _Advect.apply(phi, F, mask, mask_phi, dx, dt, nsub, recompute)


class _Reinit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, dx, band, n_iter):
        a = a.detach().contiguous()
        r = a.clone()
        ops.reinit_fsm(r, dx, band, n_iter)
        ctx.save_for_backward(a)
        ctx.args = (float(dx), float(band), n_iter)
        return r

    @staticmethod
    def backward(ctx, gout):
        (a,) = ctx.saved_tensors
        dx, band, n_iter = ctx.args
        return ops.reinit_fsm_adjoint(a, dx, band, gout.contiguous().float(), n_iter), None, \
            None, None


def reinit(a, dx, band, n_iter=None):
    """The FSM reinit of a (ops.reinit_fsm, out of place), differentiable in a: the exact
    adjoint of the map the sweeps compute, with the interface flags, the neighbour and
    active-term choices of every Godunov update, the sweeps that lower each node and the clamp
    held at the forward's (petch_reinit_fsm_adjoint). Keeps only a for the backward pass, which
    rebuilds the sweep states."""
    return _Reinit.apply(a, dx, band, n_iter)


class _Radiosity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bare, D, hitface, row_ptr, col, n_ff, beta, n_iter, tol):
        nf = bare.numel()
        b = bare.detach().contiguous()
        m, G, _ = ops.radiosity_solve(row_ptr, col, D.detach(), nf, n_ff, b, beta, n_iter, tol)
        ctx.save_for_backward(b, G, hitface)
        ctx.args = (nf, int(n_ff), float(beta), int(n_iter), float(tol))
        return m

    @staticmethod
    def backward(ctx, gm):
        b, G, hitface = ctx.saved_tensors
        nf, n_ff, beta, n_iter, tol = ctx.args
        gD, gbare, _ = ops.radiosity_adjoint(hitface, nf, n_ff, G, b, beta,
                                             gm.contiguous().float(), n_iter, tol)
        return gbare, gD, None, None, None, None, None, None, None


def radiosity(bare, D, hitface, row_ptr, col, n_ff, beta, n_iter=0, tol=1e-10):
    """m = clip(G, 0, 8) of the radiosity solve (ops.radiosity_solve on the CSR that
    ops.radiosity_assemble built from hitface), differentiable in bare and D. n_iter <= 0 solves
    to max|dG| <= tol max|G|, where the implicit adjoint is exact."""
    return _Radiosity.apply(bare, D, hitface, row_ptr, col, n_ff, beta, n_iter, tol)


class _Belen(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m_i, m_F, m_O, cos_i, iy_ie, iy_ip, chem):
        iy = (None, iy_ie, iy_ip) if iy_ie is not None else None
        bare = ops.belen_coverage(m_i, m_F.contiguous(), m_O.contiguous(), cos_i, iy, chem)
        ctx.save_for_backward(m_i, m_F, m_O, cos_i, iy_ie, iy_ip)
        ctx.chem = chem
        return bare

    @staticmethod
    def backward(ctx, gbare):
        m_i, m_F, m_O, cos_i, iy_ie, iy_ip = ctx.saved_tensors
        iy = (None, iy_ie, iy_ip) if iy_ie is not None else None
        gm_F, gm_O, gm_i, g_ie, g_ip = ops.belen_coverage_adjoint(
            m_i, m_F.contiguous(), m_O.contiguous(), cos_i, iy, ctx.chem,
            gbare.contiguous().float())
        return gm_i, gm_F, gm_O, None, g_ie, g_ip, None


def belen(m_i, m_F, m_O, cos_i, ion_yield, chem):
    """bare = clip(1 - thF - thO, 0, 1) (ops.belen_coverage), differentiable in m_F, m_O and
    the ion terms: m_i without ion yields, else the yields ion_yield[1], ion_yield[2]."""
    ie, ip = (ion_yield[1], ion_yield[2]) if ion_yield is not None else (None, None)
    return _Belen.apply(m_i, m_F, m_O, cos_i, ie, ip, chem)


def radiosity_coverage(m_i, cos_i, ion_yield, D, hitface, row_ptr, col, n_ff, chem, betaE,
                       betaO, n_fp=4, n_iter=0, tol=1e-10):
    """(bare, m_F, m_O) of the coverage fixed point (the loop of flux.radiosity_flux) on the
    form factors of hitface / (row_ptr, col, D), differentiable in m_i or the ion yields and in
    D. n_iter <= 0 solves each radiosity system to tol (where its implicit adjoint is exact)."""
    F = int(cos_i.shape[0])
    bare = torch.ones((F,), dtype=torch.float32, device=cos_i.device)
    m_F = m_O = torch.zeros((F,), dtype=torch.float32, device=cos_i.device)
    Dv = D[:F]
    for _ in range(n_fp):
        m_F = radiosity(bare, Dv, hitface, row_ptr, col, n_ff, betaE, n_iter, tol)
        m_O = radiosity(bare, Dv, hitface, row_ptr, col, n_ff, betaO, n_iter, tol)
        bare = belen(m_i, m_F, m_O, cos_i, ion_yield, chem)
    return bare, m_F, m_O



def _check_bvh(bvh, what):
    """Raises if the BVH's traversal guard tripped. The flag is cleared only by a rebuild, so one
    read after a backward pass covers its launches and those of the forward pass on the same
    BVH; a tripped guard means dropped subtrees and an incomplete gradient, which the etch loop
    treats the same way (threed.EtchLoop.end_finish)."""
    err = bvh.status()
    if err:
        raise RuntimeError(f"{what}: BVH traversal guard tripped (code {err}: 1 stack overflow, "
                           "2 corrupt tree, 3 tree too deep); the gradient is incomplete")


class _NeutralRR(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bare, bvh, box, src, beta, n, seed):
        b = bare.detach().contiguous()
        fl = ops.trace_neutral_rr(bvh, box, src, b, beta, n, seed)
        ctx.save_for_backward(b)
        ctx.args = (bvh, box, src, float(beta), int(n), int(seed))
        return fl

    @staticmethod
    def backward(ctx, gfl):
        (b,) = ctx.saved_tensors
        bvh, box, src, beta, n, seed = ctx.args
        gbare = ops.trace_neutral_rr_adjoint(bvh, box, src, b, beta, n, seed,
                                             gfl.contiguous().float())
        _check_bvh(bvh, "RR neutral trace adjoint")
        return gbare, None, None, None, None, None, None


class _NeutralRRPair(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bare, bvh, box, src, beta_a, n_a, seed_a, beta_b, n_b, seed_b):
        b = bare.detach().contiguous()
        fa, fb = ops.trace_neutral_rr_pair(bvh, box, src, b, beta_a, n_a, seed_a, beta_b, n_b,
                                           seed_b)
        ctx.save_for_backward(b)
        ctx.args = (bvh, box, src, float(beta_a), int(n_a), int(seed_a), float(beta_b), int(n_b),
                    int(seed_b))
        return fa, fb

    @staticmethod
    def backward(ctx, ga, gb):
        (b,) = ctx.saved_tensors
        bvh, box, src, beta_a, n_a, seed_a, beta_b, n_b, seed_b = ctx.args
        ga = torch.zeros_like(b) if ga is None else ga.contiguous().float()
        gb = torch.zeros_like(b) if gb is None else gb.contiguous().float()
        g = ops.trace_neutral_rr_pair_adjoint(bvh, box, src, b, beta_a, n_a, seed_a, ga, beta_b,
                                              n_b, seed_b, gb)
        _check_bvh(bvh, "paired RR neutral trace adjoint")
        return g, None, None, None, None, None, None, None, None, None


def neutral_rr_pair(bvh, box, src, bare, beta_a, n_a, seed_a, beta_b, n_b, seed_b):
    """The F and O accumulators of one paired launch (ops.trace_neutral_rr_pair, the same
    streams as two single launches), differentiable in bare (both species' path-replay adjoints
    in one launch)."""
    return _NeutralRRPair.apply(bare, bvh, box, src, beta_a, n_a, seed_a, beta_b, n_b, seed_b)


class _NeutralRRPairBeta(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bare, beta_a, beta_b, bvh, box, src, n_a, seed_a, n_b, seed_b):
        b = bare.detach().contiguous()
        ba, bb = float(beta_a), float(beta_b)
        fa, fb = ops.trace_neutral_rr_pair(bvh, box, src, b, ba, n_a, seed_a, bb, n_b, seed_b)
        ctx.save_for_backward(b, beta_a, beta_b)
        ctx.args = (bvh, box, src, ba, int(n_a), int(seed_a), bb, int(n_b), int(seed_b))
        return fa, fb

    @staticmethod
    def backward(ctx, ga, gb):
        b, ta, tb = ctx.saved_tensors
        bvh, box, src, ba, n_a, seed_a, bb, n_b, seed_b = ctx.args
        ga = torch.zeros_like(b) if ga is None else ga.contiguous().float()
        gb = torch.zeros_like(b) if gb is None else gb.contiguous().float()
        # each species' own bare gradient (the pair's streams are those of two single launches)
        g_a = ops.trace_neutral_rr_adjoint(bvh, box, src, b, ba, n_a, seed_a, ga)
        g_b = ops.trace_neutral_rr_adjoint(bvh, box, src, b, bb, n_b, seed_b, gb)
        _check_bvh(bvh, "paired RR neutral trace adjoint")
        bd = b.double()
        gba = ((bd * g_a.double()).sum() / ba).to(device=ta.device, dtype=ta.dtype)
        gbb = ((bd * g_b.double()).sum() / bb).to(device=tb.device, dtype=tb.dtype)
        return g_a + g_b, gba, gbb, None, None, None, None, None, None, None


def neutral_rr_pair_beta(bvh, box, src, bare, beta_a, n_a, seed_a, beta_b, n_b, seed_b):
    """neutral_rr_pair, differentiable also in the species' sticking factors beta_a and beta_b
    (0-d tensors): the trace sticks with S = bare * beta, and the path-replay adjoint's bare
    gradient is beta dL/dS, so dL/dbeta = sum_f bare_f g_f / beta with g the species' own bare
    gradient (two single-species adjoint launches instead of the paired one)."""
    return _NeutralRRPairBeta.apply(bare, beta_a, beta_b, bvh, box, src, n_a, seed_a, n_b,
                                    seed_b)


def neutral_rr(bvh, box, src, bare, beta, n, seed):
    """Raw RR neutral accumulator (ops.trace_neutral_rr), differentiable in bare. The BVH must
    be unchanged between the forward and the backward pass."""
    return _NeutralRR.apply(bare, bvh, box, src, beta, n, seed)


class _Finish(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fl, areas, nbr, normals, scale, lo, hi, n_smooth, alpha, blend_each):
        m = ops.normalize(fl.detach().contiguous(), areas, scale, lo, hi)
        if n_smooth > 0:
            m = ops.smooth(m, nbr, normals, n_smooth, alpha, blend_each)
        ctx.save_for_backward(fl.detach().contiguous(), areas, nbr, normals)
        ctx.args = (float(scale), float(lo), float(hi), int(n_smooth), float(alpha),
                    bool(blend_each))
        return m

    @staticmethod
    def backward(ctx, gm):
        fl, areas, nbr, normals = ctx.saved_tensors
        scale, lo, hi, n_smooth, alpha, blend_each = ctx.args
        g = gm.contiguous().float()
        if n_smooth > 0:
            g = ops.smooth_adjoint(nbr, normals, n_smooth, alpha, blend_each, g)
        return (ops.normalize_adjoint(fl, areas, scale, lo, hi, g), None, None, None, None, None,
                None, None, None, None)


def finish(fl, mesh, scale, hi, n_smooth, alpha, blend_each, lo=0.0):
    """clamp((fl / A) * scale, lo, hi) then the flux smoothing, as the coupled flux's finish
    step (lo = -inf, hi = inf and one final blend for the ion yields), differentiable in fl."""
    return _Finish.apply(fl, mesh.areas, mesh.nbr, mesh.normals, scale, lo, hi, n_smooth, alpha,
                         blend_each)


def coupled_coverage(mesh, geo, par, flags, m_i, cos_i, ion_yield, n_neu, seed, n_fp=None,
                     bare_init=None, flux_scale=None, betas=None):
    """(bare, m_F, m_O) of the neutral rounds of flux.coupled_flux_steps for given ion terms
    (m_i, cos_i, ion_yield): n_fp x {the F and O RR traces at bare (one paired launch with the
    seeds and streams of the coupled flux), finish, Belen update} from bare = 1, or from
    clamp(bare_init, 0, 1) (a warm start, itself differentiable). Differentiable in m_i (without
    ion yields) or the ion yields, through every round: the traces by their path-replay adjoint,
    finish by its transpose, Belen in closed form. flux_scale = (sF, sO), 0-d tensors, multiply
    m_F and m_O where the Belen update reads them: the F and O fluxes enter the chemistry only as
    Fflux * m_F and Oflux * m_O (belen.py:41-47), so sF = Fflux / par['Fflux'] is a change of the
    F flux, differentiable too. betas = (betaE, betaO), 0-d tensors, replace par's sticking
    factors of the traces and are differentiable (neutral_rr_pair_beta). The returned m_F, m_O
    are the unscaled normalised fluxes."""
    from .chemistry import chem_params
    from .flux import NEUTRAL_SOURCE, SEED_MASK, box_of
    F = mesh.nf
    A_src = geo['Lx'] * geo['Ly']
    betaE, betaO = par.get('betaE', 0.7), par.get('betaO', 1.0)
    n_smooth = int(par.get('flux_smooth', 1))
    alpha = float(par.get('flux_smooth_alpha', 1.0))
    blend_each = bool(par.get('device_flux', True)) and bool(par.get('flux_smooth_gpu', True))
    box = box_of(geo, par)
    chem = chem_params(par, flags)
    n_fp = int(par.get('n_fp', 4)) if n_fp is None else int(n_fp)
    if bare_init is None:
        bare = torch.ones((F,), dtype=torch.float32, device=mesh.device)
    else:
        bare = torch.clamp(bare_init.to(torch.float32), 0.0, 1.0)
    m_F = m_O = torch.zeros((F,), dtype=torch.float32, device=mesh.device)
    for it in range(n_fp):
        sF, sO = (seed * 9 + 2 + 2 * it) & SEED_MASK, (seed * 9 + 3 + 2 * it) & SEED_MASK
        if betas is None:
            fF, fO = neutral_rr_pair(mesh.bvh, box, NEUTRAL_SOURCE, bare, betaE, n_neu, sF, betaO,
                                     n_neu, sO)
        else:
            fF, fO = neutral_rr_pair_beta(mesh.bvh, box, NEUTRAL_SOURCE, bare, betas[0], n_neu,
                                          sF, betas[1], n_neu, sO)
        m_F = finish(fF, mesh, A_src / n_neu, 8.0, n_smooth, alpha, blend_each)
        m_O = finish(fO, mesh, A_src / n_neu, 8.0, n_smooth, alpha, blend_each)
        if flux_scale is None:
            bare = belen(m_i, m_F, m_O, cos_i, ion_yield, chem)
        else:
            bare = belen(m_i, m_F * flux_scale[0].float(), m_O * flux_scale[1].float(), cos_i,
                         ion_yield, chem)
    return bare, m_F, m_O


class _SurfaceRate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m_i, m_F, m_O, cos_i, iy_sp, iy_ie, iy_ip, centroids, mask_geom, chem,
                is_mask, Y):
        if Y is not None:
            chem = type(chem).from_buffer_copy(chem)
            chem.Yie, chem.Ysp, chem.Yp = (float(y) for y in Y.detach().cpu().tolist())
        iy = (iy_sp, iy_ie, iy_ip) if iy_ie is not None else None
        V, red = ops.surface_rate(m_i.contiguous(), m_F.contiguous(), m_O.contiguous(), cos_i,
                                  iy, centroids, mask_geom, chem, is_mask=is_mask)
        ctx.save_for_backward(m_i, m_F, m_O, cos_i, iy_sp, iy_ie, iy_ip, centroids, is_mask)
        ctx.args = (mask_geom, chem, (Y.device, Y.dtype) if Y is not None else None)
        return V

    @staticmethod
    def backward(ctx, gV):
        m_i, m_F, m_O, cos_i, iy_sp, iy_ie, iy_ip, centroids, is_mask = ctx.saved_tensors
        mask_geom, chem, want_y = ctx.args
        iy = (iy_sp, iy_ie, iy_ip) if iy_ie is not None else None
        out = ops.surface_rate_adjoint(
            m_i.contiguous(), m_F.contiguous(), m_O.contiguous(), cos_i, iy, centroids,
            mask_geom, chem, gV.contiguous().float(), is_mask=is_mask,
            per_face_yields=want_y is not None)
        gm_F, gm_O, gm_i, g_sp, g_ie, g_ip = out[:6]
        gY = out[6].double().sum(dim=1).to(device=want_y[0], dtype=want_y[1]) if want_y else None
        return gm_i, gm_F, gm_O, None, g_sp, g_ie, g_ip, None, None, None, None, gY


def surface_rate(m_i, m_F, m_O, cos_i, ion_yield, centroids, mask_geom, chem, is_mask=None,
                 Y=None):
    """V (ops.surface_rate), differentiable in m_F, m_O and the ion terms (m_i, or the ion
    yields when given). Y, a (3,) tensor, replaces chem's yield coefficients (Yie, Ysp, Yp) and
    is differentiable too: with Y computed from the ion energy in torch (e.g.
    A * relu(sqrt(E) - sqrt(Eth)) for the 'mean' IED, chemistry.py:10-39), a loss on the rate
    reaches the recipe's energy, as in scripts/inverse_design.py:20-45."""
    sp, ie, ip = ion_yield if ion_yield is not None else (None, None, None)
    return _SurfaceRate.apply(m_i, m_F, m_O, cos_i, sp, ie, ip, centroids, mask_geom, chem,
                              is_mask, Y)


class _Extend(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V, bvh, phi, dx, band):
        Fs = ops.extend(bvh, V.detach().contiguous(), phi, dx, band)
        ctx.save_for_backward(phi)
        ctx.args = (bvh, float(dx), float(band))
        return Fs

    @staticmethod
    def backward(ctx, gFs):
        (phi,) = ctx.saved_tensors
        bvh, dx, band = ctx.args
        g = ops.extend_adjoint(bvh, phi, dx, band, gFs.contiguous().float())
        _check_bvh(bvh, "extension adjoint")
        return g, None, None, None, None


def extend(bvh, V, phi, dx, band):
    """Fs = V[nearest face] on the band |phi| < band (ops.extend), differentiable in V. The BVH
    must be unchanged between the forward and the backward pass."""
    return _Extend.apply(V, bvh, phi, dx, band)


def advance(mesh, phi, mask, mask_phi, dx, dt, nsub, band, m_i, m_F, m_O, cos_i, ion_yield,
            mask_geom, chem):
    """The rest of an etch step after the flux (threed.py:1719-1760 without the reinit): the
    surface rate of the fluxes, its extension to the band |phi| < band and nsub upwind substeps
    of dt / nsub, differentiable in the fluxes (m_i, m_F, m_O, the ion yields) and in phi
    through the advection (the surface, its BVH and the nearest-face assignment stay fixed).
    With radiosity_coverage or coupled_coverage in front of it, a loss on the advanced level set
    reaches the ion terms:

        bare, m_F, m_O = adjoint.coupled_coverage(mesh, geo, par, flags, m_i, cos_i, None, n, s)
        phi1 = adjoint.advance(mesh, phi, mask, mask_phi, dx, dt, nsub, band, m_i, m_F, m_O,
                               cos_i, None, mask_geom, chem)
    """
    V = surface_rate(m_i, m_F, m_O, cos_i, ion_yield, mesh.centroids, mask_geom, chem)
    Fs = extend(mesh.bvh, V, phi.detach().contiguous(), dx, band)
    return advect(phi, Fs, mask, mask_phi, dx, dt / nsub, nsub)


class _IonYield(torch.autograd.Function):
    @staticmethod
    def forward(ctx, E, bvh, box, src, ion, n, seed):
        ion = type(ion).from_buffer_copy(ion)
        ion.meanE = float(E.detach().cpu())
        acc = ops.trace_ion_yield(bvh, box, src, ion, n, seed)
        ctx.args = (bvh, box, src, ion, int(n), int(seed), E.device, E.dtype)
        return acc

    @staticmethod
    def backward(ctx, gacc):
        bvh, box, src, ion, n, seed, dev, dt = ctx.args
        t = ops.trace_ion_yield_tangent(bvh, box, src, ion, n, seed)
        _check_bvh(bvh, "ion trace tangent")
        gE = (gacc[:3].double() * t.double()).sum()
        return gE.to(device=dev, dtype=dt), None, None, None, None, None, None


def ion_yield(bvh, box, src, ion, n, seed, E):
    """The faithful-ion accumulators (5, F) of ops.trace_ion_yield with the mean ion energy
    taken from the 0-d tensor E, differentiable in E: the backward pass traces the same ions
    once more in forward mode (petch_trace_ion_yield_tangent), since E is one scalar. The BVH
    must be unchanged between the forward and the backward pass."""
    return _IonYield.apply(E, bvh, box, src, ion, n, seed)


def ion_terms(mesh, geo, par, n_ion, seed, E):
    """(m_i, cos_i, ion_yield) of the faithful-ion part of flux.coupled_flux_steps
    (threed.py:806-823) with the mean ion energy from the 0-d tensor E, differentiable in E: the
    ion trace by its forward-mode derivative, the normalisation and smoothing by their
    transposes. cos_i does not depend on the energy."""
    from .flux import SEED_MASK, box_of, ion_params, ion_source
    A_src = geo['Lx'] * geo['Ly']
    n_smooth = int(par.get('flux_smooth', 1))
    alpha = float(par.get('flux_smooth_alpha', 1.0))
    acc = ion_yield(mesh.bvh, box_of(geo, par), ion_source(par), ion_params(par), n_ion,
                    (seed * 9 + 1) & SEED_MASK, E)
    norm = A_src / n_ion

    def fin(x):
        return finish(x, mesh, norm, ops.INF, n_smooth, alpha, False, lo=-ops.INF)

    iy = tuple(fin(acc[k]) for k in range(3))
    m_i = fin(acc[3])
    cos_i = ops.cos_ratio(acc[3].detach().contiguous(), acc[4].detach().contiguous())
    return m_i, cos_i, iy


class _SliceArea(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, dx, eps):
        p = phi.detach().contiguous().float()
        ctx.save_for_backward(p)
        ctx.args = (float(dx), float(eps))
        return ops.slice_gas_area(p, dx, eps)

    @staticmethod
    def backward(ctx, garea):
        (p,) = ctx.saved_tensors
        dx, eps = ctx.args
        return ops.slice_gas_area_adjoint(p, dx, eps, garea.double()), None, None


def slice_gas_area(phi, dx, eps):
    """ops.slice_gas_area, differentiable in phi."""
    return _SliceArea.apply(phi, dx, eps)


def profile_rz(phi, dx, eps=None):
    """Equivalent radius r(z) = sqrt(A(z) / pi) of every z slice (float64 (nz,)), with A the
    smoothed gas area (H_eps(-phi), eps = 1.5 dx by default), differentiable in phi. A slice
    with no gas has r = 0 and passes no gradient."""
    eps = 1.5 * dx if eps is None else eps
    A = slice_gas_area(phi, dx, eps)
    return torch.sqrt(A.clamp_min(1e-30) / math.pi)


def profile_loss(phi, dx, target, zsel=None, eps=None):
    """sum over the selected z slices of (r(z) - target(z))^2, differentiable in phi; zsel is a
    boolean mask or index tensor of slices (all slices by default)."""
    r = profile_rz(phi, dx, eps)
    t = torch.as_tensor(target, dtype=torch.float64, device=r.device)
    d = r - t
    if zsel is not None:
        d = d[torch.as_tensor(zsel, device=r.device)]
    return (d * d).sum()

