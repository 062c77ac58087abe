"""A whole etch run differentiable in the recipe: the gradient of a loss on the final level set
(typically the profile loss of adjoint.profile_loss) with respect to the mean ion energy, the F
and O fluxes and the rate scale.

The reference differentiates nothing in its etch loop; its only gradients are demo tapes on a
frozen flux (scripts/inverse_design.py:95-101, scripts/warp_3d_flux.py:106-113). Here every
step of run_etch_3d (threed.py:1671-1792) is rebuilt from the differentiable pieces of
petch.adjoint, with the seeds, warm start, fixed-point rounds, CFL substeps and reinit of
EtchLoop, so the forward is the etch loop's own (checked against run_etch_3d):

    ion terms     adjoint.ion_terms            forward-mode ion trace in the mean energy
    coverage      adjoint.coupled_coverage     RR path-replay adjoints, finish, Belen; warm
                                                start from the previous step's coverage
                                                (nearest previous face, a gather)
    rate          adjoint.surface_rate         the fluxes scaled by Fflux / Fflux0 and
                                                Oflux / Oflux0 (the chemistry reads only the
                                                products), the rate by rate_scale / rate_scale0
    extension     adjoint.extend               nearest-face scatter
    advection     adjoint.advect               upwind adjoint per substep
    reinit        adjoint.reinit               FSM sweeps transposed, choices held

What is held fixed (straight-through), and why:
- the surface geometry of each step: marching cubes, the BVH, the nearest-face assignment of
  the extension and the warm start are discontinuous in phi (topology, argmin), so each step's
  mesh is built from the detached level set and the flux's dependence on the shape of the
  surface is not differentiated;
- the CFL substep count (an integer from the host-read maximum speed).
The reinit is differentiated (adjoint.reinit: the exact adjoint of its sweeps with their
discrete choices held). Passing the gradient through it as the identity instead (the reinit
keeps the zero level set) measured 10-11 % off frozen-geometry differences over 20 steps of
the small hole; the exact adjoint is within 0.4 %.

Checkpointing: with checkpoint_every = k the run is split into segments of k steps; only the
level set, the previous step's coverage and the parameters cross a segment boundary, each
segment's graph is dropped after its forward and rebuilt from the segment's first level set in
the backward pass (torch.utils.checkpoint), which re-runs its traces with the same Philox
streams. Without it the whole graph is kept (a few MB per step at these sizes)."""
import torch

from . import adjoint, ops
from . import flux as _flux
from .params import PAR

PARAMS = ("Emean", "Fflux", "Oflux", "rate_scale")
PARAMS5 = PARAMS + ("betaE",)       # with the F sticking factor of the neutral traces


class DiffEtch:
    """run_etch_3d's loop (fsm reinit every step, GPU extension and warm start, faithful ions)
    as a function of theta = (Emean, Fflux, Oflux, rate_scale), a float64 (4,) tensor."""

    def __init__(self, Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole, t_end, n_steps,
                 par=None, flags=None, n_ion=20000, n_neu=20000, seed_offset=0, params=PARAMS):
        from .chemistry import chem_params
        from .threed import _dev, _trench_geo
        par = dict(PAR if par is None else par)
        if flags is None or not _flux._faithful(flags) or \
                getattr(flags, "chemistry", "langmuir") != "belen" or \
                not getattr(flags, "coverage_sticking", False) or \
                getattr(flags, "neutral_transport", "mc") != "mc":
            raise NotImplementedError("DiffEtch follows the faithful-ion Monte-Carlo coupled flux "
                                      "with the Belen chemistry (SF6O2's flags)")
        _dev()
        geo = _trench_geo(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole)
        self.phi0 = geo.pop('_phi_d')
        self.mask = geo.pop('_mask_d')
        self.mask_phi = self.phi0.clone()
        self.geo, self.par, self.flags = geo, par, flags
        self.dx, self.dt, self.n_steps = dx, t_end / n_steps, n_steps
        self.band = 4 * dx
        self.n_ion, self.n_neu, self.seed_offset = n_ion, n_neu, seed_offset
        self.mg = ops.MaskGeom(float(sub_top), float(mask_th), float(Lx), float(Ly),
                               float(trench_width), int(bool(hole)))
        self.chem = chem_params(par, flags)
        if tuple(params) not in (PARAMS, PARAMS5):
            raise ValueError(f"DiffEtch params must be {PARAMS} or {PARAMS5}")
        self.params = tuple(params)
        self.theta0 = torch.tensor([float(par[k]) for k in self.params], dtype=torch.float64)
        self.rays = int(n_ion) + 2 * int(n_neu) * max(1, int(par.get('n_fp', 1)))
        self.warm = getattr(flags, "warm_start_coverage", False)
        self.nsub = {}      # step -> CFL substeps of the last forward (a recompute reuses them)
        self._recompute = False
        self._frozen = self._record = None
        self._seen = dict(mesh={}, face={}, nsub=self.nsub)

    def step(self, step, phi, prev, theta):
        """One differentiable step: (phi after advection and reinit, (mesh, bare) for the next
        step's warm start)."""
        dev = phi.device
        dx = self.dx
        # a checkpointed segment's recompute replays the forward's geometry: float-atomic
        # noise in the recomputed traces can move a node of phi across zero, and the recompute
        # must build the same graph
        fz = self._frozen if self._frozen is not None else (self._seen if self._recompute
                                                             else None)
        if fz is not None:
            mesh = fz['mesh'][step]
        else:
            mesh = _flux.SurfaceMesh(dev, self.rays).build(phi.detach().contiguous(), dx)
        self._seen['mesh'][step] = mesh
        if self._record is not None:
            self._record['mesh'][step] = mesh
        if mesh.nf == 0:
            return phi, None
        th = theta.to(dev)
        E = th[0]
        sF, sO, sR = (th[1:4] / self.theta0[1:4].to(dev)).unbind()
        betas = None
        if len(self.params) > 4:
            betas = (th[4], torch.tensor(float(self.par.get('betaO', 1.0)), dtype=torch.float64,
                                         device=dev))
        seed = step + self.seed_offset
        m_i, cos_i, iy = adjoint.ion_terms(mesh, self.geo, self.par, self.n_ion, seed, E)
        bare_init, n_fp = None, int(self.par.get('n_fp', 4))
        if self.warm and prev is not None and bool(torch.isfinite(mesh.centroids).all()):
            pmesh, pbare = prev
            if fz is not None:
                face = fz['face'][step]
            else:
                face, _ = pmesh.bvh.query_point(mesh.centroids)
            self._seen['face'][step] = face
            if self._record is not None:
                self._record['face'][step] = face
            hit = face >= 0
            bare_init = torch.where(hit, pbare[face.long().clamp_min(0)],
                                    torch.ones_like(mesh.areas))
            n_fp = int(self.par.get('n_fp', 1))
        bare, m_F, m_O = adjoint.coupled_coverage(mesh, self.geo, self.par, self.flags, m_i,
                                                  cos_i, iy, self.n_neu, seed, n_fp,
                                                  bare_init=bare_init, flux_scale=(sF, sO),
                                                  betas=betas)
        V = adjoint.surface_rate(m_i, m_F * sF.float(), m_O * sO.float(), cos_i, iy,
                                 mesh.centroids, self.mg, self.chem)
        V = V * sR.float()
        nsub = fz['nsub'][step] if fz is not None else None
        if nsub is None:
            with torch.no_grad():
                vmx = float(torch.nan_to_num(V, nan=0.0, posinf=0.0, neginf=0.0).max())
            nsub, _ = _flux.cfl_substeps(vmx, self.dt, dx)
            self.nsub[step] = nsub
        Fs = adjoint.extend(mesh.bvh, V, phi.detach().contiguous(), dx, self.band)
        a = adjoint.advect(phi, Fs, self.mask, self.mask_phi, dx, self.dt / nsub, nsub,
                           recompute=True)
        return adjoint.reinit(a, dx, self.band + 2.0 * dx), (mesh, bare)

    def run(self, theta, checkpoint_every=None, record=False, frozen=None):
        """The final level set (float32, the grid's shape), differentiable in theta.

        record: keep this run's discrete choices in self.record (each step's surface mesh and
        BVH, warm-start face map and CFL substeps). frozen: a record to replay them from, so
        that the run is the smooth function of theta the adjoint differentiates (its central
        differences check the adjoint without the jumps of the geometry)."""
        theta = theta.to(torch.float64)
        phi = self.phi0.clone()
        prev = None
        self.nsub = {}
        self._recompute = False
        self._frozen = frozen
        self._record = dict(mesh={}, face={}, nsub=self.nsub) if record else None
        self._seen = dict(mesh={}, face={}, nsub=self.nsub)
        k = self.n_steps if not checkpoint_every else int(checkpoint_every)
        s0 = 0
        while s0 < self.n_steps:
            s1 = min(s0 + k, self.n_steps)
            if checkpoint_every:
                phi, prev = self._segment_checkpointed(s0, s1, phi, prev, theta)
            else:
                for s in range(s0, s1):
                    phi, prev = self.step(s, phi, prev, theta)
            s0 = s1
        if record:
            self.record = self._record
        self._record = None
        return phi

    def _segment_checkpointed(self, s0, s1, phi, prev, theta):
        from torch.utils.checkpoint import checkpoint
        pmesh = prev[0] if prev is not None else None
        pbare = prev[1] if prev is not None else torch.zeros(0, device=phi.device)
        holder = {}

        def seg(phi_in, pbare_in, th):
            # the second call of a segment is the backward pass's recompute: it replays the
            # forward's meshes, warm-start face maps and substep counts
            self._recompute = holder.get('done', False)
            holder['done'] = True
            p = (pmesh, pbare_in) if pmesh is not None else None
            x = phi_in
            for s in range(s0, s1):
                x, p = self.step(s, x, p, th)
            holder['mesh'] = p[0] if p is not None else None
            bare = p[1] if p is not None else torch.zeros(0, device=phi_in.device)
            return x, bare

        phi_out, bare_out = checkpoint(seg, phi, pbare, theta, use_reentrant=False,
                                       preserve_rng_state=False)
        self._recompute = False
        return phi_out, ((holder['mesh'], bare_out) if holder.get('mesh') is not None else None)


def profile_target(dd, theta, zsel=None, eps=None):
    """r(z) of a forward run at theta (no graph): the target of a profile loss."""
    with torch.no_grad():
        phi = dd.run(torch.as_tensor(theta, dtype=torch.float64))
        return adjoint.profile_rz(phi, dd.dx, eps)


def loss_and_gradient(dd, theta, target, zsel, eps=None, checkpoint_every=None, record=False):
    """(loss, d loss / d theta) of adjoint.profile_loss on the final level set of dd.run."""
    th = torch.as_tensor(theta, dtype=torch.float64).clone().requires_grad_(True)
    phi = dd.run(th, checkpoint_every=checkpoint_every, record=record)
    loss = adjoint.profile_loss(phi, dd.dx, target, zsel, eps)
    loss.backward()
    return float(loss.detach()), th.grad.detach().clone()


def loss_only(dd, theta, target, zsel, eps=None, frozen=None):
    """The same loss without a graph (a forward run of the same loop, or of a recorded run's
    geometry and substeps with `frozen`)."""
    with torch.no_grad():
        phi = dd.run(torch.as_tensor(theta, dtype=torch.float64), frozen=frozen)
        return float(adjoint.profile_loss(phi, dd.dx, target, zsel, eps))


def zsel_feature(dd, below=None):
    """Boolean (nz,) mask of the slices between `below` um under the substrate top (default:
    the whole substrate) and the mask top: the feature's walls."""
    zs = dd.geo['zs']
    top = dd.geo['sub_top'] + float(dd.mg.mask_th)
    lo = dd.geo['sub_top'] - (below if below is not None else dd.geo['sub_top'])
    return torch.as_tensor((zs >= lo) & (zs < top))


# The batched differentiable configuration: 256 recipe variants of one deep trench (width
# 1.5 um, 20 um of substrate under a 2 um mask, dx 0.1 um: 60 x 10 x 240 nodes), 30 steps over
# 6 s, 1M rays per species and step, the profile loss of each variant against its own target,
# gradients in five parameters (with the F sticking factor betaE), checkpoints every 6 steps.
CONFIG4 = dict(Lx=6.0, Ly=1.0, Lz=24.0, dx=0.1, trench_width=1.5, mask_th=2.0, sub_top=20.0,
               hole=False, t_end=6.0, n_steps=30)
PERTURB5 = (1.10, 0.90, 1.10, 1.10, 0.90)


def variant_recipes(n, seed=4321):
    """The variants' recipes (SF6O2 defaults, lateral periodic): mean ion energy U(50, 200) eV
    and O/F flux ratio U(0.1, 1.0), seeded (as the batched features draw theirs)."""
    import numpy as np
    from .api import SF6O2
    r = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = SF6O2(meanE=float(r.uniform(50, 200)))
        par = dict(m.par)
        par['Oflux'] = float(r.uniform(0.1, 1.0)) * par['Fflux']
        par['periodic_y'] = 1
        out.append((par, m.flags))
    return out


def variant_gradients(n_variants, n_rays=1_000_000, checkpoint_every=6, seed=4321, case=None,
                      target_seed=1234, progress=None):
    """Loss and five-parameter gradient of every variant (one after the other; each variant's
    run is its own DiffEtch): (list of (loss, gradient), seconds). The target of variant v is
    the profile of its own run with seed offset `target_seed` at the parameters PERTURB5 x its
    recipe's. progress(done, seconds) is called after each variant."""
    import time
    from .params import Flags
    case = dict(CONFIG4 if case is None else case)
    res = []
    t0 = time.time()
    for par, flags in variant_recipes(n_variants, seed):
        dd = DiffEtch(**case, par=par, flags=Flags(**flags), n_ion=n_rays, n_neu=n_rays,
                      params=PARAMS5)
        tg = DiffEtch(**case, par=par, flags=Flags(**flags), n_ion=n_rays, n_neu=n_rays,
                      seed_offset=target_seed, params=PARAMS5)
        target = profile_target(tg, tg.theta0 * torch.tensor(PERTURB5, dtype=torch.float64))
        zsel = zsel_feature(dd)
        res.append(loss_and_gradient(dd, dd.theta0, target, zsel,
                                     checkpoint_every=checkpoint_every))
        if progress is not None:
            progress(len(res), time.time() - t0)
    torch.cuda.synchronize()
    return res, time.time() - t0
