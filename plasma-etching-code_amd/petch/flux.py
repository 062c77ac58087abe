"""Device-resident flux stages: the surface mesh of a step and the three neutral-transport
orchestrators the etch loop dispatches to (threed.py:1687-1718). All tensors stay on the GPU;
every kernel is in libpetch_hip.so.

  SurfaceMesh          MC + attributes + adjacency + LBVH of one step (threed.py:1673-1675)
  coupled_flux         mc_flux_3d_coupled    threed.py:740-905
  radiosity_flux       mc_flux_3d_radiosity  threed.py:923-991 (+ _ions_deterministic 994-1029)
  uncoupled_flux       mc_flux_3d            threed.py:472-507
"""
import math

import torch

from . import charging as _charging
from . import ops
from .chemistry import chem_params
from .profiling import NULL as _NULL_TIMER

SEED_MASK = 0xFFFFFFFFFFFFFFFF


class SurfaceMesh:
    """The triangle surface of phi with everything the flux and extension stages need.

    rays_per_step (optional): the rays a step will trace on this surface. A clustered BVH
    (slower build, faster traces) pays off from ~1000 rays per face; below that the fast
    Morton build is used. Without it the BVH picks by mesh size."""

    def __init__(self, device, rays_per_step=None):
        self.device = device
        self.bufs = ops.MeshBuffers(device)
        self.bvh = ops.Bvh(device)
        self.rays_per_step = rays_per_step
        self.quality = None    # a fixed BVH build quality (petch.autotune's choice), or None
        self._quality = None
        self.nf = 0
        self.nv = 0

    def build(self, phi, dx):
        return self.build_from(*ops.marching_cubes(phi, dx, self.bufs))

    def build_from(self, v, f, bvh=True, surf=None):
        """Everything after marching cubes, from its (verts, faces) (the batched loop meshes
        all features' grids with one host read, then builds each feature's surface). bvh=False
        leaves the BVH to the caller (the batched loop builds all features' trees together with
        ops.bvh_build_batch over bvh_item()); with a list `surf` the face attributes and
        adjacency are only allocated and their item appended, for ops.surface_build_batch."""
        self.verts, self.faces = v, f
        self.nv, self.nf = int(v.shape[0]), int(f.shape[0])
        if surf is None:
            self.centroids, self.areas, self.normals = ops.mesh_attrs(v, f)
            self.nbr = ops.face_adjacency(f, self.nv)
        else:
            nf, dev = self.nf, v.device
            self.centroids = torch.empty((nf, 3), dtype=torch.float32, device=dev)
            self.areas = torch.empty((nf,), dtype=torch.float32, device=dev)
            self.normals = torch.empty((nf, 3), dtype=torch.float32, device=dev)
            self.nbr = torch.empty((nf, 3), dtype=torch.int32, device=dev)
            self.finite = torch.empty((), dtype=torch.int32, device=dev)
            surf.append((v, f, self.centroids, self.areas, self.normals, self.nbr, self.finite))
        if self.quality is not None or self.rays_per_step is not None:
            q = self.quality
            if q is None:
                q = 2 if self.rays_per_step >= 1000 * max(self.nf, 1) else 0
            if q != self._quality:
                self.bvh.set_quality(q)
                self._quality = q
        if bvh:
            self.bvh.build(v, f)
        return self

    def bvh_item(self):
        """(bvh, verts, faces) of this surface for ops.bvh_build_batch."""
        return (self.bvh, self.verts, self.faces)


def ion_params(par):
    """Faithful-ion constants as passed at threed.py:806-813 (derived at 800-802)."""
    inflect = float(par.get('inflectAngle', 1.55334303))
    minang = float(par.get('minAngle', 1.3962634))
    n_l = float(par.get('n_l', 10.0))
    A_en = 1.0 / (1.0 + n_l * (1.5707963 / inflect - 1.0))
    minE = float(min(par['Eth_ie'], par['Eth_sp']))
    return ops.IonParams(
        meanE=float(par['Emean']), sigmaE=float(par['Esig']), A_sp=float(par['A_sp']),
        Eth_sp=float(par['Eth_sp']), B_sp=float(par['B_sp']), A_ie=float(par['A_ie']),
        Eth_ie=float(par['Eth_ie']), A_p=float(par['A_p']), Eth_p=float(par['Eth_p']),
        inflect=inflect, minang=minang, n_l=n_l, A_en=float(A_en), minE=minE,
        thrmin=float(par.get('thetaRMin', 1.2217305)), thrmax=float(par.get('thetaRMax', 1.5707963)))


def ion_source(par):
    """Ion source: the reference's Gaussian tilt (sigma = ion_ang_sigma, threed.py:450-453), or
    a power-cosine cone cos^n about -z with par['ion_source'] = 'powcos' and par['ion_exponent']."""
    if par.get('ion_source', 'gauss') == 'powcos':
        return ops.make_source(ops.SRC_ION_POWCOS, 0.0, float(par.get('ion_exponent', 1000.0)))
    return ops.make_source(ops.SRC_ION_GAUSS, float(par['ion_ang_sigma']))


def box_of(geo, par):
    """Tracer box: source plane z = Lz - dx (threed.py:477), periodic wrap from par."""
    return ops.make_box(geo['Lx'], geo['Ly'], geo['Lz'], geo['Lz'] - geo['dx'],
                        int(par.get('periodic_y', 0)))


def _faithful(flags):
    return (flags is not None and getattr(flags, "ion_reflection", False)
            and getattr(flags, "chemistry", "langmuir") == "belen")


NEUTRAL_SOURCE = ops.make_source(ops.SRC_NEUTRAL)


def run_consts(geo, par, flags):
    """The per-run constants of coupled_flux_steps: the tracer box, the ion source and
    faithful-ion constants, and the chemistry constants."""
    return dict(box=box_of(geo, par), src=ion_source(par), ion=ion_params(par),
                chem=chem_params(par, flags))


def coupled_flux_steps(mesh, geo, par, n_ion, n_neu, seed, flags, bare_init=None, rounds=None,
                       phi=None, bare_in_unit=False, consts=None):
    """mc_flux_3d_coupled (threed.py:740-905) as a generator over its ray traces.

    It yields ("ion", request) and ("neutral", request) dicts and is sent the trace outputs
    back, so one driver can run it alone (run_traces) and another can advance many features in
    lockstep with one batched launch per trace (petch.batch). Its return value (StopIteration)
    is dict(m_i, m_F, m_O, cos_i, bare, ion_yield) of (F,) float32 device tensors; ion_yield =
    (sp, ie, ip) normalised (the par['_ion_yield'] side channel) or None.

    rounds: a list that, when given, receives the end of the faithful-ion trace (normalise and
    smooth the yields and arrivals, the cosine ratio; kind "ion", an ops.finish_ion_batch item)
    and of each neutral round (normalise and smooth both species, Belen update; kind "neutral",
    an ops.finish_round_batch item) with their output tensors allocated, instead of the generator
    launching them; the driver must run the collected items before the next trace it launches and
    before it reads the result. Used when smoothing takes at most one pass (the default
    flux_smooth = 1) and no charging hook needs the ion arrival at once.

    phi: the level set on the device, needed by the floor-charge throttle (its gas normals).
    bare_in_unit: bare_init is float32 and already in [0, 1] (a warm start from the previous
    step's coverage), so its clip to [0, 1] is the identity and is skipped.
    consts: the run's constants (box, ion_source, ion_params, chem) computed once from the same
    geo, par and flags (run_consts), so a loop does not rebuild them every step."""
    F = mesh.nf
    Lx, Ly = geo['Lx'], geo['Ly']
    A_src = Lx * Ly
    betaE = par.get('betaE', 0.7)
    betaO = par.get('betaO', 1.0)
    n_smooth = int(par.get('flux_smooth', 1))
    alpha = float(par.get('flux_smooth_alpha', 1.0))
    # device-resident normalize+smooth of the reference's GPU path (threed.py:757-759, 783-790)
    dev = bool(par.get('device_flux', True)) and bool(par.get('flux_smooth_gpu', True))
    if consts is None:
        consts = run_consts(geo, par, flags)
    box = consts['box']
    A = mesh.areas

    def host_smooth(x):                       # smooth_flux / smooth_flux_gpu semantics
        return ops.smooth(x, mesh.nbr, mesh.normals, n_smooth, alpha, False) if n_smooth > 0 else x

    def finish(fl, scale, hi):                # _finish (threed.py:783-790) or the host clip+smooth
        m = ops.normalize(fl, A, scale, 0.0, hi)
        if n_smooth <= 0:
            return m
        return ops.smooth(m, mesh.nbr, mesh.normals, n_smooth, alpha, dev)

    ion_seed = (seed * 9 + 1) & SEED_MASK
    if _faithful(flags):
        acc = yield ("ion", dict(bvh=mesh.bvh, box=box, src=consts['src'], ion=consts['ion'],
                                 n=n_ion, seed=ion_seed))
        norm = A_src / n_ion
        if rounds is not None and n_smooth <= 1 and not _charging.active(par, flags):
            out = torch.empty_like(acc)
            rounds.append(dict(kind="ion", acc=acc, areas=A, nbr=mesh.nbr, normals=mesh.normals,
                               scale=norm, n_smooth=max(n_smooth, 0), alpha=alpha, out=out))
            ion_yield, m_i, cos_i = (out[0], out[1], out[2]), out[3], out[4]
        else:
            ion_yield = tuple(host_smooth(ops.normalize(acc[k], A, norm, -ops.INF, ops.INF))
                              for k in range(3))
            m_i = host_smooth(ops.normalize(acc[3], A, norm, -ops.INF, ops.INF))
            cos_i = ops.cos_ratio(acc[3], acc[4])
    else:
        acc = ops.trace_legacy(mesh.bvh, box, ion_source(par), 1.0, 0, 0, 0.34, 0.8, n_ion,
                               ion_seed)
        ion_yield = None
        cos_i = ops.cos_ratio(acc[0], acc[1])
        m_i = finish(acc[0], A_src / n_ion, 1.5)

    # surface charging, alpha model (threed.py:842-868)
    if flags is not None and getattr(flags, "surface_charging", False) and \
            float(par.get('charge_alpha', 0.0)) > 0.0:
        m_i = m_i * _charging.electron_factor(mesh, box, A_src, par, m_i, n_neu, seed,
                                              host_smooth)
    # the deep-AR floor-charge throttle (threed.py:893-894)
    if flags is not None and getattr(flags, "floor_charge_throttle", False):
        if phi is None:
            raise ValueError("floor_charge_throttle needs the level set (phi) for the gas normals")
        gn = ops.gas_normals(phi, geo['dx'], mesh.centroids, mesh.normals)
        m_i, ion_yield = _charging.floor_throttle(mesh, gn, geo, m_i, ion_yield)

    cap = par.get('cov_transport', 'rr') != 'rr'
    chem = consts['chem']

    n_fp = int(par.get('n_fp', 1 if bare_init is not None else 4))
    if bare_init is None:
        bare = torch.ones((F,), dtype=torch.float32, device=mesh.device)
    else:
        bare = bare_init if bare_in_unit else torch.clamp(bare_init.to(torch.float32), 0.0, 1.0)
    if n_fp <= 0:
        m_F = m_O = torch.zeros((F,), dtype=torch.float32, device=mesh.device)
    for it in range(n_fp):
        # the F and O traces of one iteration (threed.py:876-880), seeds seed*9+2+2it and
        # seed*9+3+2it, as one launch
        sF, sO = (seed * 9 + 2 + 2 * it) & SEED_MASK, (seed * 9 + 3 + 2 * it) & SEED_MASK
        if cap:
            # cov_transport='cap' (threed.py:881-885): the capped trace, one launch per species,
            # re-emissions capped at n_reemit_cov or clip(8 / max(beta, 0.02), 24, 200)
            fF, fO = (ops.trace_neutral_cap(
                mesh.bvh, box, NEUTRAL_SOURCE, bare, b,
                int(par.get('n_reemit_cov', min(max(8.0 / max(b, 0.02), 24.0), 200.0))), n_neu, s)
                for b, s in ((betaE, sF), (betaO, sO)))
        else:
            fF, fO = yield ("neutral", dict(bvh=mesh.bvh, box=box, bare=bare, beta_a=betaE,
                                            n_a=n_neu, seed_a=sF, beta_b=betaO, n_b=n_neu,
                                            seed_b=sO))
        if rounds is not None and n_smooth <= 1:
            m_F, m_O, bare = (torch.empty_like(fF) for _ in range(3))
            rounds.append(dict(kind="neutral", flF=fF, flO=fO, areas=A, nbr=mesh.nbr,
                               normals=mesh.normals,
                               scale=A_src / n_neu, hi=8.0, n_smooth=max(n_smooth, 0),
                               alpha=alpha, m_i=m_i, cos_i=cos_i, ion_yield=ion_yield, chem=chem,
                               m_F=m_F, m_O=m_O, bare=bare))
            continue
        m_F, m_O = finish(fF, A_src / n_neu, 8.0), finish(fO, A_src / n_neu, 8.0)
        bare = ops.belen_coverage(m_i, m_F, m_O, cos_i, ion_yield, chem)
    return dict(m_i=m_i, m_F=m_F, m_O=m_O, cos_i=cos_i, bare=bare, ion_yield=ion_yield)


def run_traces(steps, timer=_NULL_TIMER):
    """Drives a coupled_flux_steps generator with single-feature traces; returns its result."""
    try:
        req = next(steps)
        while True:
            kind, r = req
            if kind == "ion":
                with timer.region("ion_trace"):
                    out = ops.trace_ion_yield(r['bvh'], r['box'], r['src'], r['ion'], r['n'],
                                              r['seed'])
            else:
                with timer.region("neutral_trace"):
                    out = ops.trace_neutral_rr_pair(r['bvh'], r['box'], NEUTRAL_SOURCE, r['bare'],
                                                    r['beta_a'], r['n_a'], r['seed_a'],
                                                    r['beta_b'], r['n_b'], r['seed_b'])
            req = steps.send(out)
    except StopIteration as stop:
        return stop.value


def coupled_flux(mesh, geo, par, n_ion, n_neu, seed, flags, bare_init=None, timer=_NULL_TIMER,
                 phi=None):
    """mc_flux_3d_coupled (threed.py:740-905) on the device, one feature.

    Returns dict(m_i, m_F, m_O, cos_i, bare, ion_yield) of (F,) float32 device tensors;
    ion_yield = (sp, ie, ip) normalised (the par['_ion_yield'] side channel) or None. phi: the
    level set on the device (the floor-charge throttle needs it)."""
    return run_traces(coupled_flux_steps(mesh, geo, par, n_ion, n_neu, seed, flags, bare_init,
                                         phi=phi), timer)


def deterministic_ions(mesh, geo, par, n_ion, seed, flags):
    """_ions_deterministic (threed.py:994-1029): faithful ion without smoothing, or the
    first-hit ion clipped at 1.5. Returns (m_i, cos_i, ion_yield)."""
    box = box_of(geo, par)
    A_src = geo['Lx'] * geo['Ly']
    A = mesh.areas
    if _faithful(flags):
        acc = ops.trace_ion_yield(mesh.bvh, box, ion_source(par), ion_params(par), n_ion,
                                  seed & SEED_MASK)
        norm = A_src / n_ion
        iy = tuple(ops.normalize(acc[k], A, norm, -ops.INF, ops.INF) for k in range(3))
        m_i = ops.normalize(acc[3], A, norm, -ops.INF, ops.INF)
        return m_i, ops.cos_ratio(acc[3], acc[4]), iy
    acc = ops.trace_legacy(mesh.bvh, box, ion_source(par), 1.0, 0, 0, 0.34, 0.8, n_ion,
                           seed & SEED_MASK)
    return ops.normalize(acc[0], A, A_src / n_ion, 0.0, 1.5), ops.cos_ratio(acc[0], acc[1]), None


def _deterministic_neutrals(mesh, phi, geo, par, n_ion, seed, flags, n_fp, neutrals,
                            ions=None):
    """The coverage loop shared by the deterministic-neutral transports (threed.py:1149-1244):
    the deterministic ion launch (then ions(m_i, ion_yield), when given), then n_fp rounds of
    {m_F, m_O = neutrals(bare); Belen coverages}. neutrals(bare) returns the two species'
    multipliers."""
    F = mesh.nf
    m_i, cos_i, ion_yield = deterministic_ions(mesh, geo, par, n_ion, seed * 9 + 1, flags)
    if ions is not None:
        m_i, ion_yield = ions(m_i, ion_yield)
    chem = chem_params(par, flags)
    bare = torch.ones((F,), dtype=torch.float32, device=mesh.device)
    m_F = m_O = None
    for _ in range(n_fp):
        m_F, m_O = neutrals(bare)
        bare = ops.belen_coverage(m_i, m_F, m_O, cos_i, ion_yield, chem)
    if m_F is None:
        m_F = m_O = torch.zeros((F,), dtype=torch.float32, device=mesh.device)
    return dict(m_i=m_i, m_F=m_F, m_O=m_O, cos_i=cos_i, bare=bare, ion_yield=ion_yield)


def knudsen_flux(mesh, phi, geo, par, n_ion, seed, flags, n_fp=4):
    """mc_flux_3d_knudsen (threed.py:1149-1208): Monte-Carlo ions, neutrals from the 1-D
    Knudsen conductance tail down the feature (knudsen.py) with the floor sink at the local
    coverage, or clamped at the open field's mean coverage (par['knudsen_sink'] = 'field')."""
    dx = geo['dx']
    gn = ops.gas_normals(phi, dx, mesh.centroids, mesh.normals)
    betaE, betaO = par.get('betaE', 0.7), par.get('betaO', 1.0)
    sink = par.get('knudsen_sink', 'local')
    if sink not in ('local', 'field'):
        raise ValueError(f"knudsen_sink must be 'local' or 'field', got {sink!r}")
    kw = dict(Ly=geo['Ly'], wall_loss_scale=float(par.get('knudsen_wall_loss_scale', 1.85)),
              front_loss=bool(getattr(flags, 'knudsen_front_loss', False)),
              band_W=float(par.get('knudsen_front_band_W', 8.0)),
              pass_frac=float(par.get('knudsen_passive_frac', 0.5)),
              ar_pass=float(par.get('knudsen_front_ar_pass', 15.0)),
              center=(0.5 * geo['Lx'], 0.5 * geo['Ly']),
              half_width=0.5 * float(geo.get('trench_width', geo['Lx'])),
              sink_field=sink == 'field')
    trench = not geo.get('hole', False)

    def neutrals(bare):
        return ops.knudsen_face_flux(phi, dx, geo['sub_top'], trench, mesh.centroids, gn,
                                     mesh.areas, bare, (betaE, betaO), **kw)

    def ions(m_i, ion_yield):                       # threed.py:1182 (surface_charging='hg')
        return _charging.hg_closure(mesh, gn, geo, par, flags, m_i, ion_yield)
    return _deterministic_neutrals(mesh, phi, geo, par, n_ion, seed, flags, n_fp, neutrals,
                                   ions)


def dda_flux(mesh, phi, geo, par, n_ion, seed, flags, n_fp=4):
    """mc_flux_3d_dda (threed.py:1211-1244): Monte-Carlo ions, neutrals from the deterministic
    discrete-ordinates march with diffuse re-emission (dda.py)."""
    dx = geo['dx']
    gn = ops.gas_normals(phi, dx, mesh.centroids, mesh.normals)
    betaE, betaO = par.get('betaE', 0.7), par.get('betaO', 1.0)
    n_dir, n_re = int(par.get('dda_n_dir', 64)), int(par.get('dda_n_reemit', 12))
    py = int(par.get('periodic_y', 0))

    def neutrals(bare):
        return ops.dda_neutral_flux(phi, dx, mesh.centroids, gn, mesh.areas, bare,
                                    (betaE, betaO), n_dir, n_re, py)
    return _deterministic_neutrals(mesh, phi, geo, par, n_ion, seed, flags, n_fp, neutrals)


def radiosity_flux(mesh, phi, geo, par, n_ion, seed, flags, n_ff=64, n_fp=4):
    """mc_flux_3d_radiosity (threed.py:923-991): form factors by n_ff first-hit cosine rays per
    face, then n_fp x {G = D + Aff diag(1-s) G solved for F and O, Belen coverages}."""
    F = mesh.nf
    gn = ops.gas_normals(phi, geo['dx'], mesh.centroids, mesh.normals)
    m_i, cos_i, ion_yield = deterministic_ions(mesh, geo, par, n_ion, seed * 9 + 1, flags)
    hit = ops.trace_ff(mesh.bvh, box_of(geo, par), mesh.centroids, gn, n_ff,
                       (seed * 9 + 5) & SEED_MASK)
    row_ptr, col, D = ops.radiosity_assemble(hit, F, n_ff)
    solver = par.get('radiosity_solver', 'jacobi')
    n_iter = 40 if solver == 'jacobi' else 0
    betaE = par.get('betaE', 0.7)
    betaO = par.get('betaO', 1.0)
    chem = chem_params(par, flags)
    bare = torch.ones((F,), dtype=torch.float32, device=mesh.device)
    m_F = m_O = torch.zeros((F,), dtype=torch.float32, device=mesh.device)
    for _ in range(n_fp):
        m_F = ops.radiosity_solve(row_ptr, col, D, F, n_ff, bare, betaE, n_iter)[0]
        m_O = ops.radiosity_solve(row_ptr, col, D, F, n_ff, bare, betaO, n_iter)[0]
        bare = ops.belen_coverage(m_i, m_F, m_O, cos_i, ion_yield, chem)
    return dict(m_i=m_i, m_F=m_F, m_O=m_O, cos_i=cos_i, bare=bare, ion_yield=ion_yield)


def uncoupled_flux(mesh, geo, par, n_ion, n_neu, seed, ion_reflection=False):
    """mc_flux_3d (threed.py:472-507): legacy ion (specular reflection with 3 re-emissions when
    ion_reflection), sticking neutrals with 12 re-emissions; no smoothing."""
    box = box_of(geo, par)
    A_src = geo['Lx'] * geo['Ly']
    A = mesh.areas
    if ion_reflection:
        acc = ops.trace_legacy(mesh.bvh, box, ion_source(par), 1.0, 3, 1, 0.34, 0.8, n_ion,
                               (seed * 9 + 1) & SEED_MASK)
    else:
        acc = ops.trace_legacy(mesh.bvh, box, ion_source(par), 1.0, 0, 0, 0.34, 0.8, n_ion,
                               (seed * 9 + 1) & SEED_MASK)
    nsrc = ops.make_source(ops.SRC_NEUTRAL)
    aF = ops.trace_legacy(mesh.bvh, box, nsrc, par['s_F'], 12, 0, 0.34, 0.8, n_neu,
                          (seed * 9 + 2) & SEED_MASK)
    aO = ops.trace_legacy(mesh.bvh, box, nsrc, par['s_O'], 12, 0, 0.34, 0.8, n_neu,
                          (seed * 9 + 3) & SEED_MASK)
    m_i = ops.normalize(acc[0], A, A_src / n_ion, 0.0, 1.5)
    m_F = ops.normalize(aF[0], A, A_src / n_neu, 0.0, 4.0)
    m_O = ops.normalize(aO[0], A, A_src / n_neu, 0.0, 4.0)
    cos_i = ops.cos_ratio(acc[0], acc[1])
    return dict(m_i=m_i, m_F=m_F, m_O=m_O, cos_i=cos_i, bare=None, ion_yield=None)


def cfl_substeps(vmax, dt, dx):
    """nsub = max(1, min(ceil(Vmax*dt/(0.4dx)), 160)) with Vmax floored at 1e-6
    (threed.py:1750-1756)."""
    Vmax = max(vmax if math.isfinite(vmax) else 0.0, 1e-6)
    return max(1, min(int(math.ceil(Vmax * dt / (0.4 * dx))), 160)), Vmax
