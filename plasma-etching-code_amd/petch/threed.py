"""3-D etch loop on the MI355X (the reference's petch.threed interface).

run_etch_3d keeps phi, the mesh, the BVH, the fluxes and the extension velocity on the GPU for
the whole run; per step the host reads back two small things (the mesh size after marching
cubes and max V for the CFL substep count). The stage functions with the reference's numpy
signatures (extract_mesh_3d, mc_flux_3d_coupled, advect_3d_gpu, reinit_fsm, ...) wrap the same
kernels for scripts written against the reference.
"""
import math
import os
import time
import warnings

import numpy as np
import torch

from . import flux as _flux
from . import ops
from .params import PAR, DEFAULT_FLAGS

DEVICE = os.environ.get("PETCH_DEVICE", "hip")


def _dev():
    if not torch.cuda.is_available():
        raise RuntimeError("petch: the etch loop runs on a ROCm GPU (libpetch_hip.so); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _t(a, dtype=torch.float32):
    if isinstance(a, torch.Tensor):
        return a.to(device=_dev(), dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=_dev())


def _np(t, dtype=np.float64):
    return t.detach().cpu().numpy().astype(dtype)


# ----------------------------------------------------------------------------- geometry
def grid_shape(Lx, Ly, Lz, dx):
    """threed.py:34."""
    return int(round(Lx / dx)), int(round(Ly / dx)), int(round(Lz / dx))


def make_trench_3d(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole=False):
    """Initial level set and mask (threed.py:32-59): exact signed distance to the CSG solid
    (substrate half-space union the masked slab), phi > 0 solid. The geo dict carries numpy
    copies like the reference plus the device tensors under '_phi_d' / '_mask_d'."""
    geo = _trench_geo(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole)
    geo['phi'] = _np(geo['_phi_d'])
    return geo


def _trench_geo(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole):
    """make_trench_3d without the host copy of the initial phi (the etch loops replace it with
    the final phi, so copying the initial one would be wasted)."""
    nx, ny, nz = grid_shape(Lx, Ly, Lz, dx)
    phi_d, mask_d = ops.init_sdf(nx, ny, nz, dx, Lx, Ly, trench_width, mask_th, sub_top, hole,
                                 device=_dev())
    xs, ys, zs = np.arange(nx) * dx, np.arange(ny) * dx, np.arange(nz) * dx
    return dict(xs=xs, ys=ys, zs=zs, dx=dx, mask=_np(mask_d, bool),
                Lx=Lx, Ly=Ly, Lz=Lz, sub_top=sub_top, trench_width=trench_width, hole=hole,
                _phi_d=phi_d, _mask_d=mask_d)


# neutral transports without Monte-Carlo neutrals (and so without the coverage warm start)
_DET_NT = ("radiosity", "knudsen", "dda")


class Mesh:
    """Triangle mesh + LBVH on the device (the wp.Mesh of threed.py:1674-1675)."""

    def __init__(self, points, indices):
        self.verts = _t(np.asarray(points, np.float32).reshape(-1, 3) if not isinstance(points, torch.Tensor) else points)
        self.faces = _t(np.asarray(indices, np.int32).reshape(-1, 3) if not isinstance(indices, torch.Tensor) else indices.reshape(-1, 3), torch.int32)
        self.nv, self.nf = int(self.verts.shape[0]), int(self.faces.shape[0])
        self.device = self.verts.device
        self.centroids, self.areas, self.normals = ops.mesh_attrs(self.verts, self.faces)
        self.nbr = ops.face_adjacency(self.faces, self.nv)
        self.bvh = ops.Bvh(self.device).build(self.verts, self.faces)


def extract_mesh_3d(phi, dx):
    """Marching cubes (threed.py:62-69): verts float32 (physical), faces int32, centroids,
    areas (float64 numpy)."""
    v, f = ops.marching_cubes(_t(phi), dx)
    cen, area, _ = ops.mesh_attrs(v, f)
    return _np(v, np.float32), _np(f, np.int32), _np(cen), _np(area)


extract_mesh_3d_gpu = extract_mesh_3d   # threed.py:75-100: the same kernel


def _edge_adjacency(faces):
    """Edge-neighbour face pairs (threed.py:103-125), from the device adjacency table."""
    f = _t(np.asarray(faces, np.int32).reshape(-1, 3), torch.int32)
    nbr = _np(ops.face_adjacency(f, int(f.max().item()) + 1 if f.numel() else 0), np.int64)
    i = np.repeat(np.arange(len(nbr)), 3)
    j = nbr.ravel()
    m = (j >= 0) & (i < j)
    return np.stack([i[m], j[m]], axis=1)


def smooth_flux(flux, normals, pairs, n_iter=1, alpha=1.0, faces=None):
    """Normal-weighted edge-neighbour smoothing (threed.py:129-148). Takes the mesh faces (the
    device kernel gathers over its own adjacency table); `pairs` is accepted for signature
    compatibility and must be the edge pairs of `faces`."""
    if faces is None:
        raise ValueError("smooth_flux on the HIP path needs faces= (the adjacency comes from them)")
    if pairs is not None and len(pairs) == 0 or n_iter <= 0 or alpha <= 0.0:
        return np.asarray(flux, np.float64)
    f = _t(np.asarray(faces, np.int32).reshape(-1, 3), torch.int32)
    nbr = ops.face_adjacency(f, int(f.max().item()) + 1)
    return _np(ops.smooth(_t(flux), nbr, _t(normals), n_iter, alpha, False))


smooth_flux_gpu = smooth_flux


def faces_in_mask(centroids, geo, mask_th, trench_width, hole=False):
    """threed.py:1608-1616 (host predicate on centroids; the loop evaluates it in the rate kernel)."""
    c = np.asarray(centroids, np.float64)
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    in_band = (z >= geo['sub_top']) & (z < geo['sub_top'] + mask_th)
    if hole:
        opening = (x - geo['Lx'] / 2) ** 2 + (y - geo['Ly'] / 2) ** 2 < (trench_width / 2) ** 2
    else:
        opening = np.abs(x - geo['Lx'] / 2) < trench_width / 2
    return in_band & (~opening)


# ----------------------------------------------------------------------------- flux stages
def _mesh_of(mesh, verts, faces):
    if isinstance(mesh, Mesh):
        return mesh
    return Mesh(verts, faces)


def mc_flux_3d_coupled(mesh, verts, faces, areas, geo, par, n_ion=20000, n_neu=20000, seed=0,
                       sampling="pseudo", flags=None, n_fp=4, bare_init=None):
    """mc_flux_3d_coupled (threed.py:740-905): returns numpy (m_i, m_F, m_O, cos_i, bare) and sets
    par['_ion_yield'] (faithful ion) or removes it, like threed.py:818/825. Sources are drawn on
    the device (the reference's gpu_source path); `sampling` is accepted and unused."""
    M = _mesh_of(mesh, verts, faces)
    bi = None if bare_init is None else _t(np.asarray(bare_init, np.float64))
    phi = geo.get('_phi_d')
    if phi is None and 'phi' in geo:
        phi = _t(geo['phi'])
    out = _flux.coupled_flux(M, geo, par, n_ion, n_neu, seed, flags, bare_init=bi, phi=phi)
    if out['ion_yield'] is not None:
        par['_ion_yield'] = tuple(_np(a) for a in out['ion_yield'])
    else:
        par.pop('_ion_yield', None)
    return tuple(_np(out[k]) for k in ("m_i", "m_F", "m_O", "cos_i", "bare"))


def mc_flux_3d_radiosity(mesh, verts, faces, centroids, areas, geo, par, n_ion=20000, n_ff=64,
                         seed=0, flags=None, n_fp=4):
    """mc_flux_3d_radiosity (threed.py:923-991): returns numpy (m_i, m_F, m_O, cos_i)."""
    M = _mesh_of(mesh, verts, faces)
    phi = geo.get('_phi_d')
    phi = _t(geo['phi']) if phi is None else phi
    out = _flux.radiosity_flux(M, phi, geo, par, n_ion, seed, flags, n_ff=n_ff, n_fp=n_fp)
    if out['ion_yield'] is not None:
        par['_ion_yield'] = tuple(_np(a) for a in out['ion_yield'])
    else:
        par.pop('_ion_yield', None)
    return tuple(_np(out[k]) for k in ("m_i", "m_F", "m_O", "cos_i"))


def _det_flux(fn, mesh, verts, faces, geo, par, seed, flags, **kw):
    M = _mesh_of(mesh, verts, faces)
    phi = geo.get('_phi_d')
    phi = _t(geo['phi']) if phi is None else phi
    out = fn(M, phi, geo, par, seed=seed, flags=flags, **kw)
    if out['ion_yield'] is not None:
        par['_ion_yield'] = tuple(_np(a) for a in out['ion_yield'])
    else:
        par.pop('_ion_yield', None)
    return tuple(_np(out[k]) for k in ("m_i", "m_F", "m_O", "cos_i"))


def mc_flux_3d_knudsen(mesh, verts, faces, centroids, areas, geo, par, n_ion=20000, seed=0,
                       flags=None, n_fp=4):
    """mc_flux_3d_knudsen (threed.py:1149-1208): returns numpy (m_i, m_F, m_O, cos_i)."""
    return _det_flux(_flux.knudsen_flux, mesh, verts, faces, geo, par, seed, flags, n_ion=n_ion,
                     n_fp=n_fp)


def mc_flux_3d_dda(mesh, verts, faces, centroids, areas, geo, par, n_ion=20000, seed=0,
                   flags=None, n_fp=4):
    """mc_flux_3d_dda (threed.py:1211-1244): returns numpy (m_i, m_F, m_O, cos_i)."""
    return _det_flux(_flux.dda_flux, mesh, verts, faces, geo, par, seed, flags, n_ion=n_ion,
                     n_fp=n_fp)


def mc_flux_3d(mesh, verts, faces, areas, geo, par, n_ion=20000, n_neu=20000, seed=0,
               sampling="pseudo", ion_reflection=False):
    """mc_flux_3d (threed.py:472-507): numpy (m_i, m_F, m_O, cos_i)."""
    M = _mesh_of(mesh, verts, faces)
    out = _flux.uncoupled_flux(M, geo, par, n_ion, n_neu, seed, ion_reflection)
    return tuple(_np(out[k]) for k in ("m_i", "m_F", "m_O", "cos_i"))


def _gas_normals(verts, faces, centroids, geo):
    """Face normals oriented into the gas (threed.py:1553-1568): numpy (F, 3)."""
    v = _t(np.asarray(verts, np.float32).reshape(-1, 3))
    f = _t(np.asarray(faces, np.int32).reshape(-1, 3), torch.int32)
    _, _, nrm = ops.mesh_attrs(v, f)
    phi = geo.get('_phi_d')
    phi = _t(geo['phi']) if phi is None else phi
    return _np(ops.gas_normals(phi, geo['dx'], _t(np.asarray(centroids).reshape(-1, 3)), nrm))


def mc_redep_3d(mesh, centroids, areas, normals, src_flux, n_redep=20000, s_redep=0.5, seed=0):
    """mc_redep_3d (threed.py:1582-1605): the redeposited flux density per face (numpy float64)
    of n_redep product rays emitted in proportion to max(src_flux * areas, 0) along the
    into-gas normals and re-sticking with probability s_redep. mesh: a Mesh of the surface."""
    if not isinstance(mesh, Mesh):
        raise TypeError("mc_redep_3d: mesh must be a petch.threed.Mesh of the surface")
    F = mesh.nf
    c, a = _t(np.asarray(centroids).reshape(-1, 3)), _t(np.asarray(areas).reshape(-1))
    n, v = _t(np.asarray(normals).reshape(-1, 3)), _t(np.asarray(src_flux).reshape(-1))
    if F == 0:
        return np.zeros(0)
    out = ops.trace_redeposition(mesh.bvh, c, n, a, v, float(s_redep), int(n_redep),
                                 (int(seed) * 7 + 5) & _flux.SEED_MASK)
    return _np(out)


def gpu_warmstart_bare(prev_mesh, prev_bare, centroids):
    """Coverage seed from the nearest face of the previous mesh (threed.py:1633-1639)."""
    return _np(ops.nearest_face_value(prev_mesh.bvh, _t(prev_bare), _t(np.asarray(centroids, np.float32)), 1.0))


def extend_velocity_gpu(mesh, V, geo, band):
    """Fs = V[nearest face] within |phi| < band (threed.py:1333-1346), numpy (nx,ny,nz)."""
    phi = geo.get('_phi_d')
    phi = _t(geo['phi']) if phi is None else phi
    return _np(ops.extend(mesh.bvh, _t(V), phi, geo['dx'], band))


def advect_3d_gpu(phi, Fspeed, mask, mask_phi, dx, dt, nsub):
    """nsub upwind substeps with mask re-pin (threed.py:1295-1307); numpy float64 out."""
    p = _t(phi)
    ops.advect(p, _t(Fspeed), _t(np.asarray(mask, np.uint8), torch.uint8), _t(mask_phi), dx, dt, nsub)
    return _np(p)


def advect_3d(phi, Fspeed, dx, dt):
    """One upwind substep without mask (threed.py:1248-1259), on the device."""
    p = _t(phi)
    z = torch.zeros(p.shape, dtype=torch.uint8, device=p.device)
    ops.advect(p, _t(Fspeed), z, p.clone(), dx, dt, 1)
    return _np(p)


def reinit_fsm(phi_np, dx, band, n_iter=None):
    """Jacobi Godunov-Eikonal reinit (threed.py:1515-1537); numpy float64 out."""
    p = _t(phi_np)
    ops.reinit_fsm(p, dx, band, n_iter)
    return _np(p)


# ----------------------------------------------------------------------------- depth metrics
def _column_floors(phi, zs, cols):
    floors = []
    for i, j in cols:
        col = phi[i, j, :] < 0
        k = len(col) - 1
        if not col[k]:
            continue
        while k > 0 and col[k - 1]:
            k -= 1
        floors.append(zs[k])
    return floors


def _depth3d(geo, half=1.0):
    """Center etch depth: median floor of the top-connected gas run over central columns
    (threed.py:1798-1834)."""
    phi, xs, ys, zs = geo['phi'], geo['xs'], geo['ys'], geo['zs']
    cx, cy = geo['Lx'] / 2, geo['Ly'] / 2
    r_feat = geo.get('trench_width', 2.0) / 2.0 - geo['dx']
    hx = max(min(half, r_feat), geo['dx'])
    ic = np.where(np.abs(xs - cx) <= hx)[0]
    if geo.get('hole', False):
        jc = np.where(np.abs(ys - cy) <= hx)[0]
        cols = [(i, j) for i in ic for j in jc if (xs[i] - cx) ** 2 + (ys[j] - cy) ** 2 <= r_feat ** 2]
    else:
        jc = np.where(np.abs(ys - cy) <= half)[0]
        cols = [(i, j) for i in ic for j in jc]
    floors = _column_floors(phi, zs, cols)
    if not floors:
        return 0.0
    return float(geo['sub_top'] - np.median(floors))


def center_depth_3d(geo, half=1.0):
    return _depth3d(geo, half)


def max_depth_3d(geo):
    """5th-percentile floor over the feature footprint (threed.py:1841-1869)."""
    xs, ys = geo['xs'], geo['ys']
    cx, cy = geo['Lx'] / 2.0, geo['Ly'] / 2.0
    r = geo.get('trench_width', 4.0) / 2.0
    if geo.get('hole', False):
        ii, jj = np.where((xs[:, None] - cx) ** 2 + (ys[None, :] - cy) ** 2 <= (r + geo['dx']) ** 2)
        cols = list(zip(ii, jj))
    else:
        ix = np.where(np.abs(xs - cx) <= r + geo['dx'])[0]
        cols = [(i, j) for i in ix for j in range(len(ys))]
    floors = _column_floors(geo['phi'], geo['zs'], cols)
    if not floors:
        return 0.0
    return float(geo['sub_top'] - np.percentile(floors, 5))


# ----------------------------------------------------------------------------- driver
_REINIT_WARNED = set()


def _reinit_args(phi_d, dx, band, method):
    """(band, n_iter) of the GPU Eikonal reinit that stands in for `method` (n_iter None: the
    default ceil(band/dx)+4 sweeps); None for 'gpu', which is the PDE reinit."""
    if method == "gpu":
        return None
    if method == "fsm":
        return band, None
    if method in ("skfmm", "skfmm_full"):
        # scikit-fmm (CPU fast marching) is not part of the GPU path: the narrow-band FSM computes
        # the same distance within the band; 'skfmm_full' sweeps until the band spans the grid.
        if method not in _REINIT_WARNED:
            warnings.warn(f"reinit_method={method!r}: using the GPU Eikonal (fsm) reinit",
                          RuntimeWarning)
            _REINIT_WARNED.add(method)
        if method == "skfmm_full":
            return dx * (max(phi_d.shape) + 4), int(max(phi_d.shape)) + 4
        return band, None
    raise NotImplementedError(f"reinit_method={method!r} is not part of the HIP etch loop "
                              "(supported: 'fsm', 'gpu', 'skfmm', 'skfmm_full')")


def _reinit(phi_d, dx, band, method):
    args = _reinit_args(phi_d, dx, band, method)
    if args is None:                      # reinit_method='gpu' (threed.py:1772-1773)
        ops.reinit_pde(phi_d, dx)
        return
    ops.reinit_fsm(phi_d, dx, *args)


class _FluxSteps:
    """A coupled-flux generator (flux.coupled_flux_steps) waiting for its traces."""

    def __init__(self, steps):
        self.steps = steps


class EtchLoop:
    """Device state and one etch step of run_etch_3d (threed.py:1671-1792).

    phi (float32), mask (uint8) and mask_phi live on the GPU; step() runs mesh -> BVH -> flux ->
    rate -> extension -> CFL -> advection -> reinit with two host reads (mesh size, max V)."""

    def __init__(self, phi_d, mask_d, geo, mask_th, t_end, n_steps, par, flags, n_ion, n_neu,
                 reinit_every=1, extend="gpu", reinit_method="fsm", seed_offset=0,
                 rays_per_point=None, etch_stop_z=None, timer=None, autotune=None,
                 surf_smooth=0.0):
        from .profiling import NULL
        from .chemistry import chem_params
        if par is None:
            par = PAR
        if flags is None:
            flags = DEFAULT_FLAGS
        self.par, self.flags = par, flags
        self.mc_par = par
        if getattr(flags, "chemistry", "langmuir") == "belen":          # threed.py:1654-1656
            self.mc_par = dict(par)
            self.mc_par['s_F'] = par['betaE']
            self.mc_par['s_O'] = par['betaO']
        self.redep = bool(getattr(flags, "redeposition", False))
        self.nt = getattr(flags, "neutral_transport", "mc")
        self.dev = phi_d.device
        self.phi = phi_d
        self.mask = mask_d
        self.mask_phi = phi_d.clone()
        self.work = torch.empty_like(phi_d)
        self.Fs = torch.empty_like(phi_d)
        self.geo = geo
        self.dx = geo['dx']
        self.n_steps = n_steps
        self.dt = t_end / n_steps
        self.band = 4 * self.dx
        self.n_ion, self.n_neu = n_ion, n_neu
        self.reinit_every, self.extend_mode, self.reinit_method = reinit_every, extend, reinit_method
        self.surf_smooth = float(surf_smooth or 0.0)
        self.seed_offset, self.rays_per_point, self.etch_stop_z = seed_offset, rays_per_point, etch_stop_z
        self.timer = NULL if timer is None else timer
        self.warm = getattr(flags, "warm_start_coverage", False)
        self.gpu_ws = bool(par.get('gpu_warmstart', True))
        self.mg = ops.MaskGeom(float(geo['sub_top']), float(mask_th), float(geo['Lx']),
                               float(geo['Ly']), float(geo['trench_width']), int(bool(geo['hole'])))
        self.chem = chem_params(par, flags)
        self.consts = _flux.run_consts(geo, par, flags)   # the coupled flux's per-run constants
        # the ray budget of a step picks the BVH build (clustered trees pay off at ~1000 rays
        # per face): ions + F and O neutrals per fixed-point iteration
        rays = None
        if rays_per_point is None:
            rays = int(n_ion) + 2 * int(n_neu) * max(1, int(par.get('n_fp', 1)))
        self.meshes = [_flux.SurfaceMesh(self.dev, rays), _flux.SurfaceMesh(self.dev, rays)]
        # autotune: a petch.autotune.TraceTuner (or True for a new one) that picks the BVH build
        # and the trace occupancy on the first surface with a Monte-Carlo neutral trace
        if autotune is True:
            from .autotune import TraceTuner
            autotune = TraceTuner()
        self.tuner = autotune or None
        self.tuned = None
        self.prev_mesh = self.prev_bare = self.prev_cen = None
        self.pt_bvh = None if self.gpu_ws else ops.Bvh(self.dev)
        self.red = torch.zeros((2,), dtype=torch.int32, device=self.dev)
        self.timings = dict(flux=0.0, extend=0.0, reinit=0.0, mesh=0.0, advect=0.0, total=0.0,
                            nsub_max=0)
        self.last = {}

    def step(self, step):
        """One etch step; the coupled flux's traces run one launch at a time."""
        job = self.begin_step(step)
        if isinstance(job, _FluxSteps):
            with self.timer.region("flux"):
                tf = time.time()
                out = _flux.run_traces(job.steps, self.timer)
                self.timings['flux'] += time.time() - tf
        else:
            out = job
        return self.end_step(step, out)

    def begin_step(self, step):
        """Mesh, BVH and warm start of one step, then the flux: returns its result (dict or
        None), or, for the coupled flux, a _FluxSteps whose trace requests the caller runs
        (alone, or batched with other features) before passing the result to end_step."""
        finite = self.begin_mesh(step)
        return self.begin_flux(step, None if finite is None else bool(finite))

    def mesh_item(self, step):
        """(phi, dx, buffers) of this step's marching cubes for ops.marching_cubes_batch."""
        return (self.phi, self.dx, self.meshes[step % 2].bufs)

    def stage_mesh(self, step, mc, surf):
        """The batched loop's first half of begin_mesh: this step's surface from its
        marching-cubes (verts, faces), its face attributes and adjacency left to one
        ops.surface_build_batch over `surf` (the list its item is appended to)."""
        tm = time.time()
        with self.timer.region("mesh"):
            self.meshes[step % 2].build_from(*mc, bvh=False, surf=surf)
        self.timings['mesh'] += time.time() - tm

    def begin_mesh(self, step, mc=None, bvh=True, staged=False):
        """Mesh + BVH of the step. Returns the device flag 'all centroids finite' that gates the
        warm start (threed.py:1699), or None when the step needs no warm start. mc: this step's
        marching-cubes (verts, faces) when the caller already extracted them (the batched loop
        meshes all features with one host read); bvh=False leaves the BVH to the caller (the
        batched loop builds all features' trees together); staged: the surface was built by
        stage_mesh and ops.surface_build_batch."""
        T, dx = self.timer, self.dx
        tm = time.time()
        with T.region("mesh"):
            if staged:
                mesh = self.meshes[step % 2]
            elif mc is None:
                mesh = self.meshes[step % 2].build(self.phi, dx)
            else:
                mesh = self.meshes[step % 2].build_from(*mc, bvh=bvh)
        self.timings['mesh'] += time.time() - tm
        self._mesh = mesh
        if self.rays_per_point is not None:
            n = max(int(self.rays_per_point) * mesh.nf, 1000)
            self._rays = (n, n)
        else:
            self._rays = (self.n_ion, self.n_neu)
        if self.tuner is not None and self.tuned is None and bvh and mesh.nf > 1 and \
                self.nt not in _DET_NT and getattr(self.flags, "coverage_sticking", False):
            from .flux import box_of
            bare = torch.ones((mesh.nf,), dtype=torch.float32, device=self.dev)
            self.tuned = self.tuner.tune(mesh, box_of(self.geo, self.par), bare,
                                         self.par.get('betaE', 0.7), self.par.get('betaO', 1.0),
                                         self._rays[1])
            for m in self.meshes:
                m.quality = self.tuned[0]
        if mesh.nf and self.nt not in _DET_NT and getattr(self.flags, "coverage_sticking", False) \
                and self.warm and self.prev_bare is not None:
            # staged: the flag the batched surface build computed (1 iff all centroids finite)
            return mesh.finite if staged else torch.isfinite(mesh.centroids).all()
        return None

    def warm_start_item(self, finite):
        """(bvh, src, points) of this step's warm-start lookup (nearest previous face of every
        new centroid, threed.py:1700-1701) for ops.nearest_face_value_batch, or None when the
        step has no such lookup."""
        if not finite or not self.gpu_ws or self._mesh.nf == 0 or self.nt in _DET_NT or \
                not getattr(self.flags, "coverage_sticking", False):
            return None
        return (self.prev_mesh.bvh, self.prev_bare, self._mesh.centroids)

    def begin_flux(self, step, finite, bare_init=None, rounds=None):
        """The flux of the step begun by begin_mesh; `finite` is that flag read on the host.
        bare_init: the warm-start values when the caller already looked them up (the batched
        loop does all features' lookups in one launch). rounds: passed to
        flux.coupled_flux_steps (the batched loop ends all features' neutral rounds in one
        launch)."""
        T, par, flags, dev = self.timer, self.par, self.flags, self.dev
        mesh = self._mesh
        n_ion_s, n_neu_s = self._rays
        seed = step + self.seed_offset
        if mesh.nf == 0:
            return None
        tf = time.time()
        try:
            with T.region("flux"):
                if self.nt in _DET_NT:              # threed.py:1688-1696
                    fn = dict(radiosity=_flux.radiosity_flux, knudsen=_flux.knudsen_flux,
                              dda=_flux.dda_flux)[self.nt]
                    return fn(mesh, self.phi, self.geo, par, n_ion_s, seed, flags)
                if getattr(flags, "coverage_sticking", False):
                    bi = None
                    if finite and bare_init is not None:
                        bi = bare_init
                    elif finite:
                        if self.gpu_ws:             # nearest previous face (threed.py:1700-1701)
                            bi = ops.nearest_face_value(self.prev_mesh.bvh, self.prev_bare,
                                                        mesh.centroids, 1.0)
                        else:                       # nearest previous centroid (threed.py:1702-1705)
                            idx = torch.arange(self.prev_cen.shape[0], dtype=torch.int32, device=dev)
                            self.pt_bvh.build(self.prev_cen, idx[:, None].repeat(1, 3).contiguous())
                            bi = ops.nearest_face_value(self.pt_bvh, self.prev_bare,
                                                        mesh.centroids, 1.0)
                    return _FluxSteps(_flux.coupled_flux_steps(mesh, self.geo, par, n_ion_s,
                                                               n_neu_s, seed, flags,
                                                               bare_init=bi, rounds=rounds,
                                                               phi=self.phi,
                                                               bare_in_unit=True,
                                                               consts=self.consts))
                return _flux.uncoupled_flux(mesh, self.geo, self.mc_par, n_ion_s, n_neu_s, seed,
                                            getattr(flags, "ion_reflection", False))
        finally:
            self.timings['flux'] += time.time() - tf

    def end_step(self, step, out):
        """Rate, extension, CFL, advection and reinit of the step begun by begin_step."""
        if self.end_rate(step, out):
            if self.redep:
                self.redeposit(step, self.red)
            vmx, nbad = ops.read_rate_reduction(self.red)
            return self.end_finish(step, vmx, nbad, self._mesh.bvh.status())
        return self.end_finish(step, 0.0, 0, 0)

    def end_rate(self, step, out, rates=None):
        """Surface rate of the step (no host read). Returns False when there is no surface.
        rates: a list that receives this step's rate as an item of ops.surface_rate_batch
        instead of running it (the batched loop rates every feature in one launch)."""
        T, flags = self.timer, self.flags
        mesh = self._mesh
        if out is not None and getattr(flags, "coverage_sticking", False) and \
                self.nt not in _DET_NT:
            self.prev_mesh, self.prev_bare, self.prev_cen = mesh, out['bare'], mesh.centroids
        self._V = None
        if out is None:
            return False
        trt = time.time()
        with T.region("rate"):
            iy = out['ion_yield'] if getattr(flags, "chemistry", "") == "belen" else None
            if rates is not None:
                self._V = torch.empty_like(out['m_F'])
                rates.append(dict(m_i=out['m_i'], m_F=out['m_F'], m_O=out['m_O'],
                                  cos_i=out['cos_i'], ion_yield=iy, centroids=mesh.centroids,
                                  mask_geom=self.mg, chem=self.chem, V=self._V, bvh=mesh.bvh))
            else:
                self._V, _ = ops.surface_rate(out['m_i'], out['m_F'], out['m_O'], out['cos_i'],
                                              iy, mesh.centroids, self.mg, self.chem,
                                              red=self.red)
        self.timings['rate'] = self.timings.get('rate', 0.0) + time.time() - trt
        return True

    def redeposit(self, step, red):
        """Etch-product redeposition (threed.py:1730-1743) on this step's rate V: n_neu product
        rays emitted in proportion to V * area (stream step * 7 + 5) give the redeposited flux
        Rf (zero on the sidewall-foot band when surface charging is on and an etch stop is set),
        then V = max(V - k_redep Rf, 0) in place and red[0] = the bits of the new max V, so the
        CFL sees the reduced rate. No host read."""
        mesh, V = self._mesh, self._V
        if V is None or mesh.nf == 0:
            return
        with self.timer.region("redep"):
            gn = ops.gas_normals(self.phi, self.dx, mesh.centroids, mesh.normals)
            Rf = ops.trace_redeposition(mesh.bvh, mesh.centroids, gn, mesh.areas, V,
                                        float(self.par.get('s_redep', 0.5)), int(self.n_neu),
                                        (step * 7 + 5) & _flux.SEED_MASK)
            if getattr(self.flags, "surface_charging", False) and self.etch_stop_z is not None:
                foot_h = max(0.3 * self.geo['trench_width'], 2.0 * self.dx)
                foot = mesh.centroids[:, 2].double() <= float(self.etch_stop_z) + foot_h
                Rf = torch.where(foot, torch.zeros_like(Rf), Rf)
            k = float(self.par.get('k_redep', 1.0))
            V.copy_((V.double() - k * Rf.double()).clamp_(min=0.0))
            red[0] = V.max().reshape(1).view(torch.int32)[0]

    def extend_item(self):
        """(bvh, V, phi, dx, band, Fs) of this step's velocity extension for ops.extend_batch,
        or None when the step extends otherwise (no surface, or the nearest-centroid mode)."""
        if self._V is None or self.extend_mode != "gpu":
            return None
        return (self._mesh.bvh, self._V, self.phi, self.dx, self.band, self.Fs)

    def end_finish(self, step, vmx, nbad, bvh_err, extended=False, deferred=None):
        """Extension, CFL, advection and reinit, given the host reads of the rate reduction
        (max V, non-finite count) and of the BVH guard flag. extended: Fs already holds this
        step's extension (the batched loop extends all features in one pass). deferred: a dict
        of 'advect' and 'reinit' lists that receive this step's advection and reinit as items of
        ops.advect_batch / ops.reinit_fsm_batch instead of running them (the batched loop runs
        every feature's in one launch per substep and sweep); work that must follow the reinit
        (the PDE reinit of reinit_method='gpu', surf_smooth) goes to its 'post' list of
        callables."""
        T, dx, dev = self.timer, self.dx, self.dev
        tim = self.timings
        mesh = self._mesh
        V = self._V
        if bvh_err:
            raise RuntimeError(f"step {step}: BVH traversal guard tripped (code {bvh_err}: "
                               "1 stack overflow, 2 corrupt tree, 3 tree too deep); the flux is "
                               "incomplete")
        if nbad:
            warnings.warn(f"step {step}: {nbad} non-finite surface velocities (NaN/inf) zeroed; "
                          "check flux normalization / coverage if this persists.", RuntimeWarning)
        if V is None:
            vmx = 0.0
        te = time.time()
        with T.region("extend"):
            if extended:
                pass
            elif V is None:
                self.Fs.zero_()
            elif self.extend_mode == "gpu":
                ops.extend(mesh.bvh, V, self.phi, dx, self.band, out=self.Fs)
            else:                                   # nearest centroid (threed.py:1310-1321)
                idx = torch.arange(mesh.nf, dtype=torch.int32, device=dev)
                cb = ops.Bvh(dev).build(mesh.centroids, idx[:, None].repeat(1, 3).contiguous())
                ops.extend(cb, V, self.phi, dx, self.band, out=self.Fs)
            if self.etch_stop_z is not None:        # threed.py:1034-1044
                below = torch.as_tensor(self.geo['zs'] < float(self.etch_stop_z), device=dev)
                self.Fs[:, :, below] = 0.0
        tim['extend'] += time.time() - te
        nsub, Vmax = _flux.cfl_substeps(vmx, self.dt, dx)
        tim['nsub_max'] = max(tim['nsub_max'], nsub)
        reinit = (step + 1) % self.reinit_every == 0 or step == self.n_steps - 1
        if deferred is not None:
            deferred['advect'].append((self.phi, self.work, self.Fs, self.mask, self.mask_phi,
                                       dx, self.dt / nsub, nsub))
            if reinit:
                args = _reinit_args(self.phi, dx, self.band + 2.0 * dx, self.reinit_method)
                if args is None:              # the PDE reinit is not batched: run it after
                    deferred.setdefault('post', []).insert(0, self._reinit_now)
                else:
                    deferred['reinit'].append((self.phi, dx) + args)
            if self.surf_smooth > 0.0:
                deferred.setdefault('post', []).append(self.smooth)
        else:
            ta = time.time()
            with T.region("advect"):
                ops.advect(self.phi, self.Fs, self.mask, self.mask_phi, dx, self.dt / nsub, nsub,
                           work=self.work)
            tim['advect'] += time.time() - ta
            if reinit:
                tr = time.time()
                with T.region("reinit"):
                    _reinit(self.phi, dx, self.band + 2.0 * dx, self.reinit_method)
                tim['reinit'] += time.time() - tr
            if self.surf_smooth > 0.0:
                self.smooth()
        self.last = dict(nf=mesh.nf, nv=mesh.nv, Vmax=Vmax, nsub=nsub, n_ion=self._rays[0],
                         n_neu=self._rays[1])
        return self.last

    def _reinit_now(self):
        _reinit(self.phi, self.dx, self.band + 2.0 * self.dx, self.reinit_method)

    def smooth(self):
        """surf_smooth (threed.py:1779-1782): phi = gaussian_filter(phi, surf_smooth), then the
        mask re-pinned, every step after the reinit."""
        with self.timer.region("smooth"):
            ops.gaussian_smooth(self.phi, self.surf_smooth, self.mask, self.mask_phi)


def run_etch_3d(Lx=10.0, Ly=4.0, Lz=14.0, dx=0.4, trench_width=4.0, mask_th=2.0,
                sub_top=10.0, t_end=2.0, n_steps=20, hole=False, par=None, flags=None,
                n_ion=20000, n_neu=20000, reinit_every=1, extend="gpu",
                reinit_method="skfmm", verbose=True, record_depth_every=0, seed_offset=0,
                rays_per_point=None, record_frames=False, surf_smooth=0.0,
                etch_stop_z=None):
    """The per-step etch loop (threed.py:1643-1795) with every stage on the GPU."""
    _dev()
    geo = _trench_geo(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole)
    phi_d, mask_d = geo.pop('_phi_d'), geo.pop('_mask_d')
    loop = EtchLoop(phi_d, mask_d, dict(geo), mask_th, t_end, n_steps, par, flags, n_ion, n_neu,
                    reinit_every=reinit_every, extend=extend, reinit_method=reinit_method,
                    seed_offset=seed_offset, rays_per_point=rays_per_point, etch_stop_z=etch_stop_z,
                    surf_smooth=surf_smooth)
    t0 = time.time()
    for step in range(n_steps):
        info = loop.step(step)
        if record_depth_every and (step % record_depth_every == 0 or step == n_steps - 1):
            geo['phi'] = _np(loop.phi)
            geo.setdefault('depth_history', []).append((step + 1, _depth3d(geo)))
            if record_frames:
                jc = geo['phi'].shape[1] // 2
                geo.setdefault('frames', []).append(
                    dict(step=step + 1, t=(step + 1) / n_steps * t_end, depth=_depth3d(geo),
                         phi_xz=geo['phi'][:, jc, :].copy()))
        if verbose and step % 5 == 0:
            geo['phi'] = _np(loop.phi)
            print(f"  step {step:3d}/{n_steps}  faces {info['nf']:5d}  depth ~ {_depth3d(geo):5.2f}  "
                  f"Vmax {info['Vmax']:.3f}")
    torch.cuda.synchronize()
    loop.timings['total'] = time.time() - t0
    geo['phi'] = _np(loop.phi)
    geo['timings'] = loop.timings
    return geo
