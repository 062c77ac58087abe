"""Tensor-level wrappers over the C ABI: each takes/returns torch tensors on the ROCm device and
enqueues on torch's current stream. The reference-shaped numpy API in threed.py is built on
these; the etch driver calls them directly so state never leaves the GPU between stages.
"""
import torch

from ._abi import Context, ptr

f32 = torch.float32
u8 = torch.uint8
i32 = torch.int32


def _ctx(t):
    return Context.get(t.device)


def _shape3(phi):
    if phi.dim() != 3:
        raise ValueError(f"level set must be 3-D (nx,ny,nz), got shape {tuple(phi.shape)}")
    return [int(s) for s in phi.shape]


def set_deterministic(on=True, device=None):
    """Deterministic mode of the ray traces on `device` (petch_set_deterministic): per-face sums
    in 64-bit fixed point, so the same inputs give bit-identical fluxes and a whole etch run
    repeats exactly. Off by default (float atomics, as the reference)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    Context.get(dev).call("petch_set_deterministic", int(bool(on)))


def set_trace_occupancy(blocks_per_cu, device=None):
    """Workgroups per CU of the persistent trace grids on `device` (petch_set_trace_occupancy);
    0 restores the kernels' occupancy limit. The fluxes do not depend on it."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    Context.get(dev).call("petch_set_trace_occupancy", int(blocks_per_cu))


def trace_occupancy_limit(device=None):
    """The neutral trace kernel's workgroups per CU at full occupancy (petch_trace_occupancy_limit)."""
    import ctypes
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    out = ctypes.c_int(0)
    Context.get(dev).call("petch_trace_occupancy_limit", ctypes.byref(out))
    return int(out.value)


def init_sdf(nx, ny, nz, dx, Lx, Ly, trench_width, mask_th, sub_top, hole, device=None):
    """Initial phi (float32) and mask (uint8) of make_trench_3d (threed.py:32-59)."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    phi = torch.empty((nx, ny, nz), dtype=f32, device=dev)
    mask = torch.empty((nx, ny, nz), dtype=u8, device=dev)
    _ctx(phi).call("petch_init_sdf", ptr(phi), ptr(mask), nx, ny, nz, float(dx), float(Lx),
                   float(Ly), float(trench_width), float(mask_th), float(sub_top), int(bool(hole)))
    return phi, mask


def advect(phi, F, mask, mask_phi, dx, dt, nsub, work=None):
    """nsub upwind substeps of size dt with mask re-pin, in place on phi (threed.py:1295-1307)."""
    nx, ny, nz = _shape3(phi)
    for t, n, d in ((F, "F", f32), (mask_phi, "mask_phi", f32), (mask, "mask", u8)):
        if tuple(t.shape) != (nx, ny, nz):
            raise ValueError(f"advect: {n} shape {tuple(t.shape)} != phi shape {(nx, ny, nz)}")
    if work is None:
        work = torch.empty_like(phi)
    _ctx(phi).call("petch_advect", ptr(phi, f32, "phi"), ptr(work, f32, "work"), ptr(F, f32, "F"),
                   ptr(mask, u8, "mask"), ptr(mask_phi, f32, "mask_phi"), nx, ny, nz, float(dx),
                   float(dt), int(nsub))
    return phi


def advect_step(phi, F, mask, mask_phi, dx, dt, out=None):
    """One substep of advect out of place (petch_advect_step): returns out."""
    nx, ny, nz = _shape3(phi)
    for t, n in ((F, "F"), (mask_phi, "mask_phi"), (mask, "mask")):
        if tuple(t.shape) != (nx, ny, nz):
            raise ValueError(f"advect_step: {n} shape {tuple(t.shape)} != phi shape {(nx, ny, nz)}")
    if out is None:
        out = torch.empty_like(phi)
    _ctx(phi).call("petch_advect_step", ptr(phi, f32, "phi"), ptr(F, f32, "F"),
                   ptr(mask, u8, "mask"), ptr(mask_phi, f32, "mask_phi"), nx, ny, nz, float(dx),
                   float(dt), ptr(out, f32, "out"))
    return out


def advect_adjoint(phi, F, mask, gout, dx, dt, gF):
    """Adjoint of one advection substep at the state phi: returns dL/dphi for
    gout = dL/d(output) and adds dL/dF to gF (petch_advect_adjoint)."""
    nx, ny, nz = _shape3(phi)
    gphi = torch.empty_like(phi)
    _ctx(phi).call("petch_advect_adjoint", ptr(phi, f32, "phi"), ptr(F, f32, "F"),
                   ptr(mask, u8, "mask"), ptr(gout, f32, "gout"), nx, ny, nz, float(dx),
                   float(dt), ptr(gphi, f32), ptr(gF, f32, "gF"))
    return gphi


def reinit_fsm(phi, dx, band, n_iter=None):
    """Jacobi Godunov-Eikonal reinit in place (threed.py:1515-1537)."""
    nx, ny, nz = _shape3(phi)
    _ctx(phi).call("petch_reinit_fsm", ptr(phi, f32, "phi"), nx, ny, nz, float(dx), float(band),
                   -1 if n_iter is None else int(n_iter))
    return phi


def reinit_fsm_adjoint(phi0, dx, band, gout, n_iter=None):
    """d loss / d phi0 of reinit_fsm(phi0, dx, band) for gout = d loss / d(output), every
    discrete choice held at the forward's (petch_reinit_fsm_adjoint); phi0 is not modified."""
    nx, ny, nz = _shape3(phi0)
    if tuple(gout.shape) != (nx, ny, nz):
        raise ValueError(f"reinit_fsm_adjoint: gout shape {tuple(gout.shape)} != {(nx, ny, nz)}")
    gphi = torch.empty_like(phi0)
    _ctx(phi0).call("petch_reinit_fsm_adjoint", ptr(phi0, f32, "phi0"), nx, ny, nz, float(dx),
                    float(band), -1 if n_iter is None else int(n_iter), ptr(gout, f32, "gout"),
                    ptr(gphi, f32))
    return gphi


def finish_round_batch(items):
    """The end of one neutral fixed-point round of many features in one launch
    (petch_finish_round_batch). items: dicts with flF, flO (raw accumulators), areas, nbr,
    normals, scale, hi, n_smooth (0 or 1), alpha, m_i, cos_i, ion_yield (sp, ie, ip) or None,
    chem, and the outputs m_F, m_O, bare (float32 (F,) tensors the call fills). Equals, per item,
    m = smooth(normalize(fl, areas, scale, 0, hi), nbr, normals, n_smooth, alpha) for both
    species and bare = belen_coverage(m_i, m_F, m_O, cos_i, ion_yield, chem)."""
    from ._abi import RoundItem
    if not items:
        return
    arr = (RoundItem * len(items))()
    for k, it in enumerate(items):
        F = int(it['flF'].shape[0])
        if int(it['n_smooth']) not in (0, 1):
            raise ValueError(f"finish_round_batch: item {k} n_smooth {it['n_smooth']} is not 0 or 1")
        for name in ("flO", "areas", "m_F", "m_O", "bare"):
            if int(it[name].shape[0]) < F:
                raise ValueError(f"finish_round_batch: item {k} {name} holds fewer than {F} values")
        iy = it.get('ion_yield')
        arr[k] = RoundItem(ptr(it['flF'], f32, "flF"), ptr(it['flO'], f32, "flO"),
                           ptr(it['areas'], f32, "areas"), ptr(it['normals'], f32, "normals"),
                           ptr(it['nbr'], i32, "nbr"), ptr(it['m_i']), ptr(it['cos_i']),
                           ptr(iy[1] if iy is not None else None),
                           ptr(iy[2] if iy is not None else None), ptr(it['m_F'], f32, "m_F"),
                           ptr(it['m_O'], f32, "m_O"), ptr(it['bare'], f32, "bare"), it['chem'],
                           F, int(it['n_smooth']), float(it['scale']), float(it['hi']),
                           float(it['alpha']))
    _ctx(items[0]['flF']).call("petch_finish_round_batch", len(items), arr)


def finish_ion_batch(items):
    """The end of the faithful-ion trace of many features in one launch (petch_finish_ion_batch).
    items: dicts with acc (the trace's (5, F) accumulator), areas, nbr, normals, scale,
    n_smooth (0 or 1), alpha and out, a (5, F) float32 tensor the call fills with (sp, ie, ip,
    m_i, cos_i): normalize(acc[k], areas, scale, -INF, INF) smoothed (one pass, the host
    semantics) for k < 4, and cos_ratio(acc[3], acc[4])."""
    from ._abi import IonFinishItem
    if not items:
        return
    arr = (IonFinishItem * len(items))()
    for k, it in enumerate(items):
        acc, out = it['acc'], it['out']
        F = int(acc.shape[1])
        if tuple(acc.shape) != (5, F) or tuple(out.shape) != (5, F):
            raise ValueError(f"finish_ion_batch: item {k} acc/out must be (5, F)")
        if int(it['n_smooth']) not in (0, 1):
            raise ValueError(f"finish_ion_batch: item {k} n_smooth {it['n_smooth']} is not 0 or 1")
        arr[k] = IonFinishItem(ptr(acc, f32, "acc"), ptr(it['areas'], f32, "areas"),
                               ptr(it['normals'], f32, "normals"), ptr(it['nbr'], i32, "nbr"),
                               ptr(out, f32, "out"), F, int(it['n_smooth']), float(it['scale']),
                               float(it['alpha']))
    _ctx(items[0]['acc']).call("petch_finish_ion_batch", len(items), arr)


def advect_batch(items):
    """advect for many grids in one launch per substep (petch_advect_batch). items: (phi, work,
    F, mask, mask_phi, dx, dt, nsub) tuples, dt the substep's step; each phi is updated in
    place as advect would."""
    import ctypes
    from ._abi import AdvectItem
    if not items:
        return
    arr = (AdvectItem * len(items))()
    for k, (phi, work, F, mask, mask_phi, dx, dt, nsub) in enumerate(items):
        nx, ny, nz = _shape3(phi)
        for t, n in ((work, "work"), (F, "F"), (mask, "mask"), (mask_phi, "mask_phi")):
            if tuple(t.shape) != (nx, ny, nz):
                raise ValueError(f"advect_batch: item {k} {n} shape {tuple(t.shape)} != phi "
                                 f"shape {(nx, ny, nz)}")
        arr[k] = AdvectItem(ptr(phi, f32, "phi"), ptr(work, f32, "work"), ptr(F, f32, "F"),
                            ptr(mask, u8, "mask"), ptr(mask_phi, f32, "mask_phi"), nx, ny, nz,
                            float(dx), float(dt), int(nsub))
    _ctx(items[0][0]).call("petch_advect_batch", len(items), ctypes.cast(arr, ctypes.c_void_p))


def reinit_fsm_batch(items):
    """reinit_fsm for many grids in one launch per stage (petch_reinit_fsm_batch). items: (phi,
    dx, band) or (phi, dx, band, n_iter) tuples; each phi is replaced in place."""
    import ctypes
    from ._abi import ReinitItem
    if not items:
        return
    arr = (ReinitItem * len(items))()
    for k, it in enumerate(items):
        phi, dx, band = it[:3]
        n_iter = it[3] if len(it) > 3 and it[3] is not None else -1
        nx, ny, nz = _shape3(phi)
        arr[k] = ReinitItem(ptr(phi, f32, "phi"), nx, ny, nz, float(dx), float(band), int(n_iter))
    _ctx(items[0][0]).call("petch_reinit_fsm_batch", len(items), ctypes.cast(arr, ctypes.c_void_p))


def slice_gas_area(phi, dx, eps):
    """Smoothed gas area of every z slice (float64 (nz,)): dx^2 * sum of H_eps(-phi) over the
    slice (petch_profile_area). The equivalent radius of a hole is sqrt(area / pi)."""
    nx, ny, nz = _shape3(phi)
    area = torch.empty((nz,), dtype=torch.float64, device=phi.device)
    _ctx(phi).call("petch_profile_area", ptr(phi, f32, "phi"), nx, ny, nz, float(dx), float(eps),
                   ptr(area))
    return area


def slice_gas_area_adjoint(phi, dx, eps, garea):
    """d loss / d phi (float32, phi's shape) of slice_gas_area for garea = d loss / d area."""
    nx, ny, nz = _shape3(phi)
    if garea.numel() != nz:
        raise ValueError(f"slice_gas_area_adjoint: garea holds {garea.numel()} values, needs {nz}")
    gphi = torch.empty_like(phi)
    _ctx(phi).call("petch_profile_area_adjoint", ptr(phi, f32, "phi"), nx, ny, nz, float(dx),
                   float(eps), ptr(garea.contiguous(), torch.float64, "garea"), ptr(gphi))
    return gphi


# ------------------------------------------------------------------ surface mesh
class MeshBuffers:
    """Grow-only device buffers for the per-step surface mesh (reused step to step)."""

    def __init__(self, device):
        self.device = device
        self.cap_v = 0
        self.cap_f = 0
        self.verts = None
        self.faces = None

    def ensure(self, nv, nf):
        if nv > self.cap_v:
            self.cap_v = max(nv + nv // 2, 1024)
            self.verts = torch.empty((self.cap_v, 3), dtype=f32, device=self.device)
        if nf > self.cap_f:
            self.cap_f = max(nf + nf // 2, 2048)
            self.faces = torch.empty((self.cap_f, 3), dtype=i32, device=self.device)


def marching_cubes(phi, dx, bufs=None):
    """Iso-0 surface of phi: (verts (nv,3) f32, faces (nf,3) i32) device views
    (threed.py:62-100). `bufs` (MeshBuffers) keeps capacity across calls."""
    import ctypes
    nx, ny, nz = _shape3(phi)
    if bufs is None:
        bufs = MeshBuffers(phi.device)
    ctx = _ctx(phi)
    nv = ctypes.c_int()
    nf = ctypes.c_int()
    from ._abi import PetchCapacityError
    for _ in range(2):
        bufs.ensure(1, 1)
        try:
            ctx.call("petch_marching_cubes", ptr(phi, f32, "phi"), nx, ny, nz, float(dx),
                     ptr(bufs.verts), ptr(bufs.faces), bufs.cap_v, bufs.cap_f,
                     ctypes.byref(nv), ctypes.byref(nf))
            break
        except PetchCapacityError:
            bufs.ensure(nv.value, nf.value)
    else:
        raise RuntimeError("marching_cubes: capacity retry failed")
    return bufs.verts[:nv.value], bufs.faces[:nf.value]


def marching_cubes_batch(items):
    """marching_cubes for many grids with one host read of their counts. items: (phi, dx, bufs)
    tuples (bufs: MeshBuffers, grown as needed); returns (verts, faces) per item."""
    import ctypes
    import numpy as np
    from ._abi import McItem, PetchCapacityError
    if not items:
        return []
    counts = np.zeros((len(items), 2), np.int32)
    base = counts.ctypes.data
    for _ in range(2):
        arr = (McItem * len(items))()
        for k, (phi, dx, bufs) in enumerate(items):
            nx, ny, nz = _shape3(phi)
            bufs.ensure(1, 1)
            arr[k] = McItem(ptr(phi, f32, "phi"), nx, ny, nz, float(dx), ptr(bufs.verts),
                            ptr(bufs.faces), bufs.cap_v, bufs.cap_f, base + 8 * k, base + 8 * k + 4)
        try:
            _ctx(items[0][0]).call("petch_marching_cubes_batch", len(items),
                                   ctypes.cast(arr, ctypes.c_void_p))
            break
        except PetchCapacityError:
            for (phi, dx, bufs), (nv, nf) in zip(items, counts.tolist()):
                bufs.ensure(nv, nf)
    else:
        raise RuntimeError("marching_cubes_batch: capacity retry failed")
    return [(bufs.verts[:nv], bufs.faces[:nf])
            for (_, _, bufs), (nv, nf) in zip(items, counts.tolist())]


def mesh_attrs(verts, faces):
    """(centroids (F,3), areas (F,), unit normals (F,3)) in float32 (threed.py:96-99, 228-234)."""
    nf = int(faces.shape[0])
    cen = torch.empty((nf, 3), dtype=f32, device=verts.device)
    area = torch.empty((nf,), dtype=f32, device=verts.device)
    nrm = torch.empty((nf, 3), dtype=f32, device=verts.device)
    if nf:
        _ctx(verts).call("petch_mesh_attrs", ptr(verts, f32, "verts"), ptr(faces, i32, "faces"), nf,
                         ptr(cen), ptr(area), ptr(nrm))
    return cen, area, nrm


def face_adjacency(faces, nv):
    """(F,3) int32 neighbour across each face edge, -1 on an open boundary (threed.py:103-125)."""
    nf = int(faces.shape[0])
    nbr = torch.empty((nf, 3), dtype=i32, device=faces.device)
    if nf:
        _ctx(faces).call("petch_face_adjacency", ptr(faces, i32, "faces"), nf, int(nv), ptr(nbr))
    return nbr


def surface_build_batch(items):
    """mesh_attrs and face_adjacency of many meshes in one launch per stage
    (petch_surface_build_batch). items: (verts, faces, centroids, areas, normals, nbr[, finite])
    with the outputs allocated as those two calls return them; they are filled in place. finite:
    an int32 device scalar set to 1 if every centroid is finite, else 0 (not written for a
    mesh without faces)."""
    from ._abi import SurfaceItem
    items = [it for it in items if int(it[1].shape[0])]
    if not items:
        return
    arr = (SurfaceItem * len(items))()
    for k, (v, f, cen, area, nrm, nbr, *fin) in enumerate(items):
        fin = fin[0] if fin else None
        if fin is not None and (fin.dtype != i32 or fin.numel() != 1):
            raise ValueError(f"surface_build_batch: item {k} finite flag must be one int32")
        nf, nv = int(f.shape[0]), int(v.shape[0])
        if tuple(cen.shape) != (nf, 3) or tuple(area.shape) != (nf,) or \
                tuple(nrm.shape) != (nf, 3) or tuple(nbr.shape) != (nf, 3):
            raise ValueError(f"surface_build_batch: item {k} outputs do not match {nf} faces")
        arr[k] = SurfaceItem(ptr(v, f32, "verts"), ptr(f, i32, "faces"), nf, nv,
                                  ptr(cen, f32, "centroids"), ptr(area, f32, "areas"),
                                  ptr(nrm, f32, "normals"), ptr(nbr, i32, "nbr"),
                                  ptr(fin))
    _ctx(items[0][0]).call("petch_surface_build_batch", len(items), arr)


def gas_normals(phi, dx, centroids, normals):
    """Unit normals oriented into the gas (threed.py:1553-1568)."""
    nx, ny, nz = _shape3(phi)
    nf = int(centroids.shape[0])
    out = torch.empty_like(normals)
    if nf:
        _ctx(phi).call("petch_gas_normals", ptr(phi, f32, "phi"), nx, ny, nz, float(dx),
                       ptr(centroids, f32), ptr(normals, f32), nf, ptr(out))
    return out


# ------------------------------------------------------------------ LBVH
class Bvh:
    """Triangle LBVH on the device (the wp.Mesh of threed.py:1674-1675). Rebuild with build()."""

    def __init__(self, device=None):
        import ctypes
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.ctx = Context.get(dev)
        self.device = dev
        h = ctypes.c_void_p()
        self.ctx.call("petch_bvh_create", ctypes.byref(h))
        self.h = h
        self.nf = 0

    def set_quality(self, quality):
        """-1 by mesh size (default), 0 fastest build ... 2 fastest traces."""
        self.ctx.call("petch_bvh_set_quality", self.h, int(quality))
        return self

    def build(self, verts, faces):
        self.nf = int(faces.shape[0])
        self.ctx.call("petch_bvh_build", self.h, ptr(verts, f32, "verts"), ptr(faces, i32, "faces"),
                      int(verts.shape[0]), self.nf)
        return self

    def query_ray(self, orig, dirs, tmax=1.0e6):
        n = int(orig.shape[0])
        face = torch.empty((n,), dtype=i32, device=self.device)
        t = torch.empty((n,), dtype=f32, device=self.device)
        if n:
            self.ctx.call("petch_bvh_query_ray", self.h, ptr(orig, f32), ptr(dirs, f32), n,
                          float(tmax), ptr(face), ptr(t))
        return face, t

    def query_point(self, pts, max_dist=1.0e6):
        n = int(pts.shape[0])
        face = torch.empty((n,), dtype=i32, device=self.device)
        dist = torch.empty((n,), dtype=f32, device=self.device)
        if n:
            self.ctx.call("petch_bvh_query_point", self.h, ptr(pts, f32), n, float(max_dist),
                          ptr(face), ptr(dist))
        return face, dist

    def status(self):
        import ctypes
        e = ctypes.c_int()
        self.ctx.call("petch_bvh_status", self.h, ctypes.byref(e))
        return e.value

    def status_into(self, dst):
        """Copies the guard flag into the int32 device tensor element dst (no host wait)."""
        self.ctx.call("petch_bvh_status_async", self.h, ptr(dst, i32))

    def root_bounds(self):
        import numpy as np
        out = np.zeros(6, np.float32)
        import ctypes
        self.ctx.call("petch_bvh_root_bounds", self.h, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None:
                from ._abi import lib
                lib().petch_bvh_destroy(self.ctx.h, self.h)
                self.h = None
        except Exception:
            pass


# ------------------------------------------------------------------ ray tracing
from ._abi import Box, IonParams, Source  # noqa: E402

SRC_NEUTRAL, SRC_ION_GAUSS, SRC_ION_POWCOS = 0, 1, 2


def make_box(Lx, Ly, Lz, z_src, periodic=0):
    return Box(float(Lx), float(Ly), float(Lz), float(z_src), int(bool(periodic)))


def make_source(kind, sigma=0.0, exponent=0.0):
    return Source(int(kind), float(sigma), float(exponent))


def gen_source(box, src, n, seed, device=None):
    import ctypes
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    o = torch.empty((n, 3), dtype=f32, device=dev)
    d = torch.empty((n, 3), dtype=f32, device=dev)
    if n:
        Context.get(dev).call("petch_gen_source", ctypes.byref(box), ctypes.byref(src), int(n),
                              int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(o), ptr(d))
    return o, d


def trace_neutral_rr(bvh, box, src, bare, beta, n, seed, out=None):
    """Russian-roulette neutral flux accumulator (F,) (threed.py:549-596)."""
    import ctypes
    fl = torch.empty((bvh.nf,), dtype=f32, device=bvh.device) if out is None else out
    bvh.ctx.call("petch_trace_neutral_rr", bvh.h, ctypes.byref(box), ctypes.byref(src),
                 ptr(bare, f32, "bare"), float(beta), int(n), int(seed) & 0xFFFFFFFFFFFFFFFF,
                 ptr(fl, f32))
    return fl


def trace_neutral_cap(bvh, box, src, bare, beta, n_reemit, n, seed):
    """Capped coverage-trace arrivals (F,) of _trace3d_cov (threed.py:511-546): 1 per hit,
    sticking with probability bare * beta, at most n_reemit re-emissions, no lateral wrap."""
    import ctypes
    if bare.numel() < bvh.nf:
        raise ValueError(f"trace_neutral_cap: bare holds {bare.numel()} values, needs {bvh.nf}")
    fl = torch.empty((bvh.nf,), dtype=f32, device=bvh.device)
    bvh.ctx.call("petch_trace_neutral_cap", bvh.h, ctypes.byref(box), ctypes.byref(src),
                 ptr(bare, f32, "bare"), float(beta), int(n_reemit), int(n),
                 int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(fl, f32))
    return fl


def trace_redeposition(bvh, centroids, gas_normals, areas, rate, s_redep, n, seed):
    """Redeposited flux density (F,) of mc_redep_3d (threed.py:1582-1605): n product rays
    emitted from the faces in proportion to max(rate * area, 0), cosine about the into-gas
    normals, re-sticking with probability s_redep (petch_trace_redeposition)."""
    F = bvh.nf
    for t, name, shape in ((centroids, "centroids", (F, 3)), (gas_normals, "gas_normals", (F, 3)),
                           (areas, "areas", (F,)), (rate, "rate", (F,))):
        if tuple(t.shape) != shape:
            raise ValueError(f"trace_redeposition: {name} has shape {tuple(t.shape)}, "
                             f"needs {shape}")
    out = torch.empty((F,), dtype=f32, device=bvh.device)
    bvh.ctx.call("petch_trace_redeposition", bvh.h, ptr(centroids, f32, "centroids"),
                 ptr(gas_normals, f32, "gas_normals"), ptr(areas, f32, "areas"),
                 ptr(rate, f32, "rate"), float(s_redep), int(n),
                 int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(out, f32))
    return out


def trace_neutral_rr_adjoint(bvh, box, src, bare, beta, n, seed, gfl):
    """d loss / d bare (F,) of trace_neutral_rr(bvh, box, src, bare, beta, n, seed) for
    gfl = d loss / d fl, paths and roulette decisions held fixed (petch_trace_neutral_rr_adjoint)."""
    import ctypes
    for t, name in ((bare, "bare"), (gfl, "gfl")):
        if t.numel() < bvh.nf:
            raise ValueError(f"trace_neutral_rr_adjoint: {name} holds {t.numel()} values, "
                             f"needs {bvh.nf}")
    gbare = torch.empty((bvh.nf,), dtype=f32, device=bvh.device)
    bvh.ctx.call("petch_trace_neutral_rr_adjoint", bvh.h, ctypes.byref(box), ctypes.byref(src),
                 ptr(bare, f32, "bare"), float(beta), int(n), int(seed) & 0xFFFFFFFFFFFFFFFF,
                 ptr(gfl, f32, "gfl"), ptr(gbare, f32))
    return gbare


def trace_neutral_rr_pair_adjoint(bvh, box, src, bare, beta_a, n_a, seed_a, gfl_a, beta_b, n_b,
                                  seed_b, gfl_b):
    """d loss / d bare (F,) of trace_neutral_rr_pair for the gradients of its two outputs, in one
    launch (petch_trace_neutral_rr_pair_adjoint)."""
    import ctypes
    for t, name in ((bare, "bare"), (gfl_a, "gfl_a"), (gfl_b, "gfl_b")):
        if t.numel() < bvh.nf:
            raise ValueError(f"trace_neutral_rr_pair_adjoint: {name} holds {t.numel()} values, "
                             f"needs {bvh.nf}")
    gbare = torch.empty((bvh.nf,), dtype=f32, device=bvh.device)
    bvh.ctx.call("petch_trace_neutral_rr_pair_adjoint", bvh.h, ctypes.byref(box),
                 ctypes.byref(src), ptr(bare, f32, "bare"), float(beta_a), int(n_a),
                 int(seed_a) & 0xFFFFFFFFFFFFFFFF, ptr(gfl_a, f32, "gfl_a"), float(beta_b),
                 int(n_b), int(seed_b) & 0xFFFFFFFFFFFFFFFF, ptr(gfl_b, f32, "gfl_b"),
                 ptr(gbare, f32))
    return gbare


def trace_neutral_rr_pair(bvh, box, src, bare, beta_a, n_a, seed_a, beta_b, n_b, seed_b):
    """Two Russian-roulette neutral species over the same surface and coverage in one launch;
    each output equals trace_neutral_rr with the same (beta, n, seed)."""
    import ctypes
    fa = torch.empty((bvh.nf,), dtype=f32, device=bvh.device)
    fb = torch.empty((bvh.nf,), dtype=f32, device=bvh.device)
    bvh.ctx.call("petch_trace_neutral_rr_pair", bvh.h, ctypes.byref(box), ctypes.byref(src),
                 ptr(bare, f32, "bare"), float(beta_a), int(n_a), int(seed_a) & 0xFFFFFFFFFFFFFFFF,
                 ptr(fa, f32), float(beta_b), int(n_b), int(seed_b) & 0xFFFFFFFFFFFFFFFF,
                 ptr(fb, f32))
    return fa, fb


def trace_neutral_rr_batch(src, items):
    """One launch for many features. items: dicts with bvh, box, bare, beta_a, n_a, seed_a,
    beta_b, n_b, seed_b; returns [(fl_a, fl_b)] per item, each equal to trace_neutral_rr_pair
    on that feature alone."""
    import ctypes
    from ._abi import NeutralFeature
    if not items:
        return []
    outs = []
    arr = (NeutralFeature * len(items))()
    for k, it in enumerate(items):
        bvh = it['bvh']
        fa = torch.empty((bvh.nf,), dtype=f32, device=bvh.device)
        fb = torch.empty((bvh.nf,), dtype=f32, device=bvh.device)
        outs.append((fa, fb))
        bare = ptr(it['bare'], f32, "bare") if bvh.nf else None
        arr[k] = NeutralFeature(bvh.h, it['box'], bare,
                                float(it['beta_a']), int(it['n_a']),
                                int(it['seed_a']) & 0xFFFFFFFFFFFFFFFF, ptr(fa, f32),
                                float(it['beta_b']), int(it['n_b']),
                                int(it['seed_b']) & 0xFFFFFFFFFFFFFFFF, ptr(fb, f32))
    items[0]['bvh'].ctx.call("petch_trace_neutral_rr_batch", ctypes.byref(src), len(items),
                             ctypes.cast(arr, ctypes.c_void_p))
    return outs


def trace_ion_yield_batch(items):
    """One launch for many features. items: dicts with bvh, box, src, ion, n, seed; returns a
    (5, F) accumulator per item, each equal to trace_ion_yield on that feature alone."""
    import ctypes
    from ._abi import IonFeature
    if not items:
        return []
    outs = []
    arr = (IonFeature * len(items))()
    for k, it in enumerate(items):
        bvh = it['bvh']
        acc = torch.empty((5, bvh.nf), dtype=f32, device=bvh.device)
        outs.append(acc)
        arr[k] = IonFeature(bvh.h, it['box'], it['src'], it['ion'], int(it['n']),
                            int(it['seed']) & 0xFFFFFFFFFFFFFFFF, ptr(acc[0]), ptr(acc[1]),
                            ptr(acc[2]), ptr(acc[3]), ptr(acc[4]))
    items[0]['bvh'].ctx.call("petch_trace_ion_yield_batch", len(items),
                             ctypes.cast(arr, ctypes.c_void_p))
    return outs


def trace_ion_yield(bvh, box, src, ion, n, seed, out=None):
    """Faithful ion accumulators (5, F): sp, ie, ip, arrival, weighted cos (threed.py:599-712)."""
    import ctypes
    acc = torch.empty((5, bvh.nf), dtype=f32, device=bvh.device) if out is None else out
    bvh.ctx.call("petch_trace_ion_yield", bvh.h, ctypes.byref(box), ctypes.byref(src),
                 ctypes.byref(ion), int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(acc[0]),
                 ptr(acc[1]), ptr(acc[2]), ptr(acc[3]), ptr(acc[4]))
    return acc


def trace_ion_yield_tangent(bvh, box, src, ion, n, seed):
    """d (sp, ie, ip) / d ion.meanE (3, F) of trace_ion_yield with the same arguments
    (petch_trace_ion_yield_tangent: forward mode along the same paths)."""
    import ctypes
    t = torch.empty((3, bvh.nf), dtype=f32, device=bvh.device)
    bvh.ctx.call("petch_trace_ion_yield_tangent", bvh.h, ctypes.byref(box), ctypes.byref(src),
                 ctypes.byref(ion), int(n), int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(t[0]), ptr(t[1]),
                 ptr(t[2]))
    return t


def trace_legacy(bvh, box, src, sticking, n_reemit, specular, cos_thr, eta, n, seed):
    """_trace3d accumulators (2, F): flux, weighted cos (threed.py:352-406)."""
    import ctypes
    acc = torch.empty((2, bvh.nf), dtype=f32, device=bvh.device)
    bvh.ctx.call("petch_trace_legacy", bvh.h, ctypes.byref(box), ctypes.byref(src),
                 float(sticking), int(n_reemit), int(specular), float(cos_thr), float(eta), int(n),
                 int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(acc[0]), ptr(acc[1]))
    return acc


def trace_ff(bvh, box, centroids, gas_normals, n_ff, seed):
    """First-hit face of n_ff cosine rays per face, -1 = escape (threed.py:908-920, 949-955)."""
    import ctypes
    nf = int(centroids.shape[0])
    hf = torch.empty((nf * int(n_ff),), dtype=i32, device=centroids.device)
    bvh.ctx.call("petch_trace_ff", bvh.h, ctypes.byref(box), ptr(centroids, f32),
                 ptr(gas_normals, f32), nf, int(n_ff), int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(hf))
    return hf


# ------------------------------------------------------------------ per-face kernels
from ._abi import ChemParams, MaskGeom  # noqa: E402

INF = float("inf")


def normalize(fl, areas, scale, lo=0.0, hi=INF, out=None):
    """clamp((fl / max(area,1e-9)) * scale, lo, hi) (threed.py:165-170)."""
    F = int(fl.shape[0])
    out = torch.empty_like(fl) if out is None else out
    if F:
        _ctx(fl).call("petch_normalize", ptr(fl, f32), ptr(areas, f32), F, float(scale),
                      float(lo), float(hi), ptr(out, f32))
    return out


def cos_ratio(fl, ang):
    F = int(fl.shape[0])
    out = torch.empty_like(fl)
    if F:
        _ctx(fl).call("petch_cos_ratio", ptr(fl, f32), ptr(ang, f32), F, ptr(out, f32))
    return out


def smooth(src, nbr, normals, n_iter=1, alpha=1.0, blend_each=False):
    """Edge-neighbour smoothing (threed.py:129-148; blend_each=True: threed.py:195-214)."""
    F = int(src.shape[0])
    out = torch.empty_like(src)
    if F:
        _ctx(src).call("petch_smooth", ptr(src, f32), ptr(nbr, i32), ptr(normals, f32), F,
                       int(n_iter), float(alpha), int(bool(blend_each)), ptr(out, f32))
    return out


def normalize_adjoint(fl, areas, scale, lo, hi, gout):
    """d loss / d fl of normalize(fl, areas, scale, lo, hi) for gout (petch_normalize_adjoint)."""
    F = int(fl.shape[0])
    for t, name in ((areas, "areas"), (gout, "gout")):
        if t.numel() < F:
            raise ValueError(f"normalize_adjoint: {name} holds {t.numel()} values, needs {F}")
    gfl = torch.empty_like(fl)
    if F:
        _ctx(fl).call("petch_normalize_adjoint", ptr(fl, f32, "fl"), ptr(areas, f32, "areas"), F,
                      float(scale), float(lo), float(hi), ptr(gout, f32, "gout"), ptr(gfl, f32))
    return gfl


def smooth_adjoint(nbr, normals, n_iter, alpha, blend_each, gout):
    """d loss / d src of smooth(src, nbr, normals, n_iter, alpha, blend_each) for gout
    (petch_smooth_adjoint; the smoothing is linear, so src is not needed)."""
    F = int(gout.shape[0])
    if nbr.numel() < 3 * F or normals.numel() < 3 * F:
        raise ValueError("smooth_adjoint: nbr and normals need 3 values per face")
    gsrc = torch.empty_like(gout)
    if F:
        _ctx(gout).call("petch_smooth_adjoint", ptr(nbr, i32, "nbr"), ptr(normals, f32, "normals"),
                        F, int(n_iter), float(alpha), int(bool(blend_each)), ptr(gout, f32, "gout"),
                        ptr(gsrc, f32))
    return gsrc


def belen_coverage(m_i, m_F, m_O, cos_i, ion_yield, chem):
    """bare = clip(1 - thF - thO, 0, 1) with the Belen coverages (threed.py:715-737, 903-904)."""
    F = int(m_F.shape[0])
    bare = torch.empty_like(m_F)
    if F:
        ie = ion_yield[1] if ion_yield is not None else None
        ip = ion_yield[2] if ion_yield is not None else None
        _ctx(m_F).call("petch_belen_coverage", ptr(m_i), ptr(m_F, f32), ptr(m_O, f32), ptr(cos_i),
                       ptr(ie), ptr(ip), ctypes_ref(chem), F, ptr(bare), ptr(None), ptr(None))
    return bare


def belen_coverage_adjoint(m_i, m_F, m_O, cos_i, ion_yield, chem, gbare):
    """Adjoint of belen_coverage (petch_belen_coverage_adjoint): (gm_F, gm_O, gm_i, g_ie, g_ip)
    for gbare = d loss / d bare; gm_i is None with ion yields, g_ie/g_ip are None without."""
    F = int(m_F.shape[0])
    for t, name in ((m_O, "m_O"), (gbare, "gbare")) + \
            (((ion_yield[1], "iy_ie"), (ion_yield[2], "iy_ip")) if ion_yield is not None
             else ((m_i, "m_i"), (cos_i, "cos_i"))):
        if t.numel() < F:
            raise ValueError(f"belen_coverage_adjoint: {name} holds {t.numel()} values, needs {F}")
    gm_F, gm_O = torch.empty_like(m_F), torch.empty_like(m_F)
    gm_i = g_ie = g_ip = None
    if ion_yield is not None:
        ie, ip = ion_yield[1], ion_yield[2]
        g_ie, g_ip = torch.empty_like(m_F), torch.empty_like(m_F)
    else:
        ie = ip = None
        gm_i = torch.empty_like(m_F)
    if F:
        _ctx(m_F).call("petch_belen_coverage_adjoint", ptr(m_i), ptr(m_F, f32, "m_F"),
                       ptr(m_O, f32, "m_O"), ptr(cos_i), ptr(ie), ptr(ip), ctypes_ref(chem), F,
                       ptr(gbare, f32, "gbare"), ptr(gm_F), ptr(gm_O), ptr(gm_i), ptr(g_ie),
                       ptr(g_ip))
    return gm_F, gm_O, gm_i, g_ie, g_ip


def surface_rate_adjoint(m_i, m_F, m_O, cos_i, ion_yield, centroids, mask_geom, chem, gV,
                         is_mask=None, per_face_yields=False):
    """Adjoint of surface_rate (petch_surface_rate_adjoint): (gm_F, gm_O, gm_i, g_sp, g_ie,
    g_ip) for gV = d loss / dV. With ion yields and the Belen chemistry the rate reads the yields
    and not m_i, so gm_i is None; without yields the g_* are None. The Langmuir rate reads m_i
    even when yields are passed (belen.py:34 reads them only for Belen), so there gm_i is the
    gradient and the g_* are zeros. With per_face_yields, a seventh value: (3, F) per-face
    shares of d loss / d (Yie, Ysp, Yp)."""
    F = int(m_F.shape[0])
    for t, name in ((m_i, "m_i"), (m_O, "m_O"), (cos_i, "cos_i"), (gV, "gV")):
        if t.numel() < F:
            raise ValueError(f"surface_rate_adjoint: {name} holds {t.numel()} values, needs {F}")
    sp, ie, ip = ion_yield if ion_yield is not None else (None, None, None)
    gm_F, gm_O = torch.empty_like(m_F), torch.empty_like(m_F)
    gm_i = g_sp = g_ie = g_ip = None
    if ion_yield is not None and chem.chemistry_belen:
        g_sp, g_ie, g_ip = (torch.empty_like(m_F) for _ in range(3))
    else:
        gm_i = torch.empty_like(m_F)
        if ion_yield is not None:
            g_sp, g_ie, g_ip = (torch.zeros_like(m_F) for _ in range(3))
    gY = torch.zeros((3, F), dtype=f32, device=m_F.device) if per_face_yields else None
    if F:
        _ctx(m_F).call("petch_surface_rate_adjoint", ptr(m_i, f32), ptr(m_F, f32), ptr(m_O, f32),
                       ptr(cos_i, f32), ptr(sp), ptr(ie), ptr(ip), ptr(centroids), ptr(is_mask),
                       ctypes_ref(mask_geom), ctypes_ref(chem), F, ptr(gV, f32, "gV"), ptr(gm_F),
                       ptr(gm_O), ptr(gm_i), ptr(g_sp), ptr(g_ie), ptr(g_ip), ptr(gY))
    if per_face_yields:
        return gm_F, gm_O, gm_i, g_sp, g_ie, g_ip, gY
    return gm_F, gm_O, gm_i, g_sp, g_ie, g_ip


def surface_rate(m_i, m_F, m_O, cos_i, ion_yield, centroids, mask_geom, chem, is_mask=None,
                 red=None):
    """(V (F,), red (2,) uint32 = [max V bits, non-finite count]) on the device."""
    F = int(m_F.shape[0])
    V = torch.empty_like(m_F)
    if red is None:
        red = torch.zeros((2,), dtype=torch.int32, device=m_F.device)
    sp, ie, ip = ion_yield if ion_yield is not None else (None, None, None)
    _ctx(m_F).call("petch_surface_rate", ptr(m_i, f32), ptr(m_F, f32), ptr(m_O, f32),
                   ptr(cos_i, f32), ptr(sp), ptr(ie), ptr(ip), ptr(centroids),
                   ptr(is_mask), ctypes_ref(mask_geom), ctypes_ref(chem), F, ptr(V), ptr(red))
    return V, red


def surface_rate_batch(items, red):
    """surface_rate of many surfaces in one launch (petch_surface_rate_batch). items: dicts with
    the surface_rate arguments (m_i, m_F, m_O, cos_i, ion_yield, centroids, mask_geom, chem,
    optional is_mask), the output V (F,) and optional bvh (a Bvh). red: (len(items), 3) int32
    device tensor; row r receives item r's reduction pair and (column 2) its BVH guard flag."""
    from ._abi import RateItem
    n = len(items)
    if tuple(red.shape) != (n, 3) or red.dtype != i32 or not red.is_contiguous():
        raise ValueError(f"surface_rate_batch: red must be a contiguous ({n}, 3) int32 tensor")
    if not n:
        return
    arr = (RateItem * n)()
    base = red.data_ptr()
    for k, it in enumerate(items):
        F = int(it['m_F'].shape[0])
        if tuple(it['V'].shape) != (F,):
            raise ValueError(f"surface_rate_batch: item {k} V does not have {F} entries")
        sp, ie, ip = it['ion_yield'] if it.get('ion_yield') is not None else (None, None, None)
        bvh = it.get('bvh')
        arr[k] = RateItem(ptr(it['m_i'], f32), ptr(it['m_F'], f32), ptr(it['m_O'], f32),
                          ptr(it['cos_i'], f32), ptr(sp), ptr(ie), ptr(ip),
                          ptr(it.get('centroids')), ptr(it.get('is_mask')), it['mask_geom'],
                          it['chem'], F, ptr(it['V'], f32), base + 12 * k,
                          bvh.h if bvh is not None else None,
                          base + 12 * k + 8 if bvh is not None else None)
    _ctx(red).call("petch_surface_rate_batch", n, arr)


def read_rate_reduction(red):
    """(Vmax, n_nonfinite) from the surface-rate reduction buffer (one device sync)."""
    return decode_rate_reduction(red.cpu().numpy())


def decode_rate_reduction(h):
    """(Vmax, n_nonfinite) from a host copy of the reduction pair [max V bits, count]."""
    import numpy as np
    h = np.asarray(h, dtype=np.int32)
    vmax = float(np.frombuffer(h[:1].tobytes(), dtype=np.float32)[0])
    return vmax, int(h[1])


def extend(bvh, V, phi, dx, band, out=None):
    """Fs = V[nearest face] where |phi| < band, else 0 (threed.py:1324-1346)."""
    nx, ny, nz = _shape3(phi)
    Fs = torch.empty_like(phi) if out is None else out
    bvh.ctx.call("petch_extend", bvh.h, ptr(V, f32), ptr(phi, f32), nx, ny, nz, float(dx),
                 float(band), ptr(Fs, f32))
    return Fs


def extend_adjoint(bvh, phi, dx, band, gFs):
    """d loss / dV (nf,) of extend(bvh, V, phi, dx, band) for gFs = d loss / dFs
    (petch_extend_adjoint)."""
    nx, ny, nz = _shape3(phi)
    if tuple(gFs.shape) != (nx, ny, nz):
        raise ValueError(f"extend_adjoint: gFs shape {tuple(gFs.shape)} != phi shape {(nx, ny, nz)}")
    gV = torch.empty((max(bvh.nf, 1),), dtype=f32, device=phi.device)
    bvh.ctx.call("petch_extend_adjoint", bvh.h, ptr(phi, f32, "phi"), nx, ny, nz, float(dx),
                 float(band), ptr(gFs, f32, "gFs"), ptr(gV, f32))
    return gV[:bvh.nf]


def extend_batch(items):
    """extend for many grids in one pass. items: (bvh, V, phi, dx, band, out) tuples; each
    out (the grid's Fs) is written as extend would write it."""
    import ctypes
    from ._abi import ExtendItem
    if not items:
        return
    arr = (ExtendItem * len(items))()
    for k, (bvh, V, phi, dx, band, out) in enumerate(items):
        nx, ny, nz = _shape3(phi)
        arr[k] = ExtendItem(bvh.h, ptr(V, f32), ptr(phi, f32), ptr(out, f32), nx, ny, nz,
                            float(dx), float(band))
    items[0][0].ctx.call("petch_extend_batch", len(items), ctypes.cast(arr, ctypes.c_void_p))


def nearest_face_value_batch(items, miss=1.0):
    """nearest_face_value for many point sets in one launch. items: (bvh, src, pts) tuples;
    returns one output per item."""
    import ctypes
    from ._abi import NearestItem
    outs = []
    arr = (NearestItem * len(items))()
    for k, (bvh, src, pts) in enumerate(items):
        n = int(pts.shape[0])
        out = torch.empty((n,), dtype=f32, device=pts.device)
        outs.append(out)
        arr[k] = NearestItem(bvh.h, ptr(src, f32) if n else None, ptr(pts, f32) if n else None,
                             n, float(miss), ptr(out, f32) if n else None)
    if items:
        items[0][0].ctx.call("petch_nearest_face_value_batch", len(items),
                             ctypes.cast(arr, ctypes.c_void_p))
    return outs


def bvh_build_batch(items):
    """Bvh.build for many meshes with their refits in one launch (petch_bvh_build_batch).
    items: (Bvh, verts, faces) tuples."""
    import ctypes
    if not items:
        return
    n = len(items)
    hs = (ctypes.c_void_p * n)()
    vs = (ctypes.c_void_p * n)()
    fs = (ctypes.c_void_p * n)()
    nv = (ctypes.c_int * n)()
    nf = (ctypes.c_int * n)()
    for k, (bvh, v, f) in enumerate(items):
        hs[k] = bvh.h.value
        vs[k] = ptr(v, f32, "verts").value
        fs[k] = ptr(f, i32, "faces").value
        nv[k], nf[k] = int(v.shape[0]), int(f.shape[0])
        bvh.nf = nf[k]
    items[0][0].ctx.call("petch_bvh_build_batch", n, ctypes.cast(hs, ctypes.c_void_p),
                         ctypes.cast(vs, ctypes.c_void_p), ctypes.cast(fs, ctypes.c_void_p),
                         ctypes.cast(nv, ctypes.c_void_p), ctypes.cast(nf, ctypes.c_void_p))


def nearest_face_value(bvh, src, pts, miss=1.0):
    """src[nearest face of each point] on `bvh` (threed.py:1619-1639)."""
    n = int(pts.shape[0])
    out = torch.empty((n,), dtype=f32, device=pts.device)
    if n:
        bvh.ctx.call("petch_nearest_face_value", bvh.h, ptr(src, f32), ptr(pts, f32), n,
                     float(miss), ptr(out))
    return out


def radiosity_assemble(hitface, nf, n_ff):
    """(row_ptr (nf+1,) i32, col (nnz,) i32, D (nf,) f64) of the form-factor matrix."""
    dev = hitface.device
    row_ptr = torch.empty((nf + 1,), dtype=i32, device=dev)
    col = torch.empty((max(nf * n_ff, 1),), dtype=i32, device=dev)
    D = torch.empty((max(nf, 1),), dtype=torch.float64, device=dev)
    _ctx(hitface).call("petch_radiosity_assemble", ptr(hitface, i32), int(nf), int(n_ff),
                       ptr(row_ptr), ptr(col), ptr(D))
    return row_ptr, col, D


def radiosity_solve(row_ptr, col, D, nf, n_ff, bare, beta, n_iter=40, tol=1e-10):
    """(m = clip(G,0,8) float32, G float64, sweeps)."""
    import ctypes
    G = torch.empty((max(nf, 1),), dtype=torch.float64, device=D.device)
    out = torch.empty((nf,), dtype=f32, device=D.device)
    it = ctypes.c_int()
    if nf:
        _ctx(D).call("petch_radiosity_solve", ptr(row_ptr, i32), ptr(col, i32),
                     ptr(D, torch.float64), int(nf), int(n_ff), ptr(bare, f32), float(beta),
                     int(n_iter), float(tol), ptr(G), ptr(out), ctypes.byref(it))
    return out, G[:nf], it.value


def radiosity_adjoint(hitface, nf, n_ff, G, bare, beta, gm, n_iter=40, tol=1e-10):
    """(gD float64 = d loss / dD, gbare float32 = d loss / d bare, sweeps) for the output gradient
    gm of radiosity_solve at its solution G (petch_radiosity_adjoint)."""
    import ctypes
    for t, n, name in ((hitface, nf * n_ff, "hitface"), (G, nf, "G"), (bare, nf, "bare"),
                       (gm, nf, "gm")):
        if t.numel() < n:
            raise ValueError(f"radiosity_adjoint: {name} holds {t.numel()} values, needs {n}")
    gD = torch.empty((max(nf, 1),), dtype=torch.float64, device=G.device)
    gbare = torch.empty((nf,), dtype=f32, device=G.device)
    it = ctypes.c_int()
    if nf:
        _ctx(G).call("petch_radiosity_adjoint", ptr(hitface, i32, "hitface"), int(nf), int(n_ff),
                     ptr(G, torch.float64, "G"), ptr(bare, f32, "bare"), float(beta),
                     ptr(gm, f32, "gm"), int(n_iter), float(tol), ptr(gD), ptr(gbare),
                     ctypes.byref(it))
    return gD[:nf], gbare, it.value


def ctypes_ref(s):
    import ctypes
    return ctypes.byref(s)


def trace_hit_counts(device=None, reset=True):
    """(neutral hits, ion hits) accumulated on the device's context since the last reset."""
    import ctypes
    import numpy as np
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    out = np.zeros(2, np.uint64)
    Context.get(dev).call("petch_trace_hit_counts", int(bool(reset)),
                          out.ctypes.data_as(ctypes.c_void_p))
    return int(out[0]), int(out[1])


# ------------------------------------------------- deterministic neutral transports ---------

def knudsen_face_flux(phi, dx, sub_top, trench, centroids, gas_normals, areas, bare, betas,
                      Ly, wall_loss_scale=1.85, front_loss=False, band_W=8.0, pass_frac=0.5,
                      ar_pass=15.0, center=None, half_width=None, sink_field=False):
    """knudsen_face_flux (knudsen.py:185-241) for the sticking coefficients bare * beta of the
    one or two betas (petch_knudsen_face_flux): one (F,) float32 multiplier per beta."""
    from ._abi import KnudsenParams
    nx, ny, nz = _shape3(phi)
    F = int(centroids.shape[0])
    betas = list(betas)
    if not 1 <= len(betas) <= 2:
        raise ValueError("knudsen_face_flux: one or two betas")
    outs = [torch.empty((F,), dtype=f32, device=phi.device) for _ in betas]
    fp = center is not None and half_width is not None
    cx, cy = (float(center[0]), float(center[1])) if fp else (0.0, 0.0)
    p = KnudsenParams(float(dx), float(sub_top), float(Ly), int(bool(trench)),
                      float(wall_loss_scale), int(bool(front_loss)), float(band_W),
                      float(pass_frac), float(ar_pass), int(fp), cx, cy,
                      float(half_width) if fp else 0.0, int(bool(sink_field)), float(betas[0]),
                      float(betas[-1]))
    if F:
        _ctx(phi).call("petch_knudsen_face_flux", ptr(phi, f32, "phi"), nx, ny, nz,
                       ptr(centroids, f32, "centroids"), ptr(gas_normals, f32, "gas_normals"),
                       ptr(areas, f32, "areas"), ptr(bare, f32, "bare"), F, ctypes_ref(p),
                       ptr(outs[0]), ptr(outs[1]) if len(outs) > 1 else None)
    return outs


def dda_neutral_flux(phi, dx, centroids, gas_normals, areas, bare, betas, n_dir=64, n_reemit=12,
                     periodic_y=0):
    """dda_neutral_flux (dda.py:138-220) for the sticking coefficients bare * beta of the one or
    two betas (petch_dda_neutral_flux): one (F,) float32 multiplier per beta."""
    from ._abi import DdaParams
    nx, ny, nz = _shape3(phi)
    F = int(centroids.shape[0])
    betas = list(betas)
    if not 1 <= len(betas) <= 2:
        raise ValueError("dda_neutral_flux: one or two betas")
    outs = [torch.empty((F,), dtype=f32, device=phi.device) for _ in betas]
    p = DdaParams(float(dx), int(n_dir), int(n_reemit), int(bool(periodic_y)), float(betas[0]),
                  float(betas[-1]))
    if F:
        _ctx(phi).call("petch_dda_neutral_flux", ptr(phi, f32, "phi"), nx, ny, nz,
                       ptr(centroids, f32, "centroids"), ptr(gas_normals, f32, "gas_normals"),
                       ptr(areas, f32, "areas"), ptr(bare, f32, "bare"), F, ctypes_ref(p),
                       ptr(outs[0]), ptr(outs[1]) if len(outs) > 1 else None)
    return outs


def gaussian_smooth(phi, sigma, mask=None, mask_phi=None, truncate=4.0):
    """In place: phi = gaussian_filter(phi, sigma) (scipy.ndimage, mode 'reflect') then the mask
    nodes re-pinned to mask_phi (threed.py:1779-1782; petch_gaussian_smooth). The weights are
    formed here exactly as scipy forms them."""
    import numpy as np
    nx, ny, nz = _shape3(phi)
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 / (float(sigma) ** 2) * x ** 2)
    w = np.ascontiguousarray(w / w.sum())
    _ctx(phi).call("petch_gaussian_smooth", ptr(phi, f32, "phi"), nx, ny, nz,
                   w.ctypes.data, r, ptr(mask), ptr(mask_phi))
    return phi


def reinit_pde(phi, dx, n_iter=24):
    """In place: reinit_gpu (threed.py:1399-1418; petch_reinit_pde), n_iter PDE steps."""
    nx, ny, nz = _shape3(phi)
    _ctx(phi).call("petch_reinit_pde", ptr(phi, f32, "phi"), nx, ny, nz, float(dx), int(n_iter))
    return phi
