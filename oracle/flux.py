"""Flux-stage and etch-loop oracle — TEST INFRASTRUCTURE ONLY (and the CPU baseline).

CPU restatement of the reference's flux orchestration and per-step driver, built from the
oracle's own pieces (oracle MC, oracle BVH + traces in C with the same Philox streams, FP64
numpy normalisation / smoothing / chemistry / advection / reinit):

  coupled_flux    mc_flux_3d_coupled    threed.py:740-905
  radiosity_flux  mc_flux_3d_radiosity  threed.py:923-991 (+ _ions_deterministic 994-1029),
                  with scipy.sparse for Aff and the 40-sweep Jacobi of threed.py:965-969
  uncoupled_flux  mc_flux_3d            threed.py:472-507
  run_etch_3d     run_etch_3d           threed.py:1643-1795
"""
import math

import numpy as np

from . import chem as oc
from . import levelset as ol
from . import mesh as om
from . import native as on


def _box(geo, par):
    return on.Box(geo['Lx'], geo['Ly'], geo['Lz'], geo['Lz'] - geo['dx'], int(par.get('periodic_y', 0)))


def _ion_src(par):
    if par.get('ion_source', 'gauss') == 'powcos':
        return on.Source(2, 0.0, float(par.get('ion_exponent', 1000.0)))
    return on.Source(1, float(par['ion_ang_sigma']), 0.0)


def _faithful(flags):
    return (flags is not None and getattr(flags, "ion_reflection", False)
            and getattr(flags, "chemistry", "langmuir") == "belen")


class OracleMesh:
    def __init__(self, verts, faces):
        self.v = np.asarray(verts, np.float32)
        self.f = np.asarray(faces, np.int32)
        self.nf = len(self.f)
        self.cen, self.area, _ = om.mesh_attrs(self.v, self.f)
        self.normals = om.face_normals(self.v, self.f)
        self.pairs = om.edge_adjacency(self.f)
        self.bvh = on.OracleBvh(self.v, self.f)


def coupled_flux(M, geo, par, n_ion, n_neu, seed, flags, bare_init=None):
    """threed.py:740-905 with the device-resident normalisation of the GPU path."""
    A_src = geo['Lx'] * geo['Ly']
    A = np.maximum(M.area, 1.0e-9)
    betaE, betaO = par.get('betaE', 0.7), par.get('betaO', 1.0)
    n_smooth = int(par.get('flux_smooth', 1))
    alpha = float(par.get('flux_smooth_alpha', 1.0))
    dev = bool(par.get('device_flux', True)) and bool(par.get('flux_smooth_gpu', True))
    box = _box(geo, par)

    def host_smooth(x):
        return om.smooth_flux(x, M.normals, M.pairs, n_smooth, alpha) if n_smooth > 0 else x

    def finish(fl, scale, hi):
        m = np.clip((fl / A) * scale, 0.0, hi)
        if n_smooth <= 0:
            return m
        if dev:
            for _ in range(n_smooth):
                m = om.smooth_flux_dev(m, M.normals, M.pairs, alpha)
            return m
        return om.smooth_flux(m, M.normals, M.pairs, n_smooth, alpha)

    if _faithful(flags):
        acc = M.bvh.trace_ion_yield(box, _ion_src(par), on.Ion(**oc.ion_params(par)), n_ion,
                                    seed * 9 + 1)
        norm = A_src / n_ion
        iy = tuple(host_smooth((acc[k] / A) * norm) for k in range(3))
        m_i = host_smooth((acc[3] / A) * norm)
        cos_i = om.cos_ratio(acc[3], acc[4])
    else:
        acc = M.bvh.trace_legacy(box, _ion_src(par), 1.0, 0, 0, 0.34, 0.8, n_ion, seed * 9 + 1)
        iy = None
        cos_i = om.cos_ratio(acc[0], acc[1])
        m_i = finish(acc[0], A_src / n_ion, 1.5)
    nsrc = on.Source(0, 0.0, 0.0)

    def neutral(beta, bare, sd):
        fl = M.bvh.trace_neutral_rr(box, nsrc, bare.astype(np.float32), beta, n_neu, sd)
        return finish(fl, A_src / n_neu, 8.0)

    n_fp = int(par.get('n_fp', 1 if bare_init is not None else 4))
    bare = np.ones(M.nf) if bare_init is None else np.clip(np.asarray(bare_init, float), 0.0, 1.0)
    m_F = m_O = np.zeros(M.nf)
    for it in range(n_fp):
        m_F = neutral(betaE, bare, seed * 9 + 2 + 2 * it)
        m_O = neutral(betaO, bare, seed * 9 + 3 + 2 * it)
        thF, thO = oc.belen_coverages(m_i, m_F, m_O, cos_i, par, flags, ion_yield=iy)
        bare = np.clip(1.0 - thF - thO, 0.0, 1.0)
    return dict(m_i=m_i, m_F=m_F, m_O=m_O, cos_i=cos_i, bare=bare, ion_yield=iy)


def _smooth_matrix(M, n_smooth, alpha, dev):
    """The linear map of finish()'s smoothing (om.smooth_flux / smooth_flux_dev) as a scipy
    matrix: one pass is P = D^-1 (I + W), W_ij = max(0, n_i.n_j) over edge pairs."""
    import scipy.sparse as sp
    F = M.nf
    I = sp.identity(F, format='csr')
    if n_smooth <= 0 or alpha <= 0.0 or len(M.pairs) == 0:
        return I
    i, j = M.pairs[:, 0], M.pairs[:, 1]
    w = np.maximum(0.0, np.sum(M.normals[i] * M.normals[j], axis=1))
    W = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([i, j]), np.concatenate([j, i]))),
                      shape=(F, F)).tocsr()
    P = sp.diags(1.0 / (1.0 + np.asarray(W.sum(axis=1)).ravel())) @ (I + W)
    if dev:
        B = (1.0 - alpha) * I + alpha * P if alpha < 1.0 else P
        T = I
        for _ in range(n_smooth):
            T = B @ T
        return T.tocsr()
    T = I
    for _ in range(n_smooth):
        T = P @ T
    return ((1.0 - alpha) * I + alpha * T).tocsr() if alpha < 1.0 else T.tocsr()


def coupled_coverage(M, geo, par, flags, m_i, cos_i, n_neu, seed, n_fp, ion_yield=None, w=None,
                     frozen=None):
    """The neutral rounds of coupled_flux (threed.py:855-904) from given ion terms: n_fp x {RR
    traces of F and O at bare, finish (clip-normalise, smooth), Belen update} from bare = 1.
    Returns (bare, m_F, m_O, bares) with bares the bare each round traced at. With w, also the
    adjoint of w . bare_final (not in the reference): grads dict(m_i or ie/ip), unrolled over the
    rounds, each trace by its path adjoint (paths and decisions fixed), finish by its transpose
    (a scipy matrix of the smoothing, transposed), Belen by belen_bare_adjoint. With frozen (a
    list of per-round bare), each round's traces keep the paths and decisions of frozen[it] and
    take their weights from the current bare (oracle_trace_neutral_rr_frozen): central
    differences of that are what the adjoint computes."""
    A_src = geo['Lx'] * geo['Ly']
    A = np.maximum(M.area, 1.0e-9)
    betaE, betaO = par.get('betaE', 0.7), par.get('betaO', 1.0)
    n_smooth = int(par.get('flux_smooth', 1))
    alpha = float(par.get('flux_smooth_alpha', 1.0))
    dev = bool(par.get('device_flux', True)) and bool(par.get('flux_smooth_gpu', True))
    box = _box(geo, par)
    nsrc = on.Source(0, 0.0, 0.0)
    T = _smooth_matrix(M, n_smooth, alpha, dev)
    scale = A_src / n_neu

    def trace(beta, bare, sd, it):
        if frozen is None:
            return M.bvh.trace_neutral_rr(box, nsrc, bare.astype(np.float32), beta, n_neu, sd)
        return M.bvh.trace_neutral_rr_frozen(box, nsrc, frozen[it].astype(np.float32),
                                             bare.astype(np.float32), beta, n_neu, sd)

    bare = np.ones(M.nf)
    bares, tape = [], []
    m_F = m_O = np.zeros(M.nf)
    for it in range(n_fp):
        bares.append(bare)
        sF, sO = seed * 9 + 2 + 2 * it, seed * 9 + 3 + 2 * it
        fF, fO = trace(betaE, bare, sF, it), trace(betaO, bare, sO, it)
        m_F = T.dot(np.clip((fF / A) * scale, 0.0, 8.0))
        m_O = T.dot(np.clip((fO / A) * scale, 0.0, 8.0))
        tape.append((bare, fF, fO, m_F, m_O, sF, sO))
        thF, thO = oc.belen_coverages(m_i, m_F, m_O, cos_i, par, flags, ion_yield=ion_yield)
        bare = np.clip(1.0 - thF - thO, 0.0, 1.0)
    if w is None:
        return bare, m_F, m_O, bares
    keys = ("ie", "ip") if ion_yield is not None else ("m_i",)
    grads = {k: np.zeros(M.nf) for k in keys}
    gbare = np.asarray(w, np.float64)
    TT = T.T.tocsr()
    for b_in, fF, fO, mF, mO, sF, sO in reversed(tape):
        g = oc.belen_bare_adjoint(m_i, mF, mO, cos_i, par, flags, gbare, ion_yield=ion_yield)
        for k in keys:
            grads[k] += g[k]
        gb = np.zeros(M.nf)
        for fl, gm, beta, sd in ((fF, g["m_F"], betaE, sF), (fO, g["m_O"], betaO, sO)):
            v = (fl / A) * scale
            gfl = np.where((v >= 0.0) & (v <= 8.0), TT.dot(gm) * scale / A, 0.0)
            gb += M.bvh.trace_neutral_rr_adjoint(box, nsrc, b_in.astype(np.float32), beta, n_neu,
                                                 sd, gfl)
        gbare = gb
    return bare, m_F, m_O, bares, grads


def radiosity_matrix(hit, nf, n_ff):
    """threed.py:956-960: D (escape fraction) and Aff[i,j] = rays j->i / n_ff (scipy CSR)."""
    import scipy.sparse as sp
    src = np.repeat(np.arange(nf), n_ff)
    esc = hit < 0
    D = np.bincount(src[esc], minlength=nf).astype(np.float64) / n_ff
    j = src[~esc]
    i = hit[~esc]
    Aff = sp.coo_matrix((np.full(len(i), 1.0 / n_ff), (i, j)), shape=(nf, nf)).tocsr()
    return D, Aff


def radiosity_solve(D, Aff, bare, beta, n_iter=40):
    """threed.py:965-972 (Jacobi from G0 = D), clipped as threed.py:985."""
    s = np.clip(np.asarray(bare, np.float64) * beta, 0.0, 1.0)
    M = Aff.multiply((1.0 - s)[None, :]).tocsr()
    G = D.copy()
    for _ in range(n_iter):
        G = D + M.dot(G)
    return np.clip(G, 0.0, 8.0), G


def radiosity_direct(D, Aff, bare, beta):
    """The exact solution (I - Aff diag(1-s)) G = D by a sparse direct solve (FP64)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    s = np.clip(np.asarray(bare, np.float64) * beta, 0.0, 1.0)
    M = Aff.multiply((1.0 - s)[None, :]).tocsr()
    return spl.spsolve((sp.identity(len(D), format='csc') - M).tocsc(), D)


def radiosity_adjoint(D, Aff, G, bare, beta, gm, n_iter=0):
    """Adjoint of radiosity_solve / radiosity_direct at the solution G (not in the reference):
    lambda solves (I - Aff diag(q))^T lambda = gG with q = 1 - clip(bare beta, 0, 1) and
    gG = gm where 0 <= G <= 8 (n_iter > 0: that many transposed Jacobi sweeps from gG; else a
    sparse direct solve). Returns (dloss/dD = lambda, dloss/dbare =
    -beta [0 <= bare beta <= 1] G (Aff^T lambda))."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    sb = np.asarray(bare, np.float64) * beta
    q = 1.0 - np.clip(sb, 0.0, 1.0)
    G = np.asarray(G, np.float64)
    gG = np.where((G >= 0.0) & (G <= 8.0), np.asarray(gm, np.float64), 0.0)
    MT = Aff.multiply(q[None, :]).T.tocsr()
    if n_iter > 0:
        lam = gG.copy()
        for _ in range(n_iter):
            lam = gG + MT.dot(lam)
    else:
        lam = spl.spsolve((sp.identity(len(gG), format='csc') - MT).tocsc(), gG)
    gq = G * Aff.T.dot(lam)
    return lam, np.where((sb >= 0.0) & (sb <= 1.0), -beta * gq, 0.0)


def radiosity_coverage(D, Aff, m_i, cos_i, par, flags, n_fp=4, ion_yield=None, w=None):
    """The coverage fixed point of radiosity_flux (threed.py:963-990) with converged solves:
    n_fp x {m_F, m_O = clip(G) of the F and O solves at bare; bare = Belen(m_i, m_F, m_O)} from
    bare = 1. With w, also the adjoint of loss = w . bare_final (not in the reference), unrolled
    over the n_fp iterations, each solve by radiosity_adjoint and each Belen step by
    belen_bare_adjoint. Returns (bare, m_F, m_O[, grads dict(m_i or ie/ip, D)])."""
    betaE, betaO = par.get('betaE', 0.7), par.get('betaO', 1.0)
    bare = np.ones(len(D))
    tape = []
    for _ in range(n_fp):
        GF = radiosity_direct(D, Aff, bare, betaE)
        GO = radiosity_direct(D, Aff, bare, betaO)
        mF, mO = np.clip(GF, 0.0, 8.0), np.clip(GO, 0.0, 8.0)
        tape.append((bare, GF, GO, mF, mO))
        thF, thO = oc.belen_coverages(m_i, mF, mO, cos_i, par, flags, ion_yield=ion_yield)
        bare = np.clip(1.0 - thF - thO, 0.0, 1.0)
    if w is None:
        return bare, mF, mO
    keys = ("ie", "ip") if ion_yield is not None else ("m_i",)
    grads = {k: np.zeros(len(D)) for k in keys}
    grads["D"] = np.zeros(len(D))
    gbare = np.asarray(w, np.float64)
    for b_in, GF, GO, mF, mO in reversed(tape):
        g = oc.belen_bare_adjoint(m_i, mF, mO, cos_i, par, flags, gbare, ion_yield=ion_yield)
        for k in keys:
            grads[k] += g[k]
        dF, bF = radiosity_adjoint(D, Aff, GF, b_in, betaE, g["m_F"])
        dO, bO = radiosity_adjoint(D, Aff, GO, b_in, betaO, g["m_O"])
        grads["D"] += dF + dO
        gbare = bF + bO
    return bare, mF, mO, grads


def deterministic_ions(M, geo, par, n_ion, seed, flags):
    """_ions_deterministic (threed.py:994-1029)."""
    box = _box(geo, par)
    A_src = geo['Lx'] * geo['Ly']
    A = np.maximum(M.area, 1.0e-9)
    if _faithful(flags):
        acc = M.bvh.trace_ion_yield(box, _ion_src(par), on.Ion(**oc.ion_params(par)), n_ion, seed)
        norm = A_src / n_ion
        iy = tuple((acc[k] / A) * norm for k in range(3))
        return (acc[3] / A) * norm, om.cos_ratio(acc[3], acc[4]), iy
    acc = M.bvh.trace_legacy(box, _ion_src(par), 1.0, 0, 0, 0.34, 0.8, n_ion, seed)
    return np.clip((acc[0] / A) / (n_ion / A_src), 0.0, 1.5), om.cos_ratio(acc[0], acc[1]), None


def radiosity_flux(M, phi, geo, par, n_ion, seed, flags, n_ff=64, n_fp=4):
    gn = om.gas_normals(M.v, M.f, M.cen, phi, geo['dx']).astype(np.float32)
    m_i, cos_i, iy = deterministic_ions(M, geo, par, n_ion, seed * 9 + 1, flags)
    hit = M.bvh.trace_ff(_box(geo, par), M.cen.astype(np.float32), gn, n_ff, seed * 9 + 5)
    D, Aff = radiosity_matrix(hit, M.nf, n_ff)
    bare = np.ones(M.nf)
    m_F = m_O = np.zeros(M.nf)
    for _ in range(n_fp):
        m_F = radiosity_solve(D, Aff, bare, par.get('betaE', 0.7))[0]
        m_O = radiosity_solve(D, Aff, bare, par.get('betaO', 1.0))[0]
        thF, thO = oc.belen_coverages(m_i, m_F, m_O, cos_i, par, flags, ion_yield=iy)
        bare = np.clip(1.0 - thF - thO, 0.0, 1.0)
    return dict(m_i=m_i, m_F=m_F, m_O=m_O, cos_i=cos_i, bare=bare, ion_yield=iy)


def uncoupled_flux(M, geo, par, n_ion, n_neu, seed, ion_reflection=False):
    box = _box(geo, par)
    A_src = geo['Lx'] * geo['Ly']
    A = np.maximum(M.area, 1.0e-9)
    if ion_reflection:
        acc = M.bvh.trace_legacy(box, _ion_src(par), 1.0, 3, 1, 0.34, 0.8, n_ion, seed * 9 + 1)
    else:
        acc = M.bvh.trace_legacy(box, _ion_src(par), 1.0, 0, 0, 0.34, 0.8, n_ion, seed * 9 + 1)
    nsrc = on.Source(0, 0.0, 0.0)
    aF = M.bvh.trace_legacy(box, nsrc, par['s_F'], 12, 0, 0.34, 0.8, n_neu, seed * 9 + 2)
    aO = M.bvh.trace_legacy(box, nsrc, par['s_O'], 12, 0, 0.34, 0.8, n_neu, seed * 9 + 3)
    return dict(m_i=np.clip((acc[0] / A) / (n_ion / A_src), 0.0, 1.5),
                m_F=np.clip((aF[0] / A) / (n_neu / A_src), 0.0, 4.0),
                m_O=np.clip((aO[0] / A) / (n_neu / A_src), 0.0, 4.0),
                cos_i=np.where(acc[0] > 0, acc[1] / np.maximum(acc[0], 1e-9), 0.0),
                bare=None, ion_yield=None)


def extend(M, V, phi, dx, band):
    """threed.py:1333-1346 with the oracle closest-point query."""
    phi = np.asarray(phi, np.float64)
    Fs = np.zeros_like(phi)
    ii, jj, kk = np.where(np.abs(phi) < band)
    if len(ii) == 0 or M.nf == 0:
        return Fs
    pts = np.stack([ii * dx, jj * dx, kk * dx], axis=1).astype(np.float32)
    face, _ = M.bvh.closest_point(pts)
    Fs[ii, jj, kk] = np.where(face >= 0, np.asarray(V)[np.maximum(face, 0)], 0.0)
    return Fs


def run_etch_3d(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, t_end, n_steps, hole, par,
                flags, n_ion, n_neu, seed_offset=0, rays_per_point=None, step_hook=None):
    """threed.py:1643-1795 on the CPU (fsm reinit, GPU-path semantics for warm start/extend)."""
    phi = ol.init_sdf(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole).astype(np.float32)
    mask = ol.trench_mask(Lx, Ly, Lz, dx, trench_width, mask_th, sub_top, hole)
    mask_phi = phi.copy()
    xs, ys, zs = ol.grid_axes(Lx, Ly, Lz, dx)
    geo = dict(xs=xs, ys=ys, zs=zs, dx=dx, Lx=Lx, Ly=Ly, Lz=Lz, sub_top=sub_top,
               trench_width=trench_width, hole=hole)
    dt = t_end / n_steps
    band = 4 * dx
    warm = getattr(flags, "warm_start_coverage", False)
    prev_M, prev_bare = None, None
    for step in range(n_steps):
        v, f = on.marching_cubes(phi, dx)
        M = OracleMesh(v, f)
        if rays_per_point is not None:
            n_ion_s = n_neu_s = max(int(rays_per_point) * M.nf, 1000)
        else:
            n_ion_s, n_neu_s = n_ion, n_neu
        seed = step + seed_offset
        nt = getattr(flags, "neutral_transport", "mc")
        if nt == "radiosity":
            out = radiosity_flux(M, phi, geo, par, n_ion_s, seed, flags)
        elif getattr(flags, "coverage_sticking", False):
            bi = None
            if warm and prev_bare is not None:
                face, _ = prev_M.bvh.closest_point(M.cen.astype(np.float32))
                bi = np.where(face >= 0, prev_bare[np.maximum(face, 0)], 1.0)
            out = coupled_flux(M, geo, par, n_ion_s, n_neu_s, seed, flags, bare_init=bi)
            prev_M, prev_bare = M, out['bare'].astype(np.float32)
        else:
            out = uncoupled_flux(M, geo, par, n_ion_s, n_neu_s, seed,
                                 getattr(flags, "ion_reflection", False))
        is_mask = oc.faces_in_mask(M.cen, geo, mask_th, trench_width, hole=hole)
        iy = out['ion_yield'] if getattr(flags, "chemistry", "") == "belen" else None
        V = oc.surface_rate(out['m_i'], out['m_F'], out['m_O'], out['cos_i'], is_mask, par,
                            flags, ion_yield=iy)
        V = np.nan_to_num(V, nan=0.0, posinf=0.0, neginf=0.0)
        Fs = extend(M, V.astype(np.float32), phi, dx, band)
        vmx = float(np.max(V)) if V.size else 0.0
        Vmax = max(vmx if np.isfinite(vmx) else 0.0, 1e-6)
        nsub = max(1, min(int(math.ceil(Vmax * dt / (0.4 * dx))), 160))
        phi = ol.advect_pinned(phi.astype(np.float64), Fs.astype(np.float32).astype(np.float64),
                               mask, mask_phi.astype(np.float64), dx, dt / nsub, nsub).astype(np.float32)
        phi = ol.reinit_fsm(phi.astype(np.float64), dx, band + 2.0 * dx).astype(np.float32)
        if step_hook is not None:
            step_hook(step, dict(phi=phi, mesh=M, out=out, V=V, nsub=nsub))
    geo['phi'] = phi.astype(np.float64)
    return geo
