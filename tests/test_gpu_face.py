"""Per-face kernels (normalise, smoothing, coverages, surface rate, extension, warm start) and
the radiosity solve against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

from oracle import chem as oc
from oracle import flux as ofl
from oracle import levelset as ol
from oracle import mesh as om
from oracle import native as on
from petch.params import PAR, Flags

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda", 0)


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device=_dev(), dtype=dtype)


def _mesh():
    """Carved hole mesh (the oracle MC equals the GPU MC bit for bit)."""
    phi = ol.init_sdf(8.0, 8.0, 14.0, 0.25, 3.0, 8.0, 4.0, True).astype(np.float32)
    v, f = on.marching_cubes(phi, 0.25)
    return phi, v, f


def test_normalize_and_cos_ratio():
    import torch
    from petch import ops
    rng = np.random.default_rng(0)
    F = 5000
    fl = rng.exponential(3.0, F).astype(np.float32)
    fl[:50] = 0.0
    area = rng.uniform(0, 0.02, F).astype(np.float32)
    area[:10] = 0.0
    ang = (fl * rng.uniform(0, 1, F)).astype(np.float32)
    m = ops.normalize(_t(fl), _t(area), 37.5, 0.0, 8.0).cpu().numpy()
    np.testing.assert_allclose(m, om.norm_clip(fl, np.maximum(area, np.float32(1e-9)), 37.5, 0.0, 8.0),
                               rtol=2e-7, atol=0)
    m = ops.normalize(_t(fl), _t(area), 37.5, -ops.INF, ops.INF).cpu().numpy()
    np.testing.assert_allclose(m, (fl / np.maximum(area, np.float32(1e-9))) * 37.5, rtol=2e-7)
    c = ops.cos_ratio(_t(fl), _t(ang)).cpu().numpy()
    np.testing.assert_allclose(c, om.cos_ratio(fl, ang), rtol=2e-7)


@pytest.mark.parametrize("n_iter,alpha", [(1, 1.0), (2, 1.0), (1, 0.5), (3, 0.7)])
def test_smoothing_both_semantics(n_iter, alpha):
    from petch import ops
    phi, v, f = _mesh()
    tv, tf = _t(v), _t(f)
    cen, area, nrm = ops.mesh_attrs(tv, tf)
    nbr = ops.face_adjacency(tf, len(v))
    x = np.random.default_rng(1).exponential(1.0, len(f)).astype(np.float32)
    normals = om.face_normals(v, f)
    pairs = om.edge_adjacency(f)
    got = ops.smooth(_t(x), nbr, nrm, n_iter, alpha, False).cpu().numpy()
    np.testing.assert_allclose(got, om.smooth_flux(x, normals, pairs, n_iter, alpha), rtol=2e-5, atol=1e-6)
    ref = x.astype(np.float64)
    for _ in range(n_iter):
        ref = om.smooth_flux_dev(ref, normals, pairs, alpha)
    got = ops.smooth(_t(x), nbr, nrm, n_iter, alpha, True).cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-6)


def _fluxes(F, seed):
    rng = np.random.default_rng(seed)
    m_i = rng.uniform(0, 1.5, F).astype(np.float32)
    m_F = rng.uniform(0, 2.0, F).astype(np.float32)
    m_O = rng.uniform(0, 2.0, F).astype(np.float32)
    cos_i = rng.uniform(0, 1, F).astype(np.float32)
    iy = tuple(rng.uniform(0, 3, F).astype(np.float32) for _ in range(3))
    m_F[:20] = 0.0
    m_O[10:30] = 0.0
    return m_i, m_F, m_O, cos_i, iy


@pytest.mark.parametrize("faithful", [True, False])
@pytest.mark.parametrize("angular", ["viennaps", "cosine"])
def test_belen_coverage(faithful, angular):
    from petch import ops
    from petch.chemistry import chem_params
    F = 4000
    m_i, m_F, m_O, cos_i, iy = _fluxes(F, 2)
    par = dict(PAR)
    par['ied_mode'] = 'gauss'
    flags = Flags(chemistry="belen", yield_angular=angular)
    bare = ops.belen_coverage(_t(m_i), _t(m_F), _t(m_O), _t(cos_i),
                              tuple(_t(a) for a in iy) if faithful else None,
                              chem_params(par, flags)).cpu().numpy()
    thF, thO = oc.belen_coverages(m_i, m_F, m_O, cos_i, par, flags, ion_yield=iy if faithful else None)
    np.testing.assert_allclose(bare, np.clip(1 - thF - thO, 0, 1), rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("model,faithful", [("belen", True), ("belen", False), ("langmuir", False)])
@pytest.mark.parametrize("hole", [True, False])
def test_surface_rate_and_mask(model, faithful, hole):
    from petch import ops
    from petch.chemistry import chem_params
    rng = np.random.default_rng(3)
    F = 6000
    m_i, m_F, m_O, cos_i, iy = _fluxes(F, 3)
    cen = np.stack([rng.uniform(0, 14, F), rng.uniform(0, 14, F), rng.uniform(15, 23, F)], 1)
    cen[:200, 2] = 18.0                                        # exactly on the band edges
    cen[200:400, 2] = 20.0
    cen = cen.astype(np.float32)
    par = dict(PAR)
    par['rate_scale'] = 0.025
    par['fnorm_E'] = 0.7                                       # surface_rate_belen divides by it
    flags = Flags(chemistry=model, yield_angular="viennaps")
    geo = dict(sub_top=18.0, Lx=14.0, Ly=14.0)
    mg = ops.MaskGeom(18.0, 2.0, 14.0, 14.0, 6.0, int(hole))
    V, red = ops.surface_rate(_t(m_i), _t(m_F), _t(m_O), _t(cos_i),
                              tuple(_t(a) for a in iy) if faithful else None, _t(cen), mg,
                              chem_params(par, flags))
    vmax, nbad = ops.read_rate_reduction(red)
    is_mask = oc.faces_in_mask(cen, geo, 2.0, 6.0, hole=hole)
    ref = oc.surface_rate(m_i, m_F, m_O, cos_i, is_mask, par, flags,
                          ion_yield=iy if faithful and model == "belen" else None)
    got = V.cpu().numpy()
    assert is_mask.sum() > 100
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-9)
    assert nbad == 0
    assert vmax == pytest.approx(float(np.float32(ref.max())))


def test_surface_rate_nonfinite_guard():
    import torch
    from petch import ops
    from petch.chemistry import chem_params
    F = 100
    m_i, m_F, m_O, cos_i, iy = _fluxes(F, 4)
    m_F[5] = np.nan
    m_O[7] = np.inf
    cen = np.zeros((F, 3), np.float32)
    mg = ops.MaskGeom(18.0, 2.0, 14.0, 14.0, 6.0, 1)
    flags = Flags(chemistry="belen", yield_angular="viennaps")
    V, red = ops.surface_rate(_t(m_i), _t(m_F), _t(m_O), _t(cos_i), tuple(_t(a) for a in iy),
                              _t(cen), mg, chem_params(dict(PAR), flags))
    v = V.cpu().numpy()
    vmax, nbad = ops.read_rate_reduction(red)
    assert np.isfinite(v).all() and v[5] == 0.0 and nbad >= 1
    assert vmax == pytest.approx(float(v.max()))


def test_batched_surface_rate_equals_single_calls():
    """surface_rate_batch over surfaces of different chemistries, masks and sizes (one empty,
    one with non-finite fluxes) gives each its single call's V and reduction pair, and copies
    each BVH's guard flag beside them."""
    import torch
    from petch import ops
    from petch.chemistry import chem_params
    rng = np.random.default_rng(11)
    cases = []
    for k, (model, faithful, hole, F) in enumerate([("belen", True, True, 6000),
                                                    ("belen", False, False, 300),
                                                    ("langmuir", False, True, 0),
                                                    ("langmuir", False, False, 2500),
                                                    ("belen", True, False, 777)]):
        m_i, m_F, m_O, cos_i, iy = _fluxes(F, 20 + k)
        if k == 4:
            m_F[5] = np.nan
            m_O[7] = np.inf
        cen = np.stack([rng.uniform(0, 14, F), rng.uniform(0, 14, F),
                        rng.uniform(15, 23, F)], 1).astype(np.float32)
        par = dict(PAR)
        par['rate_scale'] = 0.01 * (k + 1)
        par['fnorm_E'] = 0.7
        cases.append(dict(m_i=_t(m_i), m_F=_t(m_F), m_O=_t(m_O), cos_i=_t(cos_i),
                          ion_yield=tuple(_t(a) for a in iy) if faithful else None,
                          centroids=_t(cen), mask_geom=ops.MaskGeom(18.0, 2.0, 14.0, 14.0, 6.0,
                                                                    int(hole)),
                          chem=chem_params(par, Flags(chemistry=model, yield_angular="viennaps"))))
    singles = [ops.surface_rate(c['m_i'], c['m_F'], c['m_O'], c['cos_i'], c['ion_yield'],
                                c['centroids'], c['mask_geom'], c['chem'])
               if c['m_F'].numel() else (c['m_F'], torch.zeros(2, dtype=torch.int32))
               for c in cases]
    _, v, f = _mesh()
    bvh = ops.Bvh(_dev())
    bvh.build(_t(v), _t(f, torch.int32))
    for c in cases:
        c['V'] = torch.full_like(c['m_F'], -5.0)
    cases[1]['bvh'] = bvh
    red = torch.full((len(cases), 3), 77, dtype=torch.int32, device=_dev())
    ops.surface_rate_batch(cases, red)
    host = red.cpu().numpy()
    for k, (c, (V, r)) in enumerate(zip(cases, singles)):
        assert torch.equal(c['V'], V), k
        np.testing.assert_array_equal(host[k, :2], r.cpu().numpy())
    assert host[1, 2] == bvh.status() == 0
    assert (host[[0, 2, 3, 4], 2] == 77).all()            # no BVH: column left alone
    assert ops.decode_rate_reduction(host[4, :2])[1] >= 1
    assert ops.decode_rate_reduction(host[2, :2]) == (0.0, 0)


def test_extend_and_warm_start_nearest_face():
    import torch
    from petch import ops
    phi, v, f = _mesh()
    dx = 0.25
    tv, tf = _t(v), _t(f)
    bvh = ops.Bvh(_dev()).build(tv, tf)
    V = np.random.default_rng(5).uniform(0, 3, len(f)).astype(np.float32)
    band = 4 * dx
    Fs = ops.extend(bvh, _t(V), _t(phi), dx, band).cpu().numpy()
    M = ofl.OracleMesh(v, f)
    ref = ofl.extend(M, V, phi, dx, band)
    inb = np.abs(phi.astype(np.float64)) < band
    assert inb.sum() > 1000
    np.testing.assert_array_equal(Fs[~inb], 0.0)
    # equal, or the GPU picked another face at the same distance (a tie at a shared edge)
    ii, jj, kk = np.where(inb)
    pts = np.stack([ii * dx, jj * dx, kk * dx], 1).astype(np.float32)
    face_o, dist_o = M.bvh.closest_point(pts)
    gf, gd = bvh.query_point(_t(pts))
    gf, gd = gf.cpu().numpy(), gd.cpu().numpy()
    np.testing.assert_allclose(gd, dist_o, rtol=1e-4, atol=2e-6)
    np.testing.assert_array_equal(Fs[inb], V[gf])
    # a grid point whose closest surface point is a shared edge or vertex has several nearest
    # faces; both sides take the lowest face id among them, so the extensions agree except
    # where float and double rounding straddle the tie band, and there only between tied faces
    differ = Fs[inb] != ref[inb]
    assert differ.mean() < 1e-3, differ.mean()
    _assert_tie_only(v, f, pts, differ, gf, dist_o)
    # warm start: nearest face of the previous mesh for each new centroid
    cen = om.mesh_attrs(v, f)[0].astype(np.float32) + 0.03
    prev_bare = np.random.default_rng(6).uniform(0, 1, len(f)).astype(np.float32)
    ws = ops.nearest_face_value(bvh, _t(prev_bare), _t(cen), 1.0).cpu().numpy()
    fo, do = M.bvh.closest_point(cen)
    gf2, _ = bvh.query_point(_t(cen))
    gf2 = gf2.cpu().numpy()
    np.testing.assert_array_equal(ws, prev_bare[gf2])
    assert (gf2 != fo).mean() < 1e-3
    _assert_tie_only(v, f, cen, ws != prev_bare[fo], gf2, do)


def _assert_tie_only(v, f, pts, differ, gpu_face, dist_min):
    idx = np.flatnonzero(differ)
    print(f"nearest-face ties resolved differently: {len(idx)} of {len(differ)}", flush=True)
    for i in idx[:2000]:
        _, d = on.closest_point_brute(v, f[gpu_face[i]:gpu_face[i] + 1], pts[i:i + 1])
        assert abs(d[0] - dist_min[i]) <= 1e-4 * max(dist_min[i], 1.0), (i, d[0], dist_min[i])


def _ff_case(name):
    """Form-factor rays of a hole or a trench surface: (hitface, F, n_ff, row_ptr, col, D)."""
    from petch import ops
    if name == "hole":
        phi = ol.init_sdf(8.0, 8.0, 14.0, 0.25, 3.0, 8.0, 4.0, True).astype(np.float32)
        dx, periodic = 0.25, 0
        Lx, Ly, Lz = 8.0, 8.0, 14.0
    else:
        phi = ol.init_sdf(6.0, 1.5, 12.0, 0.15, 1.5, 7.0, 3.0, False).astype(np.float32)
        dx, periodic = 0.15, 1
        Lx, Ly, Lz = 6.0, 1.5, 12.0
    v, f = on.marching_cubes(phi, dx)
    tv, tf = _t(v), _t(f)
    bvh = ops.Bvh(_dev()).build(tv, tf)
    cen, area, nrm = ops.mesh_attrs(tv, tf)
    gn = ops.gas_normals(_t(phi), dx, cen, nrm)
    box = ops.make_box(Lx, Ly, Lz, Lz - dx, periodic)
    n_ff = 64
    hit = ops.trace_ff(bvh, box, cen, gn, n_ff, 17)
    F = len(f)
    row_ptr, col, D = ops.radiosity_assemble(hit, F, n_ff)
    return hit, F, n_ff, row_ptr, col, D


@pytest.mark.parametrize("name", ["hole", "trench"])
def test_radiosity_solve_fp64(name):
    """Same form factors on both sides: the GPU float64 Jacobi equals scipy's to 1e-5 relative,
    and the converged solve equals a sparse direct solve (the FP64 oracle)."""
    from petch import ops
    hit, F, n_ff, row_ptr, col, D = _ff_case(name)
    Dh, Aff = ofl.radiosity_matrix(hit.cpu().numpy(), F, n_ff)
    np.testing.assert_array_equal(D.cpu().numpy()[:F], Dh)
    bare = np.random.default_rng(7).uniform(0.2, 1.0, F).astype(np.float32)
    for beta in (0.7, 1.0):
        m, G, it = ops.radiosity_solve(row_ptr, col, D, F, n_ff, _t(bare), beta, 40)
        mref, Gref = ofl.radiosity_solve(Dh, Aff, bare.astype(np.float64), beta, 40)
        np.testing.assert_allclose(G.cpu().numpy(), Gref, rtol=1e-5, atol=1e-12)
        np.testing.assert_allclose(m.cpu().numpy(), mref, rtol=1e-5, atol=1e-7)
        m, G, it = ops.radiosity_solve(row_ptr, col, D, F, n_ff, _t(bare), beta, 0, 1e-12)
        direct = ofl.radiosity_direct(Dh, Aff, bare.astype(np.float64), beta)
        np.testing.assert_allclose(G.cpu().numpy(), direct, rtol=1e-5, atol=1e-10)
        assert it > 0


@pytest.mark.parametrize("name", ["hole", "trench"])
def test_radiosity_adjoint_matches_oracle(name):
    """The transposed solve on the GPU (gathering over hitface) against the oracle's adjoint on
    the scipy matrix of the same rays: 40 sweeps against 40 transposed sweeps, and the converged
    adjoint (through the autograd Function) against the direct sparse solve, to 1e-5 of the
    largest gradient. Some faces get bare*beta > 1 (no gradient through the clip)."""
    import torch
    from petch import adjoint, ops
    hit, F, n_ff, row_ptr, col, D = _ff_case(name)
    Dh, Aff = ofl.radiosity_matrix(hit.cpu().numpy(), F, n_ff)
    rng = np.random.default_rng(11)
    bare = rng.uniform(0.2, 1.0, F).astype(np.float32)
    bare[::17] = 1.5
    w = rng.normal(size=F).astype(np.float32)
    b64 = bare.astype(np.float64)

    def close(a, b):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-5 * max(np.abs(b).max(), 1e-30))

    for beta in (0.7, 1.0):
        m, G, _ = ops.radiosity_solve(row_ptr, col, D, F, n_ff, _t(bare), beta, 40)
        gD, gb, it = ops.radiosity_adjoint(hit, F, n_ff, G, _t(bare), beta, _t(w), 40)
        assert it == 40
        rD, rb = ofl.radiosity_adjoint(Dh, Aff, G.cpu().numpy(), b64, beta, w.astype(np.float64),
                                       n_iter=40)
        close(gD.cpu().numpy(), rD)
        close(gb.cpu().numpy(), rb)
        assert np.all(gb.cpu().numpy()[bare * beta > 1.0] == 0.0)
        x = _t(bare).requires_grad_(True)
        d = D[:F].clone().requires_grad_(True)
        out = adjoint.radiosity(x, d, hit, row_ptr, col, n_ff, beta, 0, 1e-12)
        (out * _t(w)).sum().backward()
        Gd = ofl.radiosity_direct(Dh, Aff, b64, beta)
        rD, rb = ofl.radiosity_adjoint(Dh, Aff, Gd, b64, beta, w.astype(np.float64))
        close(d.grad.cpu().numpy(), rD)
        close(x.grad.cpu().numpy(), rb)


def test_batched_extension_and_warm_start_equal_single_calls():
    """extend_batch / nearest_face_value_batch over grids of different shapes and surfaces give
    each grid exactly what a single call gives it (the same queries, one launch)."""
    import torch
    from petch import ops
    grids = [(ol.init_sdf(8.0, 8.0, 14.0, 0.25, 3.0, 8.0, 4.0, True), 0.25),
             (ol.init_sdf(6.0, 3.0, 10.0, 0.2, 2.0, 1.0, 6.0, False), 0.2),
             (ol.init_sdf(5.0, 5.0, 8.0, 0.25, 2.0, 1.0, 5.0, True), 0.25)]
    items, singles, ws_items, ws_singles = [], [], [], []
    rng = np.random.default_rng(8)
    for phi, dx in grids:
        phi = phi.astype(np.float32)
        v, f = on.marching_cubes(phi, dx)
        bvh = ops.Bvh(_dev()).build(_t(v), _t(f))
        V = _t(rng.uniform(0, 3, len(f)).astype(np.float32))
        tphi = _t(phi)
        singles.append(ops.extend(bvh, V, tphi, dx, 4 * dx).cpu().numpy())
        out = torch.full_like(tphi, -7.0)
        items.append((bvh, V, tphi, dx, 4 * dx, out))
        pts = _t((om.mesh_attrs(v, f)[0] + 0.03).astype(np.float32))
        src = _t(rng.uniform(0, 1, len(f)).astype(np.float32))
        ws_singles.append(ops.nearest_face_value(bvh, src, pts, 1.0).cpu().numpy())
        ws_items.append((bvh, src, pts))
    ops.extend_batch(items)
    for (_, _, _, _, _, out), ref in zip(items, singles):
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
    outs = ops.nearest_face_value_batch(ws_items, 1.0)
    for out, ref in zip(outs, ws_singles):
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
    assert ops.nearest_face_value_batch([], 1.0) == []


@pytest.mark.parametrize("yields", [False, True])
def test_belen_adjoint_matches_oracle(yields):
    """petch_belen_coverage_adjoint (closed form, float64 inside) against the oracle's
    complex-step derivative of the Belen bare update on the same float32 inputs."""
    from petch import ops
    from petch.chemistry import chem_params
    rng = np.random.default_rng(21)
    F = 5000
    par = dict(PAR)
    fl = Flags(chemistry="belen", ion_reflection=True)
    m_i = rng.uniform(0.05, 1.3, F).astype(np.float32)
    cos_i = rng.uniform(0.0, 1.0, F).astype(np.float32)
    m_F = rng.uniform(0.01, 1.5, F).astype(np.float32)
    m_O = rng.uniform(0.01, 1.5, F).astype(np.float32)
    iy = tuple(rng.uniform(0.0, 0.5, F).astype(np.float32) for _ in range(3)) if yields else None
    g = rng.normal(size=F).astype(np.float32)
    out = ops.belen_coverage_adjoint(_t(m_i), _t(m_F), _t(m_O), _t(cos_i),
                                     tuple(_t(a) for a in iy) if iy else None,
                                     chem_params(par, fl), _t(g))
    ref = oc.belen_bare_adjoint(m_i.astype(np.float64), m_F.astype(np.float64),
                                m_O.astype(np.float64), cos_i.astype(np.float64), par, fl,
                                g.astype(np.float64),
                                ion_yield=tuple(a.astype(np.float64) for a in iy) if iy else None)
    names = ("m_F", "m_O", "m_i", "ie", "ip")
    for name, got in zip(names, out):
        if name not in ref:
            assert got is None
            continue
        r = ref[name]
        np.testing.assert_allclose(got.cpu().numpy(), r, rtol=1e-5,
                                   atol=1e-6 * np.abs(r).max(), err_msg=name)


def test_coverage_fixed_point_adjoint_matches_oracle():
    """adjoint.radiosity_coverage (three rounds of converged F and O solves and Belen updates on
    trench form factors, backward through the HIP adjoints) against the oracle's unrolled
    float64 adjoint of the same loop: gradients of w . bare in m_i and D to 1e-4 of the largest."""
    import torch
    from petch import adjoint
    from petch.chemistry import chem_params
    hit, F, n_ff, row_ptr, col, D = _ff_case("trench")
    Dh, Aff = ofl.radiosity_matrix(hit.cpu().numpy(), F, n_ff)
    par = dict(PAR)
    fl = Flags(chemistry="belen", ion_reflection=True)
    rng = np.random.default_rng(5)
    m_i = rng.uniform(0.05, 1.2, F).astype(np.float32)
    cos_i = rng.uniform(0.0, 1.0, F).astype(np.float32)
    w = rng.normal(size=F).astype(np.float32)
    mi = _t(m_i).requires_grad_(True)
    d = D[:F].clone().requires_grad_(True)
    bare, _, _ = adjoint.radiosity_coverage(mi, _t(cos_i), None, d, hit, row_ptr, col, n_ff,
                                            chem_params(par, fl), par.get('betaE', 0.7),
                                            par.get('betaO', 1.0), n_fp=3, tol=1e-12)
    (bare * _t(w)).sum().backward()
    rb, _, _, grads = ofl.radiosity_coverage(Dh, Aff, m_i.astype(np.float64),
                                             cos_i.astype(np.float64), par, fl, n_fp=3,
                                             w=w.astype(np.float64))
    np.testing.assert_allclose(bare.detach().cpu().numpy(), rb, rtol=0, atol=1e-5)
    for got, r in ((mi.grad, grads["m_i"]), (d.grad, grads["D"])):
        np.testing.assert_allclose(got.cpu().numpy(), r, rtol=0, atol=1e-4 * np.abs(r).max())


@pytest.mark.parametrize("chem,yields", [("belen", False), ("belen", True), ("langmuir", False),
                                         ("langmuir", True)])
def test_surface_rate_adjoint_matches_oracle(chem, yields):
    """petch_surface_rate_adjoint (closed form) against the oracle's complex-step derivative of
    the same rate, with mask faces (no gradient); through the autograd Function as well. The
    Langmuir rate ignores ion yields, so with yields its gradient goes to m_i and the yields get
    zeros; the Belen rate with yields ignores m_i, whose gradient is then None."""
    import torch
    from petch import adjoint, ops
    from petch.chemistry import chem_params
    rng = np.random.default_rng(31)
    F = 4000
    par = dict(PAR)
    fl = Flags(chemistry=chem, ion_reflection=True)
    m_i = rng.uniform(0.05, 1.3, F).astype(np.float32)
    cos_i = rng.uniform(0.0, 1.0, F).astype(np.float32)
    m_F = rng.uniform(0.01, 1.5, F).astype(np.float32)
    m_O = rng.uniform(0.01, 1.5, F).astype(np.float32)
    msk = rng.random(F) < 0.1
    iy = tuple(rng.uniform(0.0, 0.5, F).astype(np.float32) for _ in range(3)) if yields else None
    g = rng.normal(size=F).astype(np.float32)
    cp = chem_params(par, fl)
    tiy = tuple(_t(a) for a in iy) if iy else None
    tm = _t(msk.astype(np.uint8))
    out = ops.surface_rate_adjoint(_t(m_i), _t(m_F), _t(m_O), _t(cos_i), tiy, None,
                                   ops.MaskGeom(), cp, _t(g), is_mask=tm)
    ref = oc.surface_rate_adjoint(m_i.astype(np.float64), m_F.astype(np.float64),
                                  m_O.astype(np.float64), cos_i.astype(np.float64), msk, par, fl,
                                  g.astype(np.float64),
                                  ion_yield=tuple(a.astype(np.float64) for a in iy) if iy else None)
    for name, got in zip(("m_F", "m_O", "m_i", "sp", "ie", "ip"), out):
        if got is None:
            assert name not in ref or not np.any(ref[name]), name
            continue
        r = ref[name]
        np.testing.assert_allclose(got.cpu().numpy(), r, rtol=1e-5, atol=1e-6 * np.abs(r).max(),
                                   err_msg=name)
        assert np.all(got.cpu().numpy()[msk] == 0.0)
    xF = _t(m_F).requires_grad_(True)
    V = adjoint.surface_rate(_t(m_i), xF, _t(m_O), _t(cos_i), tiy, None, ops.MaskGeom(), cp,
                             is_mask=tm)
    (V * _t(g)).sum().backward()
    np.testing.assert_allclose(xF.grad.cpu().numpy(), out[0].cpu().numpy(), rtol=0, atol=0)


def test_extend_adjoint_is_the_transpose():
    """petch_extend_adjoint sends each band node's gradient to its nearest face: it equals the
    scatter over the GPU's own nearest faces, and <gFs, extend(V)> = <extend_adjoint(gFs), V>
    (the extension is linear in V). Through the autograd Function as well."""
    import torch
    from petch import adjoint, ops
    phi, v, f = _mesh()
    dx = 0.25
    bvh = ops.Bvh(_dev()).build(_t(v), _t(f))
    band = 4 * dx
    rng = np.random.default_rng(12)
    gFs = rng.normal(size=phi.shape).astype(np.float32)
    gV = ops.extend_adjoint(bvh, _t(phi), dx, band, _t(gFs)).cpu().numpy().astype(np.float64)
    inb = np.abs(phi.astype(np.float64)) < band
    ii, jj, kk = np.where(inb)
    pts = np.stack([ii * dx, jj * dx, kk * dx], 1).astype(np.float32)
    gf, _ = bvh.query_point(_t(pts))
    ref = np.bincount(gf.cpu().numpy(), weights=gFs[inb].astype(np.float64), minlength=len(f))
    np.testing.assert_allclose(gV, ref, rtol=1e-5, atol=1e-4)
    V = rng.uniform(0.0, 3.0, len(f)).astype(np.float32)
    Fs = ops.extend(bvh, _t(V), _t(phi), dx, band).cpu().numpy().astype(np.float64)
    lhs, rhs = float((gFs * Fs).sum()), float(gV @ V)
    assert abs(lhs - rhs) <= 1e-5 * (np.abs(gFs * Fs).sum())
    x = _t(V).requires_grad_(True)
    (adjoint.extend(bvh, x, _t(phi), dx, band) * _t(gFs)).sum().backward()
    np.testing.assert_allclose(x.grad.cpu().numpy(), gV, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("chem", ["belen", "langmuir"])
def test_rate_gradient_in_ion_energy(chem):
    """The recipe gradient of scripts/inverse_design.py:20-45 on this path: the yields from the
    ion energy E ('mean' IED, Y = A max(sqrt(E) - sqrt(Eth), 0)) in torch, the rate on the GPU,
    d loss / dE through the rate adjoint's yield-coefficient output, against float64 central
    differences of the oracle's rate in par['Emean']."""
    import torch
    from petch import adjoint, ops
    from petch.chemistry import chem_params
    rng = np.random.default_rng(41)
    F = 3000
    par = dict(PAR)
    par['ied_mode'] = 'mean'
    par['Emean'] = 120.0
    fl = Flags(chemistry=chem)
    m_i = rng.uniform(0.05, 1.3, F).astype(np.float32)
    cos_i = rng.uniform(0.0, 1.0, F).astype(np.float32)
    m_F = rng.uniform(0.01, 1.5, F).astype(np.float32)
    m_O = rng.uniform(0.01, 1.5, F).astype(np.float32)
    msk = rng.random(F) < 0.1
    w = rng.normal(size=F)
    dev = _dev()
    E = torch.tensor(par['Emean'], dtype=torch.float64, requires_grad=True)
    A = torch.tensor([par['A_ie'], par['A_sp'], par['A_p']], dtype=torch.float64)
    Eth = torch.tensor([par['Eth_ie'], par['Eth_sp'], par['Eth_p']], dtype=torch.float64)
    Y = A * torch.relu(torch.sqrt(E) - torch.sqrt(Eth))
    V = adjoint.surface_rate(_t(m_i), _t(m_F), _t(m_O), _t(cos_i), None, None, ops.MaskGeom(),
                             chem_params(par, fl), is_mask=_t(msk.astype(np.uint8)), Y=Y)
    (V.double() * torch.from_numpy(w).to(dev)).sum().backward()

    def loss(Em):
        p = dict(par)
        p['Emean'] = Em
        return float(w @ oc.surface_rate(m_i.astype(np.float64), m_F.astype(np.float64),
                                         m_O.astype(np.float64), cos_i.astype(np.float64), msk,
                                         p, fl))

    h = 1e-3
    fd = (loss(par['Emean'] + h) - loss(par['Emean'] - h)) / (2 * h)
    print(f"{chem}: d loss / dE fd {fd:.6e} adjoint {float(E.grad):.6e}", flush=True)
    assert abs(float(E.grad) - fd) <= 1e-4 * abs(fd)


def test_batched_round_finish_equals_the_separate_calls():
    """petch_finish_round_batch (one launch for the end of a neutral round of many features:
    normalise and smooth both species, Belen update) gives exactly what petch_normalize,
    petch_smooth and petch_belen_coverage give feature by feature: three surfaces of different
    sizes, with and without ion yields, one smoothing pass blended or not, and none."""
    import torch
    from petch import ops
    from petch.chemistry import chem_params
    meshes = []
    for k, (w, d) in enumerate(((3.0, 8.0), (2.0, 6.0), (4.5, 9.0))):
        phi = ol.init_sdf(8.0, 8.0, 14.0, 0.25, w, d, 4.0, k != 1).astype(np.float32)
        v, f = on.marching_cubes(phi, 0.25)
        tv, tf = _t(v), _t(f)
        cen, area, nrm = ops.mesh_attrs(tv, tf)
        meshes.append((len(f), area, nrm, ops.face_adjacency(tf, len(v))))
    rng = np.random.default_rng(11)
    cases = [(True, 1, 1.0), (False, 1, 0.6), (True, 0, 1.0)]
    items, refs = [], []
    for k, ((F, area, nrm, nbr), (faithful, n_smooth, alpha)) in enumerate(zip(meshes, cases)):
        m_i, _, _, cos_i, iy = _fluxes(F, 20 + k)
        flF = _t(rng.exponential(2.0, F).astype(np.float32))
        flO = _t(rng.exponential(1.0, F).astype(np.float32))
        par = dict(PAR)
        par['Fflux'] *= 1.0 + 0.2 * k
        chem = chem_params(par, Flags(chemistry="belen"))
        ion_yield = tuple(_t(a) for a in iy) if faithful else None
        scale = 0.05 + 0.01 * k
        mF = ops.normalize(flF, area, scale, 0.0, 8.0)
        mO = ops.normalize(flO, area, scale, 0.0, 8.0)
        if n_smooth:
            mF = ops.smooth(mF, nbr, nrm, n_smooth, alpha, True)
            mO = ops.smooth(mO, nbr, nrm, n_smooth, alpha, True)
        bare = ops.belen_coverage(_t(m_i), mF, mO, _t(cos_i), ion_yield, chem)
        refs.append((mF, mO, bare))
        items.append(dict(flF=flF, flO=flO, areas=area, nbr=nbr, normals=nrm, scale=scale,
                          hi=8.0, n_smooth=n_smooth, alpha=alpha, m_i=_t(m_i), cos_i=_t(cos_i),
                          ion_yield=ion_yield, chem=chem, m_F=torch.empty_like(flF),
                          m_O=torch.empty_like(flF), bare=torch.empty_like(flF)))
    ops.finish_round_batch(items)
    for it, (mF, mO, bare) in zip(items, refs):
        assert torch.equal(it['m_F'], mF) and torch.equal(it['m_O'], mO)
        assert torch.equal(it['bare'], bare)


def test_batched_ion_finish_equals_the_separate_calls():
    """petch_finish_ion_batch (the end of the faithful-ion trace of many features in one launch)
    gives exactly what petch_normalize (no clip), petch_smooth (one pass, host semantics) and
    petch_cos_ratio give feature by feature, smoothed or not, blended or not."""
    import torch
    from petch import ops
    rng = np.random.default_rng(12)
    items, refs = [], []
    for k, (w, d, n_smooth, alpha) in enumerate(((3.0, 8.0, 1, 1.0), (2.0, 6.0, 1, 0.6),
                                                 (4.5, 9.0, 0, 1.0))):
        phi = ol.init_sdf(8.0, 8.0, 14.0, 0.25, w, d, 4.0, k != 1).astype(np.float32)
        v, f = on.marching_cubes(phi, 0.25)
        tv, tf = _t(v), _t(f)
        cen, area, nrm = ops.mesh_attrs(tv, tf)
        nbr = ops.face_adjacency(tf, len(v))
        F = len(f)
        acc = rng.exponential(1.0, (5, F)).astype(np.float32)
        acc[3, :40] = 0.0                                     # faces no ion reached
        acc = _t(acc)
        scale = 0.03 + 0.01 * k
        ref = [ops.normalize(acc[j], area, scale, -ops.INF, ops.INF) for j in range(4)]
        if n_smooth:
            ref = [ops.smooth(x, nbr, nrm, 1, alpha, False) for x in ref]
        ref.append(ops.cos_ratio(acc[3], acc[4]))
        refs.append(ref)
        items.append(dict(acc=acc, areas=area, nbr=nbr, normals=nrm, scale=scale,
                          n_smooth=n_smooth, alpha=alpha, out=torch.empty_like(acc)))
    ops.finish_ion_batch(items)
    for it, ref in zip(items, refs):
        for j in range(5):
            assert torch.equal(it['out'][j], ref[j]), j
