"""GPU marching cubes + mesh attributes + adjacency against the oracle, through the C ABI.
Bar: face topology bit-exact, vertices within 1e-5*dx (float64 interpolation rounded once on
both sides, so they are expected to be identical), attributes to float32 rounding."""
import numpy as np
import pytest

from oracle import levelset as ol
from oracle import mesh as om
from oracle import native as on

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda", 0)


def _fields():
    from scipy.ndimage import gaussian_filter
    out = []
    x = np.arange(40) * 0.1
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    out.append(("sphere", (1.37 - np.sqrt((X - 1.95) ** 2 + (Y - 1.95) ** 2 + (Z - 1.95) ** 2)), 0.1))
    out.append(("hole56", ol.init_sdf(14.0, 14.0, 24.0, 0.25, 6.0, 2.0, 18.0, True), 0.25))
    out.append(("trench160", ol.init_sdf(16.0, 4.0, 26.0, 0.1, 2.0, 2.0, 20.0, False), 0.1))
    for s in range(3):
        rng = np.random.default_rng(s)
        f = gaussian_filter(rng.normal(size=(33, 27, 45)), 2.5)
        out.append((f"smooth{s}", f / f.std(), 0.2))
    for s in range(3):
        rng = np.random.default_rng(100 + s)
        f = rng.normal(size=(9, 7, 11))
        if s == 2:
            f[rng.random(f.shape) < 0.1] = 0.0
        out.append((f"noise{s}", f, 0.5))
    out.append(("flat", np.ones((5, 5, 5)), 0.5))               # no surface
    out.append(("thin", np.random.default_rng(5).normal(size=(1, 6, 6)), 0.5))  # no cells
    return out


FIELDS = _fields()


@pytest.mark.parametrize("name,phi,dx", FIELDS, ids=[f[0] for f in FIELDS])
def test_marching_cubes_matches_oracle(name, phi, dx):
    import torch
    from petch import ops
    phi = np.ascontiguousarray(phi, np.float32)
    print(f"mc {name} {phi.shape}", flush=True)
    bufs = ops.MeshBuffers(_dev())                       # start empty: exercises the capacity path
    v, f = ops.marching_cubes(torch.from_numpy(phi).to(_dev()), dx, bufs)
    torch.cuda.synchronize()
    rv, rf = on.marching_cubes(phi, dx)
    assert f.shape[0] == len(rf) and v.shape[0] == len(rv)
    np.testing.assert_array_equal(f.cpu().numpy(), rf)
    np.testing.assert_allclose(v.cpu().numpy(), rv, rtol=0, atol=1e-5 * dx)
    # second call reuses the buffers
    v2, f2 = ops.marching_cubes(torch.from_numpy(phi).to(_dev()), dx, bufs)
    np.testing.assert_array_equal(f2.cpu().numpy(), rf)


@pytest.mark.parametrize("name,phi,dx", FIELDS[:6], ids=[f[0] for f in FIELDS[:6]])
def test_mesh_attrs_adjacency_gas_normals(name, phi, dx):
    import torch
    from petch import ops
    phi = np.ascontiguousarray(phi, np.float32)
    tphi = torch.from_numpy(phi).to(_dev())
    v, f = ops.marching_cubes(tphi, dx)
    cen, area, nrm = ops.mesh_attrs(v, f)
    nbr = ops.face_adjacency(f, v.shape[0])
    gn = ops.gas_normals(tphi, dx, cen, nrm)
    torch.cuda.synchronize()
    vn, fn_ = v.cpu().numpy(), f.cpu().numpy()
    rc, ra, rcross = om.mesh_attrs(vn, fn_)
    np.testing.assert_allclose(cen.cpu().numpy(), rc, rtol=0, atol=2e-6 * max(phi.shape) * dx)
    np.testing.assert_allclose(area.cpu().numpy(), ra, rtol=2e-5, atol=1e-7 * dx * dx)
    ok = ra > 1e-6 * dx * dx
    np.testing.assert_allclose(nrm.cpu().numpy()[ok], om.face_normals(vn, fn_)[ok], atol=1e-5)
    # neighbour table == the pair set of the reference's sort-based _edge_adjacency
    np.testing.assert_array_equal(om.pairs_from_nbr(nbr.cpu().numpy()),
                                  om.canonical_pairs(om.edge_adjacency(fn_)))
    rg = om.gas_normals(vn, fn_, rc, phi, dx)
    np.testing.assert_allclose(gn.cpu().numpy()[ok], rg[ok], atol=1e-4)


def test_batched_marching_cubes_equals_single_calls():
    """marching_cubes_batch over grids of different shapes (one host read for all counts) gives
    each grid exactly the single call's surface, and grows undersized buffers."""
    import torch
    from petch import ops
    d = torch.device("cuda", 0)
    grids = [(ol.init_sdf(8.0, 8.0, 14.0, 0.25, 3.0, 8.0, 4.0, True), 0.25),
             (ol.init_sdf(6.0, 3.0, 10.0, 0.2, 2.0, 1.0, 6.0, False), 0.2),
             (np.ones((4, 4, 4), np.float64), 0.5),                   # no surface
             (np.random.default_rng(5).normal(size=(1, 6, 6)), 0.5),  # no cells
             (FIELDS[3][1], FIELDS[3][2]), (FIELDS[7][1], FIELDS[7][2])]
    items, singles = [], []
    for phi, dx in grids:
        tp = torch.from_numpy(phi.astype(np.float32)).to(d)
        v, f = ops.marching_cubes(tp, dx)
        singles.append((v.cpu().numpy(), f.cpu().numpy()))
        items.append((tp, dx, ops.MeshBuffers(d)))                    # empty: forces a retry
    outs = ops.marching_cubes_batch(items)
    for (v, f), (rv, rf) in zip(outs, singles):
        np.testing.assert_array_equal(v.cpu().numpy(), rv)
        np.testing.assert_array_equal(f.cpu().numpy(), rf)
    assert len(singles[2][1]) == 0 and len(outs[2][1]) == 0
    assert len(singles[3][1]) == 0 and len(outs[3][1]) == 0
    # again with the buffers now large enough (no retry), in another order
    outs = ops.marching_cubes_batch(items[::-1])
    for (v, f), (rv, rf) in zip(outs, singles[::-1]):
        np.testing.assert_array_equal(v.cpu().numpy(), rv)
        np.testing.assert_array_equal(f.cpu().numpy(), rf)


def test_batched_surface_build_equals_single_calls():
    """surface_build_batch over meshes of different sizes (closed, open and empty) fills each
    mesh's centroids, areas, normals and neighbour table exactly as mesh_attrs and
    face_adjacency do, and a second batch reuses the workspace."""
    import torch
    from petch import ops
    d = _dev()
    meshes = []
    for name, phi, dx in FIELDS:
        v, f = ops.marching_cubes(torch.from_numpy(np.ascontiguousarray(phi, np.float32)).to(d), dx)
        meshes.append((name, v, f))
    print("surface batch", [(n, int(f.shape[0])) for n, _, f in meshes], flush=True)
    # a mesh with a non-finite vertex: its faces' attributes are NaN in both paths, and its
    # finite flag is 0
    v_bad = meshes[0][1].clone()
    v_bad[int(meshes[0][2][3, 1])] = float("nan")
    meshes.append(("nan_vertex", v_bad, meshes[0][2]))
    for order in (meshes, meshes[::-1][:4]):
        items, singles = [], []
        for _, v, f in order:
            nf = int(f.shape[0])
            singles.append(ops.mesh_attrs(v, f) + (ops.face_adjacency(f, v.shape[0]),))
            outs = (torch.full((nf, 3), -7.0, device=d), torch.full((nf,), -7.0, device=d),
                    torch.full((nf, 3), -7.0, device=d),
                    torch.full((nf, 3), -7, dtype=torch.int32, device=d),
                    torch.full((), 5, dtype=torch.int32, device=d))
            items.append((v, f) + outs)
        ops.surface_build_batch(items)
        torch.cuda.synchronize()
        for (name, _, f), it, ref in zip(order, items, singles):
            for got, want in zip(it[2:5], ref[:3]):
                torch.testing.assert_close(got, want, rtol=0, atol=0, equal_nan=True, msg=name)
            assert torch.equal(it[5], ref[3]), name
            if int(f.shape[0]):
                assert int(it[6]) == int(torch.isfinite(ref[0]).all()), name
            else:
                assert int(it[6]) == 5, name                  # no faces: flag untouched
    assert any(int(f.shape[0]) == 0 for _, _, f in meshes)
    assert int(items[0][6]) == 0                              # the NaN-vertex mesh, first now
