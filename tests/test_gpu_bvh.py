"""LBVH build + closest-hit / closest-point queries against brute force over all faces."""
import numpy as np
import pytest

from oracle import levelset as ol
from oracle import native as on

pytestmark = pytest.mark.gpu


def _dev():
    import torch
    return torch.device("cuda", 0)


def _mesh(name):
    if name == "sphere":
        x = np.arange(40) * 0.1
        X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
        phi = 1.37 - np.sqrt((X - 1.95) ** 2 + (Y - 1.95) ** 2 + (Z - 1.95) ** 2)
        return on.marching_cubes(phi.astype(np.float32), 0.1) + ((0, 3.9),)
    if name == "hole56":
        phi = ol.init_sdf(14.0, 14.0, 24.0, 0.25, 6.0, 2.0, 18.0, True)
        return on.marching_cubes(phi.astype(np.float32), 0.25) + ((0, 14.0),)
    phi = ol.init_sdf(16.0, 4.0, 26.0, 0.1, 2.0, 2.0, 20.0, False)
    return on.marching_cubes(phi.astype(np.float32), 0.1) + ((0, 16.0),)


@pytest.mark.parametrize("name,quality", [("sphere", -1), ("hole56", -1), ("trench160", -1),
                                          ("hole56", 0), ("hole56", 1), ("hole56", 2),
                                          ("sphere", 2), ("trench160", 0)])
def test_ray_queries_match_brute_force(name, quality):
    """Every build (Morton or clustered tree, fixed or area-driven 4-wide collapse) answers
    closest-hit queries like brute force over all faces."""
    import torch
    from petch import ops
    v, f, (lo, hi) = _mesh(name)
    print(f"bvh {name} nf={len(f)}", flush=True)
    rng = np.random.default_rng(3)
    n = 4000
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[: n // 4] = [0, 0, -1]                                  # axis-aligned rays too
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    dev = _dev()
    b = ops.Bvh(dev).set_quality(quality).build(torch.from_numpy(v).to(dev),
                                                 torch.from_numpy(f).to(dev))
    face, t = b.query_ray(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev))
    torch.cuda.synchronize()
    assert b.status() == 0
    face, t = face.cpu().numpy(), t.cpu().numpy()
    rf, rt = on.closest_hit_brute(v, f, o, d)
    hit, rhit = face >= 0, rf >= 0
    # hit/miss agree except rays grazing an edge (a handful at most)
    assert (hit != rhit).sum() <= max(2, n // 1000)
    both = hit & rhit
    np.testing.assert_allclose(t[both], rt[both], rtol=2e-4, atol=2e-5)
    same = (face == rf)[both]
    assert same.mean() > 0.995                                # the rest are ties at shared edges
    diff = both & (face != rf)
    np.testing.assert_allclose(t[diff], rt[diff], rtol=1e-4, atol=1e-5)
    lohi = b.root_bounds()
    np.testing.assert_allclose(lohi[:3], v.min(0), atol=1e-6)
    np.testing.assert_allclose(lohi[3:], v.max(0), atol=1e-6)


@pytest.mark.parametrize("name", ["sphere", "hole56", "trench160"])
def test_point_queries_match_brute_force(name):
    import torch
    from petch import ops
    v, f, (lo, hi) = _mesh(name)
    rng = np.random.default_rng(4)
    n = 3000
    base = v[rng.integers(0, len(v), n)]
    p = (base + rng.normal(scale=0.3, size=(n, 3))).astype(np.float32)
    dev = _dev()
    b = ops.Bvh(dev).build(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev))
    face, dist = b.query_point(torch.from_numpy(p).to(dev))
    torch.cuda.synchronize()
    face, dist = face.cpu().numpy(), dist.cpu().numpy()
    rf, rd = on.closest_point_brute(v, f, p)
    assert (face >= 0).all()
    np.testing.assert_allclose(dist, rd, rtol=1e-4, atol=2e-6)
    # both sides send nearest-face ties (shared edges and vertices) to the lowest face id; they
    # differ only where float and double rounding put a face on opposite sides of the tie band,
    # which for random points happens beside almost coplanar neighbours (a few per thousand)
    assert (face == rf).mean() > 0.99


def test_tiny_meshes():
    import torch
    from petch import ops
    dev = _dev()
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]], np.float32)
    f1 = np.array([[0, 1, 2]], np.int32)
    f2 = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    o = np.array([[0.2, 0.2, 1.0], [0.8, 0.8, 1.0], [5, 5, 1]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, -1]], np.float32)
    tv = torch.from_numpy(v).to(dev)
    b = ops.Bvh(dev)
    b.build(tv, torch.zeros((0, 3), dtype=torch.int32, device=dev))
    face, t = b.query_ray(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev))
    assert (face.cpu().numpy() == -1).all()
    b.build(tv, torch.from_numpy(f1).to(dev))
    face, t = b.query_ray(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev))
    assert face.cpu().tolist() == [0, -1, -1]
    assert t.cpu().numpy()[0] == pytest.approx(1.0)
    b.build(tv, torch.from_numpy(f2).to(dev))
    face, t = b.query_ray(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev))
    assert face.cpu().tolist() == [0, 1, -1]
    pf, pd = b.query_point(torch.tensor([[0.1, 0.1, 0.5]], device=dev))
    assert pf.cpu().tolist() == [0] and pd.cpu().numpy()[0] == pytest.approx(0.5)


def test_batched_build_equals_single_builds():
    """petch_bvh_build_batch (parallel stages per tree, the refits of all trees in one launch)
    gives trees that answer ray and point queries exactly as single builds do, with the same
    root bounds, including a tree of one face, an empty mesh and a clustered build."""
    import torch
    from petch import ops
    dev = _dev()
    meshes = [_mesh("sphere")[:2], _mesh("hole56")[:2], _mesh("trench160")[:2]]
    one = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32))
    empty = (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int32))
    meshes += [one, empty]
    tv = [(torch.from_numpy(v).to(dev), torch.from_numpy(f).to(dev)) for v, f in meshes]
    single = [ops.Bvh(dev).build(v, f) for v, f in tv]
    batched = [ops.Bvh(dev) for _ in tv]
    batched[2].set_quality(2)
    single[2].set_quality(2).build(*tv[2])
    ops.bvh_build_batch([(b, v, f) for b, (v, f) in zip(batched, tv)])
    rng = np.random.default_rng(5)
    for (v, f), a, b in zip(tv, single, batched):
        assert a.status() == 0 and b.status() == 0
        if int(f.shape[0]) > 1:
            np.testing.assert_array_equal(a.root_bounds(), b.root_bounds())
        if int(f.shape[0]) == 0:
            continue
        lo = v.min(0).values.cpu().numpy()
        hi = v.max(0).values.cpu().numpy()
        o = rng.uniform(lo - 0.5, hi + 0.5, (3000, 3)).astype(np.float32)
        d = rng.normal(size=(3000, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        to, td = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
        fa, ta = a.query_ray(to, td)
        fb, tb = b.query_ray(to, td)
        assert torch.equal(fa, fb) and torch.equal(ta, tb)
        pa, da = a.query_point(to)
        pb, db = b.query_point(to)
        assert torch.equal(pa, pb) and torch.equal(da, db)
