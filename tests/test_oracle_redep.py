"""Pins the oracle's redeposition (oracle_trace_redeposition in oracle/petch_oracle.c,
mc_redep_3d of threed.py:1582-1605) to what the mathematics fixes:

- the emission draw picks face f with probability w_f / sum w, w = max(rate * area, 0), and never
  a face with a zero, negative or non-finite weight (chi-square against the weights);
- inside a closed spherical cavity with s_redep = 1 every ray deposits exactly once, so the
  area-weighted result returns the emitted total; with s_redep = 0.3 the roulette walk deposits
  1 / 0.3 per ray on average;
- a sphere's inner surface receives a uniform flux density from a Lambertian emitter anywhere
  on it (the integrating-sphere property: dF = cos1 cos2 / (pi s^2) dA with cos1 = cos2 =
  s / 2R, so dF = dA / 4 pi R^2), which a lobe other than the cosine one, or one about the
  wrong normal, breaks;
- emission up from an open plane escapes: nothing is redeposited.
"""
import numpy as np
from scipy import stats

from oracle import mesh as om
from oracle import native as on


def _sphere_cavity(R=2.0, dx=0.25):
    """MC mesh of a gas-filled sphere of radius R in solid (phi > 0 solid), with its into-gas
    normals (threed.py:1553-1568)."""
    n = int(np.ceil(2 * R / dx)) + 5
    c = 0.5 * (n - 1) * dx
    x = np.arange(n) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    phi = (np.sqrt((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2) - R).astype(np.float32)
    v, f = on.marching_cubes(phi, dx)
    cen, area, _ = om.mesh_attrs(v, f)
    gn = om.gas_normals(v, f, cen, phi, dx)
    return v, f, cen, area, gn, np.array([c, c, c])


def test_emission_draw_follows_the_weights():
    rng = np.random.default_rng(2)
    F = 40
    rate = rng.uniform(0.0, 2.0, F).astype(np.float32)
    area = rng.uniform(0.5, 1.5, F).astype(np.float32)
    rate[[3, 7]] = 0.0
    rate[11] = -1.0
    rate[19] = np.nan
    rate[23] = np.inf
    n = 400000
    face = on.redep_faces(rate, area, n, 17)
    w = rate.astype(np.float64) * area
    w = np.where(np.isfinite(w) & (w > 0), w, 0.0)
    cnt = np.bincount(face, minlength=F)
    assert cnt[[3, 7, 11, 19, 23]].sum() == 0
    live = w > 0
    exp = n * w[live] / w.sum()
    chi2 = ((cnt[live] - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, live.sum() - 1) > 1e-3


def test_closed_cavity_returns_the_emitted_total():
    v, f, cen, area, gn, _ = _sphere_cavity()
    bvh = on.OracleBvh(v, f)
    rng = np.random.default_rng(4)
    rate = rng.uniform(0.0, 1.0, len(f)).astype(np.float32)
    tot = float((rate.astype(np.float64) * area.astype(np.float32)).sum())
    n = 20000
    out = bvh.trace_redeposition(cen, gn, area, rate, 1.0, n, 9)
    a = np.maximum(area.astype(np.float32).astype(np.float64), 1e-9)
    assert abs((out * a).sum() / tot - 1.0) < 1e-9
    out3 = bvh.trace_redeposition(cen, gn, area, rate, 0.3, n, 9)
    # per-ray deposit: mean 1 / 0.3, standard deviation below 2 (a few roulettes)
    m = (out3 * a).sum() / tot
    assert abs(m - 1.0 / 0.3) < 5 * 2.0 / np.sqrt(n)


def test_integrating_sphere_uniform_irradiance():
    v, f, cen, area, gn, c = _sphere_cavity()
    bvh = on.OracleBvh(v, f)
    rate = np.zeros(len(f), np.float32)
    cap = cen[:, 2] < c[2] - 1.5                     # emit from the bottom cap only
    rate[cap] = 1.0
    n = 200000
    out = bvh.trace_redeposition(cen, gn, area, rate, 1.0, n, 21)
    a = area.astype(np.float32).astype(np.float64)
    tot = float((rate * a).sum())
    cnt = out * a * n / tot                          # rays absorbed by each face
    exp = n * a / a.sum()
    ok = exp > 20                                    # MC leaves some slivers of tiny area
    z = (cnt[ok] - exp[ok]) / np.sqrt(exp[ok])
    print(f"sphere: faces {ok.sum()} mean z {z.mean():.3f} |z|>3 {np.mean(np.abs(z) > 3):.4f}")
    assert ok.sum() > 1000
    assert abs(z.mean()) < 0.1 and np.mean(np.abs(z) > 3) < 0.015
    top = cen[:, 2] > c[2]
    assert abs(cnt[top].sum() / exp[top].sum() - 1.0) < 0.02


def test_open_plane_redeposits_nothing():
    nx = 10
    dx = 0.25
    x = np.arange(nx) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    phi = (1.1 - Z).astype(np.float32)               # solid below z = 1.1, gas above
    v, f = on.marching_cubes(phi, dx)
    cen, area, _ = om.mesh_attrs(v, f)
    gn = om.gas_normals(v, f, cen, phi, dx)
    assert (gn[:, 2] > 0.99).all()
    out = on.OracleBvh(v, f).trace_redeposition(cen, gn, area, np.ones(len(f), np.float32), 0.5,
                                                5000, 3)
    assert (out == 0).all()


def test_nothing_to_emit_gives_zero():
    v, f, cen, area, gn, _ = _sphere_cavity(R=1.0)
    rate = np.zeros(len(f), np.float32)
    rate[0] = np.nan
    out = on.OracleBvh(v, f).trace_redeposition(cen, gn, area, rate, 0.5, 1000, 1)
    assert (out == 0).all()
