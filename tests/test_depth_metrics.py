"""The product's depth metrics (petch.threed.center_depth_3d / max_depth_3d, threed.py:1798-1869)
on synthetic level sets whose floors are known exactly: flat floors, a tilted floor whose 5th
percentile is computed by hand, a buried gas pocket below the floor (not connected to the top,
so it must not count), columns closed at the top, and a trench. These are host functions; the
GPU path hands them the level set it read back."""
import numpy as np
import pytest

from oracle import levelset as ol
from petch import threed as t3


def _geo(nx, ny, nz, dx, width, hole, sub_top):
    xs, ys, zs = (np.arange(n) * dx for n in (nx, ny, nz))
    return dict(xs=xs, ys=ys, zs=zs, dx=dx, Lx=nx * dx, Ly=ny * dx, Lz=nz * dx,
                trench_width=width, hole=hole, sub_top=sub_top)


def _carve(geo, floor_k):
    """phi < 0 (gas) above a per-column floor index: floor_k[i, j] is the lowest gas node of
    column (i, j); columns with floor_k = nz stay solid (closed at the top)."""
    nx, ny, nz = len(geo['xs']), len(geo['ys']), len(geo['zs'])
    k = np.arange(nz)[None, None, :]
    phi = np.where(k >= floor_k[:, :, None], -1.0, 1.0)
    return phi


def _footprint(geo, r):
    X, Y = np.meshgrid(geo['xs'] - geo['Lx'] / 2, geo['ys'] - geo['Ly'] / 2, indexing='ij')
    return X ** 2 + Y ** 2 <= r ** 2


def test_flat_hole_floor():
    g = _geo(40, 40, 60, 0.25, 6.0, True, 12.0)
    top = int(round(12.0 / 0.25))                                   # substrate top index
    fk = np.full((40, 40), top)
    fk[_footprint(g, 3.0 + 1e-9)] = 20                              # flat floor at z = 5.0
    g['phi'] = _carve(g, fk)
    assert t3.center_depth_3d(g) == pytest.approx(12.0 - 5.0)
    assert t3.max_depth_3d(g) == pytest.approx(12.0 - 5.0)
    assert t3.center_depth_3d(g) == ol.depth3d(g) and t3.max_depth_3d(g) == ol.max_depth_3d(g)


def test_tilted_hole_floor_percentile_and_median():
    """A floor that deepens linearly with x: the max depth is the 5th percentile of the floors
    over the footprint r <= w/2 + dx, the center depth the median over the central columns
    (|x|, |y| <= min(1, w/2 - dx), inside r <= w/2 - dx)."""
    dx, w = 0.25, 6.0
    g = _geo(40, 40, 80, dx, w, True, 15.0)
    xi = np.arange(40)
    fk = np.full((40, 40), 60)
    base = 30 - (xi - 20) // 2                                      # floor index per x column
    fk[:, :] = np.where(_footprint(g, w / 2 + dx), base[:, None], 60)
    g['phi'] = _carve(g, fk)
    zs = g['zs']
    foot = _footprint(g, w / 2 + dx)
    floors = zs[fk[foot]]
    assert t3.max_depth_3d(g) == pytest.approx(15.0 - np.percentile(floors, 5))
    X, Y = np.meshgrid(g['xs'] - g['Lx'] / 2, g['ys'] - g['Ly'] / 2, indexing='ij')
    hx = max(min(1.0, w / 2 - dx), dx)
    central = (np.abs(X) <= hx) & (np.abs(Y) <= hx) & (X ** 2 + Y ** 2 <= (w / 2 - dx) ** 2)
    assert t3.center_depth_3d(g) == pytest.approx(15.0 - np.median(zs[fk[central]]))
    assert t3.max_depth_3d(g) > t3.center_depth_3d(g)


def test_buried_pocket_and_closed_columns_are_ignored():
    g = _geo(32, 32, 60, 0.25, 4.0, True, 12.0)
    fk = np.full((32, 32), 48)
    fk[_footprint(g, 2.0 + 1e-9)] = 32                              # floor at z = 8.0
    phi = _carve(g, fk)
    phi[14:18, 14:18, 10:20] = -1.0                                 # gas pocket below the floor
    phi[15:17, 15:17, 55:] = 1.0                                    # columns closed at the top
    g['phi'] = phi
    assert t3.center_depth_3d(g) == pytest.approx(12.0 - 8.0)
    assert t3.max_depth_3d(g) == pytest.approx(12.0 - 8.0)


def test_trench_uses_the_opening_strip():
    dx, w = 0.2, 2.0
    g = _geo(30, 10, 50, dx, w, False, 8.0)
    X = g['xs'] - g['Lx'] / 2
    fk = np.full((30, 10), 40)
    fk[np.abs(X) <= w / 2 + dx, :] = 15                             # floor at z = 3.0
    fk[np.abs(X) <= 0.2, 3] = 10                                    # one deeper dimple column
    g['phi'] = _carve(g, fk)
    assert t3.center_depth_3d(g) == pytest.approx(8.0 - 3.0)         # median ignores the dimple
    floors = g['zs'][fk[np.abs(X) <= w / 2 + dx, :]]
    assert t3.max_depth_3d(g) == pytest.approx(8.0 - np.percentile(floors, 5))
    assert t3.center_depth_3d(g) == ol.depth3d(g) and t3.max_depth_3d(g) == ol.max_depth_3d(g)


def test_unetched_feature_has_zero_depth():
    g = _geo(20, 20, 40, 0.25, 3.0, True, 10.0 - 0.25)
    fk = np.full((20, 20), 40)                                      # closed everywhere
    g['phi'] = _carve(g, fk)
    assert t3.center_depth_3d(g) == 0.0 and t3.max_depth_3d(g) == 0.0
