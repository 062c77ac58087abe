"""Oracle (test infrastructure only) for the surface-charging hooks of the flux paths
(threed.py:842-868 alpha model, 1046-1109 Hwang-Giapis closure, 1112-1147 floor throttle), in
float64 numpy on host arrays.

Pinned by tests/test_oracle_charging.py: the tables' interpolation at and between their nodes
and beyond their ends, the alpha factor's limits (equal shadowing gives 1, a fully shadowed
electron arrival 1 - alpha), the throttle's and closure's floor scaling, and the closure's
conservation (the ions removed from the floor, weighted by area, equal the per-area flux
handed to the foot times the foot's area).
"""
import numpy as np

from oracle import chem as oc

# charging2d.py:1515-1518
_GATE = (np.array([0.0, 1.0, 1.2, 1.6, 2.0, 2.6, 3.0, 3.6, 4.0]),
         np.array([1.0, 0.59, 0.55, 0.47, 0.40, 0.34, 0.30, 0.26, 0.22]),
         np.array([0.0, 8.0, 10.0, 15.0, 19.0, 25.0, 28.0, 31.0, 33.0]),
         np.array([0.0, 15.0, 16.5, 17.5, 20.0, 23.0, 25.0, 26.5, 27.5]))
# charging_general.py:1038-1041
_PETCH = (np.array([1.0, 2.0, 3.0, 4.0]), np.array([0.704, 0.491, 0.406, 0.349]),
          np.array([9.5, 19.2, 27.7, 34.0]), np.array([16.6, 14.2, 18.8, 21.6]))
# charging_general.py:1070-1071
_THROTTLE = (np.array([0.0, 1.0, 2.0, 3.0, 4.0, 8.0, 15.0, 25.0]),
             np.array([1.0, 0.704, 0.491, 0.406, 0.349, 0.25, 0.60, 0.50]))


def floor_profile(AR, table="hg"):
    """(Q, Vf, E_defl) at AR (charging2d.py:1521-1529 / charging_general.py:1044-1050)."""
    t = _GATE if table == "hg" else _PETCH
    return tuple(float(np.interp(AR, t[0], col)) for col in t[1:])


def throttle_q(AR):
    """charging_general.py:1074-1080."""
    return float(np.clip(np.interp(AR, _THROTTLE[0], _THROTTLE[1], left=1.0,
                                   right=float(_THROTTLE[1][-1])), 0.0, 1.0))


def charge_factor(m_i, m_e, alpha):
    """threed.py:861-867."""
    m_i, m_e = np.asarray(m_i, float), np.asarray(m_e, float)
    ref_i = np.percentile(m_i[m_i > 1e-6], 90) if (m_i > 1e-6).any() else 1.0
    ref_e = np.percentile(m_e[m_e > 1e-6], 90) if (m_e > 1e-6).any() else 1.0
    sh_e = np.clip(m_e / max(ref_e, 1e-6), 0.0, 1.0)
    sh_i = np.clip(m_i / max(ref_i, 1e-6), 0.0, 1.0)
    return 1.0 - alpha * np.clip(1.0 - sh_e / np.maximum(sh_i, 1e-3), 0.0, 1.0)


def masks(centroids, gas_nz, geo):
    """(inside, below, floor) of threed.py:1069-1076."""
    c = np.asarray(centroids, float)
    cx, cy = 0.5 * geo['Lx'], 0.5 * geo['Ly']
    hw = 0.5 * geo['trench_width'] + 2.0 * geo['dx']
    if geo.get('hole', False):
        inside = np.hypot(c[:, 0] - cx, c[:, 1] - cy) <= hw
    else:
        inside = np.abs(c[:, 0] - cx) <= hw
    below = c[:, 2] < geo['sub_top']
    return inside, below, inside & below & (np.asarray(gas_nz, float) > 0.7)


def floor_throttle(centroids, gas_nz, geo, m_i, ion_yield):
    """threed.py:1112-1147."""
    _, _, floor = masks(centroids, gas_nz, geo)
    m_i = np.array(m_i, float)
    iy = None if ion_yield is None else tuple(np.array(a, float) for a in ion_yield)
    if not floor.any():
        return m_i, iy
    depth = geo['sub_top'] - np.asarray(centroids, float)[floor, 2].min()
    Q = throttle_q(depth / max(geo['trench_width'], 1e-9))
    m_i[floor] *= Q
    if iy is not None:
        for a in iy:
            a[floor] *= Q
    return m_i, iy


def hg_closure(centroids, gas_nz, areas, geo, par, m_i, ion_yield, table="hg"):
    """threed.py:1046-1109 (ion_yield = (sp, ie, ip)); areas floored at 1e-9 by the caller."""
    c = np.asarray(centroids, float)
    inside, below, floor = masks(c, gas_nz, geo)
    m_i = np.array(m_i, float)
    sp, ie, ip = (np.array(a, float) for a in ion_yield)
    if not floor.any():
        return m_i, (sp, ie, ip)
    zfloor = c[floor, 2].min()
    Q, Vf, E = floor_profile((geo['sub_top'] - zfloor) / max(geo['trench_width'], 1e-9), table)
    p2 = dict(par)
    p2['Emean'] = max(par['Emean'] - Vf, 0.0)
    y0, yv = oc.yields(par), oc.yields(p2)
    r = [yv[k] / y0[k] if y0[k] > 0 else 0.0 for k in range(3)]     # (ie, sp, p)
    A = np.asarray(areas, float)
    defl = float((m_i[floor] * A[floor]).sum()) * (1.0 - Q)
    ie[floor] *= Q * r[0]
    sp[floor] *= Q * r[1]
    ip[floor] *= Q * r[2]
    m_i[floor] *= Q
    foot = inside & below & (np.asarray(gas_nz, float) <= 0.7) & \
        (c[:, 2] <= zfloor + max(0.3 * geo['trench_width'], 2.0 * geo['dx']))
    if foot.any() and defl > 0.0:
        fd = defl / max(float(A[foot].sum()), 1e-12)
        sE = np.sqrt(E)
        ie[foot] += fd * par['A_ie'] * max(sE - np.sqrt(par['Eth_ie']), 0.0)
        sp[foot] += fd * par['A_sp'] * max(sE - np.sqrt(par['Eth_sp']), 0.0)
        ip[foot] += fd * par['A_p'] * max(sE - np.sqrt(par['Eth_p']), 0.0)
    return m_i, (sp, ie, ip)
