"""Surface-charging hooks of the flux paths (threed.py:842-868, 893-894, 1046-1147), on the
device. They are opt-in in the reference and here:

- the alpha model (flags.surface_charging with par['charge_alpha'] > 0, in the coupled flux):
  electrons from a Gaussian source of spread e_ang_sigma, traced with unit sticking, give each
  face its electron arrival m_e; the ion arrival is throttled by
  f = 1 - alpha clip(1 - sh_e / max(sh_i, 1e-3), 0, 1), where sh = flux / (its 90th percentile
  over the faces with flux > 1e-6), clipped to [0, 1];
- the floor-charge throttle (flags.floor_charge_throttle, in the coupled flux): the floor
  faces' ion arrival (and ion yields) times Q(AR) of a deep-AR table;
- the Hwang-Giapis closure (flags.surface_charging = 'hg' or 'petch', with the faithful ion
  yields, in the Knudsen path): the floor keeps the slice Q of its ions with yields at
  E - Vf; the removed slice lands on the sidewall foot with energy E_defl.

The floor is the faces below sub_top inside the feature footprint (|x - cx|, or the radius
from (cx, cy), <= trench_width / 2 + 2 dx) whose into-gas normal has z > 0.7; the aspect ratio
is (sub_top - lowest floor centroid) / trench_width.
"""
import numpy as np
import torch

from . import ops

# (AR, floor ion flux fraction, floor potential V, deflected-ion foot energy eV) of the
# gate-validated 2-D solver at the Hwang-Giapis reference conditions (charging2d.py:1515-1518)
GATE_AR = np.array([0.0, 1.0, 1.2, 1.6, 2.0, 2.6, 3.0, 3.6, 4.0])
GATE_FLUX = np.array([1.0, 0.59, 0.55, 0.47, 0.40, 0.34, 0.30, 0.26, 0.22])
GATE_VFLOOR = np.array([0.0, 8.0, 10.0, 15.0, 19.0, 25.0, 28.0, 31.0, 33.0])
FOOT_E = np.array([0.0, 15.0, 16.5, 17.5, 20.0, 23.0, 25.0, 26.5, 27.5])
# the same from the reference's own derived-source solver (charging_general.py:1038-1041)
PETCH_AR = np.array([1.0, 2.0, 3.0, 4.0])
PETCH_Q = np.array([0.704, 0.491, 0.406, 0.349])
PETCH_VFLOOR = np.array([9.5, 19.2, 27.7, 34.0])
PETCH_FOOT_E = np.array([16.6, 14.2, 18.8, 21.6])
# deep-AR floor-flux throttle (charging_general.py:1070-1071)
THROTTLE_AR = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 8.0, 15.0, 25.0])
THROTTLE_Q = np.array([1.0, 0.704, 0.491, 0.406, 0.349, 0.25, 0.60, 0.50])


def charging_floor_profile(AR):
    """(Q, Vf, E_defl) at aspect ratio AR, piecewise linear in the gate table
    (charging2d.py:1521-1529)."""
    return (np.interp(AR, GATE_AR, GATE_FLUX), np.interp(AR, GATE_AR, GATE_VFLOOR),
            np.interp(AR, GATE_AR, FOOT_E))


def petch_floor_profile(AR):
    """(Q, Vf, E_defl) of the derived-source table (charging_general.py:1044-1050)."""
    return (np.interp(AR, PETCH_AR, PETCH_Q), np.interp(AR, PETCH_AR, PETCH_VFLOOR),
            np.interp(AR, PETCH_AR, PETCH_FOOT_E))


def floor_charge_throttle_profile(AR):
    """Q(AR) in [0, 1], 1 below the table and its last value beyond it
    (charging_general.py:1074-1080)."""
    return np.clip(np.interp(AR, THROTTLE_AR, THROTTLE_Q, left=1.0, right=float(THROTTLE_Q[-1])),
                   0.0, 1.0)


def active(par, flags):
    """Whether a hook of the coupled flux is on (the ion finish must then run before it)."""
    if flags is None:
        return False
    alpha = float(par.get('charge_alpha', 0.0))
    return (bool(getattr(flags, "surface_charging", False)) and alpha > 0.0) or \
        bool(getattr(flags, "floor_charge_throttle", False))


def floor_mask(centroids, gas_nz, geo):
    """(F,) bool device mask of the floor faces (threed.py:1069-1076, 1130-1137)."""
    c = centroids
    cx, cy = 0.5 * geo['Lx'], 0.5 * geo['Ly']
    hw = 0.5 * geo['trench_width'] + 2.0 * geo['dx']
    if geo.get('hole', False):
        inside = torch.hypot(c[:, 0].double() - cx, c[:, 1].double() - cy) <= hw
    else:
        inside = (c[:, 0].double() - cx).abs() <= hw
    below = c[:, 2].double() < geo['sub_top']
    return inside, below, inside & below & (gas_nz.double() > 0.7)


def _aspect(centroids, floor, geo):
    """(aspect ratio, lowest floor z) of the floor (one host read), or None without a floor."""
    if not bool(floor.any()):
        return None
    zmin = float(centroids[:, 2][floor].min())
    depth = float(geo['sub_top'] - zmin)
    return depth / max(geo['trench_width'], 1e-9), zmin


def electron_factor(mesh, box, A_src, par, m_i, n_neu, seed, smooth):
    """The alpha model's ion-flux factor (threed.py:847-868): electrons from the ion-type
    Gaussian source of spread e_ang_sigma, stream seed*9+7, unit sticking; m_e normalised like
    the neutrals (times eFlux, clipped to [0, 8]) and smoothed as m_i is; returns the (F,)
    factor. A_src: the source plane's area."""
    from .flux import SEED_MASK
    F = mesh.nf
    src = ops.make_source(ops.SRC_ION_GAUSS, float(par.get('e_ang_sigma', 0.5)))
    fe = ops.trace_neutral_rr(mesh.bvh, box, src, torch.ones((F,), dtype=torch.float32,
                                                               device=mesh.device),
                              1.0, n_neu, (seed * 9 + 7) & SEED_MASK)
    m_e = smooth(ops.normalize(fe, mesh.areas, A_src / n_neu * float(par.get('eFlux', 1.0)),
                               0.0, 8.0))
    return charge_factor(m_i, m_e, float(par.get('charge_alpha', 0.0)))


def charge_factor(m_i, m_e, alpha):
    """f = 1 - alpha clip(1 - sh_e / max(sh_i, 1e-3), 0, 1) with each flux's shadow
    sh = clip(flux / its 90th percentile over the faces above 1e-6, 0, 1) (threed.py:861-867)."""
    def ref(m):
        sel = m[m > 1e-6]
        return torch.quantile(sel.double(), 0.9) if sel.numel() else torch.ones((), dtype=torch.float64, device=m.device)
    sh_e = (m_e.double() / torch.clamp(ref(m_e), min=1e-6)).clamp(0.0, 1.0)
    sh_i = (m_i.double() / torch.clamp(ref(m_i), min=1e-6)).clamp(0.0, 1.0)
    f = 1.0 - alpha * (1.0 - sh_e / torch.clamp(sh_i, min=1e-3)).clamp(0.0, 1.0)
    return f.to(torch.float32)


def floor_throttle(mesh, gas_n, geo, m_i, ion_yield):
    """flags.floor_charge_throttle (threed.py:1112-1147): m_i and the ion yields of the floor
    faces times Q(AR). Returns (m_i, ion_yield)."""
    _, _, floor = floor_mask(mesh.centroids, gas_n[:, 2], geo)
    ar = _aspect(mesh.centroids, floor, geo)
    if ar is None:
        return m_i, ion_yield
    Q = float(floor_charge_throttle_profile(ar[0]))
    scale = torch.where(floor, torch.tensor(Q, dtype=torch.float64, device=m_i.device),
                        torch.ones((), dtype=torch.float64, device=m_i.device))
    m_i = (m_i.double() * scale).to(torch.float32)
    if ion_yield is not None:
        ion_yield = tuple((a.double() * scale).to(torch.float32) for a in ion_yield)
    return m_i, ion_yield


def hg_closure(mesh, gas_n, geo, par, flags, m_i, ion_yield):
    """flags.surface_charging = 'hg' / 'petch' with ion yields (threed.py:1046-1109). Returns
    (m_i, ion_yield)."""
    from .chemistry import _yields
    mode = getattr(flags, "surface_charging", False)
    if mode not in ("hg", "petch") or ion_yield is None:
        return m_i, ion_yield
    inside, below, floor = floor_mask(mesh.centroids, gas_n[:, 2], geo)
    ar = _aspect(mesh.centroids, floor, geo)
    if ar is None:
        return m_i, ion_yield
    prof = petch_floor_profile if mode == "petch" else charging_floor_profile
    Q, Vf, E_defl = (float(v) for v in prof(ar[0]))
    p2 = dict(par)
    p2['Emean'] = max(par['Emean'] - Vf, 0.0)
    y0, yv = _yields(par), _yields(p2)
    r_ie = yv[0] / y0[0] if y0[0] > 0 else 0.0
    r_sp = yv[1] / y0[1] if y0[1] > 0 else 0.0
    r_p = yv[2] / y0[2] if y0[2] > 0 else 0.0
    dev = m_i.device
    A = mesh.areas.double().clamp(min=1e-9)
    mi = m_i.double()
    defl = (mi * A)[floor].sum() * (1.0 - Q)             # the removed slice (device scalar)
    sp, ie, ip = (a.double() for a in ion_yield)
    one = torch.ones((), dtype=torch.float64, device=dev)

    def on_floor(v):
        return torch.where(floor, torch.tensor(v, dtype=torch.float64, device=dev), one)
    ie, sp, ip = ie * on_floor(Q * r_ie), sp * on_floor(Q * r_sp), ip * on_floor(Q * r_p)
    mi = mi * on_floor(Q)
    foot_h = max(0.3 * geo['trench_width'], 2.0 * geo['dx'])
    foot = inside & below & (gas_n[:, 2].double() <= 0.7) & \
        (mesh.centroids[:, 2].double() <= ar[1] + foot_h)
    f_area = A[foot].sum()
    fdefl = torch.where(defl > 0.0, defl / torch.clamp(f_area, min=1e-12), 0.0 * defl)
    sE = float(np.sqrt(E_defl))
    zero = torch.zeros((), dtype=torch.float64, device=dev)
    for key, eth in (("A_ie", "Eth_ie"), ("A_sp", "Eth_sp"), ("A_p", "Eth_p")):
        add = torch.where(foot, fdefl * par[key] * max(sE - float(np.sqrt(par[eth])), 0.0), zero)
        if key == "A_ie":
            ie = ie + add
        elif key == "A_sp":
            sp = sp + add
        else:
            ip = ip + add
    f32 = torch.float32
    return mi.to(f32), (sp.to(f32), ie.to(f32), ip.to(f32))
