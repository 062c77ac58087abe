"""Surface-chemistry oracle (FP64 numpy) — TEST INFRASTRUCTURE ONLY.

Restates:
  ied_yield            chemistry.py:10-31   (_ied_yield: 'mean' / 'gauss' / 'bimodal' IED)
  yields               chemistry.py:34-39
  angular_factors      chemistry.py:45-59
  surface_rate_langmuir chemistry.py:62-80
  surface_rate_belen   belen.py:26-63
  belen_coverages      threed.py:715-737
  belen_bare_adjoint   derivative of clip(1 - thF - thO, 0, 1) (not in the reference), by
                       complex-step differentiation of belen_coverages
  surface_rate_adjoint derivative of surface_rate (not in the reference), by complex step
  faces_in_mask        threed.py:1608-1616
  ion_params           threed.py:800-802, 811
"""
import numpy as np


def ied_yield(A, Eth, par):
    mode = par.get('ied_mode', 'mean')
    Em = float(par['Emean'])
    sEth = np.sqrt(max(Eth, 0.0))
    if mode == 'mean':
        return float(A * max(np.sqrt(max(Em, 0.0)) - sEth, 0.0))
    if mode == 'gauss':
        Es = float(par.get('Esig', 10.0))
        z = np.linspace(-3.5, 3.5, 25)
        w = np.exp(-0.5 * z * z)
        w /= w.sum()
        E = np.clip(Em + Es * z, 0.0, None)
        return float((w * A * np.maximum(np.sqrt(E) - sEth, 0.0)).sum())
    dE = float(par.get('ied_dE', 40.0))
    phi = np.linspace(0.0, np.pi, 49)
    E = np.clip(Em + 0.5 * dE * np.cos(phi), 0.0, None)
    return float(np.mean(A * np.maximum(np.sqrt(E) - sEth, 0.0)))


def yields(par):
    return (ied_yield(par['A_ie'], par['Eth_ie'], par),
            ied_yield(par['A_sp'], par['Eth_sp'], par),
            ied_yield(par['A_p'], par['Eth_p'], par))


def angular_factors(cos_i, par, mode):
    c = np.clip(np.asarray(cos_i, np.float64), 0.0, 1.0)
    if mode == "viennaps":
        B = par.get('B_sp', 9.3)
        f_sp = np.maximum((1.0 + B * (1.0 - c * c)) * c, 0.0)
        theta = np.arccos(c)
        f_ie = np.where(c >= 0.5, 1.0, np.maximum(3.0 - 6.0 * theta / np.pi, 0.0))
        return f_ie, f_sp
    return c, c


def _mode(flags):
    return "cosine" if flags is None else getattr(flags, "yield_angular", "cosine")


def surface_rate_langmuir(m_i, m_F, m_O, cos_i, is_mask, par, flags=None):
    Yie, Ysp, Yp = yields(par)
    f_ie, f_sp = angular_factors(cos_i, par, _mode(flags))
    Fi = par['ionFlux'] * _num(m_i)
    Fev = par['Fflux'] * _num(m_F)
    Fp = par['Oflux'] * _num(m_O)
    eps = 1e-9
    Yie_a, Ysp_a, Yp_a = Yie * f_ie, Ysp * f_sp, Yp * f_ie
    rF = par['s_F'] * Fev / (Yie_a * Fi + eps)
    rO = par['s_O'] * Fp / (Yp_a * Fi + eps)
    bare = 1.0 / (1.0 + rF + rO)
    thF = rF * bare
    V = (1.0 / par['rho']) * (Yie_a * Fi * thF + Ysp_a * Fi * bare) * par['rate_scale']
    V = np.array(V, np.result_type(V, np.float64))
    V[np.asarray(is_mask, bool)] = 0.0
    return V


def surface_rate_belen(m_i, m_F, m_O, cos_i, is_mask, par, flags=None, ion_yield=None):
    """belen.py:26-63; `ion_yield` = (sp, ie, ip) stands for par['_ion_yield']."""
    Yie, Ysp, Yp = yields(par)
    f_ie, f_sp = angular_factors(cos_i, par, _mode(flags))
    Fi = par['ionFlux'] * _num(m_i)
    Yie_a, Ysp_a, Yp_a = Yie * f_ie, Ysp * f_sp, Yp * f_ie
    eps = 1e-9
    iy = ion_yield if ion_yield is not None else par.get('_ion_yield')
    if iy is not None:
        _sp, _ie, _ip = (_num(a) for a in iy)
        ionF = par['ionFlux']
        GY_ie, GY_p, sputter, ion_enh = ionF * _ie, ionF * _ip, ionF * _sp, ionF * _ie
    else:
        GY_ie, GY_p, sputter, ion_enh = Yie_a * Fi, Yp_a * Fi, Ysp_a * Fi, Yie_a * Fi
    fnE = par.get('fnorm_E', 1.0)
    fnO = par.get('fnorm_O', 1.0)
    Gb_E = par['Fflux'] * _num(m_F) / fnE + eps
    Gb_P = par['Oflux'] * _num(m_O) / fnO + eps
    a = (par['k_sigma'] + 2.0 * GY_ie) / Gb_E
    b = (par['beta_sigma'] + GY_p) / Gb_P
    thF = 1.0 / (1.0 + a * (1.0 + 1.0 / (b + eps)))
    rate = par['k_sigma'] * thF / 4.0 + sputter + thF * ion_enh
    V = (1.0 / par['rho']) * rate * par['rate_scale']
    V = np.array(V, np.result_type(V, np.float64))
    V[np.asarray(is_mask, bool)] = 0.0
    return V


def surface_rate(m_i, m_F, m_O, cos_i, is_mask, par, flags=None, ion_yield=None):
    """chemistry.py:83-92."""
    model = "langmuir" if flags is None else getattr(flags, "chemistry", "langmuir")
    if model == "langmuir":
        return surface_rate_langmuir(m_i, m_F, m_O, cos_i, is_mask, par, flags)
    if model == "belen":
        return surface_rate_belen(m_i, m_F, m_O, cos_i, is_mask, par, flags, ion_yield)
    raise ValueError(f"unknown chemistry model: {model}")


def belen_coverages(m_i, m_F, m_O, cos_i, par, flags, ion_yield=None):
    """threed.py:715-737 (note: ignores fnorm_E/O, unlike surface_rate_belen)."""
    Yie, Ysp, Yp = yields(par)
    f_ie, _ = angular_factors(cos_i, par, _mode(flags))
    Fi = par['ionFlux'] * _num(m_i)
    eps = 1e-9
    iy = ion_yield if ion_yield is not None else par.get('_ion_yield')
    if iy is not None:
        _sp, _ie, _ip = (_num(a) for a in iy)
        GY_ie = par['ionFlux'] * _ie
        GY_p = par['ionFlux'] * _ip
    else:
        GY_ie = Yie * f_ie * Fi
        GY_p = Yp * f_ie * Fi
    Gb_E = par['Fflux'] * _num(m_F) + eps
    Gb_P = par['Oflux'] * _num(m_O) + eps
    a = (par['k_sigma'] + 2.0 * GY_ie) / Gb_E
    b = (par['beta_sigma'] + GY_p) / Gb_P
    thF = 1.0 / (1.0 + a * (1.0 + 1.0 / (b + eps)))
    thO = 1.0 / (1.0 + b * (1.0 + 1.0 / (a + eps)))
    return thF, thO


def _num(x):
    """float64 array, or complex128 when x is complex (complex-step derivatives)."""
    return np.asarray(x) if np.iscomplexobj(x) else np.asarray(x, np.float64)


def belen_bare_adjoint(m_i, m_F, m_O, cos_i, par, flags, gbare, ion_yield=None):
    """d loss / d inputs of bare = clip(1 - thF - thO, 0, 1) for gbare = d loss / d bare, by
    complex-step differentiation of belen_coverages: every face's bare depends on that face's
    inputs only, so one imaginary step h on all faces at once gives the whole diagonal
    Jacobian as Im(bare) / h, exact to rounding. Returns dict(m_F, m_O, and m_i or ie, ip);
    cos_i is held constant and the clip passes the gradient where 0 <= 1 - thF - thO <= 1."""
    h = 1e-30
    ins = dict(m_i=m_i, m_F=m_F, m_O=m_O)
    if ion_yield is not None:
        ins.update(sp=ion_yield[0], ie=ion_yield[1], ip=ion_yield[2])

    def bare_of(v):
        iy = (v['sp'], v['ie'], v['ip']) if ion_yield is not None else None
        thF, thO = belen_coverages(v['m_i'], v['m_F'], v['m_O'], cos_i, par, flags, ion_yield=iy)
        return 1.0 - thF - thO

    base = {k: np.asarray(x, np.float64) for k, x in ins.items()}
    br = bare_of(base)
    g = np.where((br >= 0.0) & (br <= 1.0), np.asarray(gbare, np.float64), 0.0)
    keys = ("m_F", "m_O") + (("ie", "ip") if ion_yield is not None else ("m_i",))
    out = {}
    for k in keys:
        v = dict(base)
        v[k] = base[k] + 1j * h
        out[k] = g * bare_of(v).imag / h
    return out


def surface_rate_adjoint(m_i, m_F, m_O, cos_i, is_mask, par, flags, gV, ion_yield=None):
    """d loss / d inputs of surface_rate for gV = d loss / dV, by complex-step differentiation
    (each face's V depends on that face's inputs only): dict(m_F, m_O, m_i, and sp, ie, ip when
    yields are given; an input the rate does not read gets zeros); cos_i constant; mask faces
    pass nothing."""
    h = 1e-30
    base = dict(m_i=np.asarray(m_i, np.float64), m_F=np.asarray(m_F, np.float64),
                m_O=np.asarray(m_O, np.float64))
    if ion_yield is not None:
        base.update(sp=np.asarray(ion_yield[0], np.float64), ie=np.asarray(ion_yield[1], np.float64),
                    ip=np.asarray(ion_yield[2], np.float64))
    keys = ("m_F", "m_O", "m_i") + (("sp", "ie", "ip") if ion_yield is not None else ())
    g = np.asarray(gV, np.float64)
    out = {}
    for k in keys:
        v = dict(base)
        v[k] = base[k] + 1j * h
        iy = (v['sp'], v['ie'], v['ip']) if ion_yield is not None else None
        V = surface_rate(v['m_i'], v['m_F'], v['m_O'], cos_i, is_mask, par, flags, ion_yield=iy)
        out[k] = g * np.imag(V) / h
    return out


def faces_in_mask(centroids, geo, mask_th, trench_width, hole=False):
    c = np.asarray(centroids, np.float64)
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    in_band = (z >= geo['sub_top']) & (z < geo['sub_top'] + mask_th)
    if hole:
        opening = (x - geo['Lx'] / 2) ** 2 + (y - geo['Ly'] / 2) ** 2 < (trench_width / 2) ** 2
    else:
        opening = np.abs(x - geo['Lx'] / 2) < trench_width / 2
    return in_band & (~opening)


def ion_params(par):
    """The faithful-ion constants passed at threed.py:806-813 (derived at 800-802)."""
    inflect = float(par.get('inflectAngle', 1.55334303))
    minang = float(par.get('minAngle', 1.3962634))
    n_l = float(par.get('n_l', 10.0))
    A_en = 1.0 / (1.0 + n_l * (1.5707963 / inflect - 1.0))
    minE = float(min(par['Eth_ie'], par['Eth_sp']))
    return dict(meanE=float(par['Emean']), sigmaE=float(par['Esig']), A_sp=float(par['A_sp']),
                Eth_sp=float(par['Eth_sp']), B_sp=float(par['B_sp']), A_ie=float(par['A_ie']),
                Eth_ie=float(par['Eth_ie']), A_p=float(par['A_p']), Eth_p=float(par['Eth_p']),
                inflect=inflect, minang=minang, n_l=n_l, A_en=float(A_en), minE=minE,
                thrmin=float(par.get('thetaRMin', 1.2217305)),
                thrmax=float(par.get('thetaRMax', 1.5707963)))
