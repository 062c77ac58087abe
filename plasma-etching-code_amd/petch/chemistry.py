"""Surface chemistry -> normal velocity (the reference's chemistry.py interface).

`_ied_yield` / `_yields` are per-step scalars (an IED quadrature) evaluated on the host and
passed to the kernels; the per-face rate itself runs in libpetch_hip.so (petch_surface_rate).
The array functions accept numpy or device tensors and return numpy float64 like the
reference's (chemistry.py:45-92).
"""
import math

import numpy as np

from ._abi import ChemParams


def _ied_yield(A, Eth, par):
    """Y = A*max(sqrt(E)-sqrt(Eth),0) over the ion energy distribution (chemistry.py:10-31):
    'mean' at Emean, 'gauss' by a 25-node quadrature of N(Emean, Esig) on +-3.5 sigma,
    'bimodal' by the 49-node arcsine IED of full width ied_dE."""
    mode = par.get('ied_mode', 'mean')
    Em = float(par['Emean'])
    sEth = math.sqrt(max(Eth, 0.0))
    if mode == 'mean':
        return float(A * max(math.sqrt(max(Em, 0.0)) - sEth, 0.0))
    if mode == 'gauss':
        Es = float(par.get('Esig', 10.0))
        z = np.linspace(-3.5, 3.5, 25)
        w = np.exp(-0.5 * z * z)
        w = w / w.sum()
        E = np.clip(Em + Es * z, 0.0, None)
        return float(np.sum(w * A * np.maximum(np.sqrt(E) - sEth, 0.0)))
    dE = float(par.get('ied_dE', 40.0))
    ph = np.linspace(0.0, np.pi, 49)
    E = np.clip(Em + 0.5 * dE * np.cos(ph), 0.0, None)
    return float(np.mean(A * np.maximum(np.sqrt(E) - sEth, 0.0)))


def _yields(par):
    """(Y_ie, Y_sp, Y_p) (chemistry.py:34-39)."""
    return (_ied_yield(par['A_ie'], par['Eth_ie'], par),
            _ied_yield(par['A_sp'], par['Eth_sp'], par),
            _ied_yield(par['A_p'], par['Eth_p'], par))


def angular_factors(cos_i, par, mode):
    """(f_ie, f_sp) per face (chemistry.py:45-59): 'viennaps' sputter (1+B(1-c^2))c and
    ion-enhanced flat to 60 deg then linear to 0 at 90 deg; 'cosine' both = c."""
    c = np.clip(np.asarray(_host(cos_i), np.float64), 0.0, 1.0)
    if mode == "viennaps":
        B = par.get('B_sp', 9.3)
        f_sp = np.maximum((1.0 + B * (1.0 - c * c)) * c, 0.0)
        th = np.arccos(c)
        f_ie = np.where(c >= 0.5, 1.0, np.maximum(3.0 - 6.0 * th / np.pi, 0.0))
        return f_ie, f_sp
    return c, c


def chem_params(par, flags=None):
    """ChemParams for the kernels from PAR and Flags."""
    Yie, Ysp, Yp = _yields(par)
    mode = "cosine" if flags is None else getattr(flags, "yield_angular", "cosine")
    model = "langmuir" if flags is None else getattr(flags, "chemistry", "langmuir")
    if model not in ("langmuir", "belen"):
        raise ValueError(f"unknown chemistry model: {model}")
    return ChemParams(
        ionFlux=float(par['ionFlux']), Fflux=float(par['Fflux']), Oflux=float(par['Oflux']),
        k_sigma=float(par.get('k_sigma', 300.0)), beta_sigma=float(par.get('beta_sigma', 0.04)),
        rho=float(par['rho']), rate_scale=float(par['rate_scale']), B_sp=float(par.get('B_sp', 9.3)),
        fnorm_E=float(par.get('fnorm_E', 1.0)), fnorm_O=float(par.get('fnorm_O', 1.0)),
        Yie=Yie, Ysp=Ysp, Yp=Yp, s_F=float(par['s_F']), s_O=float(par['s_O']),
        angular_viennaps=int(mode == "viennaps"), chemistry_belen=int(model == "belen"))


def _host(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:
        pass
    return x


def _rate(m_i, m_F, m_O, cos_i, is_mask, par, flags, ion_yield):
    import torch
    from . import ops
    dev = torch.device("cuda", torch.cuda.current_device())

    def t(a):
        return torch.as_tensor(np.ascontiguousarray(_host(a), np.float32), device=dev)
    mask = torch.as_tensor(np.ascontiguousarray(np.asarray(_host(is_mask), bool), np.uint8), device=dev)
    iy = tuple(t(a) for a in ion_yield) if ion_yield is not None else None
    V, red = ops.surface_rate(t(m_i), t(m_F), t(m_O), t(cos_i), iy, None, ops.MaskGeom(),
                              chem_params(par, flags), is_mask=mask)
    return V.cpu().numpy().astype(np.float64)


def surface_rate_langmuir(m_i, m_F, m_O, cos_i, is_mask, par, flags=None):
    """Competitive-Langmuir steady state (chemistry.py:62-80)."""
    from .params import Flags
    fl = Flags(chemistry="langmuir", yield_angular=getattr(flags, "yield_angular", "cosine")
               if flags is not None else "cosine")
    return _rate(m_i, m_F, m_O, cos_i, is_mask, par, fl, None)


def surface_rate(m_i, m_F, m_O, cos_i, is_mask, par, flags=None):
    """Dispatch on flags.chemistry (chemistry.py:83-92); non-finite values are returned as is
    by the reference and zeroed here (the driver zeroes them anyway, threed.py:1723-1728)."""
    model = "langmuir" if flags is None else getattr(flags, "chemistry", "langmuir")
    if model == "langmuir":
        return surface_rate_langmuir(m_i, m_F, m_O, cos_i, is_mask, par, flags)
    if model == "belen":
        from .belen import surface_rate_belen
        return surface_rate_belen(m_i, m_F, m_O, cos_i, is_mask, par, flags)
    raise ValueError(f"unknown chemistry model: {model}")
