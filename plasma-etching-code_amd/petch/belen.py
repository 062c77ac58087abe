"""Belen/Ertl coupled-coverage SF6/O2 velocity (the reference's belen.py interface).

Reads the faithful-ion yields from par['_ion_yield'] (set by mc_flux_3d_coupled) exactly like
belen.py:34-41; the per-face arithmetic runs in libpetch_hip.so (petch_surface_rate).
"""
from .chemistry import _rate


def surface_rate_belen(m_i, m_F, m_O, cos_i, is_mask, par, flags=None):
    """V = (k_sigma*thF/4 + sputter + thF*ion_enh)/rho * rate_scale, 0 on mask faces
    (belen.py:26-63)."""
    from .params import Flags
    fl = Flags(chemistry="belen", yield_angular=getattr(flags, "yield_angular", "cosine")
               if flags is not None else "cosine")
    return _rate(m_i, m_F, m_O, cos_i, is_mask, par, fl, par.get('_ion_yield'))
