"""petch — MI355X (gfx950) implementation of petch's per-step 3-D etch loop.

Same public surface as the reference package for that path: Domain / SF6O2 / Process / Result,
PAR / Flags, and petch.threed.run_etch_3d with its stage functions. Compute runs in
libpetch_hip.so (hand-written HIP kernels) on device-resident torch tensors.
"""
from .params import PAR, Flags, DEFAULT_FLAGS
from .chemistry import surface_rate, surface_rate_langmuir
from .api import Domain, SF6O2, Process, Result

__version__ = "0.3.0+mi355x"

__all__ = ["Domain", "SF6O2", "Process", "Result", "PAR", "Flags", "DEFAULT_FLAGS",
           "surface_rate", "surface_rate_langmuir"]
