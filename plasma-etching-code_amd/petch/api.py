"""petch high-level API (the reference's api.py interface, api.py:24-171).

    import petch
    dom    = petch.Domain.hole(extent=14, dx=0.25, diameter=6, mask=2, depth=18)
    model  = petch.SF6O2()
    result = petch.Process(dom, model, duration=3.0).run(steps=40)
    result.depth
    result.save("etch.vtk")
"""
import time

import numpy as np

from .params import PAR, Flags
from . import threed as _t3

# friendly API name -> PAR key (api.py:21)
_ALIASES = {"meanE": "Emean", "sigmaE": "Esig", "iedf": "ied_mode"}


class Domain:
    """Hole or trench in a masked substrate (micrometres)."""

    def __init__(self, geo):
        self._geo = geo

    @classmethod
    def hole(cls, extent, dx, diameter, mask=2.0, depth=16.0, headroom=4.0):
        sub_top = float(depth)
        return cls(dict(Lx=extent, Ly=extent, Lz=sub_top + mask + headroom, dx=dx,
                        trench_width=diameter, mask_th=mask, sub_top=sub_top, hole=True))

    @classmethod
    def trench(cls, extent, dx, width, mask=2.0, depth=16.0, headroom=4.0):
        sub_top = float(depth)
        return cls(dict(Lx=extent, Ly=extent, Lz=sub_top + mask + headroom, dx=dx,
                        trench_width=width, mask_th=mask, sub_top=sub_top, hole=False))

    @property
    def feature_width(self):
        return self._geo['trench_width']


class SF6O2:
    """SF6/O2 process model: PAR with rate_scale / ied_mode and keyword overrides, and the
    faithful-model flags of api.py:72-73."""

    def __init__(self, rate_scale=0.025, iedf="gauss", **overrides):
        self.par = dict(PAR)
        self.par['rate_scale'] = rate_scale
        self.par['ied_mode'] = iedf
        for k, v in overrides.items():
            self.par[_ALIASES.get(k, k)] = v
        self.flags = dict(chemistry="belen", yield_angular="viennaps", coverage_sticking=True,
                          ion_reflection=True, warm_start_coverage=True, sampling="sobol")


class Result:
    """Outcome of Process.run(): depth metrics, the level set and the surface mesh."""

    def __init__(self, geo, wall_time):
        self._geo = geo
        self.wall_time = wall_time

    @property
    def depth(self):
        return _t3.center_depth_3d(self._geo)

    @property
    def max_depth(self):
        return _t3.max_depth_3d(self._geo)

    @property
    def aspect_ratio(self):
        return self.max_depth / self._geo['trench_width']

    @property
    def phi(self):
        return self._geo['phi']

    @property
    def mesh(self):
        v, f, _, _ = _t3.extract_mesh_3d(self._geo['phi'], self._geo['dx'])
        return v, f

    def plot(self, path=None, show=False):
        """x-z cross-section through the feature centre (needs matplotlib)."""
        import matplotlib
        if path is not None and not show:
            matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        g = self._geo
        jc = g['phi'].shape[1] // 2
        phi = g['phi'][:, jc, :].T
        fig, ax = plt.subplots(figsize=(5, 4))
        ax.contourf(g['xs'], g['zs'], (phi < 0).astype(float), levels=[0.5, 1.5], colors=["#cfe8ff"])
        ax.contour(g['xs'], g['zs'], phi, levels=[0.0], colors="k", linewidths=1.2)
        ax.axhline(g['sub_top'], color="0.6", lw=0.6, ls="--")
        ax.set_xlabel("x (um)")
        ax.set_ylabel("z (um)")
        ax.set_title(f"depth {self.depth:.1f} um  AR {self.aspect_ratio:.1f}  ({self.wall_time:.1f}s)")
        ax.set_aspect("equal")
        if path is not None:
            fig.tight_layout()
            fig.savefig(path, dpi=130)
        if show:
            plt.show()
        return ax

    def save(self, path):
        """Surface mesh as legacy-VTK PolyData (.vtk) or Wavefront OBJ (.obj) (api.py:130-148)."""
        v, f = self.mesh
        with open(path, "w") as fh:
            if str(path).lower().endswith(".obj"):
                fh.writelines(f"v {p[0]} {p[1]} {p[2]}\n" for p in v)
                fh.writelines(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n" for t in f)
            else:
                fh.write("# vtk DataFile Version 3.0\npetch surface\nASCII\nDATASET POLYDATA\n")
                fh.write(f"POINTS {len(v)} float\n")
                fh.writelines(f"{p[0]} {p[1]} {p[2]}\n" for p in v)
                fh.write(f"POLYGONS {len(f)} {4 * len(f)}\n")
                fh.writelines(f"3 {int(t[0])} {int(t[1])} {int(t[2])}\n" for t in f)
        return path


class Process:
    """Run `model` on `domain` for `duration` with `steps` etch steps (api.py:151-171)."""

    def __init__(self, domain, model, duration, n_ion=30000, n_neu=30000, transport="mc"):
        self.domain = domain
        self.model = model
        self.duration = float(duration)
        self.n_ion = n_ion
        self.n_neu = n_neu
        self.transport = transport

    def run(self, steps=40, verbose=False):
        flags = Flags(neutral_transport=self.transport, **self.model.flags)
        t0 = time.time()
        geo = _t3.run_etch_3d(t_end=self.duration, n_steps=steps, par=dict(self.model.par),
                              flags=flags, n_ion=self.n_ion, n_neu=self.n_neu,
                              reinit_method="fsm", verbose=verbose, **self.domain._geo)
        return Result(geo, time.time() - t0)
