"""Many features etched in lockstep on one GPU.

The reference runs one feature per run_etch_3d call. Here a BatchLoop advances any number of
features (each its own grid, mask, recipe, ray counts and seeds) step by step: every feature
builds its surface, then the ray traces of all features' coupled fluxes run as ONE launch per
trace (petch_trace_ion_yield_batch, petch_trace_neutral_rr_batch: a persistent ray queue
spanning every feature), then every feature finishes its step (rate, extension, advection,
reinit). The host reads a step needs (warm-start guard, max V, BVH guard) are gathered over all
features into one read each, instead of one per feature (the marching-cubes counts of all
features too), and the warm-start lookups and the velocity extensions of all features run as
one launch each. Each feature's rays draw exactly the Philox streams of its own run_etch_3d,
so a batched feature follows its single run up to the order of the float atomic adds.

    geos = run_etch_3d_batch([dict(Lx=14, Ly=14, Lz=24, dx=0.25, trench_width=6, mask_th=2,
                                   sub_top=18, hole=True, par=p, flags=fl, n_ion=30000,
                                   n_neu=30000), ...], t_end=3.0, n_steps=40)
"""
import time

import torch

from . import flux as _flux
from . import ops
from .threed import EtchLoop, _FluxSteps, _dev, _np, _trench_geo


class BatchLoop:
    """EtchLoops advanced together; their coupled-flux traces are batched per launch."""

    def __init__(self, loops, timer=None):
        from .profiling import NULL
        self.loops = list(loops)
        self.timer = NULL if timer is None else timer

    def step(self, step):
        loops = self.loops
        # 1. every feature's mesh (one host read for all marching-cubes counts) and BVH (the
        #    refits of all trees in one launch); the warm-start guards of all features in one read
        with self.timer.region("mesh"):
            mcs = ops.marching_cubes_batch([lp.mesh_item(step) for lp in loops])
        surf = []
        for lp, mc in zip(loops, mcs):
            lp.stage_mesh(step, mc, surf)
        with self.timer.region("mesh"):
            ops.surface_build_batch(surf)
        finite = [lp.begin_mesh(step, bvh=False, staged=True) for lp in loops]
        with self.timer.region("mesh"):
            ops.bvh_build_batch([lp._mesh.bvh_item() for lp in loops])
        need = [i for i, f in enumerate(finite) if f is not None]
        if need:
            vals = torch.stack([finite[i] for i in need]).cpu().tolist()
            for i, v in zip(need, vals):
                finite[i] = bool(v)
        # the warm-start lookups of all features in one launch
        ws_idx, ws_items = [], []
        for i, lp in enumerate(loops):
            it = lp.warm_start_item(finite[i])
            if it is not None:
                ws_idx.append(i)
                ws_items.append(it)
        warm = dict(zip(ws_idx, ops.nearest_face_value_batch(ws_items, 1.0)))
        outs = [None] * len(loops)
        pending = {}
        rounds = []     # the neutral rounds' ends of all features, run as one launch
        for i, lp in enumerate(loops):
            job = lp.begin_flux(step, finite[i], bare_init=warm.get(i), rounds=rounds)
            if isinstance(job, _FluxSteps):
                try:
                    pending[i] = (job.steps, next(job.steps))
                except StopIteration as stop:
                    outs[i] = stop.value
            else:
                outs[i] = job
        # 2. the coupled fluxes: one launch per trace for all features
        with self.timer.region("flux"):
            while pending:
                for kind in ("ion", "neutral"):
                    idx = [i for i, (_, req) in pending.items() if req[0] == kind]
                    if not idx:
                        continue
                    reqs = [pending[i][1][1] for i in idx]
                    if kind == "ion":
                        with self.timer.region("ion_trace"):
                            res = ops.trace_ion_yield_batch(reqs)
                    else:
                        with self.timer.region("neutral_trace"):
                            res = ops.trace_neutral_rr_batch(_flux.NEUTRAL_SOURCE, reqs)
                    for i, r in zip(idx, res):
                        gen = pending[i][0]
                        try:
                            pending[i] = (gen, gen.send(r))
                        except StopIteration as stop:
                            outs[i] = stop.value
                            del pending[i]
                    if rounds:
                        ops.finish_ion_batch([r for r in rounds if r['kind'] == "ion"])
                        ops.finish_round_batch([r for r in rounds if r['kind'] == "neutral"])
                        rounds.clear()
        # 3. every feature's rate; the rate reductions and BVH guards of all features in one read
        rates = []
        has = [lp.end_rate(step, out, rates=rates) for lp, out in zip(loops, outs)]
        reads = [(0.0, 0, 0)] * len(loops)
        live = [i for i, h in enumerate(has) if h]
        if live:
            buf = torch.empty((len(live), 3), dtype=torch.int32, device=loops[live[0]].dev)
            with self.timer.region("rate"):
                ops.surface_rate_batch(rates, buf)
            for r, i in enumerate(live):          # redeposition lowers V (and its max) first
                if loops[i].redep:
                    loops[i].redeposit(step, buf[r])
            host = buf.cpu().numpy()
            for r, i in enumerate(live):
                vmx, nbad = ops.decode_rate_reduction(host[r, :2])
                reads[i] = (vmx, nbad, int(host[r, 2]))
        # 4. the velocity extension of all features in one pass, every feature's CFL, then the
        #    advection substeps and the reinit sweeps of all features, one launch each
        ext = [lp.extend_item() if h and not rd[2] else None
               for lp, h, rd in zip(loops, has, reads)]
        items = [e for e in ext if e is not None]
        if items:
            with self.timer.region("extend"):
                ops.extend_batch(items)
        deferred = dict(advect=[], reinit=[])
        infos = [lp.end_finish(step, *rd, extended=e is not None, deferred=deferred)
                 for lp, rd, e in zip(loops, reads, ext)]
        with self.timer.region("advect"):
            ops.advect_batch(deferred['advect'])
        with self.timer.region("reinit"):
            ops.reinit_fsm_batch(deferred['reinit'])
        for post in deferred.get('post', ()):     # per-feature work after the batched reinit:
            post()                                # the PDE reinit, then surf_smooth
        return infos


_GEO_KEYS = ("Lx", "Ly", "Lz", "dx", "trench_width", "mask_th", "sub_top", "hole")


def run_etch_3d_batch(features, t_end=2.0, n_steps=20, reinit_method="fsm", timer=None):
    """run_etch_3d for every feature of `features` at once, in lockstep.

    features: dicts of run_etch_3d keyword arguments (geometry Lx, Ly, Lz, dx, trench_width,
    mask_th, sub_top, hole; par, flags, n_ion, n_neu, seed_offset, reinit_every, extend,
    rays_per_point, etch_stop_z, surf_smooth); t_end, n_steps and reinit_method are shared. Returns one
    geo dict per feature, as run_etch_3d does."""
    _dev()
    loops, geos = [], []
    for ft in features:
        unknown = set(ft) - set(_GEO_KEYS) - {"par", "flags", "n_ion", "n_neu", "seed_offset",
                                             "reinit_every", "extend", "rays_per_point",
                                             "etch_stop_z", "surf_smooth"}
        if unknown:
            raise TypeError(f"run_etch_3d_batch: unsupported feature keys {sorted(unknown)}")
        g = {k: ft[k] for k in _GEO_KEYS if k in ft}
        geo = _trench_geo(g.get("Lx", 10.0), g.get("Ly", 4.0), g.get("Lz", 14.0),
                          g.get("dx", 0.4), g.get("trench_width", 4.0), g.get("mask_th", 2.0),
                          g.get("sub_top", 10.0), g.get("hole", False))
        phi_d, mask_d = geo.pop('_phi_d'), geo.pop('_mask_d')
        loops.append(EtchLoop(phi_d, mask_d, dict(geo), g.get("mask_th", 2.0), t_end, n_steps,
                              ft.get("par"), ft.get("flags"), ft.get("n_ion", 20000),
                              ft.get("n_neu", 20000), reinit_every=ft.get("reinit_every", 1),
                              extend=ft.get("extend", "gpu"), reinit_method=reinit_method,
                              seed_offset=ft.get("seed_offset", 0),
                              rays_per_point=ft.get("rays_per_point"),
                              etch_stop_z=ft.get("etch_stop_z"), timer=timer,
                              surf_smooth=ft.get("surf_smooth", 0.0)))
        geos.append(geo)
    batch = BatchLoop(loops, timer)
    t0 = time.time()
    for step in range(n_steps):
        batch.step(step)
    torch.cuda.synchronize()
    total = time.time() - t0
    for geo, lp in zip(geos, loops):
        lp.timings['total'] = total
        geo['phi'] = _np(lp.phi)
        geo['timings'] = lp.timings
    return geos


def seeded_features(n_feat, seed=1234):
    """The seeded features of the batched workload (run_etch_3d_batch keyword dicts): holes and
    trenches alternating, diameter U(2,10) um, mask U(1,3) um, dx 0.25, square lateral extent
    U(12,40) um (48^2..160^2 cells), height U(20,50) um (80..200 cells), ion energy
    U(50,200) eV, O/F flux ratio U(0.1,1.0); 400k rays per feature and step split 1:1:1
    (ion : F : O). Feature f draws from the generator seeded seed + f and runs with
    seed_offset f."""
    import numpy as np
    from .api import SF6O2
    from .params import Flags
    out = []
    for f in range(n_feat):
        r = np.random.default_rng(seed + f)
        d, mask = r.uniform(2, 10), r.uniform(1, 3)
        ext = max(round(r.uniform(12, 40) * 4) / 4, round((d + 4) * 4) / 4)
        Lz = round(r.uniform(20, 50) * 4) / 4
        m = SF6O2(meanE=float(r.uniform(50, 200)))
        par = dict(m.par)
        par['Oflux'] = float(r.uniform(0.1, 1.0)) * par['Fflux']
        hole = f % 2 == 0
        par['periodic_y'] = 0 if hole else 1
        out.append(dict(Lx=ext, Ly=ext, Lz=Lz, dx=0.25, trench_width=d, mask_th=mask,
                        sub_top=Lz - mask - 2.0, hole=hole, par=par, flags=Flags(**m.flags),
                        n_ion=133334, n_neu=133333, seed_offset=f))
    return out
