"""Launch-configuration autotuner for the ray traces. No reference counterpart: the reference's
Warp launches take Warp's launch shapes and its BVH (threed.py:1674-1675, 876-880).

Two choices are tuned, per surface-size and ray-budget bucket (powers of two of the face count
and of the rays per trace):
- the BVH build (petch_bvh_set_quality): the Morton tree with the fixed 4-wide collapse (0),
  with the area-driven collapse (1), or the clustered (PLOC) tree with the area-driven collapse
  (2). A better tree costs more to build and saves trace time; a run rebuilds its tree every
  step, so a candidate's score is its build time plus its trace time;
- the persistent trace grid (petch_set_trace_occupancy): workgroups per CU, from the kernels'
  occupancy limit down by one and two (fewer resident rays share the caches among fewer).
Every candidate gives the same fluxes: the tree only changes the order in which a query meets
the boxes (the closest hit is the same; a hit at exactly equal distance from two faces aside),
and a ray's Philox stream is keyed by its id, not by the wavefront that takes it. So the choice
is made by speed alone, from HIP-event times of the caller's own surface and ray counts (the
paired F and O neutral trace of one fixed-point round, after one untimed launch per candidate).

    tuner = TraceTuner()
    q, occ = tuner.tune(mesh, box, bare, betaE, betaO, n_neu)      # applies and returns the choice

EtchLoop(..., autotune=True) tunes on the first step's surface and keeps the choice for the run.
"""
import math

import torch

from . import ops

QUALITIES = (0, 1, 2)


class TraceTuner:
    """Measured choice of BVH build quality and trace occupancy, cached per bucket."""

    def __init__(self, reps=2, device=None):
        self.reps = int(reps)
        self.device = device
        self.choices = {}      # bucket -> (quality, blocks_per_cu)
        self.last = None       # timings of the last tuning: {(quality, occ): (build_ms, trace_ms)}

    @staticmethod
    def bucket(nf, n_rays):
        return (int(math.log2(max(nf, 1))), int(math.log2(max(n_rays, 1))))

    def tune(self, mesh, box, bare, beta_a, beta_b, n_rays, seed=12345):
        """Times every candidate on `mesh` (a flux.SurfaceMesh with verts/faces) with a paired
        neutral trace of n_rays rays per species, applies the fastest (the mesh's BVH is rebuilt
        with its quality; the context's occupancy is set) and returns (quality, blocks_per_cu)."""
        key = self.bucket(mesh.nf, n_rays)
        if key in self.choices:
            return self._apply(mesh, *self.choices[key])
        dev = mesh.device
        limit = ops.trace_occupancy_limit(dev)
        occs = sorted({0} | {k for k in (limit - 1, limit - 2) if k >= 1}, reverse=False)
        from .flux import NEUTRAL_SOURCE
        timings = {}
        stream = torch.cuda.current_stream(dev)
        for q in QUALITIES:
            mesh.bvh.set_quality(q)
            mesh.bvh.build(mesh.verts, mesh.faces)          # untimed: allocation on first use
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            mesh.bvh.build(mesh.verts, mesh.faces)
            b.record(stream)
            for occ in occs:
                ops.set_trace_occupancy(occ, dev)
                ops.trace_neutral_rr_pair(mesh.bvh, box, NEUTRAL_SOURCE, bare, beta_a, n_rays,
                                          seed, beta_b, n_rays, seed + 1)
                c, d = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                c.record(stream)
                for _ in range(self.reps):
                    ops.trace_neutral_rr_pair(mesh.bvh, box, NEUTRAL_SOURCE, bare, beta_a, n_rays,
                                              seed, beta_b, n_rays, seed + 1)
                d.record(stream)
                d.synchronize()
                timings[(q, occ)] = (a.elapsed_time(b), c.elapsed_time(d) / self.reps)
        best = min(timings, key=lambda k: timings[k][0] + timings[k][1])
        self.choices[key] = best
        self.last = timings
        return self._apply(mesh, *best)

    @staticmethod
    def _apply(mesh, quality, occ):
        ops.set_trace_occupancy(occ, mesh.device)
        mesh.bvh.set_quality(quality)
        mesh.bvh.build(mesh.verts, mesh.faces)
        return quality, occ
