"""The plain small hole end to end on the HIP path against the oracle's seed distributions.

Domain.hole(extent=14, dx=0.25, diameter=d, mask=2, depth=18) (56 x 56 x 96), SF6O2 defaults,
200k ions and 200k F + 200k O neutrals per fixed-point iteration (n_fp 4 cold, 1 warm), 40 steps
over 3.0 s, fsm reinit: the run of gpu_3d_validate.py:23-28 at the reference's validation ray
counts. The oracle's final depths for eight seeds and the diameters 3, 4, 6 and 8 um were
written once on the CPU by tests/golden/make_hole_depths.py (the oracle takes about half a minute
per run there); this test runs the same cases on the GPU and compares the distributions:

- the sub-cell floor depth on the centre column, mean within 3 standard errors of the
  difference (each side's spread floored at 0.02 dx);
- the center and max depths (whole cells, threed.py:1798-1869), means within one cell;
- the normalised ARDE, depth(d) / depth(widest) (harness/benchmark.py:75-82,
  gpu_3d_validate.py:54-56), within 3 standard errors of the difference."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "hole_depths.json")
sys.path.insert(0, os.path.join(HERE, "golden"))


def _load():
    with open(FIXTURE) as fh:
        return json.load(fh)


def _run_gpu(fx, d, seed_offset):
    from make_hole_depths import floor_depth
    from petch import threed as t3
    from petch.params import PAR, Flags
    par = dict(PAR)
    par['rate_scale'] = fx['rate_scale']
    g = t3.run_etch_3d(**fx['case'], trench_width=float(d), par=par, flags=Flags(**fx['flags']),
                       n_ion=fx['n_ion'], n_neu=fx['n_neu'], reinit_method="fsm", verbose=False,
                       seed_offset=seed_offset)
    return dict(center=t3.center_depth_3d(g), max=t3.max_depth_3d(g),
                floor=floor_depth(g['phi'], fx['case']['dx'], fx['case']['sub_top']))


def _se_diff(a, b, floor):
    sa = max(np.std(a, ddof=1), floor)
    sb = max(np.std(b, ddof=1), floor)
    return np.sqrt(sa ** 2 / len(a) + sb ** 2 / len(b))


def test_hole_depths_and_arde_match_the_oracle_seed_spread():
    fx = _load()
    dx = fx['case']['dx']
    diam = sorted(fx['depths'], key=float)
    gpu, ora = {}, {}
    for d in diam:
        rows = fx['depths'][d]
        gpu[d] = [_run_gpu(fx, d, so) for so in fx['seed_offsets'][:len(rows)]]
        ora[d] = rows
        g_fl = np.array([r['floor'] for r in gpu[d]])
        o_fl = np.array([r['floor'] for r in ora[d]])
        se = _se_diff(g_fl, o_fl, 0.02 * dx)
        g_c, o_c = np.mean([r['center'] for r in gpu[d]]), np.mean([r['center'] for r in ora[d]])
        g_m, o_m = np.mean([r['max'] for r in gpu[d]]), np.mean([r['max'] for r in ora[d]])
        print(f"d {d}: floor gpu {g_fl.mean():.4f}+-{g_fl.std():.4f} oracle {o_fl.mean():.4f}"
              f"+-{o_fl.std():.4f} ({(g_fl.mean() - o_fl.mean()) / se:+.2f} se); center "
              f"{g_c:.3f}/{o_c:.3f}; max {g_m:.3f}/{o_m:.3f}", flush=True)
        assert np.all(np.isfinite(g_fl))
        assert abs(g_fl.mean() - o_fl.mean()) <= 3 * se
        assert abs(g_c - o_c) <= dx and abs(g_m - o_m) <= dx
    # normalised ARDE against the widest diameter, per seed pair of means (delta method)
    wide = diam[-1]
    for d in diam[:-1]:
        def ratio(side):
            num = np.array([r['floor'] for r in side[d]])
            den = np.array([r['floor'] for r in side[wide]])
            R = num.mean() / den.mean()
            var = R ** 2 * (num.var(ddof=1) / len(num) / num.mean() ** 2 +
                            den.var(ddof=1) / len(den) / den.mean() ** 2)
            return R, var
        Rg, vg = ratio(gpu)
        Ro, vo = ratio(ora)
        se = np.sqrt(vg + vo + (0.02 * dx / 8.0) ** 2)
        print(f"ARDE d {d}/{wide}: gpu {Rg:.4f} oracle {Ro:.4f} ({(Rg - Ro) / se:+.2f} se)",
              flush=True)
        assert abs(Rg - Ro) <= 3 * se
