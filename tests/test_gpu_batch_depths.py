"""The batched workload end to end against the oracle: four of the 64 seeded features
(petch.batch.seeded_features: the hole 16 and the trenches 21, 49, 51), six seeds each, run as
ONE batch of 24 runs for all 40 steps (3.0 s) at the workload's 400k rays per feature and step.
The oracle's final depths of the same runs were written once on the CPU by
tests/golden/make_batch_depths.py. Per feature: the sub-cell floor depth on the centre column,
mean within 3 standard errors of the difference (each side's spread floored at 0.02 dx); the
center and max depths (whole cells) within one cell."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def test_batched_features_match_the_oracle_seed_spread():
    from make_batch_depths import GEO_KEYS, N_STEPS, T_END, cases, depths
    from petch import batch as pb
    from petch import threed as t3
    with open(os.path.join(HERE, "golden", "batch_depths.json")) as fh:
        fx = json.load(fh)
    runs = cases()
    for f, s, ft in runs:
        assert ft['seed_offset'] == fx['seed_offsets'][str(f)][s]
    print(f"one batch of {len(runs)} runs, {N_STEPS} steps", flush=True)
    geos = pb.run_etch_3d_batch([ft for _, _, ft in runs], t_end=T_END, n_steps=N_STEPS)
    got = {}
    for (f, s, ft), g in zip(runs, geos):
        got.setdefault(str(f), []).append(depths(g, ft, t3.center_depth_3d, t3.max_depth_3d))
    for f in fx['features']:
        dx = runs[0][2]['dx']
        g_rows, o_rows = got[str(f)], fx['depths'][str(f)]
        gf = np.array([r['floor'] for r in g_rows])
        of = np.array([r['floor'] for r in o_rows])
        sg, so = max(gf.std(ddof=1), 0.02 * dx), max(of.std(ddof=1), 0.02 * dx)
        se = np.sqrt(sg ** 2 / len(gf) + so ** 2 / len(of))
        gc, oc = (np.mean([r['center'] for r in rows]) for rows in (g_rows, o_rows))
        gm, om = (np.mean([r['max'] for r in rows]) for rows in (g_rows, o_rows))
        print(f"feature {f}: floor gpu {gf.mean():.4f}+-{gf.std():.4f} oracle {of.mean():.4f}"
              f"+-{of.std():.4f} ({(gf.mean() - of.mean()) / se:+.2f} se); center {gc:.3f}/"
              f"{oc:.3f}; max {gm:.3f}/{om:.3f}", flush=True)
        assert np.all(np.isfinite(gf))
        assert abs(gf.mean() - of.mean()) <= 3 * se
        assert abs(gc - oc) <= dx and abs(gm - om) <= dx
