"""The trace autotuner (petch.autotune): every candidate it chooses among gives the same fluxes
(deterministic mode, so the comparison is to the bit), it applies the fastest it measured, and an
EtchLoop with autotune keeps the choice for its run."""
import numpy as np
import pytest

from oracle import levelset as ol

pytestmark = pytest.mark.gpu


@pytest.fixture
def det():
    from petch import ops
    ops.set_deterministic(True)
    yield
    ops.set_deterministic(False)
    ops.set_trace_occupancy(0)


def _trench_mesh():
    import torch
    from petch import flux, ops
    phi = ol.init_sdf(6.0, 1.5, 12.0, 0.1, 1.5, 7.0, 3.0, False).astype(np.float32)
    dev = torch.device("cuda", 0)
    mesh = flux.SurfaceMesh(dev).build(torch.from_numpy(phi).to(dev), 0.1)
    return mesh, ops.make_box(6.0, 1.5, 12.0, 11.9, 1), dev


def test_candidates_give_identical_fluxes_and_the_fastest_is_applied(det):
    import torch
    from petch import autotune, flux, ops
    mesh, box, dev = _trench_mesh()
    bare = torch.from_numpy(np.random.default_rng(3).uniform(0.2, 1, mesh.nf).astype(np.float32)).to(dev)
    outs = {}
    limit = ops.trace_occupancy_limit(dev)
    assert limit >= 1
    for q in autotune.QUALITIES:
        mesh.bvh.set_quality(q)
        mesh.bvh.build(mesh.verts, mesh.faces)
        for occ in (0, max(limit - 2, 1)):
            ops.set_trace_occupancy(occ, dev)
            outs[(q, occ)] = [x.clone() for x in ops.trace_neutral_rr_pair(
                mesh.bvh, box, flux.NEUTRAL_SOURCE, bare, 0.7, 300000, 5, 1.0, 300000, 6)]
    ref = outs[(0, 0)]
    for k, o in outs.items():
        d = max(float((a - b).abs().sum() / b.abs().sum()) for a, b in zip(o, ref))
        print(f"quality {k[0]} occupancy {k[1]}: L1 vs Morton/full occupancy {d:.1e}", flush=True)
        assert d < 1e-6          # identical but for a hit at exactly equal distance
    tuner = autotune.TraceTuner(reps=1)
    q, occ = tuner.tune(mesh, box, bare, 0.7, 1.0, 300000)
    t = tuner.last
    print({k: tuple(round(x, 3) for x in v) for k, v in t.items()}, "->", (q, occ), flush=True)
    assert (q, occ) == min(t, key=lambda k: t[k][0] + t[k][1])
    assert tuner.tune(mesh, box, bare, 0.7, 1.0, 300000) == (q, occ)     # cached per bucket


def test_etch_loop_keeps_the_tuned_build(det):
    from petch import threed as t3
    from petch.api import SF6O2
    from petch.params import Flags
    m = SF6O2()
    case = dict(Lx=6.0, Ly=1.5, Lz=12.0, dx=0.1, trench_width=1.5, mask_th=1.0, sub_top=10.0,
                hole=False)
    geo = t3._trench_geo(*(case[k] for k in ("Lx", "Ly", "Lz", "dx", "trench_width", "mask_th",
                                             "sub_top", "hole")))
    phi, mask = geo.pop('_phi_d'), geo.pop('_mask_d')
    par = dict(m.par, periodic_y=1)
    loop = t3.EtchLoop(phi, mask, dict(geo), case['mask_th'], 0.3, 3, par, Flags(**m.flags),
                       100000, 100000, autotune=True)
    for s in range(3):
        loop.step(s)
    assert loop.tuned is not None
    assert all(mm.quality == loop.tuned[0] for mm in loop.meshes)
