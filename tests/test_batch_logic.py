"""The lockstep driver on the host (no GPU): every feature's trace requests are served by one
batched call per kind, in order, and each feature gets back its own result."""
import types

import pytest
import torch

from petch import batch as B
from petch.threed import _FluxSteps


class FakeLoop:
    def __init__(self, k, n_fp):
        self.k, self.n_fp, self.log = k, n_fp, []
        self.redep = k == 1                  # this feature has flags.redeposition

    def mesh_item(self, step):
        return ("mc", self.k)

    def stage_mesh(self, step, mc, surf):
        self.mc = mc
        surf.append(("surf", self.k))

    def begin_mesh(self, step, mc=None, bvh=True, staged=False):
        assert staged and mc is None
        self.bvh_here = bvh
        self._mesh = types.SimpleNamespace(bvh_item=lambda: ("bvh", self.k),
                                           bvh=types.SimpleNamespace(status_into=lambda t: None))
        return None

    def warm_start_item(self, finite):
        return ("ws", self.k) if self.k == 1 else None

    def begin_flux(self, step, finite, bare_init=None, rounds=None):
        self.bare_init = bare_init
        self.rounds = rounds
        return self.begin_step(step)

    def end_rate(self, step, out, rates=None):
        self._out = out
        self.dev = torch.device("cpu")
        if out is not None:
            rates.append(("rate", self.k))
        return out is not None

    def redeposit(self, step, red):
        red[0] = 100 + self.k                # stands for the max of the lowered rate

    def extend_item(self):
        return ("ext", self.k)

    def end_finish(self, step, vmx, nbad, err, extended=False, deferred=None):
        self.extended = extended
        deferred['advect'].append(("adv", self.k))
        deferred['reinit'].append(("re", self.k))
        return self.end_step(step, self._out)

    def begin_step(self, step):
        def steps():
            acc = yield ("ion", dict(k=self.k, step=step))
            assert acc == ("ion", self.k)
            self.rounds.append(dict(kind="ion", tag=("ion_finish", self.k)))
            for it in range(self.n_fp):
                r = yield ("neutral", dict(k=self.k, it=it))
                assert r == ("neutral", self.k, it)
                self.rounds.append(dict(kind="neutral", tag=("round", self.k, it)))
            return dict(k=self.k, n=self.n_fp)
        if self.k == 99:
            return None                      # e.g. an empty surface: no flux job
        return _FluxSteps(steps())

    def end_step(self, step, out):
        self.log.append(out)
        return out


def test_lockstep_batches_each_kind(monkeypatch):
    calls = []

    def ion_batch(reqs):
        calls.append(("ion", [r['k'] for r in reqs]))
        return [("ion", r['k']) for r in reqs]

    def neutral_batch(src, reqs):
        calls.append(("neutral", [r['k'] for r in reqs]))
        return [("neutral", r['k'], r['it']) for r in reqs]

    def mc_batch(items):
        calls.append(("mc", list(items)))
        return [("mesh", it[1]) for it in items]

    def ws_batch(items, miss):
        calls.append(("ws", list(items)))
        return [("bare", it[1]) for it in items]

    def ext_batch(items):
        calls.append(("ext", list(items)))

    def bvh_batch(items):
        calls.append(("bvh", list(items)))

    def adv_batch(items):
        calls.append(("adv", list(items)))

    def re_batch(items):
        calls.append(("re", list(items)))

    def round_batch(items):
        calls.append(("round", [it['tag'] for it in items]))

    def ion_finish_batch(items):
        if items:
            calls.append(("ion_finish", [it['tag'] for it in items]))

    monkeypatch.setattr(B.ops, "trace_ion_yield_batch", ion_batch)
    monkeypatch.setattr(B.ops, "trace_neutral_rr_batch", neutral_batch)
    monkeypatch.setattr(B.ops, "marching_cubes_batch", mc_batch)
    monkeypatch.setattr(B.ops, "nearest_face_value_batch", ws_batch)
    monkeypatch.setattr(B.ops, "extend_batch", ext_batch)
    monkeypatch.setattr(B.ops, "bvh_build_batch", bvh_batch)
    monkeypatch.setattr(B.ops, "surface_build_batch",
                        lambda items: calls.append(("surf", list(items))))
    monkeypatch.setattr(B.ops, "advect_batch", adv_batch)
    monkeypatch.setattr(B.ops, "reinit_fsm_batch", re_batch)
    monkeypatch.setattr(B.ops, "finish_round_batch", round_batch)
    monkeypatch.setattr(B.ops, "finish_ion_batch", ion_finish_batch)
    decoded = []
    monkeypatch.setattr(B.ops, "decode_rate_reduction",
                        lambda r: decoded.append(int(r[0])) or (1.0, 0))

    def rate_batch(items, red):
        calls.append(("rate", list(items)))
        red.zero_()
    monkeypatch.setattr(B.ops, "surface_rate_batch", rate_batch)
    loops = [FakeLoop(0, 1), FakeLoop(1, 4), FakeLoop(99, 0), FakeLoop(2, 1)]
    outs = B.BatchLoop(loops).step(0)
    # one marching-cubes pass, one face attribute/adjacency build and one BVH build call for
    # all features, one warm-start launch for the features that have a lookup, one launch per
    # trace kind and fixed-point round, each neutral round ended for all its features in one
    # call before the next trace, one surface-rate launch and one extension pass for every
    # feature with a surface, then the advection and the reinit of every feature, one call each
    every = lambda tag: [(tag, k) for k in (0, 1, 99, 2)]
    assert calls == [("mc", every("mc")), ("surf", every("surf")), ("bvh", every("bvh")),
                     ("ws", [("ws", 1)]), ("ion", [0, 1, 2]),
                     ("ion_finish", [("ion_finish", 0), ("ion_finish", 1), ("ion_finish", 2)]),
                     ("round", []), ("neutral", [0, 1, 2]),
                     ("round", [("round", 0, 0), ("round", 1, 0), ("round", 2, 0)]),
                     ("neutral", [1]), ("round", [("round", 1, 1)]),
                     ("neutral", [1]), ("round", [("round", 1, 2)]),
                     ("neutral", [1]), ("round", [("round", 1, 3)]),
                     ("rate", [("rate", 0), ("rate", 1), ("rate", 2)]),
                     ("ext", [("ext", 0), ("ext", 1), ("ext", 2)]),
                     ("adv", every("adv")), ("re", every("re"))]
    assert not any(lp.bvh_here for lp in loops)
    assert [lp.mc for lp in loops] == [("mesh", 0), ("mesh", 1), ("mesh", 99), ("mesh", 2)]
    assert [lp.bare_init for lp in loops] == [None, ("bare", 1), None, None]
    assert [lp.extended for lp in loops] == [True, True, False, True]
    # the redepositing feature lowers its rate (and the max the CFL reads) after the batched
    # rate launch, before the host read
    assert decoded == [0, 101, 0]
    assert outs == [dict(k=0, n=1), dict(k=1, n=4), None, dict(k=2, n=1)]


def test_batch_rejects_unknown_feature_keys(monkeypatch):
    monkeypatch.setattr(B, "_dev", lambda: None)
    with pytest.raises(TypeError):
        B.run_etch_3d_batch([dict(Lx=1.0, bogus=3)])
