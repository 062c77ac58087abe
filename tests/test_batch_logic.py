"""The lockstep driver on the host (no GPU): every feature's trace requests are served by one
batched call per kind, in order, and each feature gets back its own result."""
import pytest

from petch import batch as B
from petch.threed import _FluxSteps


class FakeLoop:
    def __init__(self, k, n_fp):
        self.k, self.n_fp, self.log = k, n_fp, []

    def begin_mesh(self, step):
        return None

    def begin_flux(self, step, finite):
        return self.begin_step(step)

    def end_rate(self, step, out):
        self._out = out
        return False

    def end_finish(self, step, vmx, nbad, err):
        return self.end_step(step, self._out)

    def begin_step(self, step):
        def steps():
            acc = yield ("ion", dict(k=self.k, step=step))
            assert acc == ("ion", self.k)
            for it in range(self.n_fp):
                r = yield ("neutral", dict(k=self.k, it=it))
                assert r == ("neutral", self.k, it)
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

    monkeypatch.setattr(B.ops, "trace_ion_yield_batch", ion_batch)
    monkeypatch.setattr(B.ops, "trace_neutral_rr_batch", neutral_batch)
    loops = [FakeLoop(0, 1), FakeLoop(1, 4), FakeLoop(99, 0), FakeLoop(2, 1)]
    outs = B.BatchLoop(loops).step(0)
    assert calls == [("ion", [0, 1, 2]), ("neutral", [0, 1, 2]), ("neutral", [1]),
                     ("neutral", [1]), ("neutral", [1])]
    assert outs == [dict(k=0, n=1), dict(k=1, n=4), None, dict(k=2, n=1)]


def test_batch_rejects_unknown_feature_keys(monkeypatch):
    monkeypatch.setattr(B, "_dev", lambda: None)
    with pytest.raises(TypeError):
        B.run_etch_3d_batch([dict(Lx=1.0, bogus=3)])
