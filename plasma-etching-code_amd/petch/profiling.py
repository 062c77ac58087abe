"""Stage timers on the device stream: HIP events recorded on torch's current stream (the
stream every petch kernel is enqueued on) around named regions, resolved after a sync.
Disabled timers cost nothing."""
import contextlib
from collections import defaultdict

import torch


class StageTimer:
    def __init__(self, enabled=True):
        self.enabled = enabled
        self._pending = []
        self.ms = defaultdict(float)
        self.count = defaultdict(int)

    @contextlib.contextmanager
    def region(self, name):
        if not self.enabled:
            yield
            return
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        try:
            yield
        finally:
            b.record()
            self._pending.append((name, a, b))

    def resolve(self):
        """Synchronises and folds the recorded regions into ms / count."""
        if not self._pending:
            return self
        torch.cuda.synchronize()
        for name, a, b in self._pending:
            self.ms[name] += a.elapsed_time(b)
            self.count[name] += 1
        self._pending.clear()
        return self

    def reset(self):
        self._pending.clear()
        self.ms.clear()
        self.count.clear()


NULL = StageTimer(enabled=False)
