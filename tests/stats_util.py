"""Per-face statistical comparison of two Monte-Carlo estimators (batch means)."""
import numpy as np


def batch_stats(batches):
    """Mean and standard error per face from an array of independent batches (B, F)."""
    b = np.asarray(batches, np.float64)
    return b.mean(0), b.std(0, ddof=1) / np.sqrt(len(b))


def zscores(a_batches, b_batches, floor=1e-12):
    ma, sa = batch_stats(a_batches)
    mb, sb = batch_stats(b_batches)
    se = np.sqrt(sa ** 2 + sb ** 2)
    ok = se > floor
    return (ma - mb)[ok] / se[ok], ok
