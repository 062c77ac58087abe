"""The invariant-divisor division of csrc/fastdiv.h (Granlund & Montgomery 1994, Fig. 4.1),
restated with 64-bit numpy integers, equals floor division for every divisor of a grid axis
and 32-bit dividends at the edges and at random; a multiplier off by one fails."""
import numpy as np


def _magic(d):
    if d <= 1:
        return 0, 0
    l = 0
    while (1 << l) < d:
        l += 1
    return ((1 << 32) * ((1 << l) - d)) // d + 1, l


def _div(n, d, m=None):
    mm, l = _magic(d)
    m = mm if m is None else m
    if l == 0:
        return n
    t = (n * np.uint64(m)) >> np.uint64(32)
    return (t + ((n - t) >> np.uint64(1))) >> np.uint64(l - 1)


def _dividends(rng):
    edge = np.array([0, 1, 2, 3, 2 ** 16 - 1, 2 ** 16, 2 ** 24 + 1, 2 ** 31 - 1, 2 ** 31,
                     2 ** 32 - 2, 2 ** 32 - 1], np.uint64)
    return np.concatenate([edge, rng.integers(0, 2 ** 32, 20000, dtype=np.uint64)])


def test_fastdiv_equals_floor_division():
    rng = np.random.default_rng(0)
    n = _dividends(rng)
    for d in list(range(1, 2050)) + [4095, 4096, 4097, 65535, 65536, 1 << 20, 3 ** 13,
                                     2 ** 31 - 1, 2 ** 31, 2 ** 32 - 1]:
        m, _ = _magic(d)
        assert m < 2 ** 32
        # around multiples of d, where an off-by-one shows first
        k = rng.integers(0, (2 ** 32 - 1) // d + 1, 200, dtype=np.uint64)
        near = np.concatenate([k * np.uint64(d), k * np.uint64(d) - np.uint64(1)])
        near = near[near < np.uint64(2 ** 32)]
        x = np.concatenate([n, near])
        np.testing.assert_array_equal(_div(x, d), x // np.uint64(d), err_msg=f"d={d}")


def test_fastdiv_detects_a_wrong_multiplier():
    n = _dividends(np.random.default_rng(1))
    bad = [d for d in (3, 7, 160, 200, 1000) if
           not np.array_equal(_div(n, d, _magic(d)[0] - 1), n // np.uint64(d))]
    assert len(bad) == 5
