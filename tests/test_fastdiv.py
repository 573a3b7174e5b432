"""The division method the kernels use to turn a linear cell number into
(x, y, z): q = (n * m) >> s with l = ceil(log2 d), s = 29 + l,
m = ceil(2^s / d), for divisors d <= 2^29 and dividends n < 2^29.  Checked
here against integer division (exhaustively for small d, at every remainder
boundary for sampled large d); the kernels' use of it is covered by the GPU
parity tests on odd extents."""
from __future__ import annotations

import numpy as np


def magic(d):
    l = (d - 1).bit_length()
    s = 29 + l
    return -(-(1 << s) // d), s


def check(d, n):
    m, s = magic(d)
    assert m <= 1 << 30  # fits the 32-bit multiplier
    q = (n.astype(object) * m) >> s
    assert np.array_equal(q.astype(np.int64), n // d), d


def test_small_divisors_all_boundaries():
    rng = np.random.default_rng(7)
    top = (1 << 29) - 1
    for d in range(1, 2049):
        k = rng.integers(0, top // d, size=64)
        n = np.concatenate([k * d, k * d + d - 1, rng.integers(0, top, size=64), [0, top]])
        check(d, n[n <= top].astype(np.int64))


def test_large_divisors_sampled():
    rng = np.random.default_rng(11)
    top = (1 << 29) - 1
    ds = np.concatenate([rng.integers(2049, 1 << 29, size=400), [(1 << k) + o for k in range(11, 29)
                                                                     for o in (-1, 0, 1)], [1 << 29]])
    for d in ds:
        d = int(d)
        k = rng.integers(0, max(1, top // d), size=32)
        n = np.concatenate([k * d, k * d + d - 1, [top, 0, min(top, d), d - 1]])
        check(d, n[(n <= top) & (n >= 0)].astype(np.int64))
