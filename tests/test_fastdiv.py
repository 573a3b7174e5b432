"""The exact division the kernels use to turn a linear cell number into
(x, y, z): the add-back form of Granlund & Montgomery's round-up reciprocal,
l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, t = hi32(m n),
q = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0), for divisors 1 <= d < 2^32
and every dividend n < 2^32.

Two checks: the method itself against integer division (numpy, exhaustive
for small d over the remainder boundaries, sampled for large d, every power of
two and its neighbours), and the header's own make_fastdiv / fast_div compiled
for the host (tests/fastdiv_check.cpp) over the same kinds of divisors.  The
kernels' use of it is covered by the GPU parity tests on odd extents and by
the domains beyond 2^29 cells in test_gpu_large.py."""
from __future__ import annotations

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TOP = (1 << 32) - 1


def magic(d):
    l = (d - 1).bit_length()
    m = ((((1 << l) - d) << 32) // d) + 1
    return m, min(l, 1), max(l - 1, 0)


def quotient(n, d):
    m, s1, s2 = magic(d)
    assert 0 < m < 1 << 32  # a 32-bit multiplier
    t = (n * np.uint64(m)) >> np.uint64(32)  # n, m < 2^32: the product fits 64 bits
    return (t + ((n - t) >> np.uint64(s1))) >> np.uint64(s2)


def boundaries(d, rng, count=64):
    k = rng.integers(0, TOP // d + 1, size=count, dtype=np.uint64)
    d64 = np.uint64(d)
    n = np.concatenate([k * d64, k * d64 + d64 - np.uint64(1), np.maximum(k * d64, np.uint64(1)) - np.uint64(1),
                        rng.integers(0, TOP, size=count, dtype=np.uint64, endpoint=True),
                        np.array([0, TOP, TOP - 1, (TOP // d) * d, (TOP // d) * d - 1], dtype=np.uint64)])
    return n[n <= np.uint64(TOP)]


def check(d, n):
    assert np.array_equal(quotient(n, d), n // np.uint64(d)), d


def test_small_divisors_all_boundaries():
    rng = np.random.default_rng(7)
    for d in range(1, 4097):
        check(d, boundaries(d, rng))


def test_large_divisors_sampled():
    rng = np.random.default_rng(11)
    ds = [int(x) for x in rng.integers(4097, 1 << 32, size=600)]
    ds += [(1 << k) + o for k in range(12, 33) for o in (-1, 0, 1) if (1 << k) + o < 1 << 32]
    ds += [TOP, TOP - 1, (1 << 31) + 1, 3 << 30]
    for d in ds:
        check(d, boundaries(d, rng))


def test_every_dividend_at_both_ends_for_a_few_divisors():
    # every dividend in the lowest and the highest 2^24, for divisors whose
    # magic numbers sit at the extremes (largest multiplier just above a power
    # of two), where a wrong rounding of the reciprocal shows first
    for d in (3, 641, (1 << 16) + 1, (1 << 31) + 1):
        for lo in (0, (1 << 32) - (1 << 24)):
            check(d, np.arange(lo, lo + (1 << 24), dtype=np.uint64))


def test_header_code_on_the_host(tmp_path):
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++23", "--offload-arch=gfx950", "-x", "hip",
                    os.path.join(HERE, "fastdiv_check.cpp"), "-o", str(exe)], check=True, capture_output=True,
                   timeout=300)
    rng = np.random.default_rng(5)
    ds = list(range(1, 1025)) + [int(x) for x in rng.integers(1025, 1 << 32, size=300)]
    ds += [(1 << k) + o for k in range(10, 33) for o in (-1, 0, 1) if (1 << k) + o < 1 << 32]
    r = subprocess.run([str(exe)], input="\n".join(map(str, ds)) + "\n", capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    assert int(r.stdout.split()[1]) == len(ds)
