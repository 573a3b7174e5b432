"""GPU: lbc_reactive_stats_noalloc (statistics with the partial sums in the
library's static device array) returns what lbc_reactive_stats returns, bit
for bit, through the C ABI, on two calls in a row whose second input is
smaller (the array keeps the first call's partials; every slot the second call
reads must be its own), in the 2D and 3D lattices and every storage format,
and both agree with the oracle's extended-precision sums."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import cases
from lbcube import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CODES = {"D2Q5": 3, "D3Q7": 4}
STORE = {"f32": 0, "f64": 1, "f16": 2, "bf16": 3}
TORCH = {"f32": torch.float32, "f64": torch.float64, "f16": torch.float16, "bf16": torch.bfloat16}


def species(lattice, shape, dtype, seed, ref):
    g = cases.rng(seed)
    d = O.D_OF[lattice]
    u = g.uniform(-0.03, 0.03, size=(d,) + shape)
    ca = g.uniform(0.0, 1.0, size=shape)
    a = O.scalar_feq_field(lattice, dtype, ca, u, ref)
    b = O.scalar_feq_field(lattice, dtype, 1.0 - ca + 0.1 * g.uniform(size=shape), u, ref)
    return a, b


def call(fn, lattice, dtype, a, b, shape, k, ref):
    nz = shape[0] if len(shape) == 3 else 1
    ny, nx = shape[-2], shape[-1]
    out = (ctypes.c_double * 8)()
    rc = fn(CODES[lattice], STORE[dtype], a.data_ptr(), b.data_ptr(), nx, ny, nz, k, ref, ref, out,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, rc
    return np.array(out[:])


@pytest.mark.parametrize("lattice,shapes", [("D2Q5", [(300, 517), (37, 45)]), ("D3Q7", [(40, 48, 66), (5, 7, 9)])])
@pytest.mark.parametrize("dtype", ["f32", "f64", "f16", "bf16"])
def test_noalloc_stats_equal_allocating_entry_point(gpu, lattice, shapes, dtype):
    lib = _lib.lib()
    ref = 0.5 if dtype in ("f16", "bf16") else 0.0
    k = 0.05
    for call_no, shape in enumerate(shapes):
        print(f"case noalloc stats {lattice} {dtype} call {call_no} shape {shape}", flush=True)
        a_np, b_np = species(lattice, shape, dtype, seed=100 + call_no, ref=ref)
        a = torch.from_numpy(a_np.view(np.int16) if a_np.dtype == np.uint16 else a_np).to(gpu)
        b = torch.from_numpy(b_np.view(np.int16) if b_np.dtype == np.uint16 else b_np).to(gpu)
        if dtype in ("f16", "bf16"):
            a, b = a.view(TORCH[dtype]), b.view(TORCH[dtype])
        got = call(lib.lbc_reactive_stats_noalloc, lattice, dtype, a, b, shape, k, ref)
        want = call(lib.lbc_reactive_stats, lattice, dtype, a, b, shape, k, ref)
        assert np.array_equal(got, want), (got, want)
        wide = O.reactive_stats(lattice, dtype, a_np, b_np, k, wide=True, ref_a=ref, ref_b=ref)
        assert np.allclose(got, wide, rtol=1e-12, atol=1e-15)


def test_noalloc_stats_rejects_what_the_allocating_one_rejects(gpu):
    lib = _lib.lib()
    a = torch.zeros((5, 8, 8), dtype=torch.float32, device=gpu)
    out = (ctypes.c_double * 8)()
    for args in ((3, 0, None, a.data_ptr(), 8, 8, 1), (3, 0, a.data_ptr(), a.data_ptr(), 8, 0, 1),
                 (9, 0, a.data_ptr(), a.data_ptr(), 8, 8, 1), (3, 7, a.data_ptr(), a.data_ptr(), 8, 8, 1)):
        rc_new = lib.lbc_reactive_stats_noalloc(*args, 0.1, 0.0, 0.0, out, None)
        rc_old = lib.lbc_reactive_stats(*args, 0.1, 0.0, 0.0, out, None)
        assert rc_new == rc_old != 0
