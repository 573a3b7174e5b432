"""Golden FP64 fixtures (tests/golden/golden_fp64.npz, written by
tests/golden/make_golden.py from the oracle): the oracle must keep reproducing
them bit for bit, and so must the HIP kernels."""
from __future__ import annotations

import os

import numpy as np
import pytest

import cases
from golden import make_golden as G
from oracle import oracle as O

FIX = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_fp64.npz"))
FLUID = list(G.fluid_cases())


@pytest.mark.parametrize("lat,coll,shape", FLUID)
def test_oracle_reproduces_fluid_fixture(lat, coll, shape):
    f0 = G.fluid_input(lat, shape)
    assert str(FIX[f"in_sha/{lat}/{coll}"]) == G.sha(f0)
    assert np.array_equal(O.fluid_run(lat, coll, "f64", f0, G.OMEGA, G.STEPS), FIX[f"fluid/{lat}/{coll}"])


def test_oracle_reproduces_reactive_fixture():
    f, a, b = G.reactive_inputs()
    rf, ra, rb = O.reactive_run("D2Q9", "BGK", "D2Q5", "f64", f, a, b, 1 / 0.8, 1 / 0.8, 0.05, G.STEPS)
    assert np.array_equal(rf, FIX["reactive/f"])
    assert np.array_equal(ra, FIX["reactive/a"]) and np.array_equal(rb, FIX["reactive/b"])
    assert np.allclose(O.reactive_stats("D2Q5", "f64", ra, rb, 0.05, wide=True), FIX["reactive/stats"], rtol=0,
                       atol=1e-16)


@pytest.mark.gpu
@pytest.mark.parametrize("lat,coll,shape", FLUID)
def test_kernels_reproduce_fluid_fixture(gpu, lat, coll, shape):
    import torch

    import lbcube

    a = torch.from_numpy(G.fluid_input(lat, shape)).to(gpu)
    b = torch.empty_like(a)
    for _ in range(G.STEPS):
        lbcube.fluid_step(lat, coll, a, b, G.OMEGA)
        a, b = b, a
    assert np.array_equal(a.cpu().numpy(), FIX[f"fluid/{lat}/{coll}"])


@pytest.mark.gpu
def test_kernels_reproduce_reactive_fixture(gpu):
    import torch

    import lbcube

    f, a, b = (torch.from_numpy(x).to(gpu) for x in G.reactive_inputs())
    ft, at, bt = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    for _ in range(G.STEPS):
        lbcube.fluid_step("D2Q9", "BGK", f, ft, 1 / 0.8)
        lbcube.reaction_step("D2Q9", "D2Q5", a, at, b, bt, ft, 1 / 0.8, 0.05)
        f, ft, a, at, b, bt = ft, f, at, a, bt, b
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy(), FIX["reactive/f"])
    # species go through exp(): device and host exp may differ in the last ulp
    assert np.max(np.abs(a.cpu().numpy() - FIX["reactive/a"])) < 1e-14
    assert np.max(np.abs(b.cpu().numpy() - FIX["reactive/b"])) < 1e-14
    s = lbcube.reactive_stats("D2Q5", a, b, 0.05)
    got = np.array([s[k] for k in lbcube.STAT_NAMES])
    np.testing.assert_allclose(got, FIX["reactive/stats"], rtol=1e-10, atol=1e-15)
