"""GPU: the discrete adjoint of the 3D reactive loop (lbc_reactive_gradient_3d):
a D3Q19 or D3Q27 fluid (BGK, TRT, MRT, RLBM) with two D3Q7 species, A + B -> C.

  * against the oracle's FP64 CPU adjoint of the same loop (tight tolerance;
    checkpoint intervals 1, 5 and 8 over 17 steps, so the segment recompute
    and its overlap with the adjoint walk are exercised);
  * blocks of 1024 work-items (the kernels' second instantiation), 2D and 3D,
    every collision, against the same oracle;
  * the FP32 adjoint, and the FP64 adjoint of the FP32 trajectory, against the
    FP64 gradient;
  * checkpoint-interval invariance, bit for bit;
  * torch.autograd.gradcheck of the autograd wrapper in FP64 on a tiny grid;
  * a 40^3 x 200-step run of a sheared state: FP64 central differences
    along random directions of the initial populations and of tau_s.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import test_oracle_adjoint as TA
import test_oracle_adjoint_3d as TA3
from lbcube import adjoint as ADJ
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(x, gpu, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(gpu).to(dtype)


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
@pytest.mark.parametrize("interval", [1, 5, 8])
def test_adjoint_3d_matches_oracle_fp64(gpu, lattice, coll, interval):
    print(f"case adjoint 3d {lattice} {coll} interval={interval}", flush=True)
    f, a, b = TA3.small_state_3d(lattice, nx=7, ny=6, nz=5, seed=4)
    omega, omega_c, k, steps = 1 / 0.7, 1 / 0.9, 0.4, 17
    J, gf, ga, gb, dw = O.reactive_gradient(coll, "f64", f, a, b, omega, omega_c, k, steps, lattice=lattice)
    Jg, gfg, gag, gbg, dwg = ADJ.reactive_gradient(coll, dev(f, gpu), dev(a, gpu), dev(b, gpu), steps, omega, omega_c,
                                                   k, interval=interval, lattice=lattice)
    # J sums 7 x 210 species populations twice (the device in double, the
    # oracle in long double): the two agree to the sums' rounding
    assert abs(Jg - J) <= 1e-13 * abs(J)
    for got, want in ((gfg, gf), (gag, ga), (gbg, gb)):
        g = got.cpu().numpy()
        assert np.max(np.abs(g - want)) <= 1e-12 * np.max(np.abs(want))
    assert abs(dwg - dw) <= 1e-11 * abs(dw)


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
@pytest.mark.parametrize("lattice", ["D2Q9", "D3Q19", "D3Q27"])
def test_adjoint_blocks_of_1024_match_oracle(gpu, lattice, coll):
    """The kernels' 1024-work-item instantiations (the largest block the
    reference's launchers accept; they spill registers to scratch) against
    the oracle's FP64 adjoint, with checkpoints every 5 steps: blocks of
    1024 x 1 x 1 and, in 3D, 16 x 16 x 4."""
    print(f"case adjoint blocks of 1024 {lattice} {coll}", flush=True)
    if lattice == "D2Q9":
        f, a, b = TA.small_state(nx=40, ny=30, seed=4)
        blocks = [(1024, 1, 1), (32, 32, 1)]
        kw = {}
    else:
        f, a, b = TA3.small_state_3d(lattice, nx=12, ny=10, nz=9, seed=4)
        blocks = [(1024, 1, 1), (16, 16, 4)]
        kw = {"lattice": lattice}
    omega, omega_c, k, steps = 1 / 0.7, 1 / 0.9, 0.4, 11
    J, gf, ga, gb, dw = O.reactive_gradient(coll, "f64", f, a, b, omega, omega_c, k, steps, **kw)
    for block in blocks:
        Jg, gfg, gag, gbg, dwg = ADJ.reactive_gradient(coll, dev(f, gpu), dev(a, gpu), dev(b, gpu), steps, omega,
                                                       omega_c, k, interval=5, block=block, **kw)
        assert abs(Jg - J) <= 1e-13 * abs(J), block
        for got, want in ((gfg, gf), (gag, ga), (gbg, gb)):
            assert np.max(np.abs(got.cpu().numpy() - want)) <= 1e-12 * np.max(np.abs(want)), block
        assert abs(dwg - dw) <= 1e-11 * abs(dw), block


@pytest.mark.parametrize("coll", ["TRT", "RLBM"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_adjoint_3d_fp32_against_fp64(gpu, lattice, coll):
    """The FP32 adjoint is as close to the FP64 gradient as the oracle's own
    FP32 adjoint is; the FP64 adjoint of the FP32 trajectory no farther."""
    print(f"case adjoint 3d fp32 {lattice} {coll}", flush=True)
    f, a, b = TA3.small_state_3d(lattice, nx=7, ny=6, nz=5, seed=6)
    f, a, b = (x.astype(np.float32) for x in (f, a, b))
    omega, omega_c, k, steps = 1 / 0.7, 1 / 0.9, 0.4, 17
    J64, *g64 = O.reactive_gradient(coll, "f64", *(x.astype(np.float64) for x in (f, a, b)), omega, omega_c, k,
                                    steps, lattice=lattice)
    J32, *g32 = O.reactive_gradient(coll, "f32", f, a, b, omega, omega_c, k, steps, lattice=lattice)
    args = (dev(f, gpu, torch.float32), dev(a, gpu, torch.float32), dev(b, gpu, torch.float32), steps, omega,
            omega_c, k)
    Jg, gfg, gag, gbg, dwg = ADJ.reactive_gradient(coll, *args, interval=4, lattice=lattice)
    mixed = ADJ.reactive_gradient(coll, *args, interval=4, lattice=lattice, adjoint_f64=True)
    assert abs(Jg - J32) <= 1e-6 and abs(Jg - J64) <= 1e-6
    for got, gm, orc, exact in zip((gfg, gag, gbg), mixed[1:4], g32[:3], g64[:3]):
        scale = np.max(np.abs(exact))
        e_gpu = np.max(np.abs(got.cpu().numpy() - exact)) / scale
        e_mix = np.max(np.abs(gm.cpu().numpy() - exact)) / scale
        e_orc = np.max(np.abs(orc - exact)) / scale
        print(f"{lattice} {coll} fp32 adjoint: gpu {e_gpu:.2e}, FP64 adjoint of the FP32 run {e_mix:.2e}, oracle FP32 "
              f"{e_orc:.2e} (relative to the FP64 gradient)", flush=True)
        assert e_gpu <= max(4 * e_orc, 1e-5)
        assert e_mix <= 1.05 * e_gpu + 1e-7
    assert abs(dwg - g64[3]) <= max(4 * abs(g32[3] - g64[3]), 1e-5 * abs(g64[3]))


def sheared_state(n, lattice, gpu):
    """A shear layer in x with a small transverse wave, species A and B on
    either side of a smoothed plane, all at equilibrium (FP64)."""
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.stack([0.04 * np.tanh((y - n / 2) / 4.0), 0.004 * np.sin(2 * np.pi * x / n),
                  0.004 * np.cos(2 * np.pi * z / n)])
    rho = np.ones((n, n, n))
    f = O.feq_field(lattice, "f64", rho, u)
    ca = 0.5 * (1 + np.tanh((x - n / 2) / 3.0))
    a = O.scalar_feq_field("D3Q7", "f64", ca, u)
    b = O.scalar_feq_field("D3Q7", "f64", 1.0 - ca, u)
    return dev(f, gpu), dev(a, gpu), dev(b, gpu)


@pytest.mark.parametrize("coll", ["TRT", "MRT", "RLBM"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_adjoint_3d_central_differences(gpu, lattice, coll):
    """40^3, 200 steps, checkpoints every 32: directional derivatives along 6
    random directions (all populations) and dJ/d omega_c against FP64
    central differences of the forward loop."""
    print(f"case adjoint 3d central differences {lattice} {coll}", flush=True)
    n, steps, omega, omega_c, k = 40, 200, 1 / 0.6, 1 / 0.8, 0.05
    f, a, b = sheared_state(n, lattice, gpu)
    J, gf, ga, gb, dw = ADJ.reactive_gradient(coll, f, a, b, steps, omega, omega_c, k, interval=32, lattice=lattice)
    assert abs(J - ADJ.loss_of(coll, f, a, b, steps, omega, omega_c, k, lattice=lattice)) <= 1e-14
    gen = torch.Generator(device="cpu").manual_seed(7)
    worst = 0.0
    for trial in range(6):
        vf, va, vb = (torch.randn(x.shape, generator=gen, dtype=torch.float64).to(gpu) * 1e-3 for x in (f, a, b))
        # h = 1e-2 (1e-5 per population): the rounding of the loss falls as
        # 1 / h and the truncation shows only from 3e-2 on.  RLBM's loss is the
        # noisiest (its f - feq cancels): |fd - adjoint| 5e-12 at h = 1e-3,
        # 2e-13 at 3e-3, below 3e-13 for every collision at 1e-2.
        h = 1e-2
        jp = ADJ.loss_of(coll, f + h * vf, a + h * va, b + h * vb, steps, omega, omega_c, k, lattice=lattice)
        jm = ADJ.loss_of(coll, f - h * vf, a - h * va, b - h * vb, steps, omega, omega_c, k, lattice=lattice)
        fd = (jp - jm) / (2 * h)
        ad = ((gf * vf).sum() + (ga * va).sum() + (gb * vb).sum()).item()
        err = abs(fd - ad) / abs(fd)
        worst = max(worst, err)
        print(f"{lattice} {coll} direction {trial}: fd {fd:.10e} adjoint {ad:.10e} ({err:.1e})", flush=True)
    # hw = 1e-3: the difference's error is at most 1.4e-7 relative for every
    # lattice and collision (truncation ~1.1e-7).  Smaller steps pick up the
    # rounding of the reaction's closed form near the plane where A = B
    # (lattice_physics.hpp:712-743), up to 1.3e-6 at 1e-5; the adjoint's
    # omega_c term itself is pinned on the CPU against an extended-precision
    # forward.
    hw = 1e-3
    fdw = (ADJ.loss_of(coll, f, a, b, steps, omega, omega_c + hw, k, lattice=lattice) -
           ADJ.loss_of(coll, f, a, b, steps, omega, omega_c - hw, k, lattice=lattice)) / (2 * hw)
    print(f"{lattice} {coll} omega_c: fd {fdw:.10e} adjoint {dw:.10e}; worst direction {worst:.1e}", flush=True)
    assert worst <= 1e-6 and abs(fdw - dw) <= 1e-6 * abs(fdw)


def test_adjoint_3d_interval_invariance(gpu):
    f, a, b = sheared_state(24, "D3Q19", gpu)
    args = (f, a, b, 40, 1 / 0.6, 1 / 0.8, 0.05)
    ref = ADJ.reactive_gradient("BGK", *args, interval=40, lattice="D3Q19")
    for interval in (1, 7, 8, 16):
        got = ADJ.reactive_gradient("BGK", *args, interval=interval, lattice="D3Q19")
        assert got[0] == ref[0] and got[4] == ref[4]
        for x, y in zip(got[1:4], ref[1:4]):
            assert torch.equal(x, y)


@pytest.mark.parametrize("coll", ["TRT", "RLBM"])
def test_adjoint_3d_autograd_gradcheck(gpu, coll):
    print(f"case adjoint 3d gradcheck {coll}", flush=True)
    f, a, b = TA3.small_state_3d("D3Q19", nx=4, ny=3, nz=3, seed=8)
    f0, a0, b0 = (dev(x, gpu).requires_grad_(True) for x in (f, a, b))
    tau = torch.tensor(0.9, dtype=torch.float64, device=gpu, requires_grad=True)

    def fn(f0, a0, b0, tau):
        return ADJ.reactive_loss(f0, a0, b0, tau, coll, steps=3, omega=1 / 0.7, k_react=0.4, interval=2,
                                 lattice="D3Q19")

    assert torch.autograd.gradcheck(fn, (f0, a0, b0, tau), eps=1e-6, atol=1e-9, rtol=1e-6)
