"""GPU: the discrete adjoint of the 2D reactive loop (lbc_reactive_gradient).

  * against the oracle's FP64 CPU adjoint (tight tolerance, several checkpoint
    intervals so the segment recompute is exercised);
  * checkpoint-interval invariance: recomputed segments are bit-identical, so
    every interval gives the same gradient bit for bit;
  * torch.autograd.gradcheck of the autograd wrapper in FP64 on a tiny grid;
  * the 512^2 x 4096-step configuration: FP64 central differences along 32
    random directions of the initial velocity perturbation (and of tau_s).
"""
from __future__ import annotations

import math
import numpy as np
import pytest
import torch

import test_oracle_adjoint as TA
from lbcube import adjoint as ADJ
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(x, gpu, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(gpu).to(dtype)


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
@pytest.mark.parametrize("interval", [1, 5, 8, 64])
def test_adjoint_matches_oracle_fp64(gpu, coll, interval):
    # interval 8 over 17 steps: three segments (8, 8, 1), the recomputes
    # overlapped with the adjoint walks on a second stream
    print(f"case adjoint {coll} interval={interval}", flush=True)
    f, a, b = TA.small_state(nx=12, ny=10, seed=4)
    omega, omega_c, k, steps = 1 / 0.7, 1 / 0.9, 0.4, 17
    J, gf, ga, gb, dw = O.reactive_gradient(coll, "f64", f, a, b, omega, omega_c, k, steps)
    Jg, gfg, gag, gbg, dwg = ADJ.reactive_gradient(coll, dev(f, gpu), dev(a, gpu), dev(b, gpu), steps, omega, omega_c,
                                                   k, interval=interval)
    assert abs(Jg - J) <= 1e-15
    for got, want in ((gfg, gf), (gag, ga), (gbg, gb)):
        g = got.cpu().numpy()
        assert np.max(np.abs(g - want)) <= 1e-12 * np.max(np.abs(want))
    assert abs(dwg - dw) <= 1e-11 * abs(dw)


@pytest.mark.parametrize("coll", ["BGK", "RLBM"])
def test_adjoint_fp32_against_oracle(gpu, coll):
    """The FP32 adjoint is as close to the FP64 gradient as the oracle's own
    FP32 adjoint is (their operation orders differ, so not bit for bit)."""
    print(f"case adjoint fp32 {coll}", flush=True)
    f, a, b = TA.small_state(nx=12, ny=10, seed=6)
    f, a, b = (x.astype(np.float32) for x in (f, a, b))
    omega, omega_c, k, steps = 1 / 0.7, 1 / 0.9, 0.4, 17
    J64, *g64 = O.reactive_gradient(coll, "f64", *(x.astype(np.float64) for x in (f, a, b)), omega, omega_c, k,
                                    steps)
    J32, *g32 = O.reactive_gradient(coll, "f32", f, a, b, omega, omega_c, k, steps)
    Jg, gfg, gag, gbg, dwg = ADJ.reactive_gradient(coll, dev(f, gpu, torch.float32), dev(a, gpu, torch.float32),
                                                   dev(b, gpu, torch.float32), steps, omega, omega_c, k, interval=4)
    assert abs(Jg - J32) <= 1e-6 and abs(Jg - J64) <= 1e-6
    for got, orc, exact in zip((gfg, gag, gbg), g32[:3], g64[:3]):
        scale = np.max(np.abs(exact))
        e_gpu = np.max(np.abs(got.cpu().numpy() - exact)) / scale
        e_orc = np.max(np.abs(orc - exact)) / scale
        print(f"fp32 adjoint: gpu {e_gpu:.2e}, oracle {e_orc:.2e} (relative to the FP64 gradient)", flush=True)
        assert e_gpu <= max(4 * e_orc, 1e-5)
    assert abs(dwg - g64[3]) <= max(4 * abs(g32[3] - g64[3]), 1e-5 * abs(g64[3]))


def test_fp64_adjoint_of_fp32_trajectory(gpu):
    """LBC_FLAG_F64_ADJOINT: the adjoint in FP64 on the FP32 trajectory.  Over
    30 steps the FP32 trajectory's rounding moves the gradient by up to ~6e-6
    relative to the FP64 gradient (measured: BGK 2.2e-6, TRT 6.0e-6); this
    gradient must stay within 2e-5 of it, no farther than the FP32 adjoint is,
    and be bit-identical for every checkpoint interval."""
    f, a, b = TA.small_state(nx=48, ny=40, seed=9)
    args64 = (dev(f, gpu), dev(a, gpu), dev(b, gpu), 30, 1 / 0.7, 1 / 0.9, 0.4)
    args32 = tuple(x.float() for x in args64[:3]) + args64[3:]
    for coll in ("BGK", "TRT", "MRT", "RLBM"):
        print(f"case fp64 adjoint of fp32 trajectory {coll}", flush=True)
        ref = ADJ.reactive_gradient(coll, *args64, interval=8)
        g32 = ADJ.reactive_gradient(coll, *args32, interval=8)
        mixed = [ADJ.reactive_gradient(coll, *args32, interval=iv, adjoint_f64=True) for iv in (1, 8, 30)]
        for m in mixed[1:]:
            assert m[0] == mixed[0][0] and m[4] == mixed[0][4]
            assert all(torch.equal(x, y) for x, y in zip(m[1:4], mixed[0][1:4]))
        m = mixed[0]
        for x32, xm, x64 in zip(g32[1:4], m[1:4], ref[1:4]):
            scale = x64.abs().max().item()
            e32 = (x32.double() - x64).abs().max().item() / scale
            em = (xm.double() - x64).abs().max().item() / scale
            print(f"{coll}: FP32 adjoint {e32:.2e}, FP64 adjoint of the FP32 run {em:.2e} (relative to FP64)",
                  flush=True)
            assert em <= 2e-5 and em <= 1.05 * e32 + 1e-7
        assert abs(m[4] - ref[4]) <= 1e-5 * abs(ref[4])


def test_fast_forward_gradient_fp64(gpu):
    """The gradient of the fast (deviation-form, cancellation-free reaction)
    forward equals the reference-order one up to FP64 rounding."""
    f, a, b = TA.small_state(nx=40, ny=24, seed=5)
    args = (dev(f, gpu), dev(a, gpu), dev(b, gpu), 30, 1 / 0.6, 1 / 0.8, 0.2)
    ref = ADJ.reactive_gradient("TRT", *args, interval=7)
    fast = ADJ.reactive_gradient("TRT", *args, interval=7, fast=True)
    assert abs(fast[0] - ref[0]) <= 1e-12 * abs(ref[0]) and abs(fast[4] - ref[4]) <= 1e-9 * abs(ref[4])
    for x, y in zip(fast[1:4], ref[1:4]):
        assert (x - y).abs().max().item() <= 1e-9 * y.abs().max().item()


def test_checkpoint_interval_invariance(gpu):
    f, a, b = TA.small_state(nx=64, ny=48, seed=2)
    args = (dev(f, gpu), dev(a, gpu), dev(b, gpu), 50, 1 / 0.6, 1 / 0.8, 0.05)
    ref = ADJ.reactive_gradient("BGK", *args, interval=50)
    # 1 and 7: every segment on one stream; 8 and 16: recomputes overlapped
    # with the adjoint walks (interval >= 8 and more than one segment)
    for interval in (1, 7, 8, 16, 64):
        got = ADJ.reactive_gradient("BGK", *args, interval=interval)
        assert got[0] == ref[0] and got[4] == ref[4]
        for x, y in zip(got[1:4], ref[1:4]):
            assert torch.equal(x, y)


def test_overlapped_gradient_on_a_caller_stream(gpu):
    """The overlapped schedule (recomputes on the library's second stream,
    ordered by events) gives the same bits when the caller's stream is a
    non-default stream as on the default one."""
    print("case overlapped gradient on a caller stream, TRT interval 8", flush=True)
    f, a, b = TA.small_state(nx=64, ny=48, seed=3)
    args = (dev(f, gpu), dev(a, gpu), dev(b, gpu), 40, 1 / 0.6, 1 / 0.8, 0.05)
    ref = ADJ.reactive_gradient("TRT", *args, interval=8)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        got = ADJ.reactive_gradient("TRT", *args, interval=8)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert got[0] == ref[0] and got[4] == ref[4]
    for x, y in zip(got[1:4], ref[1:4]):
        assert torch.equal(x, y)


@pytest.mark.parametrize("coll", ["BGK", "RLBM"])
def test_autograd_gradcheck(gpu, coll):
    print(f"case adjoint gradcheck {coll}", flush=True)
    f, a, b = TA.small_state(nx=6, ny=5, seed=8)
    f0, a0, b0 = (dev(x, gpu).requires_grad_(True) for x in (f, a, b))
    tau = torch.tensor(0.9, dtype=torch.float64, device=gpu, requires_grad=True)

    def fn(f0, a0, b0, tau):
        return ADJ.reactive_loss(f0, a0, b0, tau, coll, steps=4, omega=1 / 0.7, k_react=0.4, interval=2)

    assert torch.autograd.gradcheck(fn, (f0, a0, b0, tau), eps=1e-6, atol=1e-9, rtol=1e-6)


def test_config4_gradients_against_central_differences(gpu):
    """512^2, 4096 steps, checkpoints every 64: dJ/d(du) and dJ/dtau_s, FP64,
    against central differences along 32 random directions; the FP32 adjoint
    is compared with the same differences."""
    n, steps, tau_s, omega, k = 512, 4096, 0.8, 1 / 0.6, 0.05
    du = ADJ.gaussian_random_field((n, n), 32.0, 0.01, seed=0x5EED1234, device=gpu)
    results = {}
    for dtype, fast, adj64 in ((torch.float64, False, False), (torch.float32, True, False),
                               (torch.float32, False, False), (torch.float32, True, True),
                               (torch.float32, False, True)):
        d = du.clone().requires_grad_(True)
        tau = torch.tensor(tau_s, dtype=torch.float64, device=gpu, requires_grad=True)
        f0, a0, b0 = ADJ.initial_state(d, dtype=dtype)
        loss = ADJ.reactive_loss(f0, a0, b0, tau, "BGK", steps=steps, omega=omega, k_react=k, interval=64,
                                 fast=fast, adjoint_f64=adj64)
        loss.backward()
        results[(dtype, fast, adj64)] = (loss.item(), d.grad.detach().clone(), tau.grad.item())
        print(f"dtype {dtype} fast={fast} adjoint_f64={adj64}: loss {loss.item():.10e}, "
              f"dJ/dtau_s {tau.grad.item():.10e}", flush=True)
    loss64, g64, gt64 = results[(torch.float64, False, False)]
    _, g32, gt32 = results[(torch.float32, True, False)]
    _, g32r, gt32r = results[(torch.float32, False, False)]
    _, g32m, gt32m = results[(torch.float32, True, True)]
    _, g32rm, gt32rm = results[(torch.float32, False, True)]

    def J(du_, tau_=tau_s):
        f0, a0, b0 = ADJ.initial_state(du_, dtype=torch.float64)
        return ADJ.loss_of("BGK", f0, a0, b0, steps, omega, 1.0 / tau_, k)

    # directions: unit-rms random fields; step h = 1e-5 (the perturbation of u
    # is 1e-5 rms against du's 1e-2; truncation ~h^2, FP64 noise ~1e-15 / h)
    gen = torch.Generator(device="cpu").manual_seed(99)
    worst64 = worst32 = worst32r = 0.0
    fds, ads32, ads32r, ads32m, ads32rm, ads64 = [], [], [], [], [], []
    for trial in range(32):
        v = ADJ.gaussian_random_field((n, n), 32.0, 1.0, seed=int(torch.randint(0, 2**31, (1,), generator=gen)),
                                      device=gpu)
        h = 1e-5
        fd = (J(du + h * v) - J(du - h * v)) / (2 * h)
        ad64 = (g64 * v).sum().item()
        ad32 = (g32.double() * v).sum().item()
        ad32r = (g32r.double() * v).sum().item()
        ad32m = (g32m.double() * v).sum().item()
        ad32rm = (g32rm.double() * v).sum().item()
        e64, e32, e32r = (abs(x - fd) / abs(fd) for x in (ad64, ad32, ad32r))
        fds.append(fd)
        ads32.append(ad32)
        ads32r.append(ad32r)
        ads32m.append(ad32m)
        ads32rm.append(ad32rm)
        ads64.append(ad64)
        worst64, worst32, worst32r = max(worst64, e64), max(worst32, e32), max(worst32r, e32r)
        print(f"direction {trial}: fd {fd:.10e} adjoint64 {ad64:.10e} ({e64:.1e}) adjoint32 fast ({e32:.1e}) "
              f"adjoint32 reference order ({e32r:.1e})", flush=True)
    ht = 1e-5
    fdt = (J(du, tau_s + ht) - J(du, tau_s - ht)) / (2 * ht)
    print(f"tau_s: fd {fdt:.10e} adjoint64 {gt64:.10e} adjoint32 fast {gt32:.10e} adjoint32 reference order "
          f"{gt32r:.10e}", flush=True)
    print(f"worst relative error: fp64 {worst64:.2e}, fp32 fast {worst32:.2e}, fp32 reference order "
          f"{worst32r:.2e}", flush=True)
    assert worst64 <= 1e-4 and abs(gt64 - fdt) <= 1e-4 * abs(fdt)
    # FP32 storage: the error of a directional derivative, relative to the rms
    # directional derivative, has two parts of about equal size (~1.0e-4 and
    # ~1.1e-4 worst over the 32 directions): the FP32 arithmetic of the adjoint
    # sweep, which the FP64 adjoint removes, and the FP32 rounding of the
    # trajectory amplified by its sensitivity, which it keeps.
    rms = math.sqrt(sum(x * x for x in fds) / len(fds))
    for name, ads, worst_bound, med_bound in (("fast", ads32, 2.5e-4, 1e-4),
                                              ("reference order", ads32r, 2.5e-4, 1e-4),
                                              ("fast, FP64 adjoint", ads32m, 1.5e-4, 6e-5),
                                              ("reference order, FP64 adjoint", ads32rm, 1.5e-4, 6e-5)):
        rel_rms = max(abs(a - f) for a, f in zip(ads, fds)) / rms
        med = sorted(abs(a - f) / abs(f) for a, f in zip(ads, fds))[len(fds) // 2]
        print(f"fp32 {name}: worst error / rms {rel_rms:.2e}, median relative error {med:.2e}", flush=True)
        assert rel_rms <= worst_bound and med <= med_bound, name
    # The split of the FP32 error: adjoint arithmetic = FP32 adjoint against the
    # FP64 adjoint of the same FP32 trajectory; trajectory rounding = that FP64
    # adjoint against the FP64 adjoint of the FP64 trajectory.
    for name, a32, a32m in (("fast", ads32, ads32m), ("reference order", ads32r, ads32rm)):
        arith = max(abs(x - y) for x, y in zip(a32, a32m)) / rms
        traj = max(abs(x - y) for x, y in zip(a32m, ads64)) / rms
        print(f"fp32 {name}: adjoint arithmetic {arith:.2e}, trajectory rounding {traj:.2e} (worst over the "
              f"directions, relative to the rms directional derivative)", flush=True)
        assert arith <= 1.5e-4 and traj <= 1.5e-4, name
    print(f"tau_s: fp32 fast FP64 adjoint {gt32m:.10e}, reference order FP64 adjoint {gt32rm:.10e}", flush=True)
    for g in (gt32, gt32r, gt32m, gt32rm):
        assert abs(g - fdt) <= 1.5e-4 * abs(fdt)
