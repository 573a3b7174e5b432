"""Gradients of the reactive loop through the HIP discrete adjoint.

The loop is D2Q9 + two D2Q5 species, or D3Q19 / D3Q27 + two D3Q7 species (BGK,
TRT, MRT or RLBM), with A + B -> C, stepped by
lbc_reactive_fused_step; the loss is the domain-mean product <C>(T) =
<A>(0) - <A>(T).  `reactive_gradient` calls lbc_reactive_gradient (2D) or
lbc_reactive_gradient_3d (forward with checkpoints every `interval` steps,
segment recompute, adjoint sweep);
`ReactiveLoss` wraps it as a torch.autograd.Function of the initial
populations and the scalar relaxation time tau_s, so gradients with respect to
anything the initial state is built from (for instance a velocity perturbation
through `initial_state`) follow from torch's chain rule.
"""
from __future__ import annotations

import ctypes
import math

import torch

from . import (LATTICES, VELOCITIES, WEIGHTS, _check_pop, _check_set, _extents, _stream, check, default_block, lib,
               reactive_fused_step)

_STORE = {torch.float32: 0, torch.float64: 1}
_COLL = {"BGK": 0, "TRT": 1, "MRT": 2, "RLBM": 3}
SPECIES = {"D2Q9": "D2Q5", "D3Q19": "D3Q7", "D3Q27": "D3Q7"}


def workspace_bytes(dtype, nx, ny, steps, interval, block, lattice="D2Q9", nz=1):
    out = ctypes.c_uint64()
    if lattice == "D2Q9":
        check(lib().lbc_reactive_gradient_workspace(_STORE[dtype], nx, ny, steps, interval, block[0], block[1],
                                                    ctypes.byref(out)), "lbc_reactive_gradient_workspace")
    else:
        check(lib().lbc_reactive_gradient_workspace_3d(LATTICES[lattice], _STORE[dtype], nx, ny, nz, steps, interval,
                                                       block[0], block[1], block[2], ctypes.byref(out)),
              "lbc_reactive_gradient_workspace_3d")
    return out.value


FLAG_F64_ADJOINT = 2


def reactive_gradient(collision, f0, a0, b0, steps, omega, omega_c, k_react, interval=64, block=None,
                      workspace=None, stream=None, fast=False, adjoint_f64=False, lattice="D2Q9"):
    """(loss, dloss/df0, dloss/da0, dloss/db0, dloss/domega_c) by the discrete adjoint;
    lattice: the fluid lattice (D2Q9 with D2Q5 species; D3Q19 or D3Q27 with
    D3Q7 species); fast=True differentiates the forward of
    reactive_fused_step(..., fast=True); adjoint_f64=True (float32 inputs)
    carries the adjoint in float64 on the float32 trajectory
    (LBC_FLAG_F64_ADJOINT)."""
    if lattice not in SPECIES:
        raise ValueError(f"the adjoint covers the fluid lattices {sorted(SPECIES)}")
    species = SPECIES[lattice]
    _check_pop(lattice, f0, "f0")
    _check_pop(species, a0, "a0")
    _check_pop(species, b0, "b0")
    _check_set([("f0", f0), ("a0", a0), ("b0", b0)])
    if f0.dtype not in _STORE:
        raise ValueError("f0, a0, b0 must share float32 or float64")
    nx, ny, nz = _extents(lattice, f0)
    block = default_block(lattice, nx, ny, nz) if block is None else tuple(block)
    nbytes = workspace_bytes(f0.dtype, nx, ny, steps, interval, block, lattice, nz)
    if workspace is None or workspace.numel() < nbytes:
        workspace = torch.empty(nbytes, dtype=torch.uint8, device=f0.device)
    gf, ga, gb = torch.empty_like(f0), torch.empty_like(a0), torch.empty_like(b0)
    loss, dw = ctypes.c_double(), ctypes.c_double()
    flags = (1 if fast else 0) | (FLAG_F64_ADJOINT if adjoint_f64 else 0)
    if lattice == "D2Q9":
        check(lib().lbc_reactive_gradient(_COLL[collision], _STORE[f0.dtype], flags, f0.data_ptr(),
                                          a0.data_ptr(), b0.data_ptr(),
                                          nx, ny, int(steps), int(interval), float(omega), float(omega_c),
                                          float(k_react), ctypes.byref(loss), gf.data_ptr(), ga.data_ptr(),
                                          gb.data_ptr(), ctypes.byref(dw), workspace.data_ptr(), workspace.numel(),
                                          block[0], block[1], _stream(stream)), "lbc_reactive_gradient")
    else:
        check(lib().lbc_reactive_gradient_3d(LATTICES[lattice], _COLL[collision], _STORE[f0.dtype], flags,
                                             f0.data_ptr(), a0.data_ptr(), b0.data_ptr(), nx, ny, nz, int(steps),
                                             int(interval), float(omega), float(omega_c), float(k_react),
                                             ctypes.byref(loss), gf.data_ptr(), ga.data_ptr(), gb.data_ptr(),
                                             ctypes.byref(dw), workspace.data_ptr(), workspace.numel(), block[0],
                                             block[1], block[2], _stream(stream)), "lbc_reactive_gradient_3d")
    return loss.value, gf, ga, gb, dw.value


def reactive_run(collision, f0, a0, b0, steps, omega, omega_c, k_react, block=None, fast=False, lattice="D2Q9"):
    """Forward only: returns the final (f, a, b) (inputs untouched)."""
    f, a, b = f0.clone(), a0.clone(), b0.clone()
    ft, at, bt = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    for _ in range(steps):
        reactive_fused_step(lattice, collision, SPECIES[lattice], f, ft, a, at, b, bt, omega, omega_c, k_react,
                            block=block, fast=fast)
        f, ft, a, at, b, bt = ft, f, at, a, bt, b
    return f, a, b


def loss_of(collision, f0, a0, b0, steps, omega, omega_c, k_react, fast=False, lattice="D2Q9"):
    """<A>(0) - <A>(T) from a forward run, accumulated in float64."""
    _, a, _ = reactive_run(collision, f0, a0, b0, steps, omega, omega_c, k_react, fast=fast, lattice=lattice)
    n = a0[0].numel()
    return (a0.double().sum() - a.double().sum()).item() / n


class ReactiveLoss(torch.autograd.Function):
    """loss = <C>(T) of the reactive loop; differentiable in (f0, a0, b0, tau_s)."""

    @staticmethod
    def forward(ctx, f0, a0, b0, tau_s, collision, steps, omega, k_react, interval, fast, adjoint_f64=False,
                lattice="D2Q9"):
        omega_c = 1.0 / float(tau_s)
        loss, gf, ga, gb, dw = reactive_gradient(collision, f0.detach().contiguous(), a0.detach().contiguous(),
                                                 b0.detach().contiguous(), steps, omega, omega_c, k_react, interval,
                                                 fast=fast, adjoint_f64=adjoint_f64, lattice=lattice)
        dtau = -dw * omega_c * omega_c
        ctx.save_for_backward(gf, ga, gb, torch.tensor(dtau, dtype=torch.float64, device=f0.device))
        return torch.tensor(loss, dtype=f0.dtype, device=f0.device)

    @staticmethod
    def backward(ctx, grad):
        gf, ga, gb, dtau = ctx.saved_tensors
        return (grad * gf, grad * ga, grad * gb, (grad.double() * dtau).to(grad.dtype), None, None, None, None, None,
                None, None, None)


def reactive_loss(f0, a0, b0, tau_s, collision="BGK", steps=4096, omega=1.0 / 0.6, k_react=0.05, interval=64,
                  fast=False, adjoint_f64=False, lattice="D2Q9"):
    tau = tau_s if torch.is_tensor(tau_s) else torch.tensor(float(tau_s), dtype=torch.float64, device=f0.device)
    return ReactiveLoss.apply(f0, a0, b0, tau, collision, steps, omega, k_react, interval, fast, adjoint_f64,
                              lattice)


def equilibrium_torch(rho, u):
    """Differentiable D2Q9 equilibrium, rho (ny, nx), u (2, ny, nx)."""
    uu = u[0] * u[0] + u[1] * u[1]
    out = []
    for (cx, cy), w in zip(VELOCITIES["D2Q9"], WEIGHTS["D2Q9"]):
        cu = cx * u[0] + cy * u[1]
        out.append(w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu))
    return torch.stack(out)


def scalar_equilibrium_torch(conc, u):
    """Differentiable D2Q5 equilibrium g_i = w_i C (1 + 3 c.u)."""
    out = []
    for (cx, cy), w in zip(VELOCITIES["D2Q5"], WEIGHTS["D2Q5"]):
        out.append(w * conc * (1.0 + 3.0 * (cx * u[0] + cy * u[1])))
    return torch.stack(out)


def gaussian_random_field(shape, corr_len, rms, seed, dtype=torch.float64, device="cuda"):
    """Two-component periodic Gaussian random field with a Gaussian spectrum of
    correlation length `corr_len` cells, scaled to the given rms per component."""
    ny, nx = shape
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    noise = torch.randn((2, ny, nx), generator=g, dtype=torch.float64)
    ky = torch.fft.fftfreq(ny, dtype=torch.float64) * 2 * math.pi
    kx = torch.fft.fftfreq(nx, dtype=torch.float64) * 2 * math.pi
    k2 = ky.view(-1, 1) ** 2 + kx.view(1, -1) ** 2
    filt = torch.exp(-0.5 * k2 * corr_len * corr_len)
    field = torch.fft.ifft2(torch.fft.fft2(noise) * filt).real
    field = field / torch.sqrt((field * field).mean(dim=(1, 2), keepdim=True)) * rms
    return field.to(dtype).to(device)


def initial_state(du, u0=0.04, dtype=None):
    """Initial (f0, a0, b0) from a velocity perturbation du (2, ny, nx): a 2D
    Taylor-Green base flow of amplitude u0 (src/simulate_reaction.cpp:173-195)
    plus du, species A on x < nx/2 and B elsewhere at the same velocity
    (src/simulate_reaction.cpp:197-232).  Differentiable in du."""
    dtype = dtype or du.dtype
    _, ny, nx = du.shape
    kx, ky = 2 * math.pi / nx, 2 * math.pi / ny
    y, x = torch.meshgrid(torch.arange(ny, dtype=torch.float64, device=du.device) * ky,
                          torch.arange(nx, dtype=torch.float64, device=du.device) * kx, indexing="ij")
    rho = 1.0 - (u0 * u0) / (4.0 / 3.0) * (torch.cos(2 * x) + torch.cos(2 * y))
    base = torch.stack([-u0 * torch.cos(x) * torch.sin(y), u0 * torch.sin(x) * torch.cos(y)])
    u = base + du.to(torch.float64)
    col = torch.arange(nx, device=du.device).view(1, -1).expand(ny, nx)
    ca = (col < nx // 2).to(torch.float64)
    f0 = equilibrium_torch(rho, u).to(dtype)
    a0 = scalar_equilibrium_torch(ca, u).to(dtype)
    b0 = scalar_equilibrium_torch(1.0 - ca, u).to(dtype)
    return f0, a0, b0
