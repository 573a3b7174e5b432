"""Device-side initial states (seeded, synthetic), built with torch on the GPU.

Populations start at the second-order equilibrium of a macroscopic field, the
way LB-Cube's drivers initialise them (src/main.cpp:368-411 via
compute_equilibrium, include/lattice_physics.hpp:267-290).  Fields are formed in
float64 and rounded once into the storage format.
"""
from __future__ import annotations

import math

import torch

from . import DIM, Q, VELOCITIES, WEIGHTS, encode


def _velocities(lattice, device):
    return torch.tensor(VELOCITIES[lattice], dtype=torch.float64, device=device)


def equilibrium(lattice, rho, u, storage="f32"):
    """feq_i = w_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u) in float64 -> storage tensor.

    rho: (...) tensor or None (= 1); u: (D, ...) tensor.
    """
    device = u.device
    c = _velocities(lattice, device)  # (Q, D)
    w = torch.tensor(WEIGHTS[lattice], dtype=torch.float64, device=device)
    if rho is None:
        rho = torch.ones(u.shape[1:], dtype=torch.float64, device=device)
    uu = (u * u).sum(0)
    out = torch.empty((Q[lattice],) + tuple(u.shape[1:]), dtype=torch.float64, device=device)
    for i in range(Q[lattice]):
        cu = sum(c[i, d] * u[d] for d in range(DIM[lattice]))
        out[i] = w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu)
    return encode(lattice, out, storage)


def scalar_equilibrium(lattice, conc, u, storage="f32", ref=0.0):
    """g_i = w_i C (1 + c.u / cs2), cs2 = 1/3 (D2Q5) or 1/4 (D3Q7); f16/bf16
    storage holds g_i - w_i ref (ref: the field's reference concentration)."""
    device = u.device
    c = _velocities(lattice, device)
    w = torch.tensor(WEIGHTS[lattice], dtype=torch.float64, device=device)
    cs2 = 1.0 / 3.0 if lattice == "D2Q5" else 0.25
    out = torch.empty((Q[lattice],) + tuple(u.shape[1:]), dtype=torch.float64, device=device)
    for i in range(Q[lattice]):
        cu = sum(c[i, d] * u[d] for d in range(DIM[lattice]))
        out[i] = w[i] * conc * (1.0 + cu / cs2)
    return encode(lattice, out, storage, ref)


def solenoidal_noise_3d(n, rms, seed, kmin=1, kmax=8, device="cuda"):
    """Divergence-free random-phase field on an n^3 periodic box.

    Fourier modes with kmin <= |k| <= kmax get unit-variance complex Gaussian
    amplitudes (torch.Generator seeded with `seed`), are projected onto the
    plane normal to k (so div u = 0), transformed back, and scaled to the
    requested rms per component.
    """
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    k1 = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64)
    kz, ky, kx = torch.meshgrid(k1, k1, k1, indexing="ij")
    kk = torch.sqrt(kx * kx + ky * ky + kz * kz)
    band = ((kk >= kmin) & (kk <= kmax)).to(torch.float64)
    amp = torch.complex(torch.randn((3, n, n, n), generator=g, dtype=torch.float64),
                        torch.randn((3, n, n, n), generator=g, dtype=torch.float64)) * band
    ks = torch.stack([kx, ky, kz])
    k2 = torch.where(kk > 0, kk * kk, torch.ones_like(kk))
    proj = (amp * ks).sum(0) / k2
    amp = amp - ks * proj
    amp = amp.to(device)
    u = torch.fft.ifftn(amp, dim=(1, 2, 3)).real
    scale = rms / torch.sqrt((u * u).mean() / 1.0)
    return (u * scale).contiguous()  # (3, z, y, x) ordered (ux, uy, uz)


def taylor_green_3d(n, u0, device="cuda", perturbation_rms=0.0, seed=0x5EED1234):
    """TGV u = (u0 sin kx cos ky cos kz, -u0 cos kx sin ky cos kz, 0), k = 2 pi / n,
    plus an optional seeded divergence-free perturbation of the given rms."""
    k = 2.0 * math.pi / n
    r = torch.arange(n, dtype=torch.float64, device=device) * k
    z, y, x = torch.meshgrid(r, r, r, indexing="ij")
    u = torch.stack([u0 * torch.sin(x) * torch.cos(y) * torch.cos(z),
                     -u0 * torch.cos(x) * torch.sin(y) * torch.cos(z), torch.zeros_like(x)])
    if perturbation_rms > 0:
        u = u + solenoidal_noise_3d(n, perturbation_rms, seed, device=device)
    return u


def tgv_re1600_tau(n, u0=0.05, re=1600.0):
    """tau for nu = u0 n / (2 pi Re) (the Re = 1600 TGV configuration)."""
    nu = u0 * n / (2.0 * math.pi * re)
    return 3.0 * nu + 0.5


def taylor_green_2d(nx, ny, u0, device="cuda"):
    """2D TGV of src/simulate_reaction.cpp:173-195 (velocity and density)."""
    kx, ky = 2.0 * math.pi / nx, 2.0 * math.pi / ny
    y, x = torch.meshgrid(torch.arange(ny, dtype=torch.float64, device=device) * ky,
                          torch.arange(nx, dtype=torch.float64, device=device) * kx, indexing="ij")
    rho = 1.0 - (u0 * u0) / (4.0 / 3.0) * (torch.cos(2 * x) + torch.cos(2 * y))
    u = torch.stack([-u0 * torch.cos(x) * torch.sin(y), u0 * torch.sin(x) * torch.cos(y)])
    return rho, u


def double_shear_layer(n, u0=0.04, delta=80.0, noise=1e-3, seed=1234, device="cuda"):
    """Periodic double shear layer: u_x = u0 tanh((y - n/4)/delta) for y < n/2 and
    u0 tanh((3n/4 - y)/delta) above, u_y = a sinusoid plus seeded white noise."""
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    y, x = torch.meshgrid(torch.arange(n, dtype=torch.float64), torch.arange(n, dtype=torch.float64), indexing="ij")
    ux = torch.where(y < n / 2, u0 * torch.tanh((y - n / 4) / delta), u0 * torch.tanh((3 * n / 4 - y) / delta))
    uy = noise * u0 * torch.sin(2.0 * math.pi * x / n) + noise * torch.randn((n, n), generator=g,
                                                                              dtype=torch.float64)
    return torch.stack([ux, uy]).to(device)


def tanh_half_planes(n, delta=80.0, device="cuda"):
    """Species A on one half plane, B on the other, tanh-smoothed interfaces."""
    y = torch.arange(n, dtype=torch.float64, device=device).view(-1, 1).expand(n, n)
    ca = 0.5 * (1.0 + torch.tanh((y - n / 4) / delta)) * 0.5 * (1.0 + torch.tanh((3 * n / 4 - y) / delta))
    return ca.contiguous(), (1.0 - ca).contiguous()
