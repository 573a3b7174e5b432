"""Initial states shared by the tests, built with the oracle's equilibrium.

Each initialiser follows a reference test or driver (cited per function); the
resulting populations are the oracle's compute_equilibrium of the macroscopic
fields, so the GPU and the oracle always start from identical bits.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def rng(seed=0x5EED1234):
    return np.random.Generator(np.random.PCG64(seed))


def random_equilibrium(lattice, shape, dtype="f64", seed=0x5EED1234, du=0.01, drho=0.05):
    """Random (rho, u) per cell then feq (cf. tests/test_validation.cpp:171-192)."""
    g = rng(seed)
    d = O.D_OF[lattice]
    rho = 1.0 + g.uniform(-drho, drho, size=shape)
    u = g.uniform(-du, du, size=(d,) + tuple(shape))
    if lattice in ("D2Q5", "D3Q7"):  # scalar lattices: rho plays the concentration
        return O.scalar_feq_field(lattice, dtype, rho, u)
    return O.feq_field(lattice, dtype, rho, u)


def random_perturbed(lattice, shape, dtype="f64", seed=7, eps=1e-3, ref=0.0):
    """Equilibrium plus a small random non-equilibrium part (exercises every
    moment); ref: a scalar field's reference concentration for f16/bf16."""
    f = random_equilibrium(lattice, shape, "f64", seed)
    g = rng(seed + 1)
    f = f + eps * g.uniform(-1, 1, size=f.shape) * np.asarray(O.WEIGHTS[lattice]).reshape((-1,) + (1,) * len(shape))
    if dtype == "f64":
        return np.ascontiguousarray(f)
    if dtype == "f32":
        return np.ascontiguousarray(f.astype(np.float32))
    return encode(lattice, dtype, f, ref)


def encode(lattice, dtype, f, ref=0.0):
    """float64 populations -> storage array of the oracle's codecs: f16/bf16
    hold f - w (fluid) or g - w ref (scalar, ref its reference concentration),
    the offset formed in float32."""
    f32 = f.astype(np.float32)
    w = np.asarray(O.WEIGHTS[lattice], dtype=np.float32).reshape((-1,) + (1,) * (f.ndim - 1))
    if lattice in ("D2Q5", "D3Q7"):
        w = (w * np.float32(ref)).astype(np.float32)
    f32 = (f32 - w).astype(np.float32)
    h, _ = O.half_roundtrip(f32.ravel(), bf16=(dtype == "bf16"))
    return h.reshape(f.shape)


def taylor_green_3d(n, u0, lattice="D3Q19", dtype="f64", rho_pert=False):
    """TGV of tests/test_validation.cpp:241-275 (rho = 1); rho_pert=True adds the
    density perturbation of src/main.cpp:368-411."""
    k = 2 * np.pi / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    px, py, pz = k * x, k * y, k * z
    u = np.stack([u0 * np.sin(px) * np.cos(py) * np.cos(pz), -u0 * np.cos(px) * np.sin(py) * np.cos(pz),
                  np.zeros_like(px)])
    rho = np.ones_like(px)
    if rho_pert:
        rho = 1.0 + (u0 * u0) / (16.0 * (1.0 / 3.0)) * (np.cos(2 * px) + np.cos(2 * py)) * (np.cos(2 * pz) + 2.0)
    return O.feq_field(lattice, dtype, rho, u)


def taylor_green_2d(n, u0, dtype="f64"):
    """2D TGV of tests/test_validation.cpp:200-222."""
    k = 2 * np.pi / n
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    px, py = k * x, k * y
    rho = 1.0 - (u0 * u0) / (4.0 / 3.0) * (np.cos(2 * px) + np.cos(2 * py))
    u = np.stack([-u0 * np.cos(px) * np.sin(py), u0 * np.sin(px) * np.cos(py)])
    return O.feq_field("D2Q9", dtype, rho, u), u


def split_species(nx, ny, u, dtype="f64"):
    """A where x < nx/2, B elsewhere, at the fluid velocity (src/simulate_reaction.cpp:197-232)."""
    x = np.broadcast_to(np.arange(nx), (ny, nx))
    ca = np.where(x < nx // 2, 1.0, 0.0)
    cb = 1.0 - ca
    return O.scalar_feq_field("D2Q5", dtype, ca, u), O.scalar_feq_field("D2Q5", dtype, cb, u)


def macro(lattice, f):
    """Density and velocity of float64 populations (numpy, for checks only)."""
    c = np.asarray({"D2Q9": [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]}.get(
        lattice, None) if lattice == "D2Q9" else _vel3(lattice))
    rho = f.sum(axis=0)
    u = np.stack([np.tensordot(c[:, d], f, axes=(0, 0)) / rho for d in range(c.shape[1])])
    return rho, u


def _vel3(lattice):
    v19 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, 1, 0),
           (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 1), (0, -1, 1),
           (0, 1, -1), (0, -1, -1)]
    if lattice == "D3Q19":
        return v19
    return v19 + [(sx, sy, sz) for sz in (1, -1) for sy in (1, -1) for sx in (1, -1)]
