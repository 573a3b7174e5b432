"""The oracle's FP64 discrete adjoint of the 2D reactive loop (D2Q9
BGK/TRT/MRT/RLBM + two D2Q5 species, A + B -> C), pinned by central finite
differences of the oracle's own forward loop (in omega_c also over 400 steps
against an extended-precision forward) and by the one-step dot-product
identity.  The reference has no adjoint; this CPU adjoint is
the checker for the HIP adjoint kernel."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from oracle import oracle as O
from test_oracle_pins import VEL


def small_state(nx=10, ny=8, seed=3):
    g = cases.rng(seed)
    u = g.uniform(-0.03, 0.03, size=(2, ny, nx))
    rho = 1.0 + g.uniform(-0.01, 0.01, size=(ny, nx))
    f = O.feq_field("D2Q9", "f64", rho, u)
    ca = 0.5 * (1 + np.tanh((np.arange(nx) - nx / 2) / 1.5)) * np.ones((ny, 1))
    a = O.scalar_feq_field("D2Q5", "f64", ca, u)
    b = O.scalar_feq_field("D2Q5", "f64", 1.0 - ca + 0.02 * g.uniform(size=(ny, nx)), u)
    return f, a, b


def J_of(coll, f, a, b, omega, omega_c, k, steps):
    _, wa, _ = O.reactive_run("D2Q9", coll, "D2Q5", "f64", f, a, b, omega, omega_c, k, steps)
    ny, nx = f.shape[1:]
    return a.sum() / (nx * ny) - wa.sum() / (nx * ny)


@pytest.mark.parametrize("dtype,ulps", [("f32", 8), ("f64", 32)])
def test_phi_slope_against_extended_precision(dtype, ulps):
    """phi(z) = expm1(z)/z and phi'(z) of the reaction Jacobian, against 40-digit
    mpmath, over |z| in [1e-7, 3] of both signs.  The closed form
    phi' = (z e^z - expm1(z))/z^2 cancels near 0 (its numerator is ~z^2/2): in
    float32 it was 7% off at z = 1e-3, the range the 512^2 adjoint's interface
    cells fall in; the series below the switch and z + expm1(z)(z - 1) above it
    stay within a few ulps (worst measured: 7 ulps in float32 at z = -1.6, 27 in
    float64 just past its switch at |z| = 0.1)."""
    import mpmath as mp

    mp.mp.dps = 40
    z = np.concatenate([s * np.logspace(-7, np.log10(3), 300) for s in (1.0, -1.0)])
    if dtype == "f32":
        z = z.astype(np.float32).astype(np.float64)
    eps = 2.0 ** -24 if dtype == "f32" else 2.0 ** -53
    phi, dphi = O.phi_slope(z, dtype)
    want_phi = np.array([float(mp.expm1(mp.mpf(x)) / mp.mpf(x)) for x in z])
    want_dphi = np.array([float((mp.mpf(x) * mp.exp(mp.mpf(x)) - mp.expm1(mp.mpf(x))) / mp.mpf(x) ** 2) for x in z])
    assert np.max(np.abs(phi - want_phi) / np.abs(want_phi)) <= ulps * eps
    assert np.max(np.abs(dphi - want_dphi) / np.abs(want_dphi)) <= ulps * eps


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_reaction_map_derivatives_against_extended_precision(dtype):
    """dS/dA and dS/dB of the exact one-step map, each relative to its own size,
    against 40-digit mpmath derivatives, at the 512^2 run's k = 0.05, at 0.4 and
    at 3.  dS/dA is formed without subtracting 1 from dA'/dA (which cost float32
    ~5e-6 of a -k B sized result at k = 0.05); measured within 5 ulps at
    k <= 0.4 and 28 at k = 3."""
    import mpmath as mp

    mp.mp.dps = 40
    g = cases.rng(12)
    a = g.uniform(0.0, 1.0, 120).astype(np.float32).astype(np.float64)
    b = np.clip(1.0 - a + g.normal(0, 0.05, a.size), 1e-3, 1.0).astype(np.float32).astype(np.float64)
    eps = 2.0 ** -24 if dtype == "f32" else 2.0 ** -53
    for k, ulps in ((float(np.float32(0.05)), 8), (float(np.float32(0.4)), 8), (3.0, 32)):
        def S(A, B):
            z = k * (B - A)
            return A / (1 + k * A * (mp.expm1(z) / z if z != 0 else 1) + mp.expm1(z)) - A

        wa = np.array([float(mp.diff(lambda x: S(x, mp.mpf(bb)), mp.mpf(aa))) for aa, bb in zip(a, b)])
        wb = np.array([float(mp.diff(lambda y: S(mp.mpf(aa), y), mp.mpf(bb))) for aa, bb in zip(a, b)])
        da, db = O.reaction_grad(a, b, k, dtype=dtype)
        assert np.max(np.abs(da - wa) / np.abs(wa)) <= ulps * eps, k
        assert np.max(np.abs(db - wb) / np.abs(wb)) <= ulps * eps, k


def test_reaction_map_derivatives_against_fd():
    g = cases.rng(11)
    a = np.concatenate([g.uniform(0.01, 1.2, 50), [0.4, 0.7]])
    b = np.concatenate([g.uniform(0.01, 1.2, 50), [0.4 + 1e-7, 0.7 - 3e-4]])
    k = 0.3
    da, db = O.reaction_grad(a, b, k)
    # the reference's closed form cancels near A = B, so the differences use a
    # step far above |A - B| (truncation error ~1e-10 here)
    h = 1e-4
    sa = (O.reaction_source("f64", a + h, b, k) - O.reaction_source("f64", a - h, b, k)) / (2 * h)
    sb = (O.reaction_source("f64", a, b + h, k) - O.reaction_source("f64", a, b - h, k)) / (2 * h)
    assert np.max(np.abs(da - sa)) < 1e-8 and np.max(np.abs(db - sb)) < 1e-8


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
def test_adjoint_matches_central_differences(coll):
    f, a, b = small_state()
    omega, omega_c, k, steps = 1 / 0.7, 1 / 0.9, 0.4, 12
    J, gf, ga, gb, dw = O.reactive_gradient(coll, "f64", f, a, b, omega, omega_c, k, steps)
    assert abs(J - J_of(coll, f, a, b, omega, omega_c, k, steps)) < 1e-15
    g = cases.rng(5)
    for trial in range(4):
        vf, va, vb = (g.standard_normal(x.shape) * 1e-3 for x in (f, a, b))
        h = 1e-4
        jp = J_of(coll, f + h * vf, a + h * va, b + h * vb, omega, omega_c, k, steps)
        jm = J_of(coll, f - h * vf, a - h * va, b - h * vb, omega, omega_c, k, steps)
        fd = (jp - jm) / (2 * h)
        ad = np.sum(gf * vf) + np.sum(ga * va) + np.sum(gb * vb)
        assert abs(fd - ad) <= 1e-7 * max(abs(fd), 1e-12), (trial, fd, ad)
    hw = 1e-6
    fd = (J_of(coll, f, a, b, omega, omega_c + hw, k, steps) - J_of(coll, f, a, b, omega, omega_c - hw, k, steps)) / (
        2 * hw)
    assert abs(fd - dw) <= 1e-6 * abs(fd), (fd, dw)


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
def test_adjoint_omega_c_long_run_extended_precision(coll):
    """dJ/d omega_c over 400 steps of a smooth 2D shear layer (32^2) against
    central differences of the forward loop in extended precision (h = 1e-5,
    exact to its truncation).  The profiles are offset by 0.3 cells so that
    no cell holds A = B, where the reaction's closed form cancels."""
    n, steps, omega, omega_c, k = 32, 400, 1 / 0.6, 1 / 0.8, 0.05
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    u = np.stack([0.04 * np.tanh((y - n / 2 - 0.3) / 3.0), 0.004 * np.sin(2 * np.pi * x / n)])
    f = O.feq_field("D2Q9", "f64", np.ones((n, n)), u)
    ca = 0.5 * (1 + np.tanh((x - n / 2 - 0.3) / 2.0))
    a = O.scalar_feq_field("D2Q5", "f64", ca, u)
    b = O.scalar_feq_field("D2Q5", "f64", 1.0 - ca, u)
    _, _, _, _, dw = O.reactive_gradient(coll, "f64", f, a, b, omega, omega_c, k, steps)
    ext = [v.astype(np.longdouble) for v in (f, a, b)]

    def minus_mean_a(oc):
        _, a1, _ = O.reactive_run("D2Q9", coll, "D2Q5", "f80", *ext, omega, oc, k, steps)
        return -a1.sum() / (n * n)

    h = 1e-5
    fd = float((minus_mean_a(omega_c + h) - minus_mean_a(omega_c - h)) / (2 * h))
    assert abs(fd - dw) <= 1e-9 * abs(dw), (fd, dw)


def to_collision_cell(lam, lattice):
    """Stored-state adjoint Lambda[i][y] -> collision-cell adjoint nu[i][x] = Lambda[i][x - c_i]."""
    nu = np.empty_like(lam)
    for i, (cx, cy) in enumerate(VEL[lattice]):
        nu[i] = np.roll(lam[i], shift=(cy, cx), axis=(0, 1))
    return nu


def to_stored(nu, lattice):
    """Collision-cell adjoint nu -> stored-state adjoint Lambda[i][y] = nu[i][y + c_i]."""
    lam = np.empty_like(nu)
    for i, (cx, cy) in enumerate(VEL[lattice]):
        lam[i] = np.roll(nu[i], shift=(-cy, -cx), axis=(0, 1))
    return lam


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
def test_adjoint_step_dot_product(coll):
    """<Lambda_1, L v> = <L^T Lambda_1, v> for one step, with random v and a
    random adjoint Lambda_1 on every population of the fluid and both species
    (so every path of the transposed step is exercised, not only those the
    loss reaches).  L v is the central difference of the forward step in
    extended precision (h = 1e-6: truncation ~1e-12, rounding ~1e-13); the
    adjoint step runs in FP64.  Also <Lambda_1, dStep/d omega_c> against the
    adjoint's accumulated dJ/d omega_c term."""
    f, a, b = small_state(nx=7, ny=6, seed=9)
    omega, omega_c, k = 1 / 0.7, 1 / 0.9, 0.4
    g = cases.rng(17)
    v = [g.standard_normal(x.shape) * 1e-2 for x in (f, a, b)]
    lam1 = [g.standard_normal(x.shape) for x in (f, a, b)]
    nu1 = [to_collision_cell(lam1[0], "D2Q9"), to_collision_cell(lam1[1], "D2Q5"), to_collision_cell(lam1[2], "D2Q5")]
    nuf0, nua0, nub0, djdw = O.reactive_adjoint_step(coll, "f64", f, a, b, *nu1, omega, omega_c, k)
    lam0 = [to_stored(nuf0, "D2Q9"), to_stored(nua0, "D2Q5"), to_stored(nub0, "D2Q5")]
    ext = [x.astype(np.longdouble) for x in (f, a, b)]
    h = np.longdouble(1e-6)

    def step(state, oc):
        return O.reactive_run("D2Q9", coll, "D2Q5", "f80", *state, omega, oc, k, 1)

    plus = step([x + h * d for x, d in zip(ext, v)], omega_c)
    minus = step([x - h * d for x, d in zip(ext, v)], omega_c)
    lv = [(p - m) / (2 * h) for p, m in zip(plus, minus)]
    lhs = sum(float(np.sum(l.astype(np.longdouble) * x)) for l, x in zip(lam1, lv))
    rhs = sum(float(np.sum(l * x)) for l, x in zip(lam0, v))
    scale = sum(float(np.sum(np.abs(l) * np.abs(x))) for l, x in zip(lam0, v))
    assert abs(lhs - rhs) <= 1e-10 * scale, (lhs, rhs, scale)
    hw = np.longdouble(2.0 ** -20)  # omega_c +- hw exact in double
    dp = step(ext, omega_c + float(hw))
    dm = step(ext, omega_c - float(hw))
    lw = sum(float(np.sum(l.astype(np.longdouble) * (p - m) / (2 * hw))) for l, p, m in zip(lam1, dp, dm))
    assert abs(lw - djdw) <= 1e-9 * max(abs(lw), 1.0), (lw, djdw)
