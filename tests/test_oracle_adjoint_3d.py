"""The oracle's FP64 discrete adjoint of the 3D reactive loop (D3Q19 or D3Q27
BGK/TRT/MRT/RLBM + two D3Q7 species, cs2 = 1/4, A + B -> C), pinned like the 2D
one: by central finite differences of the oracle's own forward loop along random
directions (and in omega_c, also over 200 steps against an extended-precision
forward), and by the one-step dot-product identity
<Lambda_1, L v> = <L^T Lambda_1, v> with random adjoint inputs on every
population.  The reference has no adjoint; this CPU adjoint is the checker for
the HIP kernel of lbc_reactive_gradient_3d."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from oracle import oracle as O

VEL3 = {name: cases._vel3(name) for name in ("D3Q19", "D3Q27")}
VEL3["D3Q7"] = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]  # lattice_traits order


def small_state_3d(lattice, nx=6, ny=5, nz=4, seed=3):
    g = cases.rng(seed)
    u = g.uniform(-0.03, 0.03, size=(3, nz, ny, nx))
    rho = 1.0 + g.uniform(-0.01, 0.01, size=(nz, ny, nx))
    f = O.feq_field(lattice, "f64", rho, u)
    ca = 0.5 * (1 + np.tanh((np.arange(nx) - nx / 2) / 1.5)) * np.ones((nz, ny, 1))
    a = O.scalar_feq_field("D3Q7", "f64", ca, u)
    b = O.scalar_feq_field("D3Q7", "f64", 1.0 - ca + 0.02 * g.uniform(size=(nz, ny, nx)), u)
    return f, a, b


def J_of(lattice, coll, f, a, b, omega, omega_c, k, steps):
    _, wa, _ = O.reactive_run(lattice, coll, "D3Q7", "f64", f, a, b, omega, omega_c, k, steps)
    cells = a[0].size
    return a.sum() / cells - wa.sum() / cells


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_adjoint_3d_matches_central_differences(lattice, coll):
    f, a, b = small_state_3d(lattice)
    omega, omega_c, k, steps = 1 / 0.7, 1 / 0.9, 0.4, 6
    J, gf, ga, gb, dw = O.reactive_gradient(coll, "f64", f, a, b, omega, omega_c, k, steps, lattice=lattice)
    assert abs(J - J_of(lattice, coll, f, a, b, omega, omega_c, k, steps)) < 1e-15
    g = cases.rng(5)
    for trial in range(3):
        vf, va, vb = (g.standard_normal(x.shape) * 1e-3 for x in (f, a, b))
        # h = 1e-3: the differences' error then sits at ~1e-12 (it grows as h
        # shrinks, 1e-11 at 1e-4 and 1e-10 at 1e-5: the rounding of the forward
        # loop's J, whose reaction formula loses digits where A ~ B, not the
        # truncation)
        h = 1e-3
        jp = J_of(lattice, coll, f + h * vf, a + h * va, b + h * vb, omega, omega_c, k, steps)
        jm = J_of(lattice, coll, f - h * vf, a - h * va, b - h * vb, omega, omega_c, k, steps)
        fd = (jp - jm) / (2 * h)
        ad = np.sum(gf * vf) + np.sum(ga * va) + np.sum(gb * vb)
        assert abs(fd - ad) <= 1e-7 * max(abs(fd), 1e-12), (trial, fd, ad)
    hw = 1e-6
    fd = (J_of(lattice, coll, f, a, b, omega, omega_c + hw, k, steps) -
          J_of(lattice, coll, f, a, b, omega, omega_c - hw, k, steps)) / (2 * hw)
    assert abs(fd - dw) <= 1e-6 * abs(fd), (fd, dw)


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_adjoint_3d_omega_c_long_run_extended_precision(lattice, coll):
    """dJ/d omega_c over 200 steps of a smooth shear layer (12^3) against
    central differences of the forward loop in extended precision: at
    h = 1e-5 the difference is exact to its truncation, -3.4e-10 relative.
    The profiles are offset by 0.3 cells so that no cell holds A = B: where
    |B - A| is a few ulps above the switch of the reaction's closed form
    (lattice_physics.hpp:712-743), that form cancels and its rounding, not the
    derivative, would dominate the differences."""
    n, steps, omega, omega_c, k = 12, 200, 1 / 0.6, 1 / 0.8, 0.05
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.stack([0.04 * np.tanh((y - n / 2 - 0.3) / 2.0), 0.004 * np.sin(2 * np.pi * x / n),
                  0.004 * np.cos(2 * np.pi * z / n)])
    f = O.feq_field(lattice, "f64", np.ones((n, n, n)), u)
    ca = 0.5 * (1 + np.tanh((x - n / 2 - 0.3) / 1.5))
    a = O.scalar_feq_field("D3Q7", "f64", ca, u)
    b = O.scalar_feq_field("D3Q7", "f64", 1.0 - ca, u)
    _, _, _, _, dw = O.reactive_gradient(coll, "f64", f, a, b, omega, omega_c, k, steps, lattice=lattice)
    ext = [v.astype(np.longdouble) for v in (f, a, b)]

    def minus_mean_a(oc):
        _, a1, _ = O.reactive_run(lattice, coll, "D3Q7", "f80", *ext, omega, oc, k, steps)
        return -a1.sum() / a[0].size

    h = 1e-5
    fd = float((minus_mean_a(omega_c + h) - minus_mean_a(omega_c - h)) / (2 * h))
    assert abs(fd - dw) <= 1e-9 * abs(dw), (fd, dw)


def roll3(x, c, sign):
    return np.roll(x, shift=(sign * c[2], sign * c[1], sign * c[0]), axis=(0, 1, 2))


def to_collision_cell(lam, lattice):
    """Stored-state adjoint Lambda[i][y] -> collision-cell adjoint nu[i][x] = Lambda[i][x - c_i]."""
    return np.stack([roll3(lam[i], c, 1) for i, c in enumerate(VEL3[lattice])])


def to_stored(nu, lattice):
    """Collision-cell adjoint nu -> stored-state adjoint Lambda[i][y] = nu[i][y + c_i]."""
    return np.stack([roll3(nu[i], c, -1) for i, c in enumerate(VEL3[lattice])])


@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_adjoint_3d_step_dot_product(lattice, coll):
    """One step: <Lambda_1, L v> (L v the central difference of the forward
    step in extended precision) against <L^T Lambda_1, v> from the FP64
    adjoint step, with random v and random Lambda_1 on the fluid and both
    species; and <Lambda_1, dStep/d omega_c> against the adjoint's term."""
    f, a, b = small_state_3d(lattice, nx=5, ny=4, nz=3, seed=9)
    omega, omega_c, k = 1 / 0.7, 1 / 0.9, 0.4
    g = cases.rng(17)
    v = [g.standard_normal(x.shape) * 1e-2 for x in (f, a, b)]
    lam1 = [g.standard_normal(x.shape) for x in (f, a, b)]
    nu1 = [to_collision_cell(lam1[0], lattice), to_collision_cell(lam1[1], "D3Q7"),
           to_collision_cell(lam1[2], "D3Q7")]
    nuf0, nua0, nub0, djdw = O.reactive_adjoint_step(coll, "f64", f, a, b, *nu1, omega, omega_c, k, lattice=lattice)
    lam0 = [to_stored(nuf0, lattice), to_stored(nua0, "D3Q7"), to_stored(nub0, "D3Q7")]
    ext = [x.astype(np.longdouble) for x in (f, a, b)]
    h = np.longdouble(1e-6)

    def step(state, oc):
        return O.reactive_run(lattice, coll, "D3Q7", "f80", *state, omega, oc, k, 1)

    plus = step([x + h * d for x, d in zip(ext, v)], omega_c)
    minus = step([x - h * d for x, d in zip(ext, v)], omega_c)
    lv = [(p - m) / (2 * h) for p, m in zip(plus, minus)]
    lhs = sum(float(np.sum(l.astype(np.longdouble) * x)) for l, x in zip(lam1, lv))
    rhs = sum(float(np.sum(l * x)) for l, x in zip(lam0, v))
    scale = sum(float(np.sum(np.abs(l) * np.abs(x))) for l, x in zip(lam0, v))
    assert abs(lhs - rhs) <= 1e-10 * scale, (lhs, rhs, scale)
    hw = np.longdouble(2.0 ** -20)
    dp = step(ext, omega_c + float(hw))
    dm = step(ext, omega_c - float(hw))
    lw = sum(float(np.sum(l.astype(np.longdouble) * (p - m) / (2 * hw))) for l, p, m in zip(lam1, dp, dm))
    assert abs(lw - djdw) <= 1e-9 * max(abs(lw), 1.0), (lw, djdw)


def test_adjoint_3d_collisions_covered(monkeypatch):
    """The 3D adjoint covers BGK, TRT, MRT and RLBM; an unknown collision id
    is refused."""
    f, a, b = small_state_3d("D3Q19", nx=3, ny=3, nz=3)
    for coll in ("BGK", "TRT", "MRT", "RLBM"):
        _, gf, ga, gb, _ = O.reactive_gradient(coll, "f64", f, a, b, 1.5, 1.2, 0.1, 1, lattice="D3Q19")
        assert all(np.all(np.isfinite(x)) for x in (gf, ga, gb))
    monkeypatch.setitem(O.COLLISIONS, "unknown", 4)
    with pytest.raises(ValueError):
        O.reactive_gradient("unknown", "f64", f, a, b, 1.5, 1.2, 0.1, 1, lattice="D3Q19")
