"""Pins of the oracle's regularized collision (RLBM) and of the D3Q7 scalar lattice
from the mathematics alone (no restated formula is reused here).

Regularized collision (include/lattice_physics.hpp:422-605).  The post-collision
state is f' = feq + (1 - omega) f_neq,reg, where f_neq,reg is the element of
w * span{polynomials of degree <= 3 in c} whose moments are those of the
recursive regularized model (Malaspinas, "Increasing stability and accuracy of
the lattice Boltzmann scheme: recursivity and regularization",
arXiv:1505.06900, 2015):
  * order 0 and 1:  sum f_neq,reg = 0,  sum f_neq,reg c = 0;
  * order 2:        sum f_neq,reg c_a c_b = Pi^neq_ab = sum (f - feq) c_a c_b;
  * order 3:        sum f_neq,reg H_abc = a3_abc = u_a Pi_bc + u_b Pi_ac + u_c Pi_ab
                    for every Hermite polynomial H_abc = c_a c_b c_c -
                    cs2 (c_a d_bc + c_b d_ac + c_c d_ab) that does not vanish on
                    the lattice (H_aaa vanishes on all three lattices because
                    c^3 = c and cs2 = 1/3; H_xyz vanishes on D3Q19, which has
                    no corner links).
Because sum w p q is an inner product on the polynomials restricted to the
lattice, these moments determine f_neq,reg uniquely; a dropped, sign-flipped or
mis-weighted third-order term (or a wrong D3Q19 pairing) changes the order-3
moments and fails.  feq is formed here from its closed form with cs2 = 1/3.

D3Q7 (include/lattice_traits.hpp:266-305, cs2 = 1/4 at :279; the scalar
equilibrium w C (1 + c.u / cs2) of include/lattice_physics.hpp:643-659).  The
lattice's own second moment sum w c c is 1/4, so the equilibrium's first
moment is C u only if the divide uses cs2 = 1/4; the advected decaying sine
then travels at u and decays with D = cs2 (tau_c - 1/2), the reference's
PassiveScalarAdvectionDiffusion test (tests/test_validation.cpp:1590-1646)
lifted to three dimensions.  Each check goes through the oracle's own scalar
step, so a wrong cs2 in the oracle fails it.
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

import cases
from oracle import oracle as O

CS2 = 1.0 / 3.0
VEL = {
    "D2Q9": np.array([(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)], dtype=float),
    "D3Q19": np.array(cases._vel3("D3Q19"), dtype=float),
    "D3Q27": np.array(cases._vel3("D3Q27"), dtype=float),
    "D2Q5": np.array([(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1)], dtype=float),
    "D3Q7": np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], dtype=float),
}


def closed_form_feq(lattice, rho, u):
    c, w = VEL[lattice], np.asarray(O.WEIGHTS[lattice])
    cu = c @ u
    return w * rho * (1 + cu / CS2 + cu * cu / (2 * CS2 * CS2) - (u @ u) / (2 * CS2))


def hermite3(c, cs2):
    """H_abc(c_i) for every index triple, shape (Q, D, D, D)."""
    eye = np.eye(c.shape[1])
    return (np.einsum("qa,qb,qc->qabc", c, c, c) -
            cs2 * (np.einsum("qa,bc->qabc", c, eye) + np.einsum("qb,ac->qabc", c, eye) +
                   np.einsum("qc,ab->qabc", c, eye)))


def monomials(c, degree):
    """Every monomial of the velocity components up to `degree`, evaluated per direction."""
    cols = [np.ones(c.shape[0])]
    for k in range(1, degree + 1):
        for idx in itertools.combinations_with_replacement(range(c.shape[1]), k):
            cols.append(np.prod(c[:, list(idx)], axis=1))
    return np.stack(cols, axis=1)


def random_cell(lattice, g, u_scale=0.1, noise=0.05):
    d = VEL[lattice].shape[1]
    rho = 1.0 + g.uniform(-0.05, 0.05)
    u = g.uniform(-u_scale, u_scale, size=d)
    w = np.asarray(O.WEIGHTS[lattice])
    return closed_form_feq(lattice, rho, u) + noise * w * g.uniform(-1, 1, size=w.size)


@pytest.mark.parametrize("omega", [1 / 0.6, 1 / 0.93, 1.95])
@pytest.mark.parametrize("lattice", ["D2Q9", "D3Q19", "D3Q27"])
def test_rlbm_hermite_moments(lattice, omega):
    c, w = VEL[lattice], np.asarray(O.WEIGHTS[lattice])
    H3 = hermite3(c, CS2)
    vanishing = np.all(np.abs(H3) < 1e-15, axis=0)
    g = cases.rng(0x4E5)
    for _ in range(6):
        f = random_cell(lattice, g)
        rho = f.sum()
        u = (c.T @ f) / rho
        eq = closed_form_feq(lattice, rho, u)
        pi = np.einsum("q,qa,qb->ab", f - eq, c, c)
        a3 = np.einsum("a,bc->abc", u, pi) + np.einsum("b,ac->abc", u, pi) + np.einsum("c,ab->abc", u, pi)
        assert np.max(np.abs(a3)) > 1e-5  # the third-order moments are far above rounding
        post = O.collide_cell(lattice, "RLBM", f, omega)
        d = post - eq
        # orders 0-2
        assert abs(d.sum()) < 1e-15
        assert np.max(np.abs(c.T @ d)) < 1e-15
        assert np.max(np.abs(np.einsum("q,qa,qb->ab", d, c, c) - (1 - omega) * pi)) < 1e-15
        # order 3: the recursive a3 on every Hermite polynomial the lattice carries
        got = np.einsum("q,qabc->abc", d, H3)
        want = np.where(vanishing, 0.0, (1 - omega) * a3)
        assert np.max(np.abs(got - want)) < 2e-16, (lattice, np.max(np.abs(got - want)))
        # f_neq,reg / w is a polynomial of degree <= 3 on the lattice
        B = monomials(c, 3)
        coef, *_ = np.linalg.lstsq(B, d / w, rcond=None)
        assert np.max(np.abs(B @ coef - d / w)) < 1e-13


def test_rlbm_d3q19_paired_projection_reproduces_each_component():
    """On D3Q19 the pairs H_xxy +- H_yzz (and the two cyclic pairs) carry the
    third order (lattice_physics.hpp:523-529); the projection must return a3 on
    each pair, hence on each of the six mixed components."""
    c = VEL["D3Q19"]
    H3 = hermite3(c, CS2)
    g = cases.rng(17)
    f = random_cell("D3Q19", g)
    rho = f.sum()
    u = (c.T @ f) / rho
    eq = closed_form_feq("D3Q19", rho, u)
    pi = np.einsum("q,qa,qb->ab", f - eq, c, c)
    a3 = np.einsum("a,bc->abc", u, pi) + np.einsum("b,ac->abc", u, pi) + np.einsum("c,ab->abc", u, pi)
    omega = 1 / 0.7
    d = O.collide_cell("D3Q19", "RLBM", f, omega) - eq
    X, Y, Z = 0, 1, 2
    for (p, q) in [((X, X, Y), (Y, Z, Z)), ((X, Z, Z), (X, Y, Y)), ((Y, Y, Z), (X, X, Z))]:
        for sgn in (1.0, -1.0):
            h = H3[(slice(None),) + p] + sgn * H3[(slice(None),) + q]
            assert abs(d @ h - (1 - omega) * (a3[p] + sgn * a3[q])) < 2e-16


# ---------------------------------------------------------------------------
# Scalar lattices: each lattice's own cs2
# ---------------------------------------------------------------------------
SCALAR_CS2 = {"D2Q5": 1.0 / 3.0, "D3Q7": 1.0 / 4.0}
FLUID_OF = {"D2Q5": "D2Q9", "D3Q7": "D3Q19"}


@pytest.mark.parametrize("sl", ["D2Q5", "D3Q7"])
def test_scalar_lattice_isotropy(sl):
    c, w = VEL[sl], np.asarray(O.WEIGHTS[sl])
    assert abs(w.sum() - 1) < 3e-16
    assert np.max(np.abs(c.T @ w)) == 0
    assert np.max(np.abs(np.einsum("q,qa,qb->ab", w, c, c) - SCALAR_CS2[sl] * np.eye(c.shape[1]))) < 1e-16


@pytest.mark.parametrize("sl", ["D2Q5", "D3Q7"])
def test_scalar_equilibrium_moments_through_the_step(sl):
    """One oracle scalar step with omega_c = 1 on a one-cell periodic domain
    (streaming is the identity there) returns the scalar equilibrium of the
    cell's concentration and the fluid's velocity: its moments must be C,
    C u and C cs2 I with the lattice's own cs2."""
    fl = FLUID_OF[sl]
    d = VEL[sl].shape[1]
    shape = (1,) * d
    g = cases.rng(41)
    for _ in range(5):
        u = g.uniform(-0.06, 0.06, size=d)
        rho = 1 + g.uniform(-0.05, 0.05)
        f = O.feq_field(fl, "f64", np.full(shape, rho), u.reshape((d,) + shape))
        rho_f, u_f, _ = O.macro_feq(fl, f.reshape(-1))
        w = np.asarray(O.WEIGHTS[sl])
        gin = (w * (0.7 + 0.3 * g.uniform(0, 1, size=w.size))).reshape((-1,) + shape)
        conc = gin.sum()
        out = O.scalar_step(fl, sl, "f64", np.ascontiguousarray(gin), f, 1.0, 0.0).reshape(-1)
        c = VEL[sl]
        assert abs(out.sum() - conc) < 1e-15
        assert np.max(np.abs(c.T @ out - conc * u_f)) < 1e-16
        assert np.max(np.abs(np.einsum("q,qa,qb->ab", out, c, c) - conc * SCALAR_CS2[sl] * np.eye(d))) < 1e-15
        # and the field initialiser writes the same equilibrium
        eqf = O.scalar_feq_field(sl, "f64", np.full(shape, conc), u_f.reshape((d,) + shape)).reshape(-1)
        assert np.max(np.abs(c.T @ eqf - conc * u_f)) < 1e-16


def test_d3q7_advected_decaying_sine():
    """test_validation.cpp:1590-1646 in 3D on D3Q7: a uniform D3Q19 flow carries a
    sine wave whose wave vector has all three components; the concentration
    must follow C0 + A sin(k.(x - u t)) exp(-D |k|^2 t) with D = cs2 (tau_c - 1/2)
    and D3Q7's cs2 = 1/4, within the reference's 1e-4 (measured 7.4e-5 at 32^3
    and 1.8e-5 at 64^3 with u halved and four times the steps: second order).
    With cs2 = 1/3 in the equilibrium the wave would travel at 3u/4, an error
    of ~3e-2 here."""
    n, steps, tau_c = 32, 200, 0.8
    u0 = np.array([0.04, -0.03, 0.02])
    kv = 2 * math.pi / n * np.array([1.0, 1.0, 1.0])
    shape = (n, n, n)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.stack([np.full(shape, v) for v in u0])
    f = O.feq_field("D3Q19", "f64", np.ones(shape), u)
    phase = kv[0] * x + kv[1] * y + kv[2] * z
    gfield = O.scalar_feq_field("D3Q7", "f64", 1.0 + 0.1 * np.sin(phase), u)
    for _ in range(steps):
        f = O.fluid_step("D3Q19", "BGK", "f64", f, 1.0, threads=8)
        gfield = O.scalar_step("D3Q19", "D3Q7", "f64", gfield, f, 1 / tau_c, 0.0, threads=8)
    conc = gfield.sum(0)
    diffusivity = SCALAR_CS2["D3Q7"] * (tau_c - 0.5)
    want = 1.0 + 0.1 * np.sin(phase - (u0 @ kv) * steps) * math.exp(-diffusivity * (kv @ kv) * steps)
    assert np.max(np.abs(conc - want)) < 1e-4
