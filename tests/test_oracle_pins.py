"""Pins for the oracle that do not go through the reference's analytic tests.

Each check is chosen so that a plausible restatement mistake (a wrong table
entry, sign, index, transposed transform or dropped term) fails it:
  * the pull permutation with unique ids (omega = 0 makes BGK the identity collision);
  * lattice isotropy of the equilibrium and the rest-state fixed point;
  * conservation of mass and momentum by every collision;
  * TRT == BGK at omega = 1 (magic parameter 1/4 then gives omega_minus = omega_plus);
  * MRT transforms against independently built numpy matrices and inverses;
  * an independent numpy D3Q19 MRT step (matrix inverse by LAPACK) against the oracle;
  * the exact A + B -> C increment against a high-accuracy ODE integration;
  * the f16 / bf16 storage rounding against numpy / torch conversions.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cases
from oracle import oracle as O

VEL = {
    "D2Q9": [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)],
    "D2Q5": [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1)],
}
VEL["D3Q19"] = cases._vel3("D3Q19")
VEL["D3Q27"] = cases._vel3("D3Q27")
FLUIDS = ["D2Q9", "D3Q19", "D3Q27"]
COLLS = ["BGK", "TRT", "MRT", "RLBM"]


@pytest.mark.parametrize("lattice", FLUIDS)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_pull_permutation_kat(lattice, dtype):
    shape = (5, 3) if lattice == "D2Q9" else (3, 5, 7)
    q = O.Q_OF[lattice]
    ids = (1.0 + np.arange(q * int(np.prod(shape)))).reshape((q,) + shape)
    ids = ids.astype(np.float32 if dtype == "f32" else np.float64)
    got = O.fluid_step(lattice, "BGK", dtype, ids, 0.0)
    axes = tuple(range(len(shape) - 1, -1, -1))
    for i, c in enumerate(VEL[lattice]):
        assert np.array_equal(got[i], np.roll(ids[i], shift=tuple(c), axis=axes)), i


@pytest.mark.parametrize("lattice", FLUIDS)
def test_opposites_and_weights(lattice):
    v = np.array(VEL[lattice])
    w = np.array(O.WEIGHTS[lattice])
    assert abs(w.sum() - 1.0) < 1e-15
    # second moment of the weights = cs2 = 1/3 (isotropy)
    for a in range(v.shape[1]):
        for b in range(v.shape[1]):
            m2 = np.sum(w * v[:, a] * v[:, b])
            assert abs(m2 - (1 / 3 if a == b else 0.0)) < 1e-15


@pytest.mark.parametrize("lattice", FLUIDS)
def test_equilibrium_isotropy(lattice):
    g = cases.rng(3)
    for _ in range(5):
        d = O.D_OF[lattice]
        rho, u = 1 + g.uniform(-0.1, 0.1), g.uniform(-0.05, 0.05, size=d)
        f = O.feq_field(lattice, "f64", np.array([rho]), u.reshape(d, 1))[:, 0]
        v = np.array(VEL[lattice])
        assert abs(f.sum() - rho) < 1e-14
        assert np.max(np.abs(v.T @ f - rho * u)) < 1e-15
        # second moment: rho cs2 delta + rho u u
        p = (v.T * f) @ v
        assert np.max(np.abs(p - (rho / 3 * np.eye(d) + rho * np.outer(u, u)))) < 1e-14
        r2, u2, eq = O.macro_feq(lattice, f)
        assert abs(r2 - rho) < 1e-14 and np.max(np.abs(u2 - u)) < 1e-14
        assert np.max(np.abs(eq - f)) < 1e-15


@pytest.mark.parametrize("coll", COLLS)
@pytest.mark.parametrize("lattice", FLUIDS)
def test_rest_state_is_fixed_point_and_conservation(lattice, coll):
    q = O.Q_OF[lattice]
    rest = np.array(O.WEIGHTS[lattice])
    out = O.collide_cell(lattice, coll, rest, 1 / 0.7)
    assert np.max(np.abs(out - rest)) < 1e-15
    g = cases.rng(5)
    v = np.array(VEL[lattice])
    for _ in range(4):
        f = rest * (1 + 0.05 * g.uniform(-1, 1, size=q))
        post = O.collide_cell(lattice, coll, f, 1 / 0.63)
        assert abs(post.sum() - f.sum()) < 4e-15
        assert np.max(np.abs(v.T @ post - v.T @ f)) < 4e-15


@pytest.mark.parametrize("lattice", FLUIDS)
def test_trt_equals_bgk_at_omega_one(lattice):
    g = cases.rng(9)
    rest = np.array(O.WEIGHTS[lattice])
    f = rest * (1 + 0.05 * g.uniform(-1, 1, size=rest.size))
    a = O.collide_cell(lattice, "TRT", f, 1.0)
    b = O.collide_cell(lattice, "BGK", f, 1.0)
    assert np.max(np.abs(a - b)) < 1e-15


@pytest.mark.parametrize("lattice", ["D2Q9", "D3Q27"])
def test_mrt_equilibrium_fixed_point(lattice):
    """D2Q9 and D3Q27 MRT equilibria are the exact moments of feq, so feq is a fixed point."""
    d = O.D_OF[lattice]
    f = O.feq_field(lattice, "f64", np.array([1.02]), np.array([0.03, -0.02, 0.01][:d]).reshape(d, 1))[:, 0]
    out = O.collide_cell(lattice, "MRT", f, 1 / 0.6)
    assert np.max(np.abs(out - f)) < 1e-15


def d2q9_basis():
    rows = []
    for cx, cy in VEL["D2Q9"]:
        c2 = cx * cx + cy * cy
        rows.append([1, -4 + 3 * c2, 4 - 21 * c2 / 2 + 9 * c2 * c2 / 2, cx, (-5 + 3 * c2) * cx, cy,
                     (-5 + 3 * c2) * cy, cx * cx - cy * cy, cx * cy])
    return np.array(rows, dtype=float).T


def d3q19_basis():
    rows = []
    for cx, cy, cz in VEL["D3Q19"]:
        c2 = cx * cx + cy * cy + cz * cz
        p, qq = 3 * cx * cx - c2, cy * cy - cz * cz
        rows.append([1, 19 * c2 - 30, (21 * c2 * c2 - 53 * c2 + 24) / 2, cx, (5 * c2 - 9) * cx, cy,
                     (5 * c2 - 9) * cy, cz, (5 * c2 - 9) * cz, p, qq, cx * cy, cy * cz, cx * cz, (3 * c2 - 5) * p,
                     (3 * c2 - 5) * qq, (cy * cy - cz * cz) * cx, (cz * cz - cx * cx) * cy, (cx * cx - cy * cy) * cz])
    return np.array(rows, dtype=float).T


def d3q27_basis():
    """Rows of lattice_mrt.hpp:246-318 from P0 = 1, P1 = c, P2 = 3c^2 - 2."""
    P = [lambda c: 1.0, lambda c: float(c), lambda c: 3.0 * c * c - 2.0]
    tens = [(1, 0, 0), (0, 1, 0), (0, 0, 1), None, None, None, (1, 1, 0), (0, 1, 1), (1, 0, 1), (2, 1, 0), (2, 0, 1),
            (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2), (1, 1, 1), (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1),
            (1, 2, 1), (1, 1, 2), (2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2)]
    M = np.zeros((27, 27))
    for i, (cx, cy, cz) in enumerate(VEL["D3Q27"]):
        b = lambda a, o, c: P[o](c)  # noqa: E731
        M[0, i] = 1.0
        r200, r020, r002 = P[2](cx), P[2](cy), P[2](cz)
        for r, t in enumerate(tens, start=1):
            if t is not None:
                M[r, i] = b(0, t[0], cx) * b(1, t[1], cy) * b(2, t[2], cz)
        M[4, i] = r200 + r020 + r002
        M[5, i] = 2 * r200 - r020 - r002
        M[6, i] = r020 - r002
    return M


@pytest.mark.parametrize("lattice,build", [("D2Q9", d2q9_basis), ("D3Q19", d3q19_basis), ("D3Q27", d3q27_basis)])
def test_mrt_transforms_against_numpy_matrices(lattice, build):
    M = build()
    q = M.shape[0]
    # orthogonal basis: M M^T is diagonal
    G = M @ M.T
    assert np.max(np.abs(G - np.diag(np.diag(G)))) == 0
    g = cases.rng(1)
    for _ in range(3):
        f = g.uniform(0.0, 1.0, size=q)
        m = O.mrt_forward(lattice, f)
        assert np.max(np.abs(m - M @ f)) < 1e-14
        assert np.max(np.abs(O.mrt_inverse(lattice, m) - np.linalg.solve(M, m))) < 1e-14


def numpy_mrt_d3q19_step(f, omega):
    """D3Q19 MRT step written independently in numpy: pull streaming by np.roll,
    inverse by LAPACK, equilibrium and rates as in lattice_mrt.hpp:908-1009."""
    M = d3q19_basis()
    Minv = np.linalg.inv(M)
    g = np.empty_like(f)
    for i, (cx, cy, cz) in enumerate(VEL["D3Q19"]):
        g[i] = np.roll(f[i], shift=(cz, cy, cx), axis=(0, 1, 2))
    rho = g.sum(0)
    v = np.array(VEL["D3Q19"], dtype=float)
    u = np.tensordot(v.T, g, axes=(1, 0)) / rho
    m = np.tensordot(M, g, axes=(1, 0))
    ux, uy, uz = u
    usq = ux * ux + uy * uy + uz * uz
    eq = np.zeros_like(m)
    eq[0], eq[1], eq[2] = rho, -11 * rho + 19 * rho * usq, 3 * rho - 5.5 * rho * usq
    eq[3], eq[5], eq[7] = rho * ux, rho * uy, rho * uz
    eq[4], eq[6], eq[8] = -2 / 3 * rho * ux, -2 / 3 * rho * uy, -2 / 3 * rho * uz
    eq[9] = 3 * rho * (2 * ux * ux - uy * uy - uz * uz)
    eq[10], eq[11], eq[12], eq[13] = rho * (uy * uy - uz * uz), rho * ux * uy, rho * uy * uz, rho * ux * uz
    s = np.array([0, 1, 1, 0, 1, 0, 1, 0, 1] + [omega] * 5 + [1] * 5, dtype=float)
    m = m - s[:, None, None, None] * (m - eq)
    return np.tensordot(Minv, m, axes=(1, 0))


def test_d3q19_mrt_against_independent_numpy():
    f = cases.random_perturbed("D3Q19", (4, 5, 6), "f64", seed=2)
    a = O.fluid_run("D3Q19", "MRT", "f64", f, 1 / 0.7, 3)
    b = f
    for _ in range(3):
        b = numpy_mrt_d3q19_step(b, 1 / 0.7)
    assert np.max(np.abs(a - b)) < 1e-14


def test_reaction_source_against_ode_integration():
    from scipy.integrate import solve_ivp

    g = cases.rng(4)
    a = np.concatenate([g.uniform(0, 1.5, 40), [0.7, 0.3, 0.0, 1.0]])
    b = np.concatenate([g.uniform(0, 1.5, 40), [0.7, 0.3 + 1e-15, 0.5, 0.0]])
    k = 0.37
    src = O.reaction_source("f64", a, b, k)
    for ai, bi, si in zip(a, b, src):
        sol = solve_ivp(lambda t, y: [-k * y[0] * y[1], -k * y[0] * y[1]], (0, 1), [ai, bi], rtol=1e-12,
                        atol=1e-14, method="DOP853")
        assert abs((ai + si) - sol.y[0, -1]) < 1e-11, (ai, bi)


def test_reaction_batch_closed_form():
    # A = B: C(t) = C0 / (1 + C0 k t) after one unit step (the |A - B| <= tol branch)
    for c0 in (0.1, 1.0, 2.5):
        s = O.reaction_source("f64", np.array([c0]), np.array([c0]), 0.01)[0]
        assert abs((c0 + s) - c0 / (1 + 0.01 * c0)) < 1e-16


def test_scalar_collision_conserves_concentration_without_source():
    g = cases.random_perturbed("D2Q5", (6, 8), "f64", seed=3, eps=0.1)
    f = cases.random_perturbed("D2Q9", (6, 8), "f64", seed=4)
    out = O.scalar_step("D2Q9", "D2Q5", "f64", g, f, 1.3, 0.0)
    assert abs(out.sum() - g.sum()) < 1e-13


def test_reactive_stats_closed_forms():
    n = 16
    u = np.zeros((2, n, n))
    a, b = cases.split_species(n, n, u)
    for wide in (False, True):
        s = O.reactive_stats("D2Q5", "f64", a, b, 0.2, wide=wide)
        assert np.allclose(s, [0.5, 0.5, 0.25, 0.25, -0.25, -1.0, 0.0, 0.05], atol=1e-15)
    c = O.scalar_feq_field("D2Q5", "f64", np.full((n, n), 0.3), u)
    s = O.reactive_stats("D2Q5", "f64", c, c, 1.0, wide=True)
    assert abs(s[2]) < 1e-16 and abs(s[4]) < 1e-16 and abs(s[5]) < 1e-15


def test_half_codecs_match_library_rounding():
    import torch

    g = cases.rng(6)
    x = np.concatenate([g.normal(0, 1, 20000) * 10.0 ** g.integers(-9, 6, 20000),
                        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e-8, 2.0 ** -25, 2.0 ** -24, 3 * 2.0 ** -26,
                         np.inf, -np.inf]]).astype(np.float32)
    h, back = O.half_roundtrip(x, bf16=False)
    with np.errstate(over="ignore"):  # values beyond 65520 round to infinity, as intended
        want = x.astype(np.float16)
    assert np.array_equal(h, want.view(np.uint16))
    assert np.array_equal(back, want.astype(np.float32))
    hb, backb = O.half_roundtrip(x, bf16=True)
    tb = torch.from_numpy(x).to(torch.bfloat16)
    assert np.array_equal(hb, tb.view(torch.int16).numpy().view(np.uint16))
    assert np.array_equal(backb, tb.to(torch.float32).numpy())


def test_reduced_storage_step_is_float_step_with_rounded_io():
    """An f16-storage step equals: decode, float step, encode (the storage contract)."""
    f = cases.random_perturbed("D3Q19", (3, 4, 5), "f64", seed=8)
    h = cases.encode("D3Q19", "f16", f)
    dec = O.decode("D3Q19", "f16", h).astype(np.float32)
    step32 = O.fluid_step("D3Q19", "BGK", "f32", dec, 1.5)
    want = cases.encode("D3Q19", "f16", step32.astype(np.float64))
    got = O.fluid_step("D3Q19", "BGK", "f16", h, 1.5)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("lattice", FLUIDS)
def test_macro_fields_recover_equilibrium_moments(lattice):
    """Per-cell fields of feq(rho, u) give back rho and u (lattice isotropy, in
    exact arithmetic; here to rounding), in [Z][Y][X] cell order, and the
    stability guard's max speed is the largest |u|."""
    shape = (6, 7) if lattice == "D2Q9" else (3, 4, 5)
    g = cases.rng(21)
    d = O.D_OF[lattice]
    rho = 1.0 + g.uniform(-0.05, 0.05, size=shape)
    u = g.uniform(-0.03, 0.03, size=(d,) + shape)
    f = O.feq_field(lattice, "f64", rho, u)
    r, v = O.macro_fields(lattice, "f64", f)
    assert r.shape == shape and v.shape == (d,) + shape
    assert np.allclose(r, rho, rtol=0, atol=1e-15) and np.allclose(v, u, rtol=0, atol=1e-15)
    speed = np.sqrt((u * u).sum(0))
    assert abs(O.max_speed(lattice, "f64", f) - speed.max()) < 1e-15
    assert O.max_speed(lattice, "f64", f) == O.fluid_diagnostics(lattice, "f64", f)[2]
    # extended precision: only the final rounding to double remains
    rl, vl = O.macro_fields(lattice, "f80", O.feq_field(lattice, "f80", rho, u))
    assert np.allclose(rl, rho, rtol=0, atol=2.3e-16) and np.allclose(vl, u, rtol=0, atol=1e-17)


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_max_speed_reports_non_finite_as_infinity(bad):
    """compute_max_velocity returns infinity for any non-finite speed (NaN
    included, which a plain max would skip)."""
    f = cases.random_perturbed("D3Q19", (3, 4, 5), seed=22)
    assert np.isfinite(O.max_speed("D3Q19", "f64", f))
    f[7, 2, 1, 3] = bad
    assert O.max_speed("D3Q19", "f64", f) == np.inf
    if np.isnan(bad):  # log_diagnostics' max keeps skipping the NaN cell
        assert np.isfinite(O.fluid_diagnostics("D3Q19", "f64", f)[2])


@pytest.mark.parametrize("sl", ["D2Q5", "D3Q7"])
def test_scalar_fields_recover_concentration(sl):
    shape = (6, 7) if sl == "D2Q5" else (3, 4, 5)
    g = cases.rng(23)
    c = g.uniform(0.0, 1.0, size=shape)
    u = g.uniform(-0.03, 0.03, size=(O.D_OF[sl],) + shape)
    p = O.scalar_feq_field(sl, "f64", c, u)
    assert np.allclose(O.scalar_fields(sl, "f64", p), c, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_scalar_storage_reference_codec(dtype):
    """F16/BF16 scalar storage holds g - w ref: a field at rest at its
    reference concentration stores exact zeros and decodes to w ref (the
    float32 product), and a field near ref round-trips more closely than with
    no reference."""
    shape = (6, 7)
    u0 = np.zeros((2,) + shape)
    ref = 0.5
    rest = O.scalar_feq_field("D2Q5", dtype, np.full(shape, ref), u0, ref)
    assert not np.any(rest)  # every stored word is +0
    w32 = np.asarray(O.WEIGHTS["D2Q5"], dtype=np.float32).reshape(-1, 1, 1)
    assert np.array_equal(O.decode("D2Q5", dtype, rest, ref), np.broadcast_to((w32 * np.float32(ref)).astype(
        np.float64), rest.shape))
    g = cases.rng(31)
    conc = ref + g.uniform(-0.05, 0.05, size=shape)
    u = g.uniform(-0.02, 0.02, size=(2,) + shape)
    exact = O.scalar_feq_field("D2Q5", "f64", conc, u)
    err = {r: np.max(np.abs(O.decode("D2Q5", dtype, O.scalar_feq_field("D2Q5", dtype, conc, u, r), r) - exact))
           for r in (0.0, ref)}
    assert err[ref] < err[0.0] / 4, err
    # the test helper's encoding (offset formed in float32 too) decodes as closely
    helper = cases.encode("D2Q5", dtype, exact, ref)
    assert np.max(np.abs(O.decode("D2Q5", dtype, helper, ref) - exact)) <= 1.01 * err[ref] + 1e-9
