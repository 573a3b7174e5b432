"""LB-Cube's own validation suite (tests/test_validation.cpp) restated against the oracle.

Each test below names the reference test it restates and keeps its workload,
parameters and tolerance.  These pin the oracle's physics to the reference's
analytic expectations (decay rates, closed forms, convergence orders); the GPU
kernels are then pinned to the oracle bit for bit in test_gpu_parity.py.

Deviations, all forced by what this environment can run:
  * random initial states use numpy's PCG64 instead of std::uniform_real_distribution
    (whose stream is implementation-defined), same seed and ranges;
  * energies and velocities are formed in numpy float64 (the reference uses
    long double accumulators); the tolerances are orders of magnitude looser.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import cases
from oracle import oracle as O

T = 8  # oracle OpenMP threads; results do not depend on it
CS2 = 1.0 / 3.0

# Three of the reference's D3Q19-MRT expectations do not hold for its own
# algorithm as written, restated here term for term (pinned to the reference's
# text by test_oracle_pins.py: literal forward rows, inverse coefficients and
# term order, and an independent numpy MRT step):
#   * angled shear and 3D convergence expect a 2-step lag of the energy decay
#     (test_validation.cpp:809-814, :989-991); the algorithm lags 0.96 steps
#     at the first sample (D3Q27 MRT: 1.0, BGK: 1.16), outside 1e-3 E0;
#   * the advected shear (Galilean invariance) misses the band by 2.7e-3 E0 at
#     step 100.  The equilibrium of moment 9 carries a factor 3 that the
#     moments of feq do not have (lattice_mrt.hpp:930, 3 rho (2ux^2 - uy^2 -
#     uz^2) against sum (3cx^2 - c^2) feq = rho (2ux^2 - uy^2 - uz^2)); without
#     it the advected case passes (-8.7e-4 E0) and the lag becomes 0.75.
# The oracle and the kernels keep the reference's algorithm; the three
# expectations are strict xfails.
D3Q19_MRT_SHIFT = pytest.mark.xfail(strict=True, reason="the reference's D3Q19 MRT expectation does not hold for "
                                                        "its own algorithm (see module comment)")


def run(lattice, coll, f, omega, steps):
    return O.fluid_run(lattice, coll, "f64", f, omega, steps, threads=T)


def mean_ke(lattice, f):
    """compute_mean_kinetic_energy (test_validation.cpp:80-139): mean of |u|^2 / 2."""
    _, u = cases.macro(lattice, f)
    return float(np.mean(0.5 * (u * u).sum(0)))


def total_ke(lattice, f):
    """kinetic_energy_d2q9 (test_validation.cpp:224-239): sum of rho |u|^2 / 2."""
    rho, u = cases.macro(lattice, f)
    return float(np.sum(0.5 * rho * (u * u).sum(0)))


def shear_wave_2d(nx, ny, u0):
    k = 2 * math.pi / ny
    y = np.arange(ny).reshape(-1, 1) * np.ones((1, nx))
    u = np.stack([u0 * np.sin(k * y), np.zeros((ny, nx))])
    return O.feq_field("D2Q9", "f64", np.ones((ny, nx)), u)


def angled_shear_3d(lattice, n, amp, background=0.0):
    k = 2 * math.pi / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    t = amp * np.sin(k * (x + y + z))
    u = np.stack([background + t, background - 0.5 * t, background - 0.5 * t])
    return O.feq_field(lattice, "f64", np.ones((n, n, n)), u)


# --- FluidDynamics3D.MassConservationD3Q19Cpu (test_validation.cpp:1283-1303) ---
def test_mass_conservation_d3q19():
    f = cases.random_equilibrium("D3Q19", (32, 32, 32), seed=0x5EED1234)
    m0 = float(np.sum(f, dtype=np.longdouble))
    g = run("D3Q19", "BGK", f, 1.0, 100)
    m1 = float(np.sum(g, dtype=np.longdouble))
    assert abs(m1 - m0) <= np.finfo(np.float64).eps * m0 * 1024.0


# --- FluidDynamics2D.ShearWaveDecay_{BGK,TRT,MRT} (test_validation.cpp:1308-1324, body :1140-1182) ---
@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT"])
def test_shear_wave_decay(coll):
    nx, ny, steps, u0, tau = 8, 64, 500, 0.05, 0.8
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / 64
    g = run("D2Q9", coll, shear_wave_2d(nx, ny, u0), 1 / tau, steps)
    rho, u = cases.macro("D2Q9", g)
    want = u0 * np.sin(k * np.arange(ny)) * math.exp(-nu * k * k * steps)
    assert np.max(np.abs(u[0] - want.reshape(-1, 1))) < 1e-4
    assert np.max(np.abs(u[1])) < 1e-7
    assert np.max(np.abs(rho - 1.0)) <= (1e-6 if coll == "MRT" else 1e-12)


# --- run_tke_shear_wave_test (test_validation.cpp:651-689): ShearWaveDecay_RLBM and TKE_ShearWave_D2Q9_* ---
@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
def test_tke_shear_wave(coll):
    n, tau, u0 = 32, 0.8, 0.01
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / n
    f = shear_wave_2d(n, n, u0)
    e0 = mean_ke("D2Q9", f)
    shift = 1.0 if coll == "MRT" else 0.5
    for step in range(100, 1001, 100):
        f = run("D2Q9", coll, f, 1 / tau, 100)
        want = e0 * math.exp(-2 * nu * k * k * (step + shift))
        assert abs(mean_ke("D2Q9", f) - want) <= 1e-3 * e0, step


# --- TaylorGreenVortex_{BGK,TRT,MRT} (test_validation.cpp:1336-1352, body :577-602) ---
@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT"])
def test_taylor_green_2d_total_energy(coll):
    n, tau, u0, steps = 64, 0.8, 0.01, 1000
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / n
    f, _ = cases.taylor_green_2d(n, u0)
    e0 = total_ke("D2Q9", f)
    g = run("D2Q9", coll, f, 1 / tau, steps)
    assert abs(total_ke("D2Q9", g) - e0 * math.exp(-4 * nu * k * k * steps)) <= 1e-5


# --- run_tke_tgv2d_test (test_validation.cpp:701-740): TaylorGreenVortex_RLBM and TKE_TaylorGreen2D_* ---
@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT", "RLBM"])
def test_tke_taylor_green_2d(coll):
    n, tau, u0 = 32, 0.8, 0.01
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / n
    f, _ = cases.taylor_green_2d(n, u0)
    e0 = mean_ke("D2Q9", f)
    for step in range(100, 1001, 100):
        f = run("D2Q9", coll, f, 1 / tau, 100)
        assert abs(mean_ke("D2Q9", f) - e0 * math.exp(-4 * nu * k * k * (step + 0.5))) <= 1e-3 * e0, step


# --- TaylorGreenVortex3D_{BGK,TRT,MRT} (test_validation.cpp:1364-1380, body :610-639) ---
@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT"])
def test_taylor_green_3d_velocity(coll):
    n, tau, u0, steps = 32, 0.8, 0.005, 20
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / n
    g = run("D3Q19", coll, cases.taylor_green_3d(n, u0), 1 / tau, steps)
    _, u = cases.macro("D3Q19", g)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d = math.exp(-3 * nu * k * k * steps)
    ua = np.stack([u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z) * d,
                   -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z) * d, np.zeros_like(x, dtype=float)])
    rel = math.sqrt(np.sum((u - ua) ** 2) / np.sum(ua ** 2))
    assert rel < 7.5e-2


# --- TKE_AngledShear3D_* / AngledShear3D_RLBM* (test_validation.cpp:1427-1478, body :756-833) ---
@pytest.mark.parametrize("lattice,coll", [("D3Q19", "BGK"), ("D3Q27", "BGK"), ("D3Q19", "TRT"), ("D3Q27", "TRT"),
                                          ("D3Q27", "MRT"), pytest.param("D3Q19", "MRT", marks=D3Q19_MRT_SHIFT),
                                          ("D3Q19", "RLBM"), ("D3Q27", "RLBM")])
def test_tke_angled_shear_3d(lattice, coll):
    n, tau, u0 = 32, 0.8, 0.01
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / n
    f = angled_shear_3d(lattice, n, u0)
    e0 = mean_ke(lattice, f)
    shift = {"BGK": 1.0, "TRT": 0.0, "RLBM": 1.0, "MRT": 2.0 if lattice == "D3Q19" else 1.0}[coll]
    for step in range(100, 1001, 100):
        f = run(lattice, coll, f, 1 / tau, 100)
        want = e0 * math.exp(-2 * nu * 3 * k * k * (step + shift))
        assert abs(mean_ke(lattice, f) - want) <= 1e-3 * e0, step


# --- AdvectedShear3D_* (test_validation.cpp:1483-1527, body :847-924); the BGK case is DISABLED there ---
@pytest.mark.parametrize("lattice,coll", [("D3Q19", "TRT"), ("D3Q27", "TRT"), ("D3Q27", "MRT"), ("D3Q27", "RLBM"),
                                          pytest.param("D3Q19", "MRT", marks=D3Q19_MRT_SHIFT), ("D3Q19", "RLBM")])
def test_advected_shear_3d(lattice, coll):
    n, tau, v0, amp = 32, 0.8, 0.02, 0.005
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / n
    f = angled_shear_3d(lattice, n, amp, background=v0)
    eb = 0.5 * 3 * v0 * v0
    e0 = mean_ke(lattice, f) - eb
    for step in range(100, 1001, 100):
        f = run(lattice, coll, f, 1 / tau, 100)
        want = e0 * math.exp(-2 * nu * 3 * k * k * step)
        assert abs(mean_ke(lattice, f) - eb - want) <= 1e-3 * e0, step


# --- MathAndTransforms.MRT_*_Transformation_Identity (test_validation.cpp:1532-1582) ---
def test_mrt_identity_d2q9():
    f = np.array([1.125, 2.25, 3.375, 4.5, 5.625, 6.75, 7.875, 9.0, 10.125])
    back = O.mrt_inverse("D2Q9", O.mrt_forward("D2Q9", f))
    # EXPECT_DOUBLE_EQ: within 4 ulps
    assert np.all(np.abs(back - f) <= 4 * np.spacing(f))


@pytest.mark.parametrize("lattice,q", [("D3Q19", 19), ("D3Q27", 27)])
def test_mrt_identity_3d(lattice, q):
    f = 0.01 * np.arange(1, q + 1)
    back = O.mrt_inverse(lattice, O.mrt_forward(lattice, f))
    assert np.max(np.abs(back - f)) <= 1e-12


# --- ReactiveTransport.PassiveScalarAdvectionDiffusion (test_validation.cpp:1590-1646) ---
def test_passive_scalar_advection_diffusion():
    nx, ny, steps, ux, tau_c = 64, 8, 200, 0.05, 0.8
    dc, k = CS2 * (tau_c - 0.5), 2 * math.pi / nx
    u = np.stack([np.full((ny, nx), ux), np.zeros((ny, nx))])
    f = O.feq_field("D2Q9", "f64", np.ones((ny, nx)), u)
    x = np.arange(nx) * np.ones((ny, 1))
    g = O.scalar_feq_field("D2Q5", "f64", 1.0 + 0.1 * np.sin(k * x), u)
    for _ in range(steps):
        f = O.fluid_step("D2Q9", "BGK", "f64", f, 1.0)
        g = O.scalar_step("D2Q9", "D2Q5", "f64", g, f, 1 / tau_c, 0.0)
    conc = g.sum(0)
    want = 1.0 + 0.1 * np.sin(k * (x - ux * steps)) * math.exp(-dc * k * k * steps)
    assert np.max(np.abs(conc - want)) < 1e-4


# --- MathAndTransforms.ReactionKineticsBatch (test_validation.cpp:1655-1702) ---
def test_reaction_kinetics_batch():
    n, steps, c0, kr, tau_c = 16, 200, 1.0, 0.005, 0.8
    u = np.zeros((2, n, n))
    f = O.feq_field("D2Q9", "f64", np.ones((n, n)), u)
    a = O.scalar_feq_field("D2Q5", "f64", np.full((n, n), c0), u)
    b = a.copy()
    for _ in range(steps):  # the fluid is never stepped in the reference test
        a, b = O.reaction_step("D2Q9", "D2Q5", "f64", a, b, f, 1 / tau_c, kr)
    want = c0 / (1.0 + c0 * kr * steps)
    assert abs(a[:, 0, 0].sum() - want) < 1e-4 and abs(b[:, 0, 0].sum() - want) < 1e-4


# --- ReactiveTransport.SegregatedReactiveMixingD2Q5Cpu (test_validation.cpp:1712-1760) ---
def test_segregated_reactive_mixing():
    n, steps, kr, tau_c = 32, 20, 0.05, 0.8
    u = np.zeros((2, n, n))
    f = O.feq_field("D2Q9", "f64", np.ones((n, n)), u)
    a, b = cases.split_species(n, n, u)
    ma0, mb0 = a.sum(), b.sum()
    for _ in range(steps):
        a, b = O.reaction_step("D2Q9", "D2Q5", "f64", a, b, f, 1 / tau_c, kr)
    ma, mb = a.sum(), b.sum()
    assert ma < ma0 and mb < mb0 and abs(ma - mb) <= 1e-10
    ca, cb = a.sum(0), b.sum(0)
    assert np.all(np.isfinite(ca)) and np.all(np.isfinite(cb))
    assert np.all(ca >= -1e-12) and np.all(cb >= -1e-12)


# --- Convergence.* (test_validation.cpp:1765-1897, bodies :941-1276) ---
def taylor_green_error(n, coll):
    q = n / 16
    ul, nul = 0.05 / q, CS2 * (0.8 - 0.5) / q
    tau = nul / CS2 + 0.5
    steps = int(round(256 * q * q))
    k = 2 * math.pi / n
    f, _ = cases.taylor_green_2d(n, ul)
    g = run("D2Q9", coll, f, 1 / tau, steps)
    _, u = cases.macro("D2Q9", g)
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    d = math.exp(-2 * nul * k * k * steps)
    sa = np.hypot(-ul * np.cos(k * x) * np.sin(k * y) * d, ul * np.sin(k * x) * np.cos(k * y) * d)
    s = np.sqrt((u * u).sum(0))
    return math.sqrt(np.sum((s - sa) ** 2) / np.sum(sa ** 2))


@pytest.mark.parametrize("coll,tol", [("BGK", 0.15), ("MRT", 0.35)])
def test_spatial_convergence_2d(coll, tol):
    e16, e32, e64 = (taylor_green_error(n, coll) for n in (16, 32, 64))
    assert math.log2(e16 / e32) > 1.5
    assert abs(math.log2(e32 / e64) - 2.0) <= tol


def shear_error_3d(lattice, coll, n):
    q = n / 16
    ul, nul = 0.01 / q, CS2 * (0.8 - 0.5) / q
    tau = 3 * nul + 0.5
    steps = int(256 * q * q)
    k = 2 * math.pi / n
    g = run(lattice, coll, angled_shear_3d(lattice, n, ul), 1 / tau, steps)
    shift = 2.0 if (coll == "MRT" and lattice == "D3Q19") else (1.0 if coll == "MRT" else 0.0)
    d = math.exp(-nul * 3 * k * k * (steps + shift))
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sa = math.sqrt(1.5) * np.abs(ul * np.sin(k * (x + y + z))) * d
    _, u = cases.macro(lattice, g)
    s = np.sqrt((u * u).sum(0))
    return math.sqrt(np.sum((s - sa) ** 2) / np.sum(sa ** 2))


@pytest.mark.parametrize("lattice,coll,floor", [pytest.param("D3Q19", "MRT", 1.85, marks=D3Q19_MRT_SHIFT),
                                                ("D3Q19", "TRT", 1.85),
                                                ("D3Q19", "RLBM", 0.80), ("D3Q27", "TRT", 1.85),
                                                ("D3Q27", "MRT", 1.85), ("D3Q27", "RLBM", 0.80)])
def test_spatial_convergence_3d(lattice, coll, floor):
    e16, e32 = shear_error_3d(lattice, coll, 16), shear_error_3d(lattice, coll, 32)
    assert math.log2(e16 / e32) >= floor, (e16, e32)


def temporal_error(coll, m):
    def go(mult):
        ul, nul = 0.08 / mult, CS2 * 0.3 / mult
        tau = 3 * nul + 0.5
        return run("D2Q9", coll, shear_wave_2d(32, 32, ul), 1 / tau, 400 * mult)

    c, r = go(m), go(2 * m)
    _, uc = cases.macro("D2Q9", c)
    _, ur = cases.macro("D2Q9", r)
    pc, pr = m * uc[0], 2 * m * ur[0]
    return math.sqrt(np.sum((pc - pr) ** 2) / np.sum(pr ** 2))


@pytest.mark.parametrize("coll,floor", [("BGK", 1.85), ("TRT", 1.85), ("MRT", 1.85), ("RLBM", 0.80)])
def test_temporal_convergence(coll, floor):
    assert math.log2(temporal_error(coll, 1) / temporal_error(coll, 2)) >= floor


def test_d3q19_mrt_observed_lag():
    """Documents the lag the restated D3Q19 MRT does show (regression pin)."""
    n, tau, u0 = 32, 0.8, 0.01
    nu, k = CS2 * (tau - 0.5), 2 * math.pi / n
    f = angled_shear_3d("D3Q19", n, u0)
    e0 = mean_ke("D3Q19", f)
    f = run("D3Q19", "MRT", f, 1 / tau, 100)
    lag = -math.log(mean_ke("D3Q19", f) / e0) / (2 * nu * 3 * k * k) - 100
    assert 0.9 < lag < 1.0, lag
