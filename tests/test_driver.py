"""The C++ drivers lb-cube_amd/lbm_sim, lbm_reaction_sim and
lbm_compare_operators (the GPU paths of src/main.cpp, src/simulate_reaction.cpp
and src/compare_operators.cpp, written against include/lattice_hip.hpp):
argument checks on the CPU; on the GPU, a run from an oracle-made state must
end as the oracle's run does (bit-identical fluid; the species within exp()
ulps) and write the oracle's diagnostics rows and fields."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import cases
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lb-cube_amd", "lbm_sim")
REACT_EXE = os.path.join(ROOT, "lb-cube_amd", "lbm_reaction_sim")


def run(*args, cwd=None, exe=EXE):
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120, cwd=cwd)


def parse_vtk(path):
    """Header lines and {field: values} of a legacy ASCII STRUCTURED_POINTS file
    (scalars as (n,), vectors as (n, 3))."""
    lines = path.read_text().split("\n")
    n = int(next(ln for ln in lines if ln.startswith("POINT_DATA")).split()[1])
    fields, i = {}, 0
    while i < len(lines):
        parts = lines[i].split()
        if parts[:1] == ["SCALARS"]:
            fields[parts[1]] = np.array([float(t) for t in lines[i + 2:i + 2 + n]])
            i += 2 + n
        elif parts[:1] == ["VECTORS"]:
            fields[parts[1]] = np.array([[float(t) for t in ln.split()] for ln in lines[i + 1:i + 1 + n]])
            i += 1 + n
        else:
            i += 1
    return lines[:8], fields


@pytest.mark.parametrize("args,message", [(["-steps", "-1"], "-steps must be non-negative"),
                                          (["-out_freq", "0"], "-out_freq must be greater than zero"),
                                          (["-bogus"], "unknown argument: -bogus"),
                                          (["-init", "vortex"], "-init must be rest or taylor_green"),
                                          (["-collision", "xyz"], "-collision must be bgk, trt, mrt or rlbm")])
def test_argument_errors(tmp_path, args, message):
    assert os.path.exists(EXE), "build first: make -C lb-cube_amd"
    r = run(*args, cwd=tmp_path)
    assert r.returncode == 1 and message in r.stderr
    assert not any(tmp_path.iterdir()), "a rejected command line must not create files"


@pytest.mark.gpu
@pytest.mark.parametrize("collision", ["bgk", "trt", "mrt", "rlbm"])
def test_run_matches_oracle(gpu, tmp_path, collision):
    print(f"case driver {collision}", flush=True)
    shape = (8, 10, 12)  # (nz, ny, nx)
    f0 = cases.random_perturbed("D3Q19", shape, "f64", seed=31)
    init = tmp_path / "init.bin"
    f0.tofile(init)
    tau, steps, freq = 0.6, 20, 5
    r = run("-nx", 12, "-ny", 10, "-nz", 8, "-steps", steps, "-out_freq", freq, "-load", init, "-dump",
            tmp_path / "out.bin", "-csv", tmp_path / "diag.csv", "-collision", collision, "-tau", tau, "-gpu",
            cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(f0.shape)
    want = O.fluid_run("D3Q19", collision.upper(), "f64", f0, 1.0 / tau, steps)
    assert np.array_equal(got, want)
    rows = (tmp_path / "diag.csv").read_text().strip().split("\n")
    assert rows[0] == "step,total_mass,total_kinetic_energy,max_velocity_magnitude"
    state = f0
    for k, row in enumerate(rows[1:]):
        step, mass, ke, umax = row.split(",")
        assert int(step) == k * freq
        ref = O.fluid_diagnostics("D3Q19", "f64", state, wide=True)
        assert np.allclose([float(mass), float(ke), float(umax)], ref, rtol=1e-13, atol=0)
        # the snapshot of the same step: fields reconstructed on the device,
        # printed with 17 digits, so they read back bit-identical
        head, fields = parse_vtk(tmp_path / f"lbm_{k * freq:06d}.vtk")
        assert head[1] == f"LBM output step {k * freq}" and head[4] == "DIMENSIONS 12 10 8"
        rho, u = O.macro_fields("D3Q19", "f64", state)
        assert np.array_equal(fields["density"], rho.ravel())
        assert np.array_equal(fields["velocity"], u.reshape(3, -1).T)
        assert np.array_equal(fields["velocity_magnitude"], np.sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]).ravel())
        state = O.fluid_run("D3Q19", collision.upper(), "f64", state, 1.0 / tau, freq)
    assert "MLUPS" in r.stdout


@pytest.mark.parametrize("args,message", [(["-tau", "0.5"], "relaxation times must be greater than 0.5"),
                                          (["-tau_c", "0.4"], "relaxation times must be greater than 0.5"),
                                          (["-k_react", "-1"], "-k_react must be non-negative"),
                                          (["-nx", "0"], "-nx must be greater than zero"),
                                          (["-ny", "12", "-tiles", "5"], "-tiles must divide -ny"),
                                          (["-bogus"], "unknown argument: -bogus")])
def test_reaction_argument_errors(tmp_path, args, message):
    assert os.path.exists(REACT_EXE), "build first: make -C lb-cube_amd"
    r = run(*args, cwd=tmp_path, exe=REACT_EXE)
    assert r.returncode == 1 and message in r.stderr
    assert not any(tmp_path.iterdir()), "a rejected command line must not create files"


def check_reactive_fields(fields, state, k):
    """VTK fields of a reactive snapshot against the oracle's state: the fluid
    velocity bit-identical; the species within exp() ulps of the oracle."""
    _, u = O.macro_fields("D2Q9", "f64", state[0])
    assert np.array_equal(fields["velocity"][:, :2], u.reshape(2, -1).T)
    assert np.all(fields["velocity"][:, 2] == 0)
    ca = O.scalar_fields("D2Q5", "f64", state[1]).ravel()
    cb = O.scalar_fields("D2Q5", "f64", state[2]).ravel()
    assert np.allclose(fields["C_A"], ca, rtol=1e-12, atol=1e-15)
    assert np.allclose(fields["C_B"], cb, rtol=1e-12, atol=1e-15)
    assert np.allclose(fields["reaction_rate"], k * ca * cb, rtol=1e-12, atol=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [False, True])
def test_reaction_run_matches_oracle(gpu, tmp_path, fused):
    print(f"case reaction driver fused={fused}", flush=True)
    n, steps, freq = 24, 16, 4
    f0, u = cases.taylor_green_2d(n, 0.03)
    a0, b0 = cases.split_species(n, n, u)
    for x, suffix in ((f0, "f"), (a0, "a"), (b0, "b")):
        np.ascontiguousarray(x, dtype=np.float64).tofile(tmp_path / f"init.{suffix}")
    tau, tau_c, k = 0.6, 0.8, 0.05
    args = ["-nx", n, "-ny", n, "-steps", steps, "-out_freq", freq, "-tau", tau, "-tau_c", tau_c, "-k_react", k,
            "-load", tmp_path / "init", "-dump", tmp_path / "out", "-csv", tmp_path / "stats.csv"]
    r = run(*args, *(["-fused"] if fused else []), cwd=tmp_path, exe=REACT_EXE)
    assert r.returncode == 0, r.stderr
    got = [np.fromfile(tmp_path / f"out.{s}", dtype=np.float64).reshape(x.shape) for s, x in
           (("f", f0), ("a", a0), ("b", b0))]
    want = O.reactive_run("D2Q9", "BGK", "D2Q5", "f64", f0, a0, b0, 1 / tau, 1 / tau_c, k, steps)
    assert np.array_equal(got[0], want[0])  # the fluid never sees exp()
    for g, w in zip(got[1:], want[1:]):
        assert np.max(np.abs(g - w)) <= 1e-13
    rows = (tmp_path / "stats.csv").read_text().strip().split("\n")
    assert rows[0].startswith("step,mean_A,mean_B,var_A")
    state = (f0, a0, b0)
    for j, row in enumerate(rows[1:]):
        vals = row.split(",")
        assert int(vals[0]) == j * freq
        ref = O.reactive_stats("D2Q5", "f64", state[1], state[2], k, wide=True)
        assert np.allclose([float(v) for v in vals[1:]], ref, rtol=1e-10, atol=1e-14)
        head, fields = parse_vtk(tmp_path / f"reaction_{j * freq:06d}.vtk")
        assert head[1] == f"LB-Cube reactive mixing step {j * freq}" and head[4] == f"DIMENSIONS {n} {n} 1"
        check_reactive_fields(fields, state, k)
        state = O.reactive_run("D2Q9", "BGK", "D2Q5", "f64", *state, 1 / tau, 1 / tau_c, k, freq)


CMP_EXE = os.path.join(ROOT, "lb-cube_amd", "lbm_compare_operators")


@pytest.mark.parametrize("args,message", [(["-operators", "bgk,xyz"], "-operators takes bgk, trt, mrt, rlbm: xyz"),
                                          (["-stats_freq", "0"], "-stats_freq must be greater than zero"),
                                          (["-vtk_freq", "-1"], "-vtk_freq must be non-negative"),
                                          (["-tau", "0.5"], "relaxation times must be greater than 0.5"),
                                          (["-steps", "-2"], "-steps must be non-negative"),
                                          (["-bogus"], "unknown argument: -bogus")])
def test_compare_operators_argument_errors(tmp_path, args, message):
    assert os.path.exists(CMP_EXE), "build first: make -C lb-cube_amd"
    r = run(*args, cwd=tmp_path, exe=CMP_EXE)
    assert r.returncode == 1 and message in r.stderr
    assert not any(tmp_path.iterdir()), "a rejected command line must not create files"


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [False, True])
def test_compare_operators_matches_oracle(gpu, tmp_path, fused):
    """Every operator's statistics rows and VTK snapshots against the oracle's
    reactive run of the same state."""
    print(f"case compare_operators fused={fused}", flush=True)
    nx, ny, steps, sfreq, vfreq = 24, 20, 12, 4, 6
    f0 = cases.random_perturbed("D2Q9", (ny, nx), seed=41)  # nx != ny: x fastest everywhere
    a0, b0 = cases.split_species(nx, ny, cases.macro("D2Q9", f0)[1])
    for x, suffix in ((f0, "f"), (a0, "a"), (b0, "b")):
        np.ascontiguousarray(x, dtype=np.float64).tofile(tmp_path / f"init.{suffix}")
    tau, tau_c, k = 0.6, 0.8, 0.05
    r = run("-nx", nx, "-ny", ny, "-steps", steps, "-stats_freq", sfreq, "-vtk_freq", vfreq, "-tau", tau,
            "-tau_c", tau_c, "-k_react", k, "-load", tmp_path / "init", "-out", tmp_path,
            *(["-fused"] if fused else []), cwd=tmp_path, exe=CMP_EXE)
    assert r.returncode == 0, r.stderr
    for coll in ("BGK", "TRT", "MRT", "RLBM"):
        assert f"{coll} complete in" in r.stdout
        rows = (tmp_path / f"output_{coll}" / f"stats_{coll}.csv").read_text().strip().split("\n")
        assert rows[0].startswith("step,mean_A,mean_B,var_A")
        assert len(rows) == 1 + steps // sfreq + 1
        state = (f0, a0, b0)
        for j, row in enumerate(rows[1:]):
            s = j * sfreq
            vals = row.split(",")
            assert int(vals[0]) == s
            ref = O.reactive_stats("D2Q5", "f64", state[1], state[2], k, wide=True)
            assert np.allclose([float(v) for v in vals[1:]], ref, rtol=1e-10, atol=1e-14), (coll, s)
            if s % vfreq == 0:
                head, fields = parse_vtk(tmp_path / f"output_{coll}" / f"comparison_{coll}_{s:06d}.vtk")
                assert head[1] == f"LB-Cube {coll} comparison step {s}" and head[4] == f"DIMENSIONS {nx} {ny} 1"
                check_reactive_fields(fields, state, k)
            state = O.reactive_run("D2Q9", coll, "D2Q5", "f64", *state, 1 / tau, 1 / tau_c, k, sfreq)


TGV_EXE = os.path.join(ROOT, "lb-cube_amd", "lbm_turbulent_tgv")


@pytest.mark.parametrize("args,message", [(["-configs", "D2Q9_BGK"], "-configs lattice must be D3Q19 or D3Q27"),
                                          (["-configs", "D3Q19_XYZ"], "-configs operator must be BGK, TRT, MRT"),
                                          (["-configs", "D3Q19"], "-configs entries are LATTICE_OPERATOR"),
                                          (["-storage", "f16"], "-storage must be f64 or f32"),
                                          (["-check_freq", "0"], "-check_freq must be greater than zero"),
                                          (["-tau", "0.5"], "-tau must be greater than 0.5"),
                                          (["-threshold", "0"], "-threshold must be positive"),
                                          (["-bogus"], "unknown argument: -bogus")])
def test_turbulent_tgv_argument_errors(tmp_path, args, message):
    assert os.path.exists(TGV_EXE), "build first: make -C lb-cube_amd"
    r = run(*args, cwd=tmp_path, exe=TGV_EXE)
    assert r.returncode == 1 and message in r.stderr
    assert not any(tmp_path.iterdir()), "a rejected command line must not create files"


@pytest.mark.gpu
@pytest.mark.parametrize("config", ["D3Q19_MRT", "D3Q27_RLBM"])
def test_turbulent_tgv_matches_oracle(gpu, tmp_path, config):
    """Stability rows (the device max-speed guard) and snapshots against the
    oracle's run of the same state, FP64 bit for bit."""
    print(f"case turbulent_tgv {config}", flush=True)
    lattice, coll = config.split("_")
    nx, ny, nz, steps, check, vfreq, tau = 12, 10, 8, 40, 10, 20, 0.52
    f0 = cases.random_perturbed(lattice, (nz, ny, nx), seed=43, eps=2e-2)
    f0.tofile(tmp_path / "init.bin")
    r = run("-nx", nx, "-ny", ny, "-nz", nz, "-steps", steps, "-check_freq", check, "-vtk_freq", vfreq, "-tau", tau,
            "-configs", config, "-load", tmp_path / "init.bin", "-out", tmp_path, cwd=tmp_path, exe=TGV_EXE)
    assert r.returncode == 0, r.stderr
    assert "terminated early" not in r.stdout and f"[{config}] elapsed=" in r.stdout
    out = tmp_path / f"output_tgv3d_{config}"
    rows = (out / f"stability_{config}.csv").read_text().strip().split("\n")
    assert rows[0] == "step,max_velocity" and len(rows) == 2 + steps // check
    state = f0
    for j, row in enumerate(rows[1:]):
        s, v = row.split(",")
        assert int(s) == j * check
        assert float(v) == O.max_speed(lattice, "f64", state)
        if j * check % vfreq == 0:
            head, fields = parse_vtk(out / f"tgv3d_{config}_{j * check:06d}.vtk")
            assert head[1] == f"LB-Cube 3D turbulent TGV {config} step {j * check}"
            rho, u = O.macro_fields(lattice, "f64", state)
            assert np.array_equal(fields["density"], rho.ravel())
            assert np.array_equal(fields["velocity"], u.reshape(3, -1).T)
        state = O.fluid_run(lattice, coll, "f64", state, 1.0 / tau, check)


@pytest.mark.gpu
def test_turbulent_tgv_guard_stops_the_run(gpu, tmp_path):
    """The guard ends a configuration at the first check whose max speed
    exceeds the threshold (forced here by a threshold below u0) or is not
    finite (a NaN population in the loaded state), and the next one runs."""
    print("case turbulent_tgv guard", flush=True)
    r = run("-nx", 16, "-ny", 16, "-nz", 16, "-steps", 100, "-check_freq", 10, "-vtk_freq", 0, "-threshold", 0.01,
            "-configs", "D3Q19_BGK,D3Q27_TRT", "-out", tmp_path, cwd=tmp_path, exe=TGV_EXE)
    assert r.returncode == 0, r.stderr
    for config in ("D3Q19_BGK", "D3Q27_TRT"):
        assert f"[{config}] Solver Crashed due to instability! step=10" in r.stdout
        rows = (tmp_path / f"output_tgv3d_{config}" / f"stability_{config}.csv").read_text().strip().split("\n")
        assert [row.split(",")[0] for row in rows[1:]] == ["0", "10"]
        assert float(rows[1].split(",")[1]) > 0.01
    assert r.stdout.count("(terminated early)") == 2
    f = cases.random_perturbed("D3Q19", (8, 8, 8), seed=44)
    f[4, 3, 2, 1] = np.nan
    f.tofile(tmp_path / "nan.bin")
    r = run("-nx", 8, "-ny", 8, "-nz", 8, "-steps", 20, "-check_freq", 5, "-vtk_freq", 0, "-configs", "D3Q19_MRT",
            "-load", tmp_path / "nan.bin", "-out", tmp_path / "nan", cwd=tmp_path, exe=TGV_EXE)
    assert r.returncode == 0, r.stderr
    rows = (tmp_path / "nan" / "output_tgv3d_D3Q19_MRT" / "stability_D3Q19_MRT.csv").read_text().strip().split("\n")
    assert rows[1] == "0,inf" and rows[2] == "5,inf"
    assert "Solver Crashed due to instability! step=5, Umax=inf" in r.stdout


SHEAR_EXE = os.path.join(ROOT, "lb-cube_amd", "lbm_shear_layer")


@pytest.mark.parametrize("args,message", [(["-operators", "bgk,foo"], "-operators takes bgk, trt, mrt, rlbm: foo"),
                                          (["-check_freq", "0"], "-check_freq must be greater than zero"),
                                          (["-tau_c", "0.5"], "relaxation times must be greater than 0.5"),
                                          (["-threshold", "-1"], "-threshold must be positive"),
                                          (["-bogus"], "unknown argument: -bogus")])
def test_shear_layer_argument_errors(tmp_path, args, message):
    assert os.path.exists(SHEAR_EXE), "build first: make -C lb-cube_amd"
    r = run(*args, cwd=tmp_path, exe=SHEAR_EXE)
    assert r.returncode == 1 and message in r.stderr
    assert not any(tmp_path.iterdir()), "a rejected command line must not create files"


@pytest.mark.gpu
def test_shear_layer_matches_oracle(gpu, tmp_path):
    """Every operator's stability rows and snapshots against the oracle's
    fluid step followed by two passive scalar steps on the new fluid state
    (all bit-identical in FP64: no exp() on this path)."""
    print("case shear_layer", flush=True)
    nx, ny, steps, check, vfreq, tau, tau_c = 16, 12, 20, 5, 10, 0.55, 0.8
    f0 = cases.random_perturbed("D2Q9", (ny, nx), seed=45, eps=1e-2)
    a0, b0 = cases.split_species(nx, ny, cases.macro("D2Q9", f0)[1])
    for x, suffix in ((f0, "f"), (a0, "a"), (b0, "b")):
        np.ascontiguousarray(x, dtype=np.float64).tofile(tmp_path / f"init.{suffix}")
    r = run("-nx", nx, "-ny", ny, "-steps", steps, "-check_freq", check, "-vtk_freq", vfreq, "-tau", tau, "-tau_c",
            tau_c, "-load", tmp_path / "init", "-out", tmp_path, cwd=tmp_path, exe=SHEAR_EXE)
    assert r.returncode == 0, r.stderr
    for coll in ("BGK", "TRT", "MRT", "RLBM"):
        name = f"D2Q9_{coll}"
        assert f"[{name}] elapsed=" in r.stdout and "terminated early" not in r.stdout
        out = tmp_path / f"output_shear2d_{name}"
        rows = (out / f"stability_{name}.csv").read_text().strip().split("\n")
        assert rows[0] == "step,max_velocity" and len(rows) == 2 + steps // check
        f, a, b = f0, a0, b0
        for s in range(steps + 1):
            if s % check == 0:
                step, v = rows[1 + s // check].split(",")
                assert int(step) == s
                want = O.max_speed("D2Q9", "f64", f)
                assert float(v) == (float(f"{want:.6g}") if s == 0 else want)  # row 0: 6 digits
            if s % vfreq == 0:
                head, fields = parse_vtk(out / f"shear_{name}_{s:06d}.vtk")
                assert head[1] == f"LB-Cube turbulent shear layer {name} step {s}"
                _, u = O.macro_fields("D2Q9", "f64", f)
                assert np.array_equal(fields["velocity"][:, :2], u.reshape(2, -1).T)
                assert np.array_equal(fields["velocity_magnitude"], np.sqrt(u[0] * u[0] + u[1] * u[1]).ravel())
                assert np.array_equal(fields["C_A"], O.scalar_fields("D2Q5", "f64", a).ravel())
                assert np.array_equal(fields["C_B"], O.scalar_fields("D2Q5", "f64", b).ravel())
            f = O.fluid_step("D2Q9", coll, "f64", f, 1 / tau)
            a = O.scalar_step("D2Q9", "D2Q5", "f64", a, f, 1 / tau_c, 0.0)
            b = O.scalar_step("D2Q9", "D2Q5", "f64", b, f, 1 / tau_c, 0.0)


@pytest.mark.gpu
def test_shear_layer_guard_writes_last_snapshot(gpu, tmp_path):
    print("case shear_layer guard", flush=True)
    r = run("-nx", 32, "-ny", 32, "-steps", 40, "-check_freq", 10, "-vtk_freq", 1000, "-threshold", 0.01,
            "-operators", "mrt", "-out", tmp_path, cwd=tmp_path, exe=SHEAR_EXE)
    assert r.returncode == 0, r.stderr
    assert "[D2Q9_MRT] Solver Crashed due to instability! step=10" in r.stdout
    assert "(terminated early)" in r.stdout
    out = tmp_path / "output_shear2d_D2Q9_MRT"
    assert sorted(p.name for p in out.glob("*.vtk")) == ["shear_D2Q9_MRT_000000.vtk", "shear_D2Q9_MRT_000010.vtk"]


# ---------------------------------------------------------------------------
# The drivers' own initial conditions (no -load)
# ---------------------------------------------------------------------------
def reference_tgv_3d(nx, ny, nz, u0):
    """initialize_taylor_green_state_3d (src/main.cpp:371-411) restated: the
    trigonometric values come from the C library (math.sin / math.cos, as
    std::sin / std::cos), each product and sum in the reference's order;
    populations from the oracle's compute_equilibrium."""
    import math

    cs2 = 1.0 / 3.0
    k = [2.0 * math.pi / n for n in (nx, ny, nz)]
    ph = [[kk * float(i) for i in range(n)] for kk, n in zip(k, (nx, ny, nz))]
    sx, cx, c2x = (np.array([f(p) for p in ph[0]]) for f in (math.sin, math.cos, lambda p: math.cos(2.0 * p)))
    sy, cy, c2y = (np.array([f(p) for p in ph[1]]) for f in (math.sin, math.cos, lambda p: math.cos(2.0 * p)))
    sz, cz, c2z = (np.array([f(p) for p in ph[2]]) for f in (math.sin, math.cos, lambda p: math.cos(2.0 * p)))
    pert = (u0 * u0) / (16.0 * cs2) * (c2x[None, None, :] + c2y[None, :, None]) * (c2z[:, None, None] + 2.0)
    rho = 1.0 + pert
    ux = u0 * sx[None, None, :] * cy[None, :, None] * cz[:, None, None]
    uy = -u0 * cx[None, None, :] * sy[None, :, None] * cz[:, None, None]
    u = np.stack([ux, uy, np.zeros_like(ux)])
    return O.feq_field("D3Q19", "f64", rho, u)


def reference_tgv_2d_split(nx, ny, u0):
    """taylor_green_macro + initialize_fields (src/simulate_reaction.cpp:173-232)
    restated the same way: fluid and both species at equilibrium."""
    import math

    cs2 = 1.0 / 3.0
    px = [2.0 * math.pi / nx * float(x) for x in range(nx)]
    py = [2.0 * math.pi / ny * float(y) for y in range(ny)]
    cx, sx, c2x = (np.array([f(p) for p in px]) for f in (math.cos, math.sin, lambda p: math.cos(2.0 * p)))
    cy, sy, c2y = (np.array([f(p) for p in py]) for f in (math.cos, math.sin, lambda p: math.cos(2.0 * p)))
    rho = 1.0 - (u0 * u0) / (4.0 * cs2) * (c2x[None, :] + c2y[:, None])
    u = np.stack([-u0 * cx[None, :] * sy[:, None], u0 * sx[None, :] * cy[:, None]])
    ca = np.broadcast_to((np.arange(nx) < nx // 2).astype(np.float64), (ny, nx))
    return (O.feq_field("D2Q9", "f64", rho, u), O.scalar_feq_field("D2Q5", "f64", ca, u),
            O.scalar_feq_field("D2Q5", "f64", 1.0 - ca, u))


@pytest.mark.gpu
def test_lbm_sim_reference_default_run(gpu, tmp_path):
    """lbm_sim without -load at the reference's literal default: 64^3, tau = 0.8,
    u0 = 0.01, the density-perturbed Taylor-Green vortex, BGK, 1000 steps,
    diagnostics every 100 (src/main.cpp:149-195, :368-411, :542).  The initial
    populations, every diagnostics row and the final populations must be the
    oracle's, started from the restated initial state (populations bit for
    bit; FP64 sums to 1e-13)."""
    print("case lbm_sim reference default, 64^3 x 1000 steps", flush=True)
    want0 = reference_tgv_3d(64, 64, 64, 0.01)
    r = run("-steps", 0, "-dump", tmp_path / "init.bin", "-csv", tmp_path / "d0.csv", "-no_vtk", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    got0 = np.fromfile(tmp_path / "init.bin", dtype=np.float64).reshape(want0.shape)
    assert np.array_equal(got0, want0)
    r = run("-dump", tmp_path / "out.bin", "-csv", tmp_path / "diag.csv", "-no_vtk", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    rows = (tmp_path / "diag.csv").read_text().strip().split("\n")
    assert len(rows) == 1 + 11
    state = want0
    threads = O.max_threads()
    for k, row in enumerate(rows[1:]):
        step, mass, ke, umax = row.split(",")
        assert int(step) == 100 * k
        ref = O.fluid_diagnostics("D3Q19", "f64", state, wide=True)
        assert np.allclose([float(mass), float(ke), float(umax)], ref, rtol=1e-13, atol=0), k
        if k < 10:
            state = O.fluid_run("D3Q19", "BGK", "f64", state, 1.0 / 0.8, 100, threads=threads)
    got = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(want0.shape)
    assert np.array_equal(got, state)


@pytest.mark.gpu
def test_lbm_sim_rest_initial_state(gpu, tmp_path):
    """-init rest: every cell at feq(1, 0) = w_i (initialize_rest_state,
    src/main.cpp:295-312), then three steps as the oracle takes them (the
    density sums the rounded weights, so the state moves by ulps)."""
    common = ["-nx", 12, "-ny", 10, "-nz", 8, "-init", "rest", "-csv", tmp_path / "d.csv", "-no_vtk"]
    r = run(*common, "-steps", 0, "-dump", tmp_path / "init.bin", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    got0 = np.fromfile(tmp_path / "init.bin", dtype=np.float64).reshape(19, 8, 10, 12)
    want0 = O.feq_field("D3Q19", "f64", np.ones((8, 10, 12)), np.zeros((3, 8, 10, 12)))
    w = np.asarray(O.WEIGHTS["D3Q19"]).reshape(19, 1, 1, 1)
    assert np.array_equal(want0, np.broadcast_to(w, want0.shape)) and np.array_equal(got0, want0)
    r = run(*common, "-steps", 3, "-dump", tmp_path / "out.bin", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "out.bin", dtype=np.float64).reshape(19, 8, 10, 12)
    want = O.fluid_run("D3Q19", "BGK", "f64", want0, 1.0 / 0.8, 3)
    assert np.array_equal(got, want), float(np.max(np.abs(got - want)))


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [False, True])
def test_reaction_sim_own_initial_state(gpu, tmp_path, fused):
    """lbm_reaction_sim without -load: the 2D Taylor-Green fluid and the species
    split at x = nx/2 (src/simulate_reaction.cpp:173-232) at the reference's
    defaults (u0 = 0.04, tau = 0.6, tau_c = 0.8, k = 0.05) on a 64 x 48 grid.
    The initial populations bit for bit; after 200 steps the fluid bit for bit
    and the species within exp() ulps; every statistics row against the
    oracle."""
    print(f"case lbm_reaction_sim own initial state fused={fused}", flush=True)
    nx, ny, steps, freq = 64, 48, 200, 50
    want0 = reference_tgv_2d_split(nx, ny, 0.04)
    common = ["-nx", nx, "-ny", ny, "-no_vtk", *(["-fused"] if fused else [])]
    r = run(*common, "-steps", 0, "-dump", tmp_path / "init", "-csv", tmp_path / "s0.csv", cwd=tmp_path,
            exe=REACT_EXE)
    assert r.returncode == 0, r.stderr
    for suffix, w in zip("fab", want0):
        assert np.array_equal(np.fromfile(tmp_path / f"init.{suffix}", dtype=np.float64).reshape(w.shape), w), suffix
    r = run(*common, "-steps", steps, "-out_freq", freq, "-dump", tmp_path / "out", "-csv", tmp_path / "stats.csv",
            cwd=tmp_path, exe=REACT_EXE)
    assert r.returncode == 0, r.stderr
    rows = (tmp_path / "stats.csv").read_text().strip().split("\n")
    assert len(rows) == 1 + steps // freq + 1
    state = want0
    for j, row in enumerate(rows[1:]):
        vals = row.split(",")
        assert int(vals[0]) == j * freq
        ref = O.reactive_stats("D2Q5", "f64", state[1], state[2], 0.05, wide=True)
        assert np.allclose([float(v) for v in vals[1:]], ref, rtol=1e-10, atol=1e-14), j
        if j * freq < steps:
            state = O.reactive_run("D2Q9", "BGK", "D2Q5", "f64", *state, 1 / 0.6, 1 / 0.8, 0.05, freq)
    got = [np.fromfile(tmp_path / f"out.{s}", dtype=np.float64).reshape(w.shape) for s, w in zip("fab", want0)]
    assert np.array_equal(got[0], state[0])
    for g, w in zip(got[1:], state[1:]):
        assert np.max(np.abs(g - w)) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("collision", ["trt", "mrt"])
def test_lbm_sim_slab_order_equals_planes(gpu, tmp_path, collision):
    """lbm_sim -slabs (the device populations in the slab order, diagnostics,
    snapshots and the dump converted back): every diagnostics row, every VTK
    field and the final populations equal the run on the reference's planes
    byte for byte."""
    print(f"case lbm_sim -slabs {collision}", flush=True)
    outs = []
    for extra in ([], ["-slabs"]):
        d = tmp_path / ("slabs" if extra else "planes")
        d.mkdir()
        r = run("-nx", 12, "-ny", 10, "-nz", 8, "-steps", 20, "-out_freq", 5, "-collision", collision, "-dump",
                d / "out.bin", "-csv", d / "d.csv", "-vtk_dir", d, *extra, cwd=d)
        assert r.returncode == 0, r.stderr
        outs.append(d)
    planes, slabs = outs
    assert (planes / "out.bin").read_bytes() == (slabs / "out.bin").read_bytes()
    assert (planes / "d.csv").read_text() == (slabs / "d.csv").read_text()
    vtks = sorted(p.name for p in planes.glob("*.vtk"))
    assert len(vtks) == 5 and vtks == sorted(p.name for p in slabs.glob("*.vtk"))
    for name in vtks:
        assert (planes / name).read_text() == (slabs / name).read_text(), name


@pytest.mark.gpu
@pytest.mark.parametrize("tiles", [1, 4, 48])
def test_reaction_sim_tiles_equal_fused(gpu, tmp_path, tiles):
    """lbm_reaction_sim -tiles R (the three population sets in tiles of R
    rows: the reference's two launches as tiled launches, or with -fused the
    tiled fused step; statistics on the tiles): every statistics row, every
    VTK field and the final populations equal the -fused run on the
    reference's planes byte for byte."""
    outs = []
    for name, extra in (("fused", ["-fused"]), ("tiles", ["-tiles", tiles]),
                        ("tiles_fused", ["-tiles", tiles, "-fused"])):
        print(f"case lbm_reaction_sim {' '.join(map(str, extra))}", flush=True)
        d = tmp_path / name
        d.mkdir()
        r = run("-nx", 40, "-ny", 48, "-steps", 60, "-out_freq", 20, "-dump", d / "out", "-csv", d / "s.csv",
                "-vtk_dir", d, *extra, cwd=d, exe=REACT_EXE)
        assert r.returncode == 0, r.stderr
        outs.append(d)
    fused = outs[0]
    vtks = sorted(p.name for p in fused.glob("*.vtk"))
    assert len(vtks) == 4
    for tiled in outs[1:]:
        for suffix in "fab":
            assert (fused / f"out.{suffix}").read_bytes() == (tiled / f"out.{suffix}").read_bytes(), (tiled, suffix)
        assert (fused / "s.csv").read_text() == (tiled / "s.csv").read_text()
        assert vtks == sorted(p.name for p in tiled.glob("*.vtk"))
        for name in vtks:
            assert (fused / name).read_text() == (tiled / name).read_text(), (tiled, name)
