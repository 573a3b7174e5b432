"""The C++ drivers lb-cube_amd/lbm_sim and lbm_reaction_sim (the GPU paths of
src/main.cpp and src/simulate_reaction.cpp, written against
include/lattice_hip.hpp): argument checks on the CPU; on the GPU, a run from an
oracle-made state must end as the oracle's run does (bit-identical fluid; the
species within exp() ulps) and write the oracle's diagnostics rows."""
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


@pytest.mark.parametrize("args,message", [(["-steps", "-1"], "-steps must be non-negative"),
                                          (["-out_freq", "0"], "-out_freq must be greater than zero"),
                                          (["-bogus"], "unknown argument: -bogus"),
                                          (["-init", "vortex"], "-init must be rest or taylor_green"),
                                          (["-collision", "xyz"], "-collision must be bgk, trt, mrt or rlbm")])
def test_argument_errors(args, message):
    assert os.path.exists(EXE), "build first: make -C lb-cube_amd"
    r = run(*args)
    assert r.returncode == 1 and message in r.stderr


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
        state = O.fluid_run("D3Q19", collision.upper(), "f64", state, 1.0 / tau, freq)
    assert "MLUPS" in r.stdout


@pytest.mark.parametrize("args,message", [(["-tau", "0.5"], "relaxation times must be greater than 0.5"),
                                          (["-tau_c", "0.4"], "relaxation times must be greater than 0.5"),
                                          (["-k_react", "-1"], "-k_react must be non-negative"),
                                          (["-nx", "0"], "-nx must be greater than zero"),
                                          (["-bogus"], "unknown argument: -bogus")])
def test_reaction_argument_errors(args, message):
    assert os.path.exists(REACT_EXE), "build first: make -C lb-cube_amd"
    r = run(*args, exe=REACT_EXE)
    assert r.returncode == 1 and message in r.stderr


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
        state = O.reactive_run("D2Q9", "BGK", "D2Q5", "f64", *state, 1 / tau, 1 / tau_c, k, freq)
