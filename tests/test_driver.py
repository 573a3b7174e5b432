"""The C++ driver lb-cube_amd/lbm_sim (the GPU path of src/main.cpp written
against include/lattice_hip.hpp): argument checks on the CPU; on the GPU, a run
from an oracle-made state must end bit-identical to the oracle's run and write
the oracle's diagnostics rows."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import cases
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lb-cube_amd", "lbm_sim")


def run(*args, cwd=None):
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120, cwd=cwd)


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
