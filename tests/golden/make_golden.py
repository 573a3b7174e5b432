"""Writes tests/golden/golden_fp64.npz: small FP64 outputs of the oracle.

Inputs are regenerated from fixed seeds by tests/cases.py (their SHA-256 is
stored next to the outputs so a changed generator is caught).  Run from the
repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

import cases  # noqa: E402
from oracle import oracle as O  # noqa: E402

SHAPES = {"D2Q9": (9, 11), "D3Q19": (4, 5, 6), "D3Q27": (4, 5, 6)}
OMEGA = 1.0 / 0.58
STEPS = 10


def fluid_cases():
    for lat, shape in SHAPES.items():
        for coll in ("BGK", "TRT", "MRT", "RLBM"):
            yield lat, coll, shape


def fluid_input(lat, shape):
    return cases.random_perturbed(lat, shape, "f64", seed=101)


def reactive_inputs():
    n = 16
    f, u = cases.taylor_green_2d(n, 0.03)
    a, b = cases.split_species(n, n, u)
    return f, a, b


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    out = {}
    for lat, coll, shape in fluid_cases():
        f0 = fluid_input(lat, shape)
        out[f"in_sha/{lat}/{coll}"] = np.array(sha(f0))
        out[f"fluid/{lat}/{coll}"] = O.fluid_run(lat, coll, "f64", f0, OMEGA, STEPS)
    f, a, b = reactive_inputs()
    out["in_sha/reactive"] = np.array(sha(np.concatenate([f.ravel(), a.ravel(), b.ravel()])))
    rf, ra, rb = O.reactive_run("D2Q9", "BGK", "D2Q5", "f64", f, a, b, 1 / 0.8, 1 / 0.8, 0.05, STEPS)
    out["reactive/f"], out["reactive/a"], out["reactive/b"] = rf, ra, rb
    out["reactive/stats"] = O.reactive_stats("D2Q5", "f64", ra, rb, 0.05, wide=True)
    np.savez_compressed(os.path.join(HERE, "golden_fp64.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
