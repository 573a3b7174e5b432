"""GPU: batched ensembles (lbc_ensemble_*).  Every instance must advance bit
for bit like lbc_reactive_fused_step on that instance alone, for every tiling
candidate, with and without hipGraph replay."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import lbcube
from lbcube.ensemble import Ensemble, reactive_mixing_ensemble
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [(64, 48), (100, 37), (130, 64), (17, 9), (256, 31), (70, 70), (300, 5)]


def fill(ens, storage, seed, refs=None):
    g = cases.rng(seed)
    for i, (nx, ny) in enumerate(ens.sizes):
        u = g.uniform(-0.03, 0.03, size=(2, ny, nx))
        rho = 1.0 + g.uniform(-0.01, 0.01, size=(ny, nx))
        ca = g.uniform(0, 1, size=(ny, nx))
        dt = "f64" if storage in ("f16", "bf16") else storage
        f = O.feq_field("D2Q9", dt, rho, u)
        a = O.scalar_feq_field("D2Q5", dt, ca, u)
        b = O.scalar_feq_field("D2Q5", dt, 1.0 - ca, u)
        tf, ta, tb = (torch.from_numpy(x).to("cuda") for x in (f, a, b))
        if storage in ("f16", "bf16"):
            ra, rb = refs[i] if refs else (0.0, 0.0)
            tf = lbcube.encode("D2Q9", tf, storage)
            ta = lbcube.encode("D2Q5", ta, storage, ra)
            tb = lbcube.encode("D2Q5", tb, storage, rb)
        ens.fluid(i).copy_(tf)
        sa, sb = ens.species(i)
        sa.copy_(ta)
        sb.copy_(tb)


def standalone(ens, i, coll, omega, omega_c, k, steps, fast=False, refs=(0.0, 0.0)):
    f = ens.fluid(i).clone()
    a, b = (x.clone() for x in ens.species(i))
    ft, at, bt = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    for _ in range(steps):
        lbcube.reactive_fused_step("D2Q9", coll, "D2Q5", f, ft, a, at, b, bt, omega, omega_c, k, fast=fast,
                                   ref_a=refs[0], ref_b=refs[1])
        f, ft, a, at, b, bt = ft, f, at, a, bt, b
    return f, a, b


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
@pytest.mark.parametrize("coll", ["BGK", "TRT"])
def test_instances_match_standalone_runs(gpu, coll, storage, fast):
    print(f"case ensemble {coll} {storage} fast={fast}", flush=True)
    n = len(SIZES)
    omega = [1 / (0.55 + 0.05 * i) for i in range(n)]
    omega_c = [1 / (0.7 + 0.03 * i) for i in range(n)]
    k = [0.01 * (i + 1) for i in range(n)]
    ens = Ensemble(coll, storage, SIZES, omega, omega_c, k, fast=fast)
    fill(ens, storage, seed=21)
    want = [standalone(ens, i, coll, omega[i], omega_c[i], k[i], 6, fast=fast) for i in range(n)]
    start = [(ens.fluid(i).clone(),) + tuple(x.clone() for x in ens.species(i)) for i in range(n)]
    for threads in (128, 256, 512, 1024):
        for graph in (False, True):
            for i in range(n):  # restore the initial state
                ens.fluid(i).copy_(start[i][0])
                sa, sb = ens.species(i)
                sa.copy_(start[i][1])
                sb.copy_(start[i][2])
            ens.set_block(threads)
            ens.run(6, graph=graph)
            torch.cuda.synchronize()
            for i in range(n):
                sa, sb = ens.species(i)
                for got, exp in ((ens.fluid(i), want[i][0]), (sa, want[i][1]), (sb, want[i][2])):
                    assert torch.equal(got.contiguous().view(torch.uint8), exp.contiguous().view(torch.uint8)), (
                        threads, graph, i)
    ens.close()


def test_autotune_and_graph_replay(gpu):
    sizes = [(int(s), int(s)) for s in (64, 96, 200, 333, 512)]
    ens = Ensemble("TRT", "f16", sizes, [1.8] * 5, [1.8] * 5, [0.01] * 5, fast=True)
    fill(ens, "f16", seed=3)
    chosen, timings = ens.autotune(trial_steps=2)
    assert chosen in (128, 256, 512, 1024)
    assert len(timings) == 4 and min(timings.values()) > 0
    lv = ens.level
    snap = [x[lv].clone() for x in (ens.F, ens.A, ens.B)]
    ens.run(130, graph=True)  # two 64-step graph replays, then two direct steps
    torch.cuda.synchronize()
    with_graph = [x[ens.level].clone() for x in (ens.F, ens.A, ens.B)]
    for buf, s0 in zip((ens.F, ens.A, ens.B), snap):
        buf[lv].copy_(s0)
    ens.level = lv
    ens.run(130, graph=False)
    torch.cuda.synchronize()
    for buf, g in zip((ens.F, ens.A, ens.B), with_graph):
        assert torch.equal(buf[ens.level].view(torch.uint8), g.view(torch.uint8))
    ens.close()


def test_config5_generator():
    n, omega, omega_c, k = reactive_mixing_ensemble()
    assert len(n) == 512 and n.min() >= 64 and n.max() <= 1024
    tau = 1 / omega
    assert np.all(tau > 0.5) and np.all(k > 0)


@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_instances_with_species_references(gpu, storage, fast):
    """Per-instance reference concentrations of the species storage: each
    instance still advances bit for bit like the fused step on it alone with
    the same references (graph replay included)."""
    print(f"case ensemble references {storage} fast={fast}", flush=True)
    n = len(SIZES)
    omega = [1 / (0.55 + 0.05 * i) for i in range(n)]
    omega_c = [1 / (0.7 + 0.03 * i) for i in range(n)]
    k = [0.01 * (i + 1) for i in range(n)]
    refs = [(0.5, 0.5 - 0.05 * i) for i in range(n)]
    ens = Ensemble("TRT", storage, SIZES, omega, omega_c, k, fast=fast, ref_a=[r[0] for r in refs],
                   ref_b=[r[1] for r in refs])
    fill(ens, storage, seed=23, refs=refs)
    want = [standalone(ens, i, "TRT", omega[i], omega_c[i], k[i], 6, fast=fast, refs=refs[i]) for i in range(n)]
    ens.run(6, graph=True)
    torch.cuda.synchronize()
    for i in range(n):
        sa, sb = ens.species(i)
        for got, exp in ((ens.fluid(i), want[i][0]), (sa, want[i][1]), (sb, want[i][2])):
            assert torch.equal(got.contiguous().view(torch.uint8), exp.contiguous().view(torch.uint8)), i
    ens.close()

