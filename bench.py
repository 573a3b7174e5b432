#!/usr/bin/env python3
"""Headline benchmark: D3Q19 fused stream-collide on a 256^3 periodic TGV.

One step = one lbc_fluid_step launch over the whole 256^3 domain (the
Re = 1600 Taylor-Green configuration with a seeded divergence-free
perturbation), populations resident in HBM.  The headline `value` is the TRT
collision with FP32 storage; the same loop is also timed with FP16 storage,
with BGK and MRT, and reported under "variants".  MLUPS = cells * steps / t
(LB-Cube's definition, src/main.cpp:620-625).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]
Only one GPU is supported: --gpus > 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lb-cube_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E spec peak
BYTES_PER_CELL = {("D2Q9", "f32"): 72, ("D2Q9", "f16"): 36,
                  ("D3Q19", "f32"): 152, ("D3Q19", "f16"): 76, ("D3Q19", "bf16"): 76, ("D3Q19", "f64"): 304,
                  ("D3Q27", "f32"): 216, ("D3Q27", "f16"): 108, ("D3Q27", "bf16"): 108, ("D3Q27", "f64"): 432}


def time_variant(lbcube, torch, lattice, coll, storage, u, omega, steps, warmup, block=None, fast=False,
                 state=None, mfma=False, layout="planes"):
    """MLUPS of `steps` fluid steps from the equilibrium of velocity field u
    (or from `state`, a population tensor the call takes over); mfma: the
    deviation-form MRT's transforms on the matrix cores; layout "slabs": the
    populations stored slab by slab, [Z][Q][Y][X] (lbc_fluid_step_strided),
    converted before the warm-up and back after the timed steps."""
    from lbcube import init as lbinit

    a = lbinit.equilibrium(lattice, None, u, storage) if state is None else state
    del state
    cells = a[0].numel()
    step = lbcube.fluid_step
    if layout == "slabs":
        a = lbcube.to_slabs(lattice, a)
        step = lbcube.fluid_step_slabs
    b = torch.empty_like(a)
    for _ in range(warmup):
        step(lattice, coll, a, b, omega, block=block, fast=fast, mfma=mfma)
        a, b = b, a
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        step(lattice, coll, a, b, omega, block=block, fast=fast, mfma=mfma)
        a, b = b, a
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    mlups = cells / (ms * 1e-3) / 1e6
    gbs = BYTES_PER_CELL[(lattice, storage)] * cells / (ms * 1e-3) / 1e9
    if layout == "slabs":
        del b
        a = lbcube.from_slabs(lattice, a)
        b = None
    diag = lbcube.fluid_diagnostics(lattice, a)
    del a, b
    torch.cuda.empty_cache()
    return {"ms_per_step": ms, "mlups": mlups, "achieved_gbs": gbs, "roofline_frac": gbs / HBM_PEAK_GBS,
            "bytes_per_cell": BYTES_PER_CELL[(lattice, storage)], "final_mass": diag[0], "final_ke": diag[1]}


def copy_ceiling(lbcube, torch, nbytes, reps=20):
    """Measured HBM ceiling: streaming copy of the same bytes one D3Q19 FP32 step moves."""
    src = torch.empty(nbytes // 2 // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    lbcube.stream_copy(src, dst)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        lbcube.stream_copy(src, dst)
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / reps
    del src, dst
    torch.cuda.empty_cache()
    return nbytes / (ms * 1e-3) / 1e9


def _events_ms(torch, fn, steps, warmup):
    """Average milliseconds of fn() over `steps` calls, timed with events on
    the current stream after `warmup` untimed calls."""
    for _ in range(warmup):
        fn()
    stream = torch.cuda.current_stream()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def pattern_ceiling(lbcube, torch, u, steps=20, warmup=3, lattice="D3Q19", storage="f32", shift=True,
                    layout="planes"):
    """GB/s of the pull-streaming step alone (lbc_pull_stream: the fluid
    kernels' loads, stores, addressing and residency, no collision) on the
    lattice's buffers (default: the headline's D3Q19 FP32): the ceiling of
    its access pattern.  shift=False: the same plane streams unshifted
    (lbc_plane_copy).  layout "slabs": the [Z][Q][Y][X] order
    (lbc_pull_stream_strided)."""
    from lbcube import init as lbinit

    bufs = [lbinit.equilibrium(lattice, None, u, storage)]
    cells = bufs[0][0].numel()
    if layout == "slabs":
        bufs[0] = lbcube.to_slabs(lattice, bufs[0])
    bufs.append(torch.empty_like(bufs[0]))

    def step():
        if layout == "slabs":
            ext = (bufs[0].shape[3], bufs[0].shape[2], bufs[0].shape[0])
            lbcube.pull_stream_strided(lattice, bufs[0], bufs[1], ext, *lbcube.slab_strides(lattice, ext),
                                       shift=shift)
        else:
            lbcube.pull_stream(lattice, bufs[0], bufs[1], shift=shift)
        bufs.reverse()

    ms = _events_ms(torch, step, steps, warmup)
    del bufs
    torch.cuda.empty_cache()
    return BYTES_PER_CELL[(lattice, storage)] * cells / (ms * 1e-3) / 1e9, ms


def pitched_variants(lbcube, torch, u, omega, steps, warmup, pads=(256, 1024, 4096, 16384)):
    """D3Q19 TRT FP32 and the pull-stream step on buffers whose population
    planes are padded by `pad` elements (lbc_fluid_step_pitched): the
    reference's contiguous layout puts all 38 streams on a common 64 MiB
    stride at 256^3.  Reported beside the headline, which keeps the
    contiguous layout."""
    from lbcube import init as lbinit

    a0 = lbinit.equilibrium("D3Q19", None, u, "f32")
    cells = a0[0].numel()
    ext = tuple(reversed(a0.shape[1:]))
    out = {}
    for pad in pads:
        pitch = cells + pad
        bufs = [torch.empty(19 * pitch, dtype=torch.float32, device="cuda") for _ in range(2)]
        bufs[0].view(19, pitch)[:, :cells].copy_(a0.view(19, cells))

        def step():
            lbcube.fluid_step_pitched("D3Q19", "TRT", bufs[0], bufs[1], ext, pitch, omega)
            bufs.reverse()

        def pull():
            lbcube.pull_stream("D3Q19", bufs[0], bufs[1], extents=ext, plane_pitch=pitch)
            bufs.reverse()

        ms = _events_ms(torch, step, steps, warmup)
        ms_pull = _events_ms(torch, pull, steps, warmup)
        out[f"pad{pad}"] = {"ms_per_step": ms, "mlups": cells / (ms * 1e-3) / 1e6,
                            "achieved_gbs": 152 * cells / (ms * 1e-3) / 1e9,
                            "pattern_gbs": 152 * cells / (ms_pull * 1e-3) / 1e9}
        print(f"# trt_f32 planes padded by {pad}: {out[f'pad{pad}']['mlups']:.0f} MLUPS, "
              f"pull-stream {out[f'pad{pad}']['pattern_gbs']:.0f} GB/s", file=sys.stderr, flush=True)
        del bufs
    del a0
    torch.cuda.empty_cache()
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_D3Q19_TRT_f32_256.md")


def measured_traffic():
    """HBM bytes per launch of the headline kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (counters cannot be read inside the timed
    run); None when the summary is absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def lattice_sweep(lbcube, torch, steps=10, warmup=3):
    """Taylor-Green Re = 1600 on D3Q19 and D3Q27 at 256^3 and 384^3, TRT and
    MRT, FP32 storage and FP16 / BF16 deviation storage (deviation-form TRT;
    MRT in both forms, and the deviation form with its transforms on the
    matrix cores), with the pull-stream pattern ceilings of FP32 and FP16;
    TRT FP32 and FP16 deviation form again in the slab order [Z][Q][Y][X]
    (keys *_slabs); RLBM in FP32."""
    from lbcube import init as lbinit

    out = {}
    for n in (256, 384):
        u = lbinit.taylor_green_3d(n, 0.05, device="cuda", perturbation_rms=0.01 * 0.05, seed=0x5EED1234)
        omega = 1.0 / lbinit.tgv_re1600_tau(n)
        for lattice in ("D3Q19", "D3Q27"):
            combos = [("TRT", "f32", False, False, "planes"), ("TRT", "f16", True, False, "planes"),
                      ("MRT", "f32", False, False, "planes"), ("MRT", "f32", True, False, "planes"),
                      ("MRT", "f16", True, False, "planes"), ("MRT", "f16", True, True, "planes"),
                      ("MRT", "bf16", True, False, "planes"), ("MRT", "bf16", True, True, "planes"),
                      ("TRT", "f32", False, False, "slabs"), ("TRT", "f16", True, False, "slabs"),
                      ("RLBM", "f32", False, False, "planes")]
            for coll, storage, fast, mfma, layout in combos:
                key = (f"{lattice}_{n}_{coll.lower()}_{storage}{'fast' if fast else ''}{'mfma' if mfma else ''}"
                       f"{'_slabs' if layout == 'slabs' else ''}")
                r = time_variant(lbcube, torch, lattice, coll, storage, u, omega, steps, warmup, fast=fast,
                                 mfma=mfma, layout=layout)
                out[key] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
                print(f"# {key}: {r['mlups']:.0f} MLUPS, {r['achieved_gbs']:.0f} GB/s", file=sys.stderr,
                      flush=True)
            for storage in ("f32", "f16"):
                key = f"{lattice}_{n}_pull_{storage}"
                gbs, ms = pattern_ceiling(lbcube, torch, u, steps=steps, warmup=warmup, lattice=lattice,
                                          storage=storage)
                out[key] = {"ms_per_step": round(ms, 4), "achieved_gbs": round(gbs, 1),
                            "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
                print(f"# {key}: pull-stream pattern {gbs:.0f} GB/s", file=sys.stderr, flush=True)
        del u
        torch.cuda.empty_cache()
    return out


def tgv_state_chunked(torch, lattice, n, storage, u0=0.05, chunk=32):
    """Equilibrium populations of the Re = 1600 TGV velocity on an n^3 box
    (no perturbation), built z-slab by z-slab so no full-size float64
    temporary exists: the populations themselves are the only large
    allocation."""
    from lbcube import Q
    from lbcube import init as lbinit

    dt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16, "f64": torch.float64}[storage]
    a = torch.empty((Q[lattice], n, n, n), dtype=dt, device="cuda")
    r = torch.arange(n, dtype=torch.float64, device="cuda") * (2.0 * math.pi / n)
    for z0 in range(0, n, chunk):
        z, y, x = torch.meshgrid(r[z0:z0 + chunk], r, r, indexing="ij")
        u = torch.stack([u0 * torch.sin(x) * torch.cos(y) * torch.cos(z),
                         -u0 * torch.cos(x) * torch.sin(y) * torch.cos(z), torch.zeros_like(x)])
        a[:, z0:z0 + chunk] = lbinit.equilibrium(lattice, None, u, storage)
        del z, y, x, u
    return a


def large_domain_rate(lbcube, torch, steps=10, warmup=2):
    """D3Q19 TRT TGV (Re = 1600) on the largest cubes one MI355X holds:
    1024^3 FP32 (2^30 cells, 4 GiB planes, 32-bit offsets), 1152^3 FP32
    (1.53e9 cells, 232 GB of populations, 6.1 GB planes: the 64-bit-address
    kernels) and 1440^3 FP16 deviation storage (2.99e9 cells, 227 GB)."""
    from lbcube import init as lbinit

    out = {}
    for n, storage, fast in ((1024, "f32", False), (1152, "f32", False), (1440, "f16", True)):
        key = f"D3Q19_{n}_trt_{storage}{'fast' if fast else ''}"
        bpc = BYTES_PER_CELL[("D3Q19", storage)]  # both buffers: 2 x 19 values per cell
        need = bpc * n ** 3 + (16 << 30)
        torch.cuda.empty_cache()
        free, _ = torch.cuda.mem_get_info()
        if free < need:
            out[key] = {"skipped": f"needs {need / 1e9:.0f} GB free, {free / 1e9:.0f} GB free"}
            continue
        state = tgv_state_chunked(torch, "D3Q19", n, storage)
        r = time_variant(lbcube, torch, "D3Q19", "TRT", storage, None, 1.0 / lbinit.tgv_re1600_tau(n), steps, warmup,
                         fast=fast, state=state)
        del state
        r["cells"] = n ** 3
        r["population_gb"] = round(bpc * n ** 3 / 1e9, 1)
        r["plane_bytes"] = n ** 3 * bpc // 38
        out[key] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        print(f"# {key}: {r['mlups']:.0f} MLUPS, {r['achieved_gbs']:.0f} GB/s", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    return out


def _time(torch, fn, reps=1):
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps


def _graph_ms(torch, step, steps, chunk=50):
    """ms per call of `step` (a launch sequence advancing ping-pong buffers by
    one step) replayed from a CUDA graph of `chunk` calls (chunk even, so the
    buffers return to their roles): no host launch overhead in the timing."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(chunk):
            step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(chunk):
            step()
    graph.replay()
    ms = _time(torch, graph.replay, reps=max(1, steps // chunk)) / chunk
    del graph
    return ms


def adjoint_ratio(torch, n=512, steps=4096, interval=64):
    """Differentiated 2D reactive run (D2Q9 BGK + two D2Q5, A+B->C), FP32 with
    the fast (deviation-form, cancellation-free reaction) forward:
    time of the forward loop vs time of the full gradient (forward with
    checkpoints, segment recompute, adjoint sweep), per step.  The adjoint
    time is the gradient time minus one forward run (recompute included)."""
    from lbcube import adjoint as ADJ

    du = ADJ.gaussian_random_field((n, n), 32.0, 0.01, seed=0x5EED1234)
    f0, a0, b0 = ADJ.initial_state(du, dtype=torch.float32)
    omega, omega_c, k = 1 / 0.6, 1 / 0.8, 0.05
    # forward: 64 steps captured in a CUDA graph (no per-step host overhead),
    # replayed steps / 64 times
    import lbcube
    st = [f0.clone(), torch.empty_like(f0), a0.clone(), torch.empty_like(a0), b0.clone(), torch.empty_like(b0)]

    def fwd_chunk():
        for t in range(64):
            c, x = (0, 1) if t % 2 == 0 else (1, 0)
            lbcube.reactive_fused_step("D2Q9", "BGK", "D2Q5", st[c], st[x], st[2 + c], st[2 + x], st[4 + c],
                                       st[4 + x], omega, omega_c, k, fast=True)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fwd_chunk()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        fwd_chunk()
    graph.replay()
    t_fwd = _time(torch, lambda: [graph.replay() for _ in range(steps // 64)])
    block = lbcube.default_block("D2Q9", n, n)
    del graph, st

    def gradient_time(iv):
        ws = torch.empty(ADJ.workspace_bytes(torch.float32, n, n, steps, iv, block), dtype=torch.uint8,
                         device="cuda")
        ADJ.reactive_gradient("BGK", f0, a0, b0, 64, omega, omega_c, k, iv, workspace=ws, fast=True)
        out = {}
        t = _time(torch, lambda: out.update(r=ADJ.reactive_gradient("BGK", f0, a0, b0, steps, omega, omega_c, k, iv,
                                                                    workspace=ws, fast=True)))
        gb = ws.numel() / 1e9
        del ws
        torch.cuda.empty_cache()
        return t, out["r"][0], gb

    t_grad, loss, ws_gb = gradient_time(interval)
    # every state kept (interval 1: no recompute), 4096 x 19.9 MB in HBM
    t_all, _, ws_all_gb = gradient_time(1)
    return {"workload": f"D2Q9 BGK + 2xD2Q5 A+B->C, {n}^2, {steps} steps, FP32 (fast forward), checkpoint every "
                        f"{interval}",
            "forward_ms_per_step": t_fwd / steps, "gradient_ms_per_step": t_grad / steps,
            "adjoint_over_forward": (t_grad - t_fwd) / t_fwd, "gradient_over_forward": t_grad / t_fwd,
            "workspace_gb": ws_gb, "loss": loss,
            "every_state_kept": {"workspace_gb": ws_all_gb, "gradient_ms_per_step": t_all / steps,
                                 "adjoint_over_forward": (t_all - t_fwd) / t_fwd}}


def adjoint_3d_ratio(torch, n=128, steps=256, interval=64, collision="TRT"):
    """Differentiated 3D reactive run (D3Q19 + two D3Q7, A+B->C; TRT unless
    `collision` says otherwise), FP32 with the fast forward: a sheared layer
    with species on either side of a smoothed plane; time of the forward steps
    (graph replay) against the full gradient (forward with checkpoints,
    segment recompute, adjoint sweep)."""
    import lbcube
    from lbcube import adjoint as ADJ
    from lbcube import init as lbinit

    r = torch.arange(n, dtype=torch.float64, device="cuda")
    z, y, x = torch.meshgrid(r, r, r, indexing="ij")
    u = torch.stack([0.04 * torch.tanh((y - n / 2) / 4.0), 0.004 * torch.sin(2 * math.pi * x / n),
                     0.004 * torch.cos(2 * math.pi * z / n)])
    ca = 0.5 * (1 + torch.tanh((x - n / 2) / 3.0))
    f0 = lbinit.equilibrium("D3Q19", None, u, "f32")
    a0 = lbinit.scalar_equilibrium("D3Q7", ca, u, "f32")
    b0 = lbinit.scalar_equilibrium("D3Q7", 1.0 - ca, u, "f32")
    del z, y, x, u, ca
    omega, omega_c, k = 1 / 0.6, 1 / 0.8, 0.05
    st = [f0.clone(), torch.empty_like(f0), a0.clone(), torch.empty_like(a0), b0.clone(), torch.empty_like(b0)]

    def step():
        lbcube.reactive_fused_step("D3Q19", collision, "D3Q7", st[0], st[1], st[2], st[3], st[4], st[5], omega,
                                   omega_c, k, fast=True)
        st[:] = [st[1], st[0], st[3], st[2], st[5], st[4]]

    t_fwd = _graph_ms(torch, step, steps, chunk=64) * steps
    del st
    block = lbcube.default_block("D3Q19", n, n, n)
    ws = torch.empty(ADJ.workspace_bytes(torch.float32, n, n, steps, interval, block, "D3Q19", n), dtype=torch.uint8,
                     device="cuda")
    ADJ.reactive_gradient(collision, f0, a0, b0, 64, omega, omega_c, k, interval, workspace=ws, fast=True,
                          lattice="D3Q19")
    out = {}
    t_grad = _time(torch, lambda: out.update(r=ADJ.reactive_gradient(collision, f0, a0, b0, steps, omega, omega_c,
                                                                     k, interval, workspace=ws, fast=True,
                                                                     lattice="D3Q19")))
    ws_gb = ws.numel() / 1e9
    del ws, f0, a0, b0
    torch.cuda.empty_cache()
    return {"workload": f"D3Q19 {collision} + 2xD3Q7 A+B->C, {n}^3, {steps} steps, FP32 (fast forward), checkpoint "
                        f"every {interval}",
            "forward_ms_per_step": t_fwd / steps, "gradient_ms_per_step": t_grad / steps,
            "adjoint_over_forward": (t_grad - t_fwd) / t_fwd, "gradient_over_forward": t_grad / t_fwd,
            "workspace_gb": ws_gb, "loss": out["r"][0]}


def ensemble_rate(torch, steps=1000):
    """512 independent 2D reactive instances (D2Q9 TRT + 2xD2Q5), edges
    log-uniform in 64..1024, FP16 storage with FP32 deviation-form arithmetic,
    autotuned block size, hipGraph replay."""
    from lbcube import init as lbinit
    from lbcube.ensemble import Ensemble, reactive_mixing_ensemble

    n, omega, omega_c, k = reactive_mixing_ensemble()
    half = [0.5] * len(n)  # species stored as g - w/2: the reference concentration of 0..1 fields
    ens = Ensemble("TRT", "f16", [(int(x), int(x)) for x in n], omega, omega_c, k, fast=True, ref_a=half,
                   ref_b=half)
    for i, ni in enumerate(n):
        rho, u = lbinit.taylor_green_2d(int(ni), int(ni), 0.04)
        ca, cb = lbinit.tanh_half_planes(int(ni), delta=max(2.0, ni / 32))
        ens.fluid(i).copy_(lbinit.equilibrium("D2Q9", rho, u, "f16"))
        a, b = ens.species(i)
        a.copy_(lbinit.scalar_equilibrium("D2Q5", ca, u, "f16", 0.5))
        b.copy_(lbinit.scalar_equilibrium("D2Q5", cb, u, "f16", 0.5))
    chosen, timings = ens.autotune(trial_steps=3)
    ens.run(2, graph=True)
    ms = _time(torch, lambda: ens.run(steps, graph=True))
    cells = ens.cells
    ens.close()
    gbs = 76 * cells * steps / (ms * 1e-3) / 1e9
    return {"workload": "512 instances D2Q9 TRT + 2xD2Q5 A+B->C, N log-uniform 64..1024, FP16 storage "
                        "(FP32 deviation-form arithmetic), "
                        f"{steps} steps, one launch per step, hipGraph replay",
            "cells": cells, "mlups": cells * steps / (ms * 1e-3) / 1e6, "achieved_gbs": gbs,
            "bytes_per_cell": 76, "threads_per_block": chosen,
            "autotune_ms_per_step": {str(t): round(v, 4) for t, v in timings.items()}}


def adr_3d_rate(lbcube, torch, n=256, steps=40):
    """The 3D advection-diffusion kernels on the 256^3 TGV (D3Q19 fluid, D3Q7
    scalars, FP32): the passive-scalar step (reads the fluid, pulls one
    scalar: (19 + 2 x 7) x 4 = 132 B/cell), the A + B -> C reaction step
    ((19 + 4 x 7) x 4 = 188 B/cell) and the fused TRT + reaction step in both
    arithmetic forms ((2 x 19 + 4 x 7) x 4 = 264 B/cell), on the reference's
    planes and in the slab order.  CUDA-graph replays of 10-step chunks."""
    from lbcube import init as lbinit

    u = lbinit.taylor_green_3d(n, 0.05, device="cuda")
    x = torch.arange(n, dtype=torch.float64, device="cuda") * (2.0 * math.pi / n)
    ca = (0.5 + 0.2 * torch.sin(x)).expand(n, n, n).contiguous()
    cb = (0.5 - 0.2 * torch.sin(x)).view(n, 1, 1).expand(n, n, n).contiguous()
    f = lbinit.equilibrium("D3Q19", None, u, "f32")
    a = lbinit.scalar_equilibrium("D3Q7", ca, u, "f32")
    b = lbinit.scalar_equilibrium("D3Q7", cb, u, "f32")
    del ca, cb, u
    ft, at, bt = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    state = [f, ft, a, at, b, bt]
    omega, omega_c, k = 1.0 / 0.6, 1.0 / 0.7, 0.005
    cells = n ** 3

    def scalar():
        lbcube.scalar_step("D3Q19", "D3Q7", state[2], state[3], state[0], omega_c, source=1e-6)
        state[2], state[3] = state[3], state[2]

    def reaction():
        lbcube.reaction_step("D3Q19", "D3Q7", state[2], state[3], state[4], state[5], state[0], omega_c, k)
        state[2], state[3], state[4], state[5] = state[3], state[2], state[5], state[4]

    def fused(fast):
        f, ft, a, at, b, bt = state
        lbcube.reactive_fused_step("D3Q19", "TRT", "D3Q7", f, ft, a, at, b, bt, omega, omega_c, k, fast=fast)
        state[:] = [ft, f, at, a, bt, b]

    # the fused step with the three population sets in the slab order
    # (lbc_reactive_fused_step_tiled with tiles of ny rows: each xy-slab holds
    # its Q planes back to back), converted once before the timing
    slabs = [lbcube.to_tiles(lat, x, n) for lat, x in (("D3Q19", f), ("D3Q7", a), ("D3Q7", b))]
    sstate = [slabs[0], torch.empty_like(slabs[0]), slabs[1], torch.empty_like(slabs[1]), slabs[2],
              torch.empty_like(slabs[2])]
    del slabs

    def fused_slabs(fast):
        f, ft, a, at, b, bt = sstate
        lbcube.reactive_fused_step_tiled("D3Q19", "TRT", "D3Q7", f, ft, a, at, b, bt, omega, omega_c, k, fast=fast)
        sstate[:] = [ft, f, at, a, bt, b]

    # the reference's separate scalar and reaction launches in the slab order
    def scalar_slabs():
        lbcube.scalar_step_tiled("D3Q19", "D3Q7", sstate[2], sstate[3], sstate[0], omega_c, source=1e-6)
        sstate[2], sstate[3] = sstate[3], sstate[2]

    def reaction_slabs():
        lbcube.reaction_step_tiled("D3Q19", "D3Q7", sstate[2], sstate[3], sstate[4], sstate[5], sstate[0], omega_c,
                                   k)
        sstate[2], sstate[3], sstate[4], sstate[5] = sstate[3], sstate[2], sstate[5], sstate[4]

    def pattern(st):
        lbcube.reactive_pull_stream("D3Q19", "D3Q7", *st, tiled=st is sstate)
        st[:] = [st[1], st[0], st[3], st[2], st[5], st[4]]

    out = {"workload": f"D3Q19 TGV {n}^3 (u0 = 0.05) + two D3Q7 species (C = 0.5 +- 0.2 sin), FP32"}
    for name, fn, bpc in (("fused_pattern", lambda: pattern(state), 264),
                          ("fused_pattern_slabs", lambda: pattern(sstate), 264),
                          ("scalar_step", scalar, 132), ("reaction_step", reaction, 188),
                          ("scalar_step_slabs", scalar_slabs, 132), ("reaction_step_slabs", reaction_slabs, 188),
                          ("fused_trt", lambda: fused(False), 264), ("fused_trt_fast", lambda: fused(True), 264),
                          ("fused_trt_slabs", lambda: fused_slabs(False), 264),
                          ("fused_trt_fast_slabs", lambda: fused_slabs(True), 264)):
        ms = _graph_ms(torch, fn, steps, chunk=10)
        out[name] = {"mlups": cells / (ms * 1e-3) / 1e6, "achieved_gbs": bpc * cells / (ms * 1e-3) / 1e9,
                     "bytes_per_cell": bpc, "ms_per_step": ms}
        print(f"# adr 3d {name}: {out[name]['mlups']:.0f} MLUPS, {out[name]['achieved_gbs']:.0f} GB/s",
              file=sys.stderr, flush=True)
    del state, f, ft, a, at, b, bt, sstate
    torch.cuda.empty_cache()
    return out


def tiles_3d_rate(lbcube, torch, n=512, steps=10, warmup=3, row_set=(512, 64, 16)):
    """D3Q19 TRT at 512^3, FP32 and FP16 deviation form, on the reference's
    planes and in 3D tiles of 512 (the slab order), 64 and 16 rows
    (lbc_fluid_step_tiled), converted once before the timing."""
    from lbcube import init as lbinit

    u = lbinit.taylor_green_3d(n, 0.05, device="cuda", perturbation_rms=0.01 * 0.05, seed=0x5EED1234)
    omega = 1.0 / lbinit.tgv_re1600_tau(n)
    out = {"workload": f"D3Q19 TRT periodic TGV {n}^3, Re=1600, seeded perturbation"}
    for storage, fast in (("f32", False), ("f16", True)):
        r = time_variant(lbcube, torch, "D3Q19", "TRT", storage, u, omega, steps, warmup, fast=fast)
        line = {"planes": r["mlups"]}
        for rows in row_set:
            a = lbcube.to_tiles("D3Q19", lbinit.equilibrium("D3Q19", None, u, storage), rows)
            b = torch.empty_like(a)
            bufs = [a, b]
            del a, b

            def step(bufs=bufs, fast=fast):
                lbcube.fluid_step_tiled("D3Q19", "TRT", bufs[0], bufs[1], omega, fast=fast)
                bufs.reverse()

            ms = _events_ms(torch, step, steps, warmup)
            line[f"rows{rows}"] = n ** 3 / (ms * 1e-3) / 1e6
            del bufs
            torch.cuda.empty_cache()
        key = f"trt_{storage}{'fast' if fast else ''}"
        out[key] = {k: round(v, 1) for k, v in line.items()}
        print(f"# tiles 3D {n}^3 {key}: " + ", ".join(f"{k} {v:.0f}" for k, v in line.items()) + " MLUPS",
              file=sys.stderr, flush=True)
    del u
    torch.cuda.empty_cache()
    return out


def shard_rate(lbcube, torch, u, omega, steps=50, warmup=5, gpus=8):
    """The steps at one GPU's share of a z-slab (2D: y-slab) decomposition
    over `gpus` GPUs: D3Q19 TRT at 256 x 256 x (256 / gpus) (the first slabs
    of the headline's velocity field), FP32 and FP16 deviation form, on the
    reference's planes and in the slab order; and the fused 2D reactive MRT
    step at 2048 x (2048 / gpus).  Set against each step, the time the halo
    exchange of that shard would take on one xGMI link (~153 GB/s): per face
    the populations entering across it (D3Q19: 5 per cell of the face; D2Q9
    + 2 D2Q5: 3 + 1 + 1 per cell of the edge row)."""
    from lbcube import init as lbinit

    out = {}
    n = u.shape[1]
    nz = n // gpus
    us = u[:, :nz].contiguous()
    link = 153e9
    for storage, fast in (("f32", False), ("f16", True)):
        for layout in ("planes", "slabs"):
            r = time_variant(lbcube, torch, "D3Q19", "TRT", storage, us, omega, steps, warmup, fast=fast,
                             layout=layout)
            es = 4 if storage == "f32" else 2
            halo_us = 5 * n * n * es / link * 1e6  # one face
            key = f"d3q19_{n}x{n}x{nz}_trt_{storage}{'fast' if fast else ''}_{layout}"
            out[key] = {"mlups": r["mlups"], "achieved_gbs": r["achieved_gbs"], "roofline_frac": r["roofline_frac"],
                        "step_us": r["ms_per_step"] * 1e3, "halo_face_us_one_link": halo_us}
            print(f"# shard {key}: {r['mlups']:.0f} MLUPS, {r['achieved_gbs']:.0f} GB/s, step "
                  f"{r['ms_per_step'] * 1e3:.1f} us, halo per face {halo_us:.1f} us", file=sys.stderr, flush=True)
    del us
    m = 2048
    ny = m // gpus
    uu = lbinit.double_shear_layer(m, u0=0.04, delta=80.0)[:, :ny].contiguous()
    ca, cb = (c[:ny].contiguous() for c in lbinit.tanh_half_planes(m, delta=80.0))
    k = 10.0 * 0.04 / 80.0
    for storage, fast, ref in (("f32", False, 0.0), ("f16", True, 0.5)):
        st = [lbinit.equilibrium("D2Q9", None, uu, storage)]
        st.append(torch.empty_like(st[0]))
        for c in (ca, cb):
            st.append(lbinit.scalar_equilibrium("D2Q5", c, uu, storage, ref))
            st.append(torch.empty_like(st[-1]))

        def fused(st=st, fast=fast, ref=ref):
            f, ft, a, at, b, bt = st
            lbcube.reactive_fused_step("D2Q9", "MRT", "D2Q5", f, ft, a, at, b, bt, 1 / 0.6, 1 / 0.6, k, fast=fast,
                                       ref_a=ref, ref_b=ref)
            st[:] = [ft, f, at, a, bt, b]

        for _ in range(4):
            fused()
        ms = _graph_ms(torch, fused, 100)
        es = st[0].element_size()
        halo_us = 5 * m * es / link * 1e6  # one edge row
        key = f"c3_{m}x{ny}_mrt_{storage}{'fast' if fast else ''}"
        out[key] = {"mlups": m * ny / (ms * 1e-3) / 1e6, "achieved_gbs": 38 * es * m * ny / (ms * 1e-3) / 1e9,
                    "step_us": ms * 1e3, "halo_face_us_one_link": halo_us}
        print(f"# shard {key}: {out[key]['mlups']:.0f} MLUPS, step {ms * 1e3:.1f} us, halo per face {halo_us:.2f} us",
              file=sys.stderr, flush=True)
        del st
    torch.cuda.empty_cache()
    out["note"] = (f"one GPU's share of a {gpus}-way slab decomposition; periodic in the slab direction here "
                   "(the halo exchange is estimated, not run: one GPU)")
    return out


def reactive_2d_rate(lbcube, torch, n=2048, steps=100):
    """2048^2 D2Q9 MRT double shear layer + two D2Q5 species (Da = 10), FP32
    and FP64: the fused single-pass step (152 / 304 B per cell), the reference's
    two launches per step (fluid step, then reaction step: 188 / 376 B per
    cell), and the whole loop with the reactive statistics every 50 steps; the
    fused step in the deviation form (FP32, on the same buffers), and both
    forms with FP16 storage (76 B).  Step rates are timed from CUDA-graph
    replays of 50-step chunks (a 2048^2 step is ~0.12 ms, close to the host's
    per-launch cost); the loop with statistics runs eagerly."""
    from lbcube import init as lbinit

    out = {}
    u = lbinit.double_shear_layer(n, u0=0.04, delta=80.0)
    ca, cb = lbinit.tanh_half_planes(n, delta=80.0)
    k = 10.0 * 0.04 / 80.0
    for storage, es in (("f32", 4), ("f64", 8)):
        f = lbinit.equilibrium("D2Q9", None, u, storage)
        a = lbinit.scalar_equilibrium("D2Q5", ca, u, storage)
        b = lbinit.scalar_equilibrium("D2Q5", cb, u, storage)
        ft, at, bt = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
        state = [f, ft, a, at, b, bt]

        def fused(state=state, fast=False):
            f, ft, a, at, b, bt = state
            lbcube.reactive_fused_step("D2Q9", "MRT", "D2Q5", f, ft, a, at, b, bt, 1 / 0.6, 1 / 0.6, k, fast=fast)
            state[:] = [ft, f, at, a, bt, b]

        def unfused(state=state):
            f, ft, a, at, b, bt = state
            lbcube.fluid_step("D2Q9", "MRT", f, ft, 1 / 0.6)
            lbcube.reaction_step("D2Q9", "D2Q5", a, at, b, bt, ft, 1 / 0.6, k)
            state[:] = [ft, f, at, a, bt, b]

        def loop(state=state):
            for s in range(1, steps + 1):
                fused(state)
                if s % 50 == 0:
                    lbcube.reactive_stats("D2Q5", state[2], state[4], k)

        for _ in range(4):
            fused()
        ms = _graph_ms(torch, fused, steps)
        ms_unfused = _graph_ms(torch, unfused, steps)
        ms_loop = _time(torch, loop) / steps  # eager launches and the statistics' synchronisations
        stats = lbcube.reactive_stats("D2Q5", state[2], state[4], k)
        ms_stats = _time(torch, lambda: lbcube.reactive_stats("D2Q5", state[2], state[4], k), reps=5)
        cells = n * n
        out[storage] = {"mlups": cells / (ms * 1e-3) / 1e6, "achieved_gbs": 38 * es * cells / (ms * 1e-3) / 1e9,
                        "bytes_per_cell": 38 * es,
                        "unfused_mlups": cells / (ms_unfused * 1e-3) / 1e6,
                        "unfused_achieved_gbs": 47 * es * cells / (ms_unfused * 1e-3) / 1e9,
                        "unfused_bytes_per_cell": 47 * es,
                        "loop_with_stats_every_50_mlups": cells / (ms_loop * 1e-3) / 1e6,
                        "stats_ms": ms_stats, "segregation_intensity": stats["segregation_intensity"]}
        if storage == "f32":  # the deviation form (LBC_FLAG_FAST) on the same buffers
            ms_fast = _graph_ms(torch, lambda: fused(fast=True), steps)
            out["f32_fast"] = {"mlups": cells / (ms_fast * 1e-3) / 1e6,
                               "achieved_gbs": 38 * es * cells / (ms_fast * 1e-3) / 1e9, "bytes_per_cell": 38 * es}
        del f, ft, a, at, b, bt, state
        torch.cuda.empty_cache()
    # FP16 storage of the deviations (species about their reference 1/2): the
    # fused step in both arithmetic forms
    f = lbinit.equilibrium("D2Q9", None, u, "f16")
    a = lbinit.scalar_equilibrium("D2Q5", ca, u, "f16", 0.5)
    b = lbinit.scalar_equilibrium("D2Q5", cb, u, "f16", 0.5)
    state = [f, torch.empty_like(f), a, torch.empty_like(a), b, torch.empty_like(b)]

    def fused16(fast):
        f, ft, a, at, b, bt = state
        lbcube.reactive_fused_step("D2Q9", "MRT", "D2Q5", f, ft, a, at, b, bt, 1 / 0.6, 1 / 0.6, k, fast=fast,
                                   ref_a=0.5, ref_b=0.5)
        state[:] = [ft, f, at, a, bt, b]

    for fast in (False, True):
        for _ in range(4):
            fused16(fast)
        ms = _graph_ms(torch, lambda: fused16(fast), steps)
        out["f16_fast" if fast else "f16"] = {"mlups": n * n / (ms * 1e-3) / 1e6,
                                             "achieved_gbs": 38 * 2 * n * n / (ms * 1e-3) / 1e9, "bytes_per_cell": 76}
    del f, a, b, state
    torch.cuda.empty_cache()
    out["workload"] = f"D2Q9 MRT + 2xD2Q5 A+B->C, {n}^2 double shear layer, Da=10 (k = Da u0 / delta = 0.005)"
    # the D2Q9 plane streams of the same domain, with and without the pull
    # shifts, in both element sizes: what the shifted 2- and 4-byte loads cost
    for storage in ("f32", "f16"):
        for shift in (True, False):
            gbs, ms = pattern_ceiling(lbcube, torch, u, steps=50, warmup=5, lattice="D2Q9", storage=storage,
                                      shift=shift)
            key = f"d2q9_{'pull' if shift else 'plane_copy'}_{storage}"
            out[key] = {"achieved_gbs": round(gbs, 1), "ms_per_step": round(ms, 4)}
            print(f"# reactive 2D {key}: {gbs:.0f} GB/s", file=sys.stderr, flush=True)
    out["d2q9_plane_copy_by_row_length"] = plane_copy_by_row_length(lbcube, torch)
    out.update(reactive_2d_large(lbcube, torch))
    out["tiles"] = reactive_2d_tiles(lbcube, torch, u, ca, cb, k, n, steps)
    return out


def reactive_2d_tiles(lbcube, torch, u, ca, cb, k, n, steps, row_set=(2, 4, 8, 32)):
    """The same fused 2048^2 step with the three population sets stored in
    tiles of R rows (lbc_reactive_fused_step_tiled: each tile holds the Q
    planes of its rows back to back), FP32 and FP16 in both arithmetic forms
    and FP64,
    graph-replayed like the lines above; and the FP32 loop with the reactive
    statistics every 50 steps in tiles of 4 rows (lbc_reactive_stats_tiled).
    The conversion from the reference's planes (lbc_relayout_tiled) happens
    once, before the timing."""
    from lbcube import init as lbinit

    out = {}
    for storage, fast, ref in (("f32", False, 0.0), ("f32", True, 0.0), ("f16", False, 0.5), ("f16", True, 0.5),
                               ("f64", False, 0.0)):
        planes = [lbinit.equilibrium("D2Q9", None, u, storage), lbinit.scalar_equilibrium("D2Q5", ca, u, storage, ref),
                  lbinit.scalar_equilibrium("D2Q5", cb, u, storage, ref)]
        es = planes[0].element_size()
        for rows in row_set:
            f, a, b = (lbcube.to_tiles(lat, x, rows) for lat, x in zip(("D2Q9", "D2Q5", "D2Q5"), planes))
            state = [f, torch.empty_like(f), a, torch.empty_like(a), b, torch.empty_like(b)]

            def fused(state=state, fast=fast, ref=ref):
                f, ft, a, at, b, bt = state
                lbcube.reactive_fused_step_tiled("D2Q9", "MRT", "D2Q5", f, ft, a, at, b, bt, 1 / 0.6, 1 / 0.6, k,
                                                 fast=fast, ref_a=ref, ref_b=ref)
                state[:] = [ft, f, at, a, bt, b]

            for _ in range(4):
                fused()
            ms = _graph_ms(torch, fused, steps)
            key = f"{storage}{'_fast' if fast else ''}_rows{rows}"
            out[key] = {"mlups": n * n / (ms * 1e-3) / 1e6, "achieved_gbs": 38 * es * n * n / (ms * 1e-3) / 1e9,
                        "bytes_per_cell": 38 * es}
            if storage == "f32" and not fast and rows == 4:
                def loop(state=state):
                    for s in range(1, steps + 1):
                        fused(state)
                        if s % 50 == 0:
                            lbcube.reactive_stats_tiled("D2Q5", state[2], state[4], k)

                ms_loop = _time(torch, loop) / steps
                out[key]["loop_with_stats_every_50_mlups"] = n * n / (ms_loop * 1e-3) / 1e6
            del f, a, b, state
        del planes
        torch.cuda.empty_cache()
    # the reference's two launches per step (fluid step, then reaction step:
    # 188 B/cell in FP32) in tiles of 4 rows
    planes = [lbinit.equilibrium("D2Q9", None, u, "f32"), lbinit.scalar_equilibrium("D2Q5", ca, u, "f32"),
              lbinit.scalar_equilibrium("D2Q5", cb, u, "f32")]
    tiles = [lbcube.to_tiles(lat, x, 4) for lat, x in zip(("D2Q9", "D2Q5", "D2Q5"), planes)]
    del planes
    tstate = [tiles[0], torch.empty_like(tiles[0]), tiles[1], torch.empty_like(tiles[1]), tiles[2],
              torch.empty_like(tiles[2])]
    del tiles

    def two_launches(st=tstate):
        f, ft, a, at, b, bt = st
        lbcube.fluid_step_tiled("D2Q9", "MRT", f, ft, 1 / 0.6)
        lbcube.reaction_step_tiled("D2Q9", "D2Q5", a, at, b, bt, ft, 1 / 0.6, k)
        st[:] = [ft, f, at, a, bt, b]

    ms = _graph_ms(torch, two_launches, steps)
    out["two_launches_f32_rows4"] = {"mlups": n * n / (ms * 1e-3) / 1e6,
                                     "achieved_gbs": 47 * 4 * n * n / (ms * 1e-3) / 1e9, "bytes_per_cell": 188}
    print(f"# reactive 2D two launches in tiles of 4 rows (FP32): {out['two_launches_f32_rows4']['mlups']:.0f} MLUPS",
          file=sys.stderr, flush=True)
    del tstate
    torch.cuda.empty_cache()
    # the fused step's pattern ceiling (lbc_reactive_pull_stream: its loads,
    # stores, addressing and block order without arithmetic) on the planes and
    # in tiles of 4 and 8 rows
    for storage in ("f32", "f16"):
        planes = [lbinit.equilibrium("D2Q9", None, u, storage), lbinit.scalar_equilibrium("D2Q5", ca, u, storage),
                  lbinit.scalar_equilibrium("D2Q5", cb, u, storage)]
        es = planes[0].element_size()
        for rows in (0, 4, 8):
            cur = ([lbcube.to_tiles(lat, x, rows) for lat, x in zip(("D2Q9", "D2Q5", "D2Q5"), planes)] if rows
                   else [x.clone() for x in planes])
            bufs = [cur[0], torch.empty_like(cur[0]), cur[1], torch.empty_like(cur[1]), cur[2], torch.empty_like(cur[2])]

            def pattern(bufs=bufs, rows=rows):
                lbcube.reactive_pull_stream("D2Q9", "D2Q5", *bufs, tiled=bool(rows))
                bufs[:] = [bufs[1], bufs[0], bufs[3], bufs[2], bufs[5], bufs[4]]

            ms = _graph_ms(torch, pattern, steps)
            key = f"pattern_{storage}_{'planes' if not rows else f'rows{rows}'}"
            out[key] = {"achieved_gbs": 38 * es * n * n / (ms * 1e-3) / 1e9, "ms_per_step": ms}
            del cur, bufs
        del planes
        torch.cuda.empty_cache()
    print("# reactive 2D pattern ceilings (GB/s): " + ", ".join(f"{k[8:]} {v['achieved_gbs']:.0f}"
                                                                for k, v in out.items() if k.startswith("pattern_")),
          file=sys.stderr, flush=True)
    best = {}
    for key, v in out.items():
        if key.startswith("pattern_"):
            continue
        form = key.rsplit("_rows", 1)[0]
        if form not in best or v["mlups"] > out[best[form]]["mlups"]:
            best[form] = key
    print("# reactive 2D in tiles: " + ", ".join(f"{k} {out[v]['mlups']:.0f} MLUPS ({v})" for k, v in best.items()),
          file=sys.stderr, flush=True)
    return out


def plane_copy_by_row_length(lbcube, torch, cells=1 << 22, reps=50):
    """GB/s of the unshifted D2Q9 plane copy (lbc_plane_copy) at a fixed cell
    count with rows of 256 to 8,192 cells, FP32 and FP16."""
    out = {}
    for storage, dt in (("f32", torch.float32), ("f16", torch.float16)):
        for nx in (256, 1024, 2048, 8192):
            bufs = [torch.zeros((9, cells // nx, nx), dtype=dt, device="cuda")]
            bufs.append(torch.empty_like(bufs[0]))

            def step():
                lbcube.pull_stream("D2Q9", bufs[0], bufs[1], shift=False)
                bufs.reverse()

            ms = _events_ms(torch, step, reps, 5)
            out[f"{storage}_nx{nx}"] = round(2 * 9 * bufs[0].element_size() * cells / (ms * 1e-3) / 1e9, 1)
            del bufs
    print(f"# reactive 2D plane copy by row length: {out}", file=sys.stderr, flush=True)
    return out


def reactive_2d_large(lbcube, torch, n=8192, steps=20):
    """The same fused step on an 8192^2 double shear layer (16x the cells of
    the 2048^2 line, ~0.9 ms per FP16 step against ~0.06): the rate of the
    kernel when the fixed start and drain of a launch no longer count; on the
    reference's planes and in tiles of 4 and 8 rows."""
    from lbcube import init as lbinit

    out = {}
    delta = 80.0 * n / 2048
    u = lbinit.double_shear_layer(n, u0=0.04, delta=delta)
    ca, cb = lbinit.tanh_half_planes(n, delta=delta)
    k = 10.0 * 0.04 / delta
    for storage, fast, ref in (("f32", False, 0.0), ("f16", False, 0.5), ("f16", True, 0.5)):
        state = [lbinit.equilibrium("D2Q9", None, u, storage)]
        state.append(torch.empty_like(state[0]))
        for c in (ca, cb):
            state.append(lbinit.scalar_equilibrium("D2Q5", c, u, storage, ref))
            state.append(torch.empty_like(state[-1]))

        def fused(state=state, fast=fast, ref=ref):
            f, ft, a, at, b, bt = state
            lbcube.reactive_fused_step("D2Q9", "MRT", "D2Q5", f, ft, a, at, b, bt, 1 / 0.6, 1 / 0.6, k, fast=fast,
                                       ref_a=ref, ref_b=ref)
            state[:] = [ft, f, at, a, bt, b]

        for _ in range(2):
            fused()
        ms = _graph_ms(torch, fused, steps)
        bpc = 38 * (4 if storage == "f32" else 2)
        key = f"{storage}{'_fast' if fast else ''}_{n}"
        out[key] = {"mlups": n * n / (ms * 1e-3) / 1e6, "achieved_gbs": bpc * n * n / (ms * 1e-3) / 1e9,
                    "bytes_per_cell": bpc}
        print(f"# reactive 2D {key}: {out[key]['mlups']:.0f} MLUPS, {out[key]['achieved_gbs']:.0f} GB/s",
              file=sys.stderr, flush=True)
        planes = state[0::2]
        del state
        for rows in (4, 8):
            tiled = [lbcube.to_tiles(lat, x, rows) for lat, x in zip(("D2Q9", "D2Q5", "D2Q5"), planes)]
            tstate = [tiled[0], torch.empty_like(tiled[0]), tiled[1], torch.empty_like(tiled[1]), tiled[2],
                      torch.empty_like(tiled[2])]
            del tiled

            def fused_tiles(tstate=tstate, fast=fast, ref=ref):
                f, ft, a, at, b, bt = tstate
                lbcube.reactive_fused_step_tiled("D2Q9", "MRT", "D2Q5", f, ft, a, at, b, bt, 1 / 0.6, 1 / 0.6, k,
                                                 fast=fast, ref_a=ref, ref_b=ref)
                tstate[:] = [ft, f, at, a, bt, b]

            for _ in range(2):
                fused_tiles()
            ms = _graph_ms(torch, fused_tiles, steps)
            tkey = f"{key}_rows{rows}"
            out[tkey] = {"mlups": n * n / (ms * 1e-3) / 1e6, "achieved_gbs": bpc * n * n / (ms * 1e-3) / 1e9,
                         "bytes_per_cell": bpc}
            print(f"# reactive 2D {tkey}: {out[tkey]['mlups']:.0f} MLUPS, {out[tkey]['achieved_gbs']:.0f} GB/s",
                  file=sys.stderr, flush=True)
            del tstate
            torch.cuda.empty_cache()
        del planes
        torch.cuda.empty_cache()
    del u, ca, cb
    torch.cuda.empty_cache()
    return out


def cpu_baseline(u_host, omega, n):
    """The oracle (CPU restatement of step_cpu) timed on this host: one step of
    the same 256^3 D3Q19 TRT FP32 workload on all host threads, and one on a
    single thread."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import oracle as O

    # The library is built beforehand (__graft_entry__.build()); building it here
    # would start a compiler from a process that holds the GPU.
    if not os.path.exists(O.LIB_PATH):
        raise RuntimeError(f"{O.LIB_PATH} is not built: run __graft_entry__.build() before the benchmark")
    rho = np.ones(u_host.shape[1:])
    f = O.feq_field("D3Q19", "f32", rho, u_host)
    threads = O.max_threads()
    t0 = time.perf_counter()
    O.fluid_step("D3Q19", "TRT", "f32", f, omega, threads=threads)
    t_all = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.fluid_step("D3Q19", "TRT", "f32", f, omega, threads=1)
    t_one = time.perf_counter() - t0
    cells = n ** 3
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"value": cells / t_all / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"1 step of the D3Q19 TRT FP32 {n}^3 TGV, oracle, {threads} OpenMP threads",
            "single_thread_value": cells / t_one / 1e6, "host_cpus": os.cpu_count(), "cpu_model": model}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variants", default="trt_f32,trt_f16fast,trt_f32fast,trt_f16,bgk_f32,bgk_f16fast,mrt_f32,"
                                          "mrt_f16,mrt_f32fast,mrt_f16fast,trt_bf16fast,trt_f32_slabs,"
                                          "trt_f32fast_slabs,trt_f16fast_slabs,mrt_f16fast_slabs,rlbm_f32,rlbm_f16",
                    help="coll_storage[fast][_slabs]: _slabs stores the populations slab by slab, [Z][Q][Y][X]")
    ap.add_argument("--block", default=None, help="block shape bx,by,bz (default: the library's)")
    ap.add_argument("--no-extras", action="store_true", help="skip the adjoint, ensemble and 2D reactive lines")
    ap.add_argument("--extras", default="pitched,adjoint,reactive_2d,adr_3d,ensemble,lattices,large,tiles_3d,shard",
                    help="which configuration lines to run (comma-separated subset)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the lattice Boltzmann path here is single-device "
              "(multi-GPU slab decomposition is designed, not built). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import lbcube
    from lbcube import init as lbinit

    n = args.n
    tau = lbinit.tgv_re1600_tau(n)
    omega = 1.0 / tau
    u = lbinit.taylor_green_3d(n, 0.05, device="cuda", perturbation_rms=0.01 * 0.05, seed=0x5EED1234)

    block = tuple(int(v) for v in args.block.split(",")) if args.block else None
    variants = {}
    for name in args.variants.split(","):
        coll, storage, *layout = name.split("_")
        fast = storage.endswith("fast")
        storage = storage[:-4] if fast else storage
        variants[name] = time_variant(lbcube, torch, "D3Q19", coll.upper(), storage, u, omega, args.steps,
                                      args.warmup, block=block, fast=fast, layout=layout[0] if layout else "planes")
        print(f"# {name}: {variants[name]['mlups']:.0f} MLUPS, {variants[name]['achieved_gbs']:.0f} GB/s",
              file=sys.stderr, flush=True)
    head = variants["trt_f32"]
    copy_gbs = copy_ceiling(lbcube, torch, 152 * n ** 3)
    pattern_gbs, pattern_ms = pattern_ceiling(lbcube, torch, u)
    print(f"# pull-stream ceiling: {pattern_gbs:.0f} GB/s ({pattern_ms:.4f} ms/step)", file=sys.stderr, flush=True)
    planes_gbs, planes_ms = pattern_ceiling(lbcube, torch, u, shift=False)
    print(f"# unshifted 38-plane copy: {planes_gbs:.0f} GB/s ({planes_ms:.4f} ms/step)", file=sys.stderr, flush=True)
    slab_pattern_gbs, _ = pattern_ceiling(lbcube, torch, u, layout="slabs")
    slab_planes_gbs, _ = pattern_ceiling(lbcube, torch, u, shift=False, layout="slabs")
    print(f"# slab order [Z][Q][Y][X]: pull-stream {slab_pattern_gbs:.0f} GB/s, unshifted copy {slab_planes_gbs:.0f} "
          "GB/s", file=sys.stderr, flush=True)
    chosen = set() if args.no_extras else set(args.extras.split(","))
    pitched = pitched_variants(lbcube, torch, u, omega, args.steps, args.warmup) if "pitched" in chosen else None
    extras = {}
    if chosen - {"pitched"}:
        del u
        torch.cuda.empty_cache()
        if "adjoint" in chosen:
            extras["adjoint"] = adjoint_ratio(torch)
            extras["adjoint_3d"] = adjoint_3d_ratio(torch)
            extras["adjoint_3d_rlbm"] = adjoint_3d_ratio(torch, collision="RLBM")
            for name in ("adjoint_3d", "adjoint_3d_rlbm"):
                a3 = extras[name]
                print(f"# {name}: forward {a3['forward_ms_per_step']:.3f} ms/step, gradient "
                      f"{a3['gradient_ms_per_step']:.3f} ms/step, adjoint / forward "
                      f"{a3['adjoint_over_forward']:.2f}", file=sys.stderr, flush=True)
            print(f"# adjoint/forward: {extras['adjoint']['adjoint_over_forward']:.2f}", file=sys.stderr, flush=True)
        if "reactive_2d" in chosen:
            extras["reactive_2d"] = reactive_2d_rate(lbcube, torch)
            r2 = extras["reactive_2d"]
            print(f"# reactive 2048^2 MRT f32: {r2['f32']['mlups']:.0f} MLUPS, f32 fast {r2['f32_fast']['mlups']:.0f}, "
                  f"f16 fast {r2['f16_fast']['mlups']:.0f}", file=sys.stderr, flush=True)
        if "adr_3d" in chosen:
            extras["adr_3d"] = adr_3d_rate(lbcube, torch)
        if "ensemble" in chosen:
            extras["ensemble"] = ensemble_rate(torch)
            print(f"# ensemble: {extras['ensemble']['mlups']:.0f} MLUPS", file=sys.stderr, flush=True)
        if "lattices" in chosen:
            extras["lattices"] = lattice_sweep(lbcube, torch)
        if "large" in chosen:
            extras["large"] = large_domain_rate(lbcube, torch)
        if "tiles_3d" in chosen:
            extras["tiles_3d"] = tiles_3d_rate(lbcube, torch)
        if "shard" in chosen:
            u = lbinit.taylor_green_3d(n, 0.05, device="cuda", perturbation_rms=0.01 * 0.05, seed=0x5EED1234)
            extras["shard"] = shard_rate(lbcube, torch, u, omega)
            del u
        u = lbinit.taylor_green_3d(n, 0.05, device="cuda", perturbation_rms=0.01 * 0.05, seed=0x5EED1234)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(u.cpu().numpy(), omega, n)
    out = {
        "metric": "D3Q19 MLUPS & %HBM roofline @256^3, fp32 vs fp16-store; adjoint/forward time",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"D3Q19 TRT periodic TGV {n}^3, Re=1600, u0=0.05, tau={tau:.6f}, FP32 storage, "
                               "seeded divergence-free perturbation rms 0.01*u0",
                   "cells": n ** 3, "lattice": "D3Q19", "collision": "TRT", "storage": "f32"},
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(head["roofline_frac"], 4),
                     "traffic": measured_traffic() if (n == 256 and block is None) else None,
                     "algorithmic_bytes": 152 * n ** 3,
                     "kernel": "fluid_step_kernel<D3Q19, TRT, F32>", "bytes_per_cell": 152,
                     "copy_ceiling_gbs": round(copy_gbs, 1), "frac_of_copy": round(head["achieved_gbs"] / copy_gbs, 4),
                     "pattern_ceiling_gbs": round(pattern_gbs, 1),
                     "frac_of_pattern": round(head["achieved_gbs"] / pattern_gbs, 4),
                     "pattern_kernel": "pull_stream_kernel<D3Q19, F32> (lbc_pull_stream, same buffers)",
                     "plane_streams_ceiling_gbs": round(planes_gbs, 1),
                     "plane_streams_kernel": "pull_stream_kernel<D3Q19, F32, unshifted> (lbc_plane_copy, same buffers)"},
        "cpu_baseline": cpu,
        "variants": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                     for k, v in variants.items()},
    }
    if "trt_f32_slabs" in variants:
        sl = variants["trt_f32_slabs"]
        out["slab_order"] = {
            "layout": "[Z][Q][Y][X]: the 19 planes of each xy-slab side by side (lbc_fluid_step_strided)",
            "kernel": "fluid_step_kernel<D3Q19, TRT, F32>", "mlups": round(sl["mlups"], 1),
            "achieved_gbs": round(sl["achieved_gbs"], 1), "frac": round(sl["roofline_frac"], 4),
            "pattern_ceiling_gbs": round(slab_pattern_gbs, 1), "plane_streams_ceiling_gbs": round(slab_planes_gbs, 1),
            "over_reference_layout": round(sl["mlups"] / head["mlups"], 4)}
    if pitched:
        out["pitched_planes"] = {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in pitched.items()}
    if "trt_f16fast" in variants:
        out["fp16_over_fp32"] = round(variants["trt_f16fast"]["mlups"] / head["mlups"], 3)
    if extras:
        if "adjoint" in extras:
            out["adjoint_over_forward"] = round(extras["adjoint"]["adjoint_over_forward"], 3)
        out["configs"] = extras
    print(json.dumps(out))


if __name__ == "__main__":
    main()
