"""GPU parity: the HIP kernels (through the C ABI) against the oracle.

The kernels evaluate the reference's per-cell operation sequence with
-ffp-contract=off, and the oracle does the same on the CPU, so fluid and
passive-scalar steps must agree BIT FOR BIT in every storage format.  The
reaction step calls exp(), whose device and host implementations may differ
in the last ulp; it is held to a relative tolerance instead (stated per test).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import lbcube
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FLUIDS = ["D2Q9", "D3Q19", "D3Q27"]
COLLS = ["BGK", "TRT", "MRT", "RLBM"]
STORES = ["f64", "f32", "f16", "bf16"]
SHAPE = {"D2Q9": (11, 13), "D3Q19": (5, 7, 9), "D3Q27": (5, 7, 9)}  # odd extents: every wrap path
TORCH_DT = {"f64": torch.float64, "f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}


def to_dev(a, storage, device):
    a = np.ascontiguousarray(a)
    if storage in ("f16", "bf16"):
        return torch.from_numpy(a.view(np.int16)).view(TORCH_DT[storage]).to(device)
    return torch.from_numpy(a).to(device)


def to_host(t, storage):
    if storage in ("f16", "bf16"):
        return t.cpu().view(torch.int16).numpy().view(np.uint16)
    return t.cpu().numpy()


def gpu_run(lattice, coll, storage, f0, omega, steps, device, block=None):
    a = to_dev(f0, storage, device)
    b = torch.empty_like(a)
    for _ in range(steps):
        lbcube.fluid_step(lattice, coll, a, b, omega, block=block)
        a, b = b, a
    torch.cuda.synchronize()
    return to_host(a, storage)


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("coll", COLLS)
@pytest.mark.parametrize("lattice", FLUIDS)
def test_fluid_step_bit_exact(gpu, lattice, coll, storage):
    print(f"case fluid {lattice} {coll} {storage}", flush=True)
    f0 = cases.random_perturbed(lattice, SHAPE[lattice], storage, seed=11)
    omega = 1.0 / 0.62
    got = gpu_run(lattice, coll, storage, f0, omega, 6, gpu)
    want = O.fluid_run(lattice, coll, storage, f0, omega, 6)
    assert got.dtype == want.dtype
    assert np.array_equal(got, want), f"max |diff| {np.max(np.abs(got.astype(np.float64) - want))}"


@pytest.mark.parametrize("block", [(16, 8, 1), (32, 32, 1), (64, 1, 1), (7, 3, 2), (1, 1, 1)])
def test_block_shape_does_not_change_results(gpu, block):
    print(f"case block {block}", flush=True)
    f0 = cases.random_perturbed("D3Q19", (6, 10, 70), "f32", seed=3)
    want = O.fluid_run("D3Q19", "TRT", "f32", f0, 1.7, 3)
    got = gpu_run("D3Q19", "TRT", "f32", f0, 1.7, 3, gpu, block=block)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("lattice", FLUIDS)
def test_fluid_kernels_blocks_of_1024(gpu, lattice, storage):
    """The fluid step's 1024-work-item instantiations, every collision and
    arithmetic form (and the MFMA form where it applies), on the reference's
    planes and in tiles: blocks of 1024 x 1 x 1 and 32 x 32 x 1 / 16 x 16 x 4,
    several per launch, give the default launch's results bit for bit."""
    print(f"case fluid blocks of 1024 {lattice} {storage}", flush=True)
    two_d = lattice == "D2Q9"
    shape = (70, 90) if two_d else (12, 14, 18)
    f = to_dev(cases.random_perturbed(lattice, shape, storage, seed=6), storage, gpu)
    blocks = [(1024, 1, 1), (32, 32, 1) if two_d else (16, 16, 4)]
    rows_set = (7, 70) if two_d else (7, 14)

    def run(block):
        out = []
        for coll in COLLS:
            for fast in ((False, True) if coll != "RLBM" else (False,)):
                n = torch.empty_like(f)
                lbcube.fluid_step(lattice, coll, f, n, 1.3, block=block, fast=fast)
                out.append(n)
        if not two_d and storage in ("f16", "bf16"):
            n = torch.empty_like(f)
            lbcube.fluid_step(lattice, "MRT", f, n, 1.3, block=block, fast=True, mfma=True)
            out.append(n)
        for rows in rows_set:
            t = lbcube.to_tiles(lattice, f, rows)
            for coll in ("TRT", "MRT"):
                n = torch.empty_like(t)
                lbcube.fluid_step_tiled(lattice, coll, t, n, 1.3, block=block, fast=True)
                out.append(lbcube.from_tiles(lattice, n))
        torch.cuda.synchronize()
        return out

    want = run(None)
    for block in blocks:
        for x, y in zip(run(block), want):
            assert torch.equal(x.view(torch.uint8), y.view(torch.uint8)), block


@pytest.mark.parametrize("lattice,shape", [("D3Q19", (3, 5, 263)), ("D2Q9", (70001, 3)), ("D3Q27", (2, 3, 1000)),
                                           ("D2Q9", (2, 4099))])
def test_linear_cell_mapping_odd_extents(gpu, lattice, shape):
    """Cells are numbered linearly across row ends: extents that are not
    multiples of the block, rows shorter than a wave, ny above 65535."""
    print(f"case linear {lattice} {shape}", flush=True)
    f0 = cases.random_perturbed(lattice, shape, "f32", seed=17)
    want = O.fluid_run(lattice, "TRT", "f32", f0, 1.6, 2)
    for block in (None, (16, 8, 1), (1024, 1, 1)):
        assert np.array_equal(gpu_run(lattice, "TRT", "f32", f0, 1.6, 2, gpu, block=block), want), block


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("lattice", FLUIDS)
def test_streaming_permutation_kat(gpu, lattice, storage):
    """BGK with omega = 0 is the pure pull permutation next[i][x] = cur[i][x - c_i]."""
    print(f"case kat {lattice} {storage}", flush=True)
    shape = SHAPE[lattice]
    q = O.Q_OF[lattice]
    n = int(np.prod(shape))
    ids = (1.0 + np.arange(q * n, dtype=np.float64)).reshape((q,) + shape)
    if storage == "f32":
        ids = ids.astype(np.float32)
    got = gpu_run(lattice, "BGK", storage, ids, 0.0, 1, gpu)
    vel = lbcube.VELOCITIES[lattice]
    want = np.empty_like(ids)
    for i, c in enumerate(vel):
        axes = tuple(range(len(shape) - 1, -1, -1))  # x is the last array axis
        want[i] = np.roll(ids[i], shift=tuple(c), axis=axes)
    assert np.array_equal(got, want)


def test_config1_tgv_64cubed_1000_steps(gpu):
    """D3Q19 BGK, 64^3 periodic TGV, FP64, tau = 0.6, u0 = 0.05, 1000 steps."""
    print("case config1 D3Q19 BGK f64 64^3 x 1000", flush=True)
    f0 = cases.taylor_green_3d(64, 0.05)
    omega = 1.0 / 0.6
    got = gpu_run("D3Q19", "BGK", "f64", f0, omega, 1000, gpu)
    want = O.fluid_run("D3Q19", "BGK", "f64", f0, omega, 1000, threads=O.max_threads())
    rel = np.max(np.abs(got - want) / np.abs(want))
    assert rel <= 1e-12, rel  # the stated tolerance; the kernels are in fact bit-exact here
    assert np.array_equal(got, want)


@pytest.mark.parametrize("ref", [0.0, 0.75])
@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("pair", [("D2Q9", "D2Q5"), ("D3Q19", "D3Q7"), ("D3Q27", "D3Q7")])
def test_scalar_step_bit_exact(gpu, pair, storage, ref):
    """ref: the scalar field's reference concentration (f16/bf16 store g - w ref)."""
    fl, sl = pair
    print(f"case scalar {fl} {sl} {storage} ref {ref}", flush=True)
    shape = SHAPE[fl]
    f = cases.random_perturbed(fl, shape, storage, seed=5)
    g = cases.random_perturbed(sl, shape, storage, seed=9, eps=0.05, ref=ref)
    want = O.scalar_step(fl, sl, storage, g, f, 1.25, 3e-4, ref=ref)
    gd, fd = to_dev(g, storage, gpu), to_dev(f, storage, gpu)
    out = torch.empty_like(gd)
    lbcube.scalar_step(fl, sl, gd, out, fd, 1.25, 3e-4, ref=ref)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(out, storage), want)


def _species(sl, shape, storage, seed, refs=(0.0, 0.0)):
    g = cases.rng(seed)
    d = O.D_OF[sl]
    ca = g.uniform(0.0, 1.0, size=shape)
    cb = 1.0 - ca + g.uniform(-0.05, 0.05, size=shape)
    cb[..., 0] = ca[..., 0]  # exercise the |A - B| <= tol branch
    u = g.uniform(-0.02, 0.02, size=(d,) + tuple(shape))
    fa = O.scalar_feq_field(sl, storage if storage in ("f32", "f64") else "f64", ca, u)
    fb = O.scalar_feq_field(sl, storage if storage in ("f32", "f64") else "f64", cb, u)
    if storage in ("f16", "bf16"):
        fa, fb = cases.encode(sl, storage, fa, refs[0]), cases.encode(sl, storage, fb, refs[1])
    return fa, fb


# Reaction tolerances: exp() of the device library may differ from the host's
# by an ulp; relative error of the populations after the steps below.
REACT_TOL = {"f64": 1e-13, "f32": 2e-6, "f16": 2e-3, "bf16": 1.6e-2}


@pytest.mark.parametrize("refs", [(0.0, 0.0), (0.5, 0.3)])
@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("pair", [("D2Q9", "D2Q5"), ("D3Q19", "D3Q7"), ("D3Q27", "D3Q7")])
def test_reaction_step_parity(gpu, pair, storage, refs):
    fl, sl = pair
    print(f"case reaction {fl} {sl} {storage} refs {refs}", flush=True)
    shape = SHAPE[fl]
    f = cases.random_perturbed(fl, shape, storage, seed=21)
    a, b = _species(sl, shape, storage, seed=4, refs=refs)
    k = 0.05
    wa, wb = O.reaction_step(fl, sl, storage, a, b, f, 1.1, k, ref_a=refs[0], ref_b=refs[1])
    ad, bd, fd = to_dev(a, storage, gpu), to_dev(b, storage, gpu), to_dev(f, storage, gpu)
    na, nb = torch.empty_like(ad), torch.empty_like(bd)
    lbcube.reaction_step(fl, sl, ad, na, bd, nb, fd, 1.1, k, ref_a=refs[0], ref_b=refs[1])
    torch.cuda.synchronize()
    ga, gb = to_host(na, storage), to_host(nb, storage)
    for got, want, r in ((ga, wa, refs[0]), (gb, wb, refs[1])):
        gd_, wd_ = O.decode(sl, storage, got, r), O.decode(sl, storage, want, r)
        err = np.max(np.abs(gd_ - wd_) / np.maximum(np.abs(wd_), 1e-3))
        assert err <= REACT_TOL[storage], err
        # nearly everything is bit-identical; the exceptions are exp() ulps
        assert np.mean(got == want) > 0.5


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("pair", [("D2Q9", "D2Q5"), ("D3Q19", "D3Q7"), ("D3Q27", "D3Q7")])
def test_reactive_kernels_blocks_of_1024(gpu, pair, storage):
    """The scalar, reaction and fused steps' 1024-work-item instantiations
    (blocks of 1024 x 1 x 1 and 32 x 32 x 1 / 16 x 16 x 4, several blocks per
    launch) give the default launch's results bit for bit, every collision and
    both arithmetic forms."""
    fl, sl = pair
    print(f"case reactive blocks of 1024 {fl} {storage}", flush=True)
    shape = (70, 90) if fl == "D2Q9" else (12, 14, 18)
    refs = (0.5, 0.25) if storage in ("f16", "bf16") else (0.0, 0.0)
    f = to_dev(cases.random_perturbed(fl, shape, storage, seed=2), storage, gpu)
    a0, b0 = _species(sl, shape, storage, seed=8, refs=refs)
    a, b = to_dev(a0, storage, gpu), to_dev(b0, storage, gpu)
    blocks = [(1024, 1, 1), (32, 32, 1) if fl == "D2Q9" else (16, 16, 4)]

    def run(block):
        out = []
        s = torch.empty_like(a)
        lbcube.scalar_step(fl, sl, a, s, f, 1.25, 3e-4, block=block, ref=refs[0])
        out.append(s)
        na, nb = torch.empty_like(a), torch.empty_like(b)
        lbcube.reaction_step(fl, sl, a, na, b, nb, f, 1.1, 0.05, block=block, ref_a=refs[0], ref_b=refs[1])
        out += [na, nb]
        for coll in COLLS:
            for fast in ((False, True) if coll != "RLBM" else (False,)):
                nf, na, nb = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
                lbcube.reactive_fused_step(fl, coll, sl, f, nf, a, na, b, nb, 1.3, 1.05, 0.02, block=block,
                                           fast=fast, ref_a=refs[0], ref_b=refs[1])
                out += [nf, na, nb]
        torch.cuda.synchronize()
        return out

    want = run(None)
    for block in blocks:
        for x, y in zip(run(block), want):
            assert torch.equal(x.view(torch.uint8), y.view(torch.uint8)), block


@pytest.mark.parametrize("refs", [(0.0, 0.0), (0.5, 0.25)])
@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("coll", COLLS)
def test_fused_reactive_equals_two_launches(gpu, coll, storage, refs):
    print(f"case fused D2Q9 {coll} {storage} refs {refs}", flush=True)
    shape = SHAPE["D2Q9"]
    f = to_dev(cases.random_perturbed("D2Q9", shape, storage, seed=2), storage, gpu)
    a0, b0 = _species("D2Q5", shape, storage, seed=8, refs=refs)
    a, b = to_dev(a0, storage, gpu), to_dev(b0, storage, gpu)
    omega, omega_c, k = 1.3, 1.05, 0.02
    ra, rb = refs
    # two launches, reference order
    f1, a1, b1 = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    lbcube.fluid_step("D2Q9", coll, f, f1, omega)
    lbcube.reaction_step("D2Q9", "D2Q5", a, a1, b, b1, f1, omega_c, k, ref_a=ra, ref_b=rb)
    # one pass
    f2, a2, b2 = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    lbcube.reactive_fused_step("D2Q9", coll, "D2Q5", f, f2, a, a2, b, b2, omega, omega_c, k, ref_a=ra, ref_b=rb)
    torch.cuda.synchronize()
    for x, y in ((f1, f2), (a1, a2), (b1, b2)):
        assert torch.equal(x.view(torch.uint8), y.view(torch.uint8))


def test_reactive_run_against_oracle(gpu):
    """16 steps of the 2D reactive loop (fluid BGK + A+B->C), FP64."""
    print("case reactive run D2Q9/D2Q5 f64", flush=True)
    n = 32
    f0, u = cases.taylor_green_2d(n, 0.02)
    a0, b0 = cases.split_species(n, n, u)
    omega, omega_c, k = 1.0 / 0.8, 1.0 / 0.8, 0.05
    wf, wa, wb = O.reactive_run("D2Q9", "BGK", "D2Q5", "f64", f0, a0, b0, omega, omega_c, k, 16)
    f, a, b = (to_dev(x, "f64", gpu) for x in (f0, a0, b0))
    ft, at, bt = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    for _ in range(16):
        lbcube.reactive_fused_step("D2Q9", "BGK", "D2Q5", f, ft, a, at, b, bt, omega, omega_c, k)
        f, ft, a, at, b, bt = ft, f, at, a, bt, b
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy(), wf)  # the fluid never sees exp()
    for got, want in ((a, wa), (b, wb)):
        g = got.cpu().numpy()
        assert np.max(np.abs(g - want)) <= 1e-13


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("lattice", FLUIDS)
def test_fluid_diagnostics(gpu, lattice, storage):
    f = cases.random_perturbed(lattice, SHAPE[lattice], storage, seed=13)
    want = O.fluid_diagnostics(lattice, storage, f, wide=True)
    got = lbcube.fluid_diagnostics(lattice, to_dev(f, storage, gpu))
    tol = 1e-13 if storage == "f64" else 1e-6
    assert abs(got[0] - want[0]) <= tol * abs(want[0])
    assert abs(got[1] - want[1]) <= tol * abs(want[1]) + 1e-300
    assert got[2] == want[2]  # max of correctly rounded sqrt: exact


@pytest.mark.parametrize("refs", [(0.0, 0.0), (0.5, 0.4)])
@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("sl", ["D2Q5", "D3Q7"])
def test_reactive_stats(gpu, sl, storage, refs):
    shape = SHAPE["D2Q9"] if sl == "D2Q5" else SHAPE["D3Q19"]
    a, b = _species(sl, shape, storage, seed=17, refs=refs)
    want = O.reactive_stats(sl, storage, a, b, 0.01, wide=True, ref_a=refs[0], ref_b=refs[1])
    got = lbcube.reactive_stats(sl, to_dev(a, storage, gpu), to_dev(b, storage, gpu), 0.01, ref_a=refs[0],
                                ref_b=refs[1])
    got = np.array([got[k] for k in lbcube.STAT_NAMES])
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-14)


def test_stats_closed_forms(gpu):
    """Complementary half planes: var = 1/4, cov = -1/4, segregation = -1 (rest fluid)."""
    n = 64
    u = np.zeros((2, n, n))
    a, b = cases.split_species(n, n, u)
    s = lbcube.reactive_stats("D2Q5", to_dev(a, "f64", gpu), to_dev(b, "f64", gpu), 0.5)
    assert abs(s["mean_A"] - 0.5) < 1e-15 and abs(s["mean_B"] - 0.5) < 1e-15
    assert abs(s["var_A"] - 0.25) < 1e-15 and abs(s["covariance"] + 0.25) < 1e-15
    assert abs(s["segregation_intensity"] + 1.0) < 1e-14
    assert abs(s["true_reaction_rate"]) < 1e-15 and abs(s["mixed_reaction_rate"] - 0.125) < 1e-15


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("lattice", FLUIDS)
def test_macro_fields(gpu, lattice, storage):
    """Per-cell density and velocity (the VTK writers' fields): bit-identical
    to the oracle's compute_macro_state in the storage's arithmetic."""
    print(f"case macro_fields {lattice} {storage}", flush=True)
    f = cases.random_perturbed(lattice, SHAPE[lattice], storage, seed=19)
    want_rho, want_u = O.macro_fields(lattice, storage, f)
    rho, u = lbcube.macro_fields(lattice, to_dev(f, storage, gpu))
    assert np.array_equal(rho.cpu().numpy(), want_rho)
    assert np.array_equal(u.cpu().numpy(), want_u)


@pytest.mark.parametrize("ref", [0.0, 0.6])
@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("sl", ["D2Q5", "D3Q7"])
def test_scalar_fields(gpu, sl, storage, ref):
    print(f"case scalar_fields {sl} {storage} ref {ref}", flush=True)
    shape = SHAPE["D2Q9"] if sl == "D2Q5" else SHAPE["D3Q19"]
    a, _ = _species(sl, shape, storage, seed=20, refs=(ref, 0.0))
    got = lbcube.scalar_fields(sl, to_dev(a, storage, gpu), ref=ref)
    assert np.array_equal(got.cpu().numpy(), O.scalar_fields(sl, storage, a, ref))


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("lattice", FLUIDS)
def test_max_speed(gpu, lattice, storage):
    """Stability guard: the oracle's max |u| exactly (max of correctly
    rounded square roots), and +inf once a population is NaN or infinite."""
    print(f"case max_speed {lattice} {storage}", flush=True)
    f = cases.random_perturbed(lattice, SHAPE[lattice], storage, seed=24)
    assert lbcube.max_speed(lattice, to_dev(f, storage, gpu)) == O.max_speed(lattice, storage, f)
    if storage in ("f64", "f32"):
        for bad in (np.nan, np.inf):
            g = f.copy()
            g[3].flat[g[3].size // 2] = bad
            assert lbcube.max_speed(lattice, to_dev(g, storage, gpu)) == np.inf


def test_max_speed_large_domain(gpu):
    """Many blocks: one NaN cell far into the domain still reports +inf."""
    f = torch.full((19, 64, 64, 64), 1.0 / 36.0, dtype=torch.float32, device=gpu)
    for i, w in enumerate(O.WEIGHTS["D3Q19"]):
        f[i] = w
    f[1, 50, 33, 17] += 0.01
    assert abs(lbcube.max_speed("D3Q19", f) - 0.01 / 1.01) < 1e-7
    f[5, 63, 63, 63] = float("nan")
    assert lbcube.max_speed("D3Q19", f) == np.inf
