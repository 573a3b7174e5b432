"""GPU: strided and tiled population layouts (lbc_fluid_step_strided,
lbc_pull_stream_strided, lbc_relayout; lbc_fluid_step_tiled,
lbc_reactive_fused_step_tiled, lbc_reactive_stats_tiled, lbc_relayout_tiled),
above all the slab order [Z][Q][Y][X], in which the Q planes of each xy-slab
lie side by side, and 2D tiles of rows [ny / R][Q][R][X].

The layout changes where a population is stored, never its value:
  * the reference-order step in the slab order equals the oracle bit for bit
    (every lattice, collision and storage, odd extents);
  * the deviation form and the matrix-core MRT equal their contiguous-layout
    runs bit for bit;
  * padded strides (planes and slabs spaced apart) leave the padding as it
    was;
  * lbc_relayout equals torch's permutation of the axes, and the pull step in
    the slab order is the pull permutation;
  * at 256^3 (the benchmark's domain) the slab order's TRT step equals the
    contiguous one after 4 steps;
  * 2D tiles (1, 4, a ragged divisor and all rows per tile): the D2Q9 step
    equals the oracle bit for bit; the fused fluid + A + B -> C step (2D and
    3D, both arithmetic forms, every storage) and the reactive statistics
    equal their runs on the reference's planes bit for bit, also on the
    2048^2 domain of the 2D reactive benchmark.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import lbcube
from oracle import oracle as O
from test_gpu_parity import COLLS, SHAPE, STORES, _species, to_dev, to_host

pytestmark = pytest.mark.gpu

LATTICES3 = ["D3Q19", "D3Q27"]


def run_slabs(lattice, coll, f, omega, steps, fast=False, mfma=False, block=None):
    """`steps` steps of a (Q, nz, ny, nx) population tensor, run in the slab
    order and converted back."""
    a = lbcube.to_slabs(lattice, f)
    b = torch.empty_like(a)
    for _ in range(steps):
        lbcube.fluid_step_slabs(lattice, coll, a, b, omega, fast=fast, mfma=mfma, block=block)
        a, b = b, a
    return lbcube.from_slabs(lattice, a)


def run_planes(lattice, coll, f, omega, steps, fast=False, mfma=False, block=None):
    a = f.clone()
    b = torch.empty_like(a)
    for _ in range(steps):
        lbcube.fluid_step(lattice, coll, a, b, omega, fast=fast, mfma=mfma, block=block)
        a, b = b, a
    return a


def bits(t):
    return t.view(torch.int16) if t.element_size() == 2 else t


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("coll", COLLS)
@pytest.mark.parametrize("lattice", LATTICES3)
def test_slab_order_reference_step_matches_oracle(gpu, lattice, coll, storage):
    print(f"case slabs {lattice} {coll} {storage}", flush=True)
    f0 = cases.random_perturbed(lattice, SHAPE[lattice], storage, seed=11)
    omega = 1.0 / 0.62
    got = to_host(run_slabs(lattice, coll, to_dev(f0, storage, gpu), omega, 6), storage)
    want = O.fluid_run(lattice, coll, storage, f0, omega, 6)
    assert np.array_equal(got, want), f"max |diff| {np.max(np.abs(got.astype(np.float64) - want))}"


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("coll", ["BGK", "TRT", "MRT"])
@pytest.mark.parametrize("lattice", LATTICES3)
def test_slab_order_fast_step_matches_planes(gpu, lattice, coll, storage):
    print(f"case slabs fast {lattice} {coll} {storage}", flush=True)
    f = to_dev(cases.random_perturbed(lattice, SHAPE[lattice], storage, seed=12), storage, gpu)
    for block in (None, (64, 1, 1), (7, 3, 2)):
        got = run_slabs(lattice, coll, f, 1.0 / 0.57, 5, fast=True, block=block)
        want = run_planes(lattice, coll, f, 1.0 / 0.57, 5, fast=True, block=block)
        assert torch.equal(bits(got), bits(want)), block


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("lattice", LATTICES3)
def test_slab_order_mfma_matches_planes(gpu, lattice, storage):
    print(f"case slabs mfma {lattice} {storage}", flush=True)
    f = to_dev(cases.random_perturbed(lattice, (6, 5, 13), storage, seed=13), storage, gpu)
    for block in (None, (64, 1, 1)):
        got = run_slabs(lattice, "MRT", f, 1.0 / 0.6, 4, fast=True, mfma=True, block=block)
        want = run_planes(lattice, "MRT", f, 1.0 / 0.6, 4, fast=True, mfma=True, block=block)
        assert torch.equal(bits(got), bits(want)), block


def test_relayout_is_the_axis_permutation(gpu):
    for lattice in ("D3Q19", "D3Q27", "D2Q9"):
        for storage in STORES:
            shape = SHAPE[lattice]
            f = to_dev(cases.random_perturbed(lattice, shape, storage, seed=5), storage, gpu)
            if lattice == "D2Q9":
                # one slab: the slab order is the reference's
                h = torch.empty_like(f)
                lbcube.relayout(lattice, f, h, (shape[1], shape[0], 1), (0, 0),
                                lbcube.slab_strides(lattice, (shape[1], shape[0])))
                assert torch.equal(bits(h), bits(f))
                continue
            h = lbcube.to_slabs(lattice, f)
            assert torch.equal(bits(h), bits(f.permute(1, 0, 2, 3).contiguous())), (lattice, storage)
            assert torch.equal(bits(lbcube.from_slabs(lattice, h)), bits(f))


@pytest.mark.parametrize("lattice", LATTICES3)
def test_padded_strides_leave_padding_untouched(gpu, lattice):
    """Planes spaced nx ny + 37 apart within a slab, slabs Q (nx ny + 37) + 101
    apart: the step equals the dense slab order's, the padding keeps its
    sentinel."""
    print(f"case padded strides {lattice}", flush=True)
    shape = (5, 7, 9)
    nz, ny, nx = shape
    q = lbcube.Q[lattice]
    f = to_dev(cases.random_perturbed(lattice, shape, "f32", seed=19), "f32", gpu)
    want = run_slabs(lattice, "TRT", f, 1.0 / 0.7, 1, fast=True)
    plane, slab = nx * ny + 37, q * (nx * ny + 37) + 101
    size = (q - 1) * plane + (nz - 1) * slab + nx * ny + 64
    bufs = []
    for sentinel in (-7.0, 13.0):
        buf = torch.full((size,), sentinel, dtype=torch.float32, device=gpu)
        bufs.append(buf)
    lbcube.relayout(lattice, f, bufs[0], (nx, ny, nz), (0, 0), (plane, slab))
    mask = torch.ones(size, dtype=torch.bool, device=gpu)
    for z in range(nz):
        for i in range(q):
            base = z * slab + i * plane
            mask[base:base + nx * ny] = False
    assert torch.all(bufs[0][mask] == -7.0)
    lbcube.fluid_step_strided(lattice, "TRT", bufs[0], bufs[1], (nx, ny, nz), plane, slab, 1.0 / 0.7, fast=True)
    torch.cuda.synchronize()
    assert torch.all(bufs[1][mask] == 13.0)
    back = torch.empty_like(f)
    lbcube.relayout(lattice, bufs[1], back, (nx, ny, nz), (plane, slab), (0, 0))
    assert torch.equal(back, want)


@pytest.mark.parametrize("lattice", LATTICES3)
def test_pull_stream_in_slab_order_is_the_pull_permutation(gpu, lattice):
    shape = SHAPE[lattice]
    nz, ny, nx = shape
    q = lbcube.Q[lattice]
    ids = (1.0 + torch.arange(q * nz * ny * nx, dtype=torch.float64, device=gpu)).reshape((q,) + shape)
    h = lbcube.to_slabs(lattice, ids)
    out = torch.empty_like(h)
    strides = lbcube.slab_strides(lattice, (nx, ny))
    lbcube.pull_stream_strided(lattice, h, out, (nx, ny, nz), *strides)
    got = lbcube.from_slabs(lattice, out)
    want = torch.stack([torch.roll(ids[i], shifts=(c[2], c[1], c[0]), dims=(0, 1, 2))
                        for i, c in enumerate(lbcube.VELOCITIES[lattice])])
    assert torch.equal(got, want)
    lbcube.pull_stream_strided(lattice, h, out, (nx, ny, nz), *strides, shift=False)
    assert torch.equal(out, h)


def test_slab_order_at_256_cubed(gpu):
    """The benchmark's domain: 4 TRT steps of the Taylor-Green state in FP32,
    reference order and deviation form, equal in both layouts."""
    print("case slabs 256^3", flush=True)
    from lbcube import init as lbinit
    u = lbinit.taylor_green_3d(256, 0.05, device=gpu, perturbation_rms=0.01 * 0.05, seed=0x5EED1234)
    f = lbinit.equilibrium("D3Q19", None, u, "f32")
    del u
    for fast in (False, True):
        got = run_slabs("D3Q19", "TRT", f, 1 / 0.50382, 4, fast=fast)
        want = run_planes("D3Q19", "TRT", f, 1 / 0.50382, 4, fast=fast)
        assert torch.equal(got, want), fast
        del got, want
    torch.cuda.empty_cache()


def test_slab_order_and_tiles_near_the_element_limit(gpu):
    """608^3 cells, D3Q19 TRT, FP32 deviation form: a buffer in the slab order
    or in tiles holds 19 x 608^3 = 4,270,358,528 elements, 0.6% below 2^32
    (the limit of these layouts; 610^3 is refused), and takes the
    64-bit-address kernels.  Two steps in the slab order and in tiles of 16
    rows equal the reference's planes bit for bit."""
    print("case slabs and tiles 608^3", flush=True)
    from lbcube import init as lbinit
    n, omega = 608, 1 / 0.51
    assert 19 * n ** 3 < 1 << 32 < 19 * 610 ** 3
    u = lbinit.taylor_green_3d(n, 0.05, device=gpu)
    f = lbinit.equilibrium("D3Q19", None, u, "f32")
    del u
    torch.cuda.empty_cache()
    want = run_planes("D3Q19", "TRT", f, omega, 2, fast=True)
    torch.cuda.empty_cache()
    got = run_slabs("D3Q19", "TRT", f, omega, 2, fast=True)
    assert torch.equal(got, want)
    del got
    torch.cuda.empty_cache()
    print("case tiles of 16 rows 608^3", flush=True)
    a = lbcube.to_tiles("D3Q19", f, 16)
    b = torch.empty_like(a)
    for _ in range(2):
        lbcube.fluid_step_tiled("D3Q19", "TRT", a, b, omega, fast=True)
        a, b = b, a
    del b
    assert torch.equal(lbcube.from_tiles("D3Q19", a), want)
    del a, f, want
    torch.cuda.empty_cache()


def test_relayout_tiled_is_the_tile_permutation(gpu):
    """2D tiles of R rows: tile t holds rows tR..tR+R-1 of every plane; 3D
    tiles are the xy-slabs.  Fluid and species lattices, every storage."""
    for lattice, shape in (("D2Q9", (12, 7)), ("D2Q5", (12, 7)), ("D3Q19", (3, 4, 5)), ("D3Q7", (3, 4, 5))):
        for storage in STORES:
            q = lbcube.Q[lattice]
            f = to_dev(cases.random_perturbed("D2Q9" if q in (9, 5) else "D3Q19", shape, storage, seed=3)[:q],
                       storage, gpu)
            for rows in ((1, 3, 4, 12) if len(shape) == 2 else (shape[1],)):
                t = lbcube.to_tiles(lattice, f, rows)
                if len(shape) == 2:
                    want = f.reshape(q, shape[0] // rows, rows, shape[1]).permute(1, 0, 2, 3).contiguous()
                else:
                    want = f.permute(1, 0, 2, 3).contiguous()
                assert torch.equal(bits(t), bits(want)), (lattice, storage, rows)
                assert torch.equal(bits(lbcube.from_tiles(lattice, t)), bits(f))


def run_tiled_fluid(lattice, coll, f, omega, steps, rows, fast=False, block=None, mfma=False):
    a = lbcube.to_tiles(lattice, f, rows)
    b = torch.empty_like(a)
    for _ in range(steps):
        lbcube.fluid_step_tiled(lattice, coll, a, b, omega, fast=fast, block=block, mfma=mfma)
        a, b = b, a
    return lbcube.from_tiles(lattice, a)


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("coll", COLLS)
def test_tiled_d2q9_step_matches_oracle(gpu, coll, storage):
    """Odd nx, ny = 12 in tiles of 1, 3, 4 and 12 rows: every y step across a
    tile boundary and the periodic wrap across all tiles."""
    print(f"case tiles D2Q9 {coll} {storage}", flush=True)
    f0 = cases.random_perturbed("D2Q9", (12, 13), storage, seed=11)
    omega = 1.0 / 0.62
    want = O.fluid_run("D2Q9", coll, storage, f0, omega, 6)
    f = to_dev(f0, storage, gpu)
    for rows in (1, 3, 4, 12):
        got = to_host(run_tiled_fluid("D2Q9", coll, f, omega, 6, rows), storage)
        assert np.array_equal(got, want), rows
    if coll != "RLBM":
        for rows in (1, 4):
            got = run_tiled_fluid("D2Q9", coll, f, omega, 5, rows, fast=True, block=(64, 1, 1))
            assert torch.equal(bits(got), bits(run_planes("D2Q9", coll, f, omega, 5, fast=True, block=(64, 1, 1))))


def run_fused(fl, coll, sl, f, a, b, steps, fast, refs, rows=None, block=None):
    """`steps` fused steps on the reference's planes (rows None) or in tiles."""
    omega, omega_c, k = 1.3, 1.05, 0.02
    if rows is not None:
        f, a, b = lbcube.to_tiles(fl, f, rows), lbcube.to_tiles(sl, a, rows), lbcube.to_tiles(sl, b, rows)
        step = lbcube.reactive_fused_step_tiled
    else:
        f, a, b = f.clone(), a.clone(), b.clone()
        step = lbcube.reactive_fused_step
    ft, at, bt = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
    for _ in range(steps):
        step(fl, coll, sl, f, ft, a, at, b, bt, omega, omega_c, k, fast=fast, ref_a=refs[0], ref_b=refs[1],
             block=block)
        f, ft, a, at, b, bt = ft, f, at, a, bt, b
    if rows is not None:
        return lbcube.from_tiles(fl, f), lbcube.from_tiles(sl, a), lbcube.from_tiles(sl, b)
    return f, a, b


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("pair", [("D2Q9", "D2Q5"), ("D3Q19", "D3Q7"), ("D3Q27", "D3Q7")])
def test_tiled_fused_step_matches_planes(gpu, pair, storage):
    fl, sl = pair
    shape = (12, 13) if fl == "D2Q9" else (4, 6, 7)
    refs = (0.5, 0.25) if storage in ("f16", "bf16") else (0.0, 0.0)
    f = to_dev(cases.random_perturbed(fl, shape, storage, seed=2), storage, gpu)
    a0, b0 = _species(sl, shape, storage, seed=8, refs=refs)
    a, b = to_dev(a0, storage, gpu), to_dev(b0, storage, gpu)
    row_set = (1, 4, 12) if fl == "D2Q9" else (shape[1],)
    for coll in ("BGK", "TRT", "MRT", "RLBM"):
        for fast in ((False, True) if coll != "RLBM" else (False,)):
            print(f"case tiles fused {fl} {coll} {storage} fast={fast}", flush=True)
            want = run_fused(fl, coll, sl, f, a, b, 4, fast, refs)
            for rows in row_set:
                got = run_fused(fl, coll, sl, f, a, b, 4, fast, refs, rows=rows)
                for x, y in zip(got, want):
                    assert torch.equal(bits(x), bits(y)), (coll, fast, rows)


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("pair", [("D2Q9", "D2Q5"), ("D3Q19", "D3Q7"), ("D3Q27", "D3Q7")])
def test_tiled_scalar_and_reaction_steps_match_planes(gpu, pair, storage):
    """The reference's separate launches in tiles (lbc_scalar_step_tiled,
    lbc_reaction_ab_step_tiled): three alternating fluid / scalar / reaction
    steps on tiled buffers equal the same steps on the reference's planes bit
    for bit, for tiles of several heights (3D: the slab order and tiles of
    fewer rows)."""
    fl, sl = pair
    shape = (12, 13) if fl == "D2Q9" else (4, 6, 7)
    refs = (0.5, 0.25) if storage in ("f16", "bf16") else (0.0, 0.0)
    f = to_dev(cases.random_perturbed(fl, shape, storage, seed=2), storage, gpu)
    a0, b0 = _species(sl, shape, storage, seed=8, refs=refs)
    a, b = to_dev(a0, storage, gpu), to_dev(b0, storage, gpu)
    omega, omega_c, k = 1.3, 1.05, 0.02

    def run(rows):
        tiled = rows is not None
        cf, ca, cb = ((lbcube.to_tiles(fl, f, rows), lbcube.to_tiles(sl, a, rows), lbcube.to_tiles(sl, b, rows))
                      if tiled else (f.clone(), a.clone(), b.clone()))
        nf, na, nb, ns = (torch.empty_like(x) for x in (cf, ca, cb, ca))
        scal = lbcube.scalar_step_tiled if tiled else lbcube.scalar_step
        react = lbcube.reaction_step_tiled if tiled else lbcube.reaction_step
        fluid = lbcube.fluid_step_tiled if tiled else lbcube.fluid_step
        for _ in range(3):
            fluid(fl, "TRT", cf, nf, omega)
            scal(fl, sl, ca, ns, nf, omega_c, 3e-4, ref=refs[0])
            react(fl, sl, ca, na, cb, nb, nf, omega_c, k, ref_a=refs[0], ref_b=refs[1])
            cf, nf, ca, na, cb, nb = nf, cf, na, ca, nb, cb
        out = (cf, ca, cb, ns)
        if tiled:
            out = tuple(lbcube.from_tiles(lat, x) for lat, x in zip((fl, sl, sl, sl), out))
        torch.cuda.synchronize()
        return out

    want = run(None)
    for rows in ((1, 4, 12) if fl == "D2Q9" else (6, 3, 2)):
        print(f"case tiles scalar and reaction {fl} {storage} rows {rows}", flush=True)
        for x, y in zip(run(rows), want):
            assert torch.equal(bits(x), bits(y)), rows


@pytest.mark.parametrize("pair", [("D2Q5", (12, 13), (1, 4, 12)), ("D3Q7", (4, 6, 7), (6,))])
def test_tiled_reactive_stats_equal_planes(gpu, pair):
    sl, shape, row_set = pair
    for storage in STORES:
        refs = (0.5, 0.25) if storage in ("f16", "bf16") else (0.0, 0.0)
        a0, b0 = _species(sl, shape, storage, seed=5, refs=refs)
        a, b = to_dev(a0, storage, gpu), to_dev(b0, storage, gpu)
        want = lbcube.reactive_stats(sl, a, b, 0.03, ref_a=refs[0], ref_b=refs[1])
        for rows in row_set:
            got = lbcube.reactive_stats_tiled(sl, lbcube.to_tiles(sl, a, rows), lbcube.to_tiles(sl, b, rows), 0.03,
                                              ref_a=refs[0], ref_b=refs[1])
            assert got == want, (storage, rows)


def test_tiled_fused_step_at_2048_squared(gpu):
    """The 2D reactive benchmark's domain: 3 fused MRT steps of a seeded
    state, FP32 reference order and FP16 deviation form, in tiles of 4 and 64
    rows, equal the reference's planes bit for bit."""
    print("case tiles fused 2048^2", flush=True)
    from lbcube import init as lbinit
    n = 2048
    rho, u = lbinit.taylor_green_2d(n, n, 0.04)
    ca, cb = lbinit.tanh_half_planes(n, delta=40.0)
    for storage, fast, refs in (("f32", False, (0.0, 0.0)), ("f16", True, (0.5, 0.5))):
        f = lbinit.equilibrium("D2Q9", rho, u, storage)
        a = lbinit.scalar_equilibrium("D2Q5", ca, u, storage, refs[0])
        b = lbinit.scalar_equilibrium("D2Q5", cb, u, storage, refs[1])
        want = run_fused("D2Q9", "MRT", "D2Q5", f, a, b, 3, fast, refs)
        for rows in (4, 64):
            got = run_fused("D2Q9", "MRT", "D2Q5", f, a, b, 3, fast, refs, rows=rows)
            for x, y in zip(got, want):
                assert torch.equal(bits(x), bits(y)), (storage, rows)
        del f, a, b, want, got
    torch.cuda.empty_cache()


def test_tiled_fused_step_near_the_element_limit(gpu):
    """21824^2 cells, FP16 deviation form: the tiled fluid buffer holds
    9 x 476,286,976 = 4,286,582,784 elements, 0.2% below 2^32 (the tiled
    limit; 21888^2 is refused), and takes the 64-bit-address kernels.  Two
    fused MRT steps in tiles of 4 and 64 rows equal the reference's planes
    bit for bit."""
    print("case tiles fused 21824^2 f16 fast", flush=True)
    from lbcube import init as lbinit
    n = 21824
    rho, u = lbinit.taylor_green_2d(n, n, 0.04)
    f = lbinit.equilibrium("D2Q9", rho, u, "f16")
    ca, cb = lbinit.tanh_half_planes(n, delta=400.0)
    a = lbinit.scalar_equilibrium("D2Q5", ca, u, "f16", 0.5)
    b = lbinit.scalar_equilibrium("D2Q5", cb, u, "f16", 0.5)
    del rho, u, ca, cb
    torch.cuda.empty_cache()
    assert 9 * n * n < 1 << 32 < 9 * 21888 * 21888
    want = run_fused("D2Q9", "MRT", "D2Q5", f, a, b, 2, True, (0.5, 0.5))
    torch.cuda.empty_cache()
    for rows in (4, 64):
        print(f"case tiles fused 21824^2 rows {rows}", flush=True)
        got = run_fused("D2Q9", "MRT", "D2Q5", f, a, b, 2, True, (0.5, 0.5), rows=rows)
        for x, y in zip(got, want):
            assert torch.equal(bits(x), bits(y)), rows
        del got
        torch.cuda.empty_cache()
    del f, a, b, want
    torch.cuda.empty_cache()


@pytest.mark.parametrize("pair", [("D2Q9", "D2Q5", (12, 13), (0, 4)), ("D3Q19", "D3Q7", (4, 6, 7), (0, 6)),
                                  ("D3Q27", "D3Q7", (4, 6, 7), (0, 6))])
def test_reactive_pull_stream_is_the_pull_permutation(gpu, pair):
    """lbc_reactive_pull_stream, the fused step's pattern ceiling: on the
    reference's planes and in tiles, every fluid and species population moves
    by the pull permutation (shift) or stays (no shift)."""
    fl, sl, shape, row_set = pair
    qf, qs = lbcube.Q[fl], lbcube.Q[sl]
    n = int(np.prod(shape))
    ids = [(1.0 + torch.arange(q * n, dtype=torch.float64, device=gpu) + off).reshape((q,) + shape)
           for q, off in ((qf, 0), (qs, 1e6), (qs, 2e6))]
    dims = tuple(range(len(shape)))

    def rolled(x, lattice):
        vel = lbcube.VELOCITIES[lattice]
        return torch.stack([torch.roll(x[i], shifts=tuple(reversed(c[:len(shape)])), dims=dims)
                            for i, c in enumerate(vel)])

    want = [rolled(ids[0], fl), rolled(ids[1], sl), rolled(ids[2], sl)]
    for rows in row_set:
        for shift in (True, False):
            if rows:
                cur = [lbcube.to_tiles(lat, x, rows) for lat, x in zip((fl, sl, sl), ids)]
            else:
                cur = [x.clone() for x in ids]
            nxt = [torch.empty_like(x) for x in cur]
            lbcube.reactive_pull_stream(fl, sl, cur[0], nxt[0], cur[1], nxt[1], cur[2], nxt[2], tiled=bool(rows),
                                        shift=shift)
            got = [lbcube.from_tiles(lat, x) if rows else x for lat, x in zip((fl, sl, sl), nxt)]
            for g, w, x in zip(got, want, ids):
                assert torch.equal(g, w if shift else x), (rows, shift)


@pytest.mark.parametrize("storage", STORES)
@pytest.mark.parametrize("lattice", LATTICES3)
def test_3d_tiles_of_rows_match_oracle(gpu, lattice, storage):
    """3D tiles of fewer rows than a slab (ny = 6 in tiles of 1, 2 and 3
    rows): y steps across tile boundaries, z steps across whole slabs of
    tiles; the reference-order step equals the oracle bit for bit and the
    deviation form equals its run on the planes."""
    print(f"case 3d tiles {lattice} {storage}", flush=True)
    f0 = cases.random_perturbed(lattice, (4, 6, 7), storage, seed=23)
    want = O.fluid_run(lattice, "TRT", storage, f0, 1.0 / 0.62, 4)
    f = to_dev(f0, storage, gpu)
    for rows in (1, 2, 3):
        got = to_host(run_tiled_fluid(lattice, "TRT", f, 1.0 / 0.62, 4, rows), storage)
        assert np.array_equal(got, want), rows
        fast = run_tiled_fluid(lattice, "MRT", f, 1.0 / 0.6, 3, rows, fast=True)
        assert torch.equal(bits(fast), bits(run_planes(lattice, "MRT", f, 1.0 / 0.6, 3, fast=True))), rows
        if storage in ("f16", "bf16"):  # the matrix-core MRT through the same tile addressing
            mf = run_tiled_fluid(lattice, "MRT", f, 1.0 / 0.6, 3, rows, fast=True, mfma=True, block=(64, 1, 1))
            want_mf = run_planes(lattice, "MRT", f, 1.0 / 0.6, 3, fast=True, mfma=True, block=(64, 1, 1))
            assert torch.equal(bits(mf), bits(want_mf)), rows


def test_3d_tiles_fused_and_stats_match_planes(gpu):
    """The fused reactive step and the statistics in 3D tiles of 2 and 3 rows
    (D3Q19 + two D3Q7) equal their runs on the planes bit for bit."""
    shape = (4, 6, 7)
    for storage in ("f32", "f16"):
        refs = (0.5, 0.25) if storage == "f16" else (0.0, 0.0)
        f = to_dev(cases.random_perturbed("D3Q19", shape, storage, seed=2), storage, gpu)
        a0, b0 = _species("D3Q7", shape, storage, seed=8, refs=refs)
        a, b = to_dev(a0, storage, gpu), to_dev(b0, storage, gpu)
        for fast in (False, True):
            want = run_fused("D3Q19", "TRT", "D3Q7", f, a, b, 3, fast, refs)
            for rows in (2, 3):
                got = run_fused("D3Q19", "TRT", "D3Q7", f, a, b, 3, fast, refs, rows=rows)
                for x, y in zip(got, want):
                    assert torch.equal(bits(x), bits(y)), (storage, fast, rows)
        want = lbcube.reactive_stats("D3Q7", a, b, 0.03, ref_a=refs[0], ref_b=refs[1])
        for rows in (2, 3):
            got = lbcube.reactive_stats_tiled("D3Q7", lbcube.to_tiles("D3Q7", a, rows), lbcube.to_tiles("D3Q7", b, rows),
                                              0.03, ref_a=refs[0], ref_b=refs[1])
            assert got == want, (storage, rows)
        t = lbcube.to_tiles("D3Q19", f, 2)
        assert t.shape == (4, 3, 19, 2, 7)
        want_t = f.reshape(19, 4, 3, 2, 7).permute(1, 2, 0, 3, 4).contiguous()
        assert torch.equal(bits(t), bits(want_t))
