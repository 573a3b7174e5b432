"""GPU: strided population layouts (lbc_fluid_step_strided, lbc_pull_stream_strided,
lbc_relayout), above all the slab order [Z][Q][Y][X], in which the Q planes of
each xy-slab lie side by side.

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
    contiguous one after 4 steps.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import lbcube
from oracle import oracle as O
from test_gpu_parity import COLLS, SHAPE, STORES, TORCH_DT, to_dev, to_host

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
