"""The Python wrappers refuse tensors that do not belong to one domain and one
storage format before any kernel is enqueued.

The kernels take the extents from one tensor and the storage code from one
dtype; a fluid tensor of another shape, or a species stored in another format,
would be read or written past the end of its allocation.  These checks run on
the CPU: the CUDA-placement check is relaxed and the kernel library is replaced
by a stub that fails the test if any entry point is reached.
"""
from __future__ import annotations

import pytest
import torch

import lbcube
from lbcube import adjoint


class _NoKernel:
    def __getattr__(self, name):
        if name == "lbc_default_block":
            return lambda lattice, nx, ny, nz, out: None
        raise AssertionError(f"{name} was reached with mismatched tensors")


@pytest.fixture
def cpu_wrappers(monkeypatch):
    def check_pop(lattice, t, name):  # the shape/dtype part of _check_pop, without the CUDA placement
        if t.dim() != lbcube.DIM[lattice] + 1 or t.shape[0] != lbcube.Q[lattice]:
            raise ValueError(f"{name}: bad population shape")

    def check_storage(t, name):  # _check_storage without the CUDA placement
        if not t.is_contiguous() or t.dtype not in lbcube.STORAGE:
            raise ValueError(f"{name} must be a contiguous tensor of a storage dtype")

    stub = _NoKernel()
    monkeypatch.setattr(lbcube, "_check_pop", check_pop)
    monkeypatch.setattr(lbcube, "_check_storage", check_storage)
    monkeypatch.setattr(lbcube, "lib", lambda: stub)
    monkeypatch.setattr(adjoint, "_check_pop", check_pop)
    monkeypatch.setattr(adjoint, "lib", lambda: stub)
    return lbcube


def pops(q, shape, dtype=torch.float32):
    return torch.zeros((q,) + tuple(shape), dtype=dtype)


def test_fluid_pair_must_match(cpu_wrappers):
    with pytest.raises(ValueError, match="covers cells"):
        lbcube.fluid_step("D3Q19", "BGK", pops(19, (4, 4, 4)), pops(19, (4, 4, 5)), 1.0, block=(64, 1, 1))
    with pytest.raises(ValueError, match="storage format"):
        lbcube.fluid_step("D3Q19", "BGK", pops(19, (4, 4, 4)), pops(19, (4, 4, 4), torch.float16), 1.0,
                          block=(64, 1, 1))


def test_scalar_step_fluid_of_another_shape(cpu_wrappers):
    s = pops(5, (8, 8))
    with pytest.raises(ValueError, match="cur_f covers cells"):
        lbcube.scalar_step("D2Q9", "D2Q5", s, s.clone(), pops(9, (8, 16)), 1.0, block=(64, 1, 1))
    with pytest.raises(ValueError, match="storage format"):
        lbcube.scalar_step("D2Q9", "D2Q5", s, s.clone(), pops(9, (8, 8), torch.float64), 1.0, block=(64, 1, 1))


def test_reaction_step_mismatches(cpu_wrappers):
    a, f = pops(5, (8, 8)), pops(9, (8, 8))
    with pytest.raises(ValueError, match="nxt_b covers cells"):
        lbcube.reaction_step("D2Q9", "D2Q5", a, a.clone(), a.clone(), pops(5, (8, 4)), f, 1.0, 0.1,
                             block=(64, 1, 1))
    with pytest.raises(ValueError, match="cur_f covers cells"):
        lbcube.reaction_step("D2Q9", "D2Q5", a, a.clone(), a.clone(), a.clone(), pops(9, (4, 8)), 1.0, 0.1,
                             block=(64, 1, 1))
    with pytest.raises(ValueError, match="storage format"):
        lbcube.reaction_step("D2Q9", "D2Q5", a, a.clone(), pops(5, (8, 8), torch.float16), a.clone(), f, 1.0, 0.1,
                             block=(64, 1, 1))


def test_fused_step_mismatches(cpu_wrappers):
    f, a = pops(9, (16, 16)), pops(5, (16, 16))
    with pytest.raises(ValueError, match="storage format"):  # f16 species with an f32 fluid
        lbcube.reactive_fused_step("D2Q9", "BGK", "D2Q5", f, f.clone(), pops(5, (16, 16), torch.float16),
                                   pops(5, (16, 16), torch.float16), a, a.clone(), 1.0, 1.0, 0.1, block=(64, 1, 1))
    with pytest.raises(ValueError, match="cur_a covers cells"):  # species smaller than the fluid
        lbcube.reactive_fused_step("D2Q9", "BGK", "D2Q5", f, f.clone(), pops(5, (8, 8)), pops(5, (8, 8)), a,
                                   a.clone(), 1.0, 1.0, 0.1, block=(64, 1, 1))


def test_stats_pair_must_match(cpu_wrappers):
    with pytest.raises(ValueError, match="covers cells"):
        lbcube.reactive_stats("D2Q5", pops(5, (8, 8)), pops(5, (8, 9)), 0.1)


def test_gradient_species_shape(cpu_wrappers):
    f = pops(9, (64, 64))
    with pytest.raises(ValueError, match="a0 covers cells"):
        adjoint.reactive_gradient("BGK", f, pops(5, (32, 64)), pops(5, (64, 64)), 8, 1.0, 1.0, 0.1,
                                  block=(256, 1))
    with pytest.raises(ValueError, match="storage format"):
        adjoint.reactive_gradient("BGK", f, pops(5, (64, 64)), pops(5, (64, 64), torch.float64), 8, 1.0, 1.0, 0.1,
                                  block=(256, 1))


def test_strided_and_tiled_layout_mismatches(cpu_wrappers):
    """The strided, slab-order and tiled wrappers refuse buffers too small for
    their layout, pairs of different shape or dtype, tiles of another height,
    and rows that do not divide ny, before any kernel is reached."""
    q = 19
    with pytest.raises(ValueError, match="the layout needs"):  # 19 planes of 4 x 4 in slabs of 4: 1216 elements
        lbcube.fluid_step_strided("D3Q19", "TRT", torch.zeros(1215), torch.zeros(1216), (4, 4, 4), 16, 19 * 16, 1.0,
                                  block=(64, 1, 1))
    with pytest.raises(ValueError, match="share a shape"):
        lbcube.fluid_step_slabs("D3Q19", "TRT", torch.zeros(4, q, 4, 4), torch.zeros(5, q, 4, 4), 1.0,
                                block=(64, 1, 1))
    with pytest.raises(ValueError, match="shape"):
        lbcube.fluid_step_slabs("D3Q19", "TRT", torch.zeros(4, 9, 4, 4), torch.zeros(4, 9, 4, 4), 1.0,
                                block=(64, 1, 1))
    with pytest.raises(ValueError, match="rows must divide"):
        lbcube.to_tiles("D2Q9", pops(9, (12, 8)), rows=5)
    f, a = torch.zeros(3, 9, 4, 8), torch.zeros(3, 5, 4, 8)
    with pytest.raises(ValueError, match="match cur_a"):  # a fluid in tiles of 2 rows against species in 4
        lbcube.reactive_fused_step_tiled("D2Q9", "MRT", "D2Q5", torch.zeros(6, 9, 2, 8), torch.zeros(6, 9, 2, 8), a,
                                         a.clone(), a.clone(), a.clone(), 1.0, 1.0, 0.1, block=(64, 1, 1))
    with pytest.raises(ValueError, match="match cur_a"):
        lbcube.reactive_fused_step_tiled("D2Q9", "MRT", "D2Q5", f, f.clone(), a, a.clone(), a.clone(),
                                         a.to(torch.float16), 1.0, 1.0, 0.1, block=(64, 1, 1))
    with pytest.raises(ValueError, match="must have shape"):
        lbcube.reactive_fused_step_tiled("D2Q9", "MRT", "D2Q5", a, a.clone(), a, a.clone(), a.clone(), a.clone(),
                                         1.0, 1.0, 0.1, block=(64, 1, 1))
    with pytest.raises(ValueError, match="match cur_a"):
        lbcube.reactive_pull_stream("D2Q9", "D2Q5", f, f.clone(), a, a.clone(), a.clone(), torch.zeros(2, 5, 6, 8),
                                    tiled=True, block=(64, 1, 1))
    with pytest.raises(ValueError, match="match a's tiles"):
        lbcube.reactive_stats_tiled("D2Q5", a, torch.zeros(6, 5, 2, 8), 0.1)
    with pytest.raises(ValueError, match="match cur_s"):  # the fluid in tiles of 2 rows, the scalar in 4
        lbcube.scalar_step_tiled("D2Q9", "D2Q5", a, a.clone(), torch.zeros(6, 9, 2, 8), 1.0, block=(64, 1, 1))
    with pytest.raises(ValueError, match="must not alias"):
        lbcube.scalar_step_tiled("D2Q9", "D2Q5", a, a, f, 1.0, block=(64, 1, 1))
    with pytest.raises(ValueError, match="match cur_a"):
        lbcube.reaction_step_tiled("D2Q9", "D2Q5", a, a.clone(), a.clone(), a.to(torch.float16), f, 1.0, 0.1,
                                   block=(64, 1, 1))
    with pytest.raises(ValueError, match="must have shape"):  # a fluid lattice's tiles for a species
        lbcube.reaction_step_tiled("D2Q9", "D2Q5", f, f.clone(), f.clone(), f.clone(), f, 1.0, 0.1,
                                   block=(64, 1, 1))
