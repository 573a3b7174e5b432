"""GPU: the pull-streaming step alone (lbc_pull_stream), the unshifted plane
copy (lbc_plane_copy) and the fluid step on pitched population buffers
(lbc_fluid_step_pitched).

  * lbc_pull_stream is the periodic shift next[i][x] = cur[i][x - c_i]
    (include/lattice_core.hpp:70-78) bit for bit, on odd extents, in every
    storage format, contiguous and pitched;
  * a pitched step equals the contiguous step plane for plane bit for bit, for
    every lattice, both arithmetic forms and every storage format, and leaves
    the padding between planes untouched.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import lbcube
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VEL = {k: v for k, v in lbcube.VELOCITIES.items()}
TORCH = {"f32": torch.float32, "f64": torch.float64, "f16": torch.float16, "bf16": torch.bfloat16}
SHAPES = {"D2Q9": (5, 7), "D3Q19": (3, 5, 7), "D3Q27": (4, 3, 9)}


def to_device(arr, storage, gpu):
    if storage in ("f16", "bf16"):
        return torch.from_numpy(np.ascontiguousarray(arr).view(np.int16)).to(gpu).view(TORCH[storage])
    return torch.from_numpy(np.ascontiguousarray(arr)).to(gpu)


def bits(t):
    if t.dtype in (torch.float16, torch.bfloat16):
        return t.view(torch.int16).cpu().numpy()
    return t.cpu().numpy()


def pitched(t, pitch, sentinel):
    """Copy of a (Q, ...) tensor into a flat buffer of Q planes of `pitch`
    elements, the padding filled with `sentinel` bits."""
    q = t.shape[0]
    cells = t[0].numel()
    flat = torch.empty(q * pitch, dtype=t.dtype, device=t.device)
    flat.view(torch.int16 if t.element_size() == 2 else (torch.int32 if t.element_size() == 4 else torch.int64)) \
        .fill_(sentinel)
    flat.view(q, pitch)[:, :cells].copy_(t.reshape(q, cells))
    return flat


@pytest.mark.parametrize("storage", ["f32", "f64", "f16"])
@pytest.mark.parametrize("lattice", ["D2Q9", "D3Q19", "D3Q27"])
def test_pull_stream_is_the_periodic_shift(gpu, lattice, storage):
    print(f"case pull_stream {lattice} {storage}", flush=True)
    shape = SHAPES[lattice]
    q = O.Q_OF[lattice]
    ids = (1.0 + np.arange(q * int(np.prod(shape)))).reshape((q,) + shape)
    if storage == "f16":
        ids = (ids % 2000).astype(np.float16).view(np.uint16)  # integers below 2048 are exact
    else:
        ids = ids.astype(np.float32 if storage == "f32" else np.float64)
    cur = to_device(ids, storage, gpu)
    nxt = torch.empty_like(cur)
    lbcube.pull_stream(lattice, cur, nxt)
    torch.cuda.synchronize()
    got = bits(nxt)
    want_src = ids.view(np.int16) if storage == "f16" else ids
    axes = tuple(range(len(shape) - 1, -1, -1))
    for i, c in enumerate(VEL[lattice]):
        assert np.array_equal(got[i], np.roll(want_src[i], shift=tuple(c), axis=axes)), i
    # pitched: the same shift, the padding untouched
    cells = int(np.prod(shape))
    pitch = cells + 37
    fc, fn = pitched(cur, pitch, 0x5A5A), pitched(cur, pitch, 0x3C3C)
    ext = tuple(reversed(shape)) + ((1,) if len(shape) == 2 else ())
    lbcube.pull_stream(lattice, fc, fn, extents=ext, plane_pitch=pitch)
    torch.cuda.synchronize()
    planes = bits(fn).reshape(q, pitch)
    assert np.array_equal(planes[:, :cells].reshape(got.shape), got)
    pad = planes[:, cells:]
    ref = bits(pitched(cur, pitch, 0x3C3C)).reshape(q, pitch)[:, cells:]
    assert np.array_equal(pad, ref)
    # lbc_plane_copy: the same plane streams without the shifts (the identity)
    same = torch.empty_like(cur)
    lbcube.pull_stream(lattice, cur, same, shift=False)
    fc, fn = pitched(cur, pitch, 0x5A5A), pitched(cur, pitch, 0x3C3C)
    lbcube.pull_stream(lattice, fc, fn, extents=ext, plane_pitch=pitch, shift=False)
    torch.cuda.synchronize()
    assert np.array_equal(bits(same), bits(cur))
    planes = bits(fn).reshape(q, pitch)
    assert np.array_equal(planes[:, :cells].reshape(got.shape), bits(cur))
    assert np.array_equal(planes[:, cells:], ref)


CASES = [("D3Q19", "TRT", "f32", False), ("D3Q19", "TRT", "f32", True), ("D3Q19", "TRT", "f16", True),
         ("D3Q19", "MRT", "f64", False), ("D3Q27", "MRT", "f32", False), ("D3Q27", "RLBM", "bf16", False),
         ("D2Q9", "BGK", "f64", False), ("D2Q9", "MRT", "f32", True)]


@pytest.mark.parametrize("lattice,coll,storage,fast", CASES)
def test_pitched_step_equals_contiguous(gpu, lattice, coll, storage, fast):
    print(f"case pitched step {lattice} {coll} {storage} fast={fast}", flush=True)
    shape = {"D2Q9": (33, 70), "D3Q19": (6, 9, 70), "D3Q27": (5, 7, 66)}[lattice]
    f = cases.random_perturbed(lattice, shape, storage, seed=61)
    cur = to_device(f, storage, gpu)
    want = torch.empty_like(cur)
    lbcube.fluid_step(lattice, coll, cur, want, 1 / 0.62, fast=fast)
    cells = int(np.prod(shape))
    ext = tuple(reversed(shape)) + ((1,) if len(shape) == 2 else ())
    for pitch in (cells, cells + 1000, cells + 4096 + 3):
        fc, fn = pitched(cur, pitch, 0x7777), pitched(cur, pitch, 0x1111)
        lbcube.fluid_step_pitched(lattice, coll, fc, fn, ext, pitch, 1 / 0.62, fast=fast)
        torch.cuda.synchronize()
        q = cur.shape[0]
        planes = bits(fn).reshape(q, pitch)
        assert np.array_equal(planes[:, :cells].reshape(bits(want).shape), bits(want)), pitch
        ref = bits(pitched(cur, pitch, 0x1111)).reshape(q, pitch)[:, cells:]
        assert np.array_equal(planes[:, cells:], ref), pitch
