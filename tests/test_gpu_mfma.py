"""GPU: the deviation-form D3Q19 / D3Q27 MRT with its moment transforms on the
matrix cores (LBC_FLAG_MFMA with LBC_FLAG_FAST, FP16 and BF16 storage).

It computes h' = h - M^T R (M h - dmeq) with both products as
v_mfma_f32_32x32x16 over tiles of 32 cells; only the relaxed moments R(...)
are rounded to the 16-bit format.  Held to the tolerances of the
deviation form:
  * against the reference-order oracle after 6 steps (the deviation form's
    tolerance), and against the VALU deviation form of the same run;
  * on domains whose last wave and last block are partial (dead lanes take
    part in the products and must store nothing: pitched buffers with
    sentinel padding show every plane's padding untouched);
  * conservation of mass and momentum to the storage's rounding;
  * the precision study of the reduced-storage formats (Taylor-Green,
    256^3, 2000 steps against FP64) in test_gpu_fast.py.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import lbcube
from oracle import oracle as O
from test_gpu_fast import FAST_TOL
from test_gpu_parity import SHAPE, to_dev, to_host

pytestmark = pytest.mark.gpu


def run_raw(lattice, storage, f0, omega, steps, gpu, mfma, block=None):
    a = to_dev(f0, storage, gpu)
    b = torch.empty_like(a)
    for _ in range(steps):
        lbcube.fluid_step(lattice, "MRT", a, b, omega, fast=True, mfma=mfma, block=block)
        a, b = b, a
    torch.cuda.synchronize()
    return to_host(a, storage)


def run(lattice, storage, f0, omega, steps, gpu, mfma, block=None):
    return O.decode(lattice, storage, run_raw(lattice, storage, f0, omega, steps, gpu, mfma, block))


def deviations64(storage, raw):
    """The stored deviations f - w of 16-bit populations, in float64."""
    if storage == "f16":
        return raw.view(np.float16).astype(np.float64)
    return (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_mfma_mrt_against_oracle(gpu, lattice, storage):
    print(f"case mfma {lattice} MRT {storage}", flush=True)
    f0 = cases.random_perturbed(lattice, SHAPE[lattice], storage, seed=13)
    for omega in (1.0 / 0.57, 1.0 / 0.9):
        got = run(lattice, storage, f0, omega, 6, gpu, mfma=True)
        want = O.decode(lattice, storage, O.fluid_run(lattice, "MRT", storage, f0, omega, 6))
        valu = run(lattice, storage, f0, omega, 6, gpu, mfma=False)
        err, err_valu = np.max(np.abs(got - want)), np.max(np.abs(valu - want))
        diff = np.max(np.abs(got - valu))
        print(f"{lattice} {storage} omega {omega:.4f}: max |mfma - oracle| {err:.2e}, |valu - oracle| "
              f"{err_valu:.2e}, |mfma - valu| {diff:.2e}", flush=True)
        assert err <= FAST_TOL[storage], (omega, err)
        assert diff <= FAST_TOL[storage], (omega, diff)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_mfma_partial_waves_leave_padding_untouched(gpu, lattice, storage):
    """Domains of 105 and 1 cells with blocks of 64 and 256 work-items (a
    partial last wave, waves wholly past the end): pitched planes whose
    padding holds a sentinel; the padding must stay as it was and the cells
    must equal the contiguous step's."""
    for shape, block in (((3, 5, 7), (256, 1, 1)), ((3, 5, 7), (64, 1, 1)), ((1, 1, 1), (128, 1, 1))):
        print(f"case mfma partial {lattice} {storage} shape {shape} block {block}", flush=True)
        f0 = cases.random_perturbed(lattice, shape, storage, seed=21)
        cur = to_dev(f0, storage, gpu)
        want = torch.empty_like(cur)
        lbcube.fluid_step(lattice, "MRT", cur, want, 1.0 / 0.6, fast=True, mfma=True, block=block)
        q = cur.shape[0]
        cells = int(np.prod(shape))
        pitch = cells + 77
        planes = []
        for sentinel in (0x3C3C, 0x5A5A):
            flat = torch.empty(q * pitch, dtype=cur.dtype, device=gpu)
            flat.view(torch.int16).fill_(sentinel)
            flat.view(q, pitch)[:, :cells].copy_(cur.reshape(q, cells))
            planes.append(flat)
        src, dst = planes
        lbcube.fluid_step_pitched(lattice, "MRT", src, dst, tuple(reversed(shape)), pitch, 1.0 / 0.6, fast=True,
                                  mfma=True, block=block)
        torch.cuda.synchronize()
        out = dst.view(torch.int16).cpu().numpy().reshape(q, pitch)
        assert np.array_equal(out[:, :cells].reshape((q,) + shape), want.view(torch.int16).cpu().numpy())
        assert np.all(out[:, cells:] == np.int16(0x5A5A))


def storage_ulp(storage, h):
    """Spacing of the 16-bit format at each stored deviation."""
    if storage == "f16":
        return np.spacing(np.abs(h).astype(np.float16)).astype(np.float64)
    return np.spacing(np.abs(h).astype(np.float32)).astype(np.float64) * 2.0 ** 16


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("lattice", ["D3Q19", "D3Q27"])
def test_mfma_conserves_mass_and_momentum(gpu, lattice, storage):
    """100 steps of a random state: the relaxation leaves rho and j alone, so
    mass and momentum move only by the rounding of the stored deviations to
    16 bits, a random walk of unbiased steps of about an ulp each: its
    standard deviation over the run is sqrt(steps * sum(ulp^2) / 12).  Both
    forms stay within six standard deviations.  (Sums of the deviations in
    float64: mass - cells = sum(f - w), momentum = sum c (f - w).)"""
    print(f"case mfma conservation {lattice} {storage}", flush=True)
    shape, steps = (12, 12, 12), 100
    f0 = cases.random_perturbed(lattice, shape, storage, seed=31)
    vel = np.asarray(lbcube.VELOCITIES[lattice], dtype=np.float64)

    def moments(h):
        flat = h.reshape(h.shape[0], -1)
        return flat.sum(), vel.T @ flat.sum(1)

    start = deviations64(storage, f0)
    m0, j0 = moments(start)
    ulp = storage_ulp(storage, start).reshape(start.shape[0], -1)
    sigma_m = np.sqrt(steps * (ulp ** 2).sum() / 12.0)
    sigma_j = np.sqrt(steps * ((vel[:, 0:1] != 0) * ulp ** 2).sum() / 12.0)
    for mfma in (True, False):
        got = deviations64(storage, run_raw(lattice, storage, f0, 1.0 / 0.6, steps, gpu, mfma=mfma))
        m1, j1 = moments(got)
        dm, dj = abs(m1 - m0), np.max(np.abs(j1 - j0))
        print(f"{lattice} {storage} mfma={mfma}: mass drift {dm:.2e} ({dm / sigma_m:.2f} sigma), momentum drift "
              f"{dj:.2e} ({dj / sigma_j:.2f} sigma)", flush=True)
        assert dm <= 6 * sigma_m and dj <= 6 * sigma_j, mfma
