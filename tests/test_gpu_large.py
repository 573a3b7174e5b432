"""Domains of more than 2^29 cells and population planes of more than 4 GiB
(the kernels' 64-bit-address instantiations), at their real sizes.

A periodic domain whose content repeats with period P along its slowest axis
evolves exactly as the P-long domain does, tile by tile.  Each test tiles a
small seeded state `reps` times, steps the large domain, and requires every
tile of the result to equal, bit for bit, the small domain's result, which is
itself compared with the oracle (bit-exact for the fluid step; the reaction
step to the tolerance of its exp(), as in the parity tests).  The statistics
of the large domain are the small domain's, summed reps times.

The tests need up to 232 GB of device memory (one MI355X has 288 GiB) and skip
where the device has less free.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import cases
import lbcube
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TORCH_DT = {"f64": torch.float64, "f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}
BITS = {torch.float64: torch.int64, torch.float32: torch.int32, torch.float16: torch.int16,
        torch.bfloat16: torch.int16}
REPS = 1025


def need(gpu, nbytes):
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info(gpu)
    if free < nbytes + (4 << 30):
        pytest.skip(f"needs {nbytes / 2**30:.0f} GiB of free device memory, {free / 2**30:.0f} GiB free")


@pytest.fixture
def release():
    yield
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def to_dev(a, storage, device):
    a = np.ascontiguousarray(a)
    if storage in ("f16", "bf16"):
        return torch.from_numpy(a.view(np.int16)).view(TORCH_DT[storage]).to(device)
    return torch.from_numpy(a).to(device)


def tile(small, reps):
    """Repeat along the slowest spatial axis (axis 1 of [Q, (Z,) Y, X])."""
    return small.repeat(1, reps, *([1] * (small.dim() - 2)))


def assert_tiles(big, want, reps, what):
    q = big.shape[0]
    bits = BITS[big.dtype]
    bv = big.view(bits).view(q, reps, -1)
    wv = want.view(bits).reshape(q, 1, -1)
    for i in range(q):
        bad = (bv[i] != wv[i]).any(dim=1)
        if bool(bad.any()):
            t = int(torch.nonzero(bad)[0])
            raise AssertionError(f"{what}: population {i} differs from the small domain in tile {t} of {reps}")


def plane_bytes(t):
    return t[0].numel() * t.element_size()


def steps(lattice, coll, a, b, omega, n, fast=False):
    for _ in range(n):
        lbcube.fluid_step(lattice, coll, a, b, omega, fast=fast)
        a, b = b, a
    return a, b


@pytest.mark.parametrize("lattice,coll,storage,small_shape,fast", [
    ("D2Q9", "TRT", "f32", (256, 4096), False),       # 2^30 + 2^20 cells, 4.004 GiB planes
    ("D2Q9", "MRT", "f64", (128, 4096), False),       # 2^29 + 2^19 cells, 4.004 GiB planes
    ("D3Q19", "TRT", "f32", (256, 64, 64), False),    # 2^30 + 2^20 cells, 4.004 GiB planes, 163 GB
    ("D3Q19", "TRT", "f16", (256, 64, 64), True),     # 2^30 + 2^20 cells, 2 GiB planes (32-bit offsets)
    ("D3Q27", "MRT", "f32", (256, 64, 64), False),    # 2^30 + 2^20 cells, 4.004 GiB planes, 232 GB
])
def test_fluid_large_domain_tiles(gpu, release, lattice, coll, storage, small_shape, fast):
    print(f"case large fluid {lattice} {coll} {storage} fast={fast} {small_shape} x {REPS}", flush=True)
    q = len(O.WEIGHTS[lattice])
    esize = torch.empty(0, dtype=TORCH_DT[storage]).element_size()
    cells = int(np.prod(small_shape)) * REPS
    assert cells > 1 << 29
    need(gpu, 2 * q * cells * esize)
    omega = 1.0 / 0.62
    f0 = cases.random_perturbed(lattice, small_shape, storage, seed=29)
    small = to_dev(f0, storage, gpu)
    s_out, s_tmp = steps(lattice, coll, small.clone(), torch.empty_like(small), omega, 2, fast)
    if not fast:
        want = O.fluid_run(lattice, coll, storage, f0, omega, 2)
        got = s_out.cpu().view(BITS[s_out.dtype]).numpy()
        assert np.array_equal(got, np.ascontiguousarray(want).view(got.dtype)), "small domain against the oracle"
    big = tile(small, REPS)
    assert big.shape[1] == small_shape[0] * REPS
    if storage in ("f32", "f64"):
        assert plane_bytes(big) > 1 << 32  # the 64-bit-address instantiation
    tmp = torch.empty_like(big)
    out, tmp = steps(lattice, coll, big, tmp, omega, 2, fast)
    torch.cuda.synchronize()
    del tmp
    assert_tiles(out, s_out, REPS, f"{lattice} {coll} {storage}")
    # statistics over the whole domain: reps times the small domain's
    m_big, e_big, u_big = lbcube.fluid_diagnostics(lattice, out)
    m_s, e_s, u_s = lbcube.fluid_diagnostics(lattice, s_out)
    assert abs(m_big - REPS * m_s) <= 2e-12 * abs(REPS * m_s)
    assert abs(e_big - REPS * e_s) <= 1e-10 * abs(REPS * e_s)
    assert u_big == u_s
    assert lbcube.max_speed(lattice, out) == lbcube.max_speed(lattice, s_out)


def test_macro_fields_large_domain(gpu, release):
    print("case large macro_fields D2Q9 f32 (256, 4096) x 1025", flush=True)
    lattice, storage, small_shape = "D2Q9", "f32", (256, 4096)
    cells = small_shape[0] * small_shape[1] * REPS
    need(gpu, 9 * cells * 4 + 3 * cells * 8)
    f0 = cases.random_perturbed(lattice, small_shape, storage, seed=31)
    small = to_dev(f0, storage, gpu)
    big = tile(small, REPS)
    assert plane_bytes(big) > 1 << 32
    rho_s, u_s = lbcube.macro_fields(lattice, small)
    rho, u = lbcube.macro_fields(lattice, big)
    del big
    assert torch.equal(rho.view(REPS, -1), rho_s.reshape(1, -1).expand(REPS, -1))
    assert torch.equal(u.view(2, REPS, -1), u_s.reshape(2, 1, -1).expand(2, REPS, -1))
    want_rho, want_u = O.macro_fields(lattice, storage, f0)
    assert np.array_equal(rho_s.cpu().numpy().ravel(), np.asarray(want_rho).ravel())
    assert np.array_equal(u_s.cpu().numpy().ravel(), np.asarray(want_u).ravel())


@pytest.mark.parametrize("fast", [False, True])
def test_reactive_fused_large_domain_tiles(gpu, release, fast):
    print(f"case large reactive fused D2Q9 MRT + 2 D2Q5 f32 fast={fast} (256, 4096) x 1025", flush=True)
    storage, shape = "f32", (256, 4096)
    cells = shape[0] * shape[1] * REPS
    need(gpu, (2 * 9 + 4 * 5) * cells * 4)
    # B - A stays near 0.5: the reference's FP32 reaction formula cancels at
    # A ~ B, where exp() ulps of device and host grow past the tolerance
    # below (the parity tests cover that branch on small domains)
    g = cases.rng(37)
    ca = g.uniform(0.05, 0.4, size=shape)
    cb = ca + 0.5 + g.uniform(-0.05, 0.05, size=shape)
    u = g.uniform(-0.02, 0.02, size=(2,) + shape)
    fa = O.scalar_feq_field("D2Q5", storage, ca, u)
    fb = O.scalar_feq_field("D2Q5", storage, cb, u)
    ff = cases.random_perturbed("D2Q9", shape, storage, seed=41)
    omega, omega_c, k = 1.0 / 0.6, 1.0 / 0.8, 0.3

    def run(f, a, b, n=2):
        nf, na, nb = torch.empty_like(f), torch.empty_like(a), torch.empty_like(b)
        for _ in range(n):
            lbcube.reactive_fused_step("D2Q9", "MRT", "D2Q5", f, nf, a, na, b, nb, omega, omega_c, k, fast=fast)
            f, nf, a, na, b, nb = nf, f, na, a, nb, b
        del nf, na, nb
        return f, a, b

    sf, sa, sb = (to_dev(x, storage, gpu) for x in (ff, fa, fb))
    rf, ra, rb = run(sf.clone(), sa.clone(), sb.clone())
    if not fast:
        wf, wa, wb = O.reactive_run("D2Q9", "MRT", "D2Q5", storage, ff, fa, fb, omega, omega_c, k, 2)
        assert np.array_equal(rf.cpu().numpy(), wf), "small domain fluid against the oracle"
        for got, want in ((ra, wa), (rb, wb)):  # exp() of device and host may differ by an ulp
            np.testing.assert_allclose(got.cpu().numpy(), want, rtol=2e-6, atol=1e-7)
    big = [tile(x, REPS) for x in (sf, sa, sb)]
    assert plane_bytes(big[0]) > 1 << 32
    del sf, sa, sb
    of, oa, ob = run(*big)
    torch.cuda.synchronize()
    del big
    assert_tiles(of, rf, REPS, "fluid")
    assert_tiles(oa, ra, REPS, "species A")
    assert_tiles(ob, rb, REPS, "species B")
    st_big = lbcube.reactive_stats("D2Q5", oa, ob, k)
    st_s = lbcube.reactive_stats("D2Q5", ra, rb, k)
    for name in ("mean_A", "mean_B", "var_A", "var_B", "true_reaction_rate"):
        assert abs(st_big[name] - st_s[name]) <= 1e-11 * max(abs(st_s[name]), 1e-30), name


def _separated_species(shape, storage, seed):
    """A and B about 0.5 apart (no cancellation in the reference's reaction
    formula), advected by a small random velocity."""
    g = cases.rng(seed)
    ca = g.uniform(0.05, 0.4, size=shape)
    cb = ca + 0.5 + g.uniform(-0.05, 0.05, size=shape)
    u = g.uniform(-0.02, 0.02, size=(2,) + shape)
    return O.scalar_feq_field("D2Q5", storage, ca, u), O.scalar_feq_field("D2Q5", storage, cb, u)


@pytest.mark.parametrize("react", [False, True])
def test_scalar_and_reaction_steps_large_domain_tiles(gpu, release, react):
    """The unfused ADR kernels (passive scalar; A + B -> C) on 4.004 GiB planes."""
    what = "reaction" if react else "scalar"
    print(f"case large {what} step D2Q9/D2Q5 f32 (256, 4096) x 1025", flush=True)
    storage, shape = "f32", (256, 4096)
    cells = shape[0] * shape[1] * REPS
    need(gpu, (9 + (4 if react else 2) * 5) * cells * 4)
    ff = cases.random_perturbed("D2Q9", shape, storage, seed=43)
    fa, fb = _separated_species(shape, storage, 47)
    omega_c, k, src = 1.0 / 0.8, 0.3, 1e-4

    def run(f, a, b, n=2):
        na, nb = torch.empty_like(a), (torch.empty_like(b) if react else None)
        for _ in range(n):
            if react:
                lbcube.reaction_step("D2Q9", "D2Q5", a, na, b, nb, f, omega_c, k)
                a, na, b, nb = na, a, nb, b
            else:
                lbcube.scalar_step("D2Q9", "D2Q5", a, na, f, omega_c, source=src)
                a, na = na, a
        del na, nb
        return a, b

    sf, sa, sb = (to_dev(x, storage, gpu) for x in (ff, fa, fb))
    ra, rb = run(sf, sa.clone(), sb.clone() if react else None)
    if react:
        wa, wb = O.reaction_step("D2Q9", "D2Q5", storage, fa, fb, ff, omega_c, k)
        wa, wb = O.reaction_step("D2Q9", "D2Q5", storage, wa, wb, ff, omega_c, k)
        for got, want in ((ra, wa), (rb, wb)):  # exp() of device and host may differ by an ulp
            np.testing.assert_allclose(got.cpu().numpy(), want, rtol=2e-6, atol=1e-7)
    else:
        want = O.scalar_step("D2Q9", "D2Q5", storage, fa, ff, omega_c, source=src)
        want = O.scalar_step("D2Q9", "D2Q5", storage, want, ff, omega_c, source=src)
        assert np.array_equal(ra.cpu().numpy(), want), "small domain against the oracle"
    big_f, big_a = tile(sf, REPS), tile(sa, REPS)
    big_b = tile(sb, REPS) if react else None
    assert plane_bytes(big_a) > 1 << 32 and plane_bytes(big_f) > 1 << 32
    del sf, sa, sb
    oa, ob = run(big_f, big_a, big_b)
    torch.cuda.synchronize()
    del big_f, big_a, big_b
    assert_tiles(oa, ra, REPS, f"{what} A")
    if react:
        assert_tiles(ob, rb, REPS, f"{what} B")
