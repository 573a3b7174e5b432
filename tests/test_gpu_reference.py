"""GPU: LB-Cube's own validation suite (tests/test_validation.cpp) run on the
HIP kernels.

test_oracle_reference.py restates the reference's 47 tests against the oracle.
This module collects the same test functions again and runs them with every
step taken on the device through the C ABI, in FP64 storage (the reference's
tests use double): the fluid loops through lbc_fluid_step, the passive-scalar
test through lbc_scalar_step, the reaction tests through
lbc_reaction_step.  Each test runs twice: in the reference's evaluation order
(bit-exact to the oracle) and with the fluid step's deviation-form
arithmetic (LBC_FLAG_FAST; RLBM has none).  The oracle only builds the
initial fields here.  So the
kernels meet the reference's analytic expectations (decay rates, closed
forms, convergence orders) at the reference's tolerances directly, not only
through their bit-parity with the oracle.  The MRT transform-identity tests
exercise no kernel and stay in the CPU suite; the three D3Q19 MRT
expectations that do not hold for the reference's own algorithm remain strict
xfails.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import lbcube
import test_oracle_reference as TR
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_DEV = torch.device("cuda", 0)


def _to(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(_DEV)


class _DeviceSteps:
    """The oracle module with its stepping functions replaced by the HIP
    kernels (FP64 storage); everything else (field builders) delegates.
    `launches` counts the device steps a test took."""

    launches = 0
    fast = False

    def __getattr__(self, name):
        return getattr(O, name)

    @staticmethod
    def fluid_run(lattice, collision, dtype, f, omega, steps, threads=None):
        assert dtype == "f64"
        _DeviceSteps.launches += steps
        a = _to(f)
        b = torch.empty_like(a)
        for _ in range(steps):
            lbcube.fluid_step(lattice, collision, a, b, omega, fast=_DeviceSteps.fast)
            a, b = b, a
        return a.cpu().numpy()

    @staticmethod
    def fluid_step(lattice, collision, dtype, f, omega, threads=None):
        return _DeviceSteps.fluid_run(lattice, collision, dtype, f, omega, 1)

    @staticmethod
    def scalar_step(fluid_lattice, scalar_lattice, dtype, g, f, omega_c, source, ref=0.0):
        assert dtype == "f64"
        _DeviceSteps.launches += 1
        gd, fd = _to(g), _to(f)
        out = torch.empty_like(gd)
        lbcube.scalar_step(fluid_lattice, scalar_lattice, gd, out, fd, omega_c, source, ref=ref)
        return out.cpu().numpy()

    @staticmethod
    def reaction_step(fluid_lattice, scalar_lattice, dtype, a, b, f, omega_c, k, ref_a=0.0, ref_b=0.0):
        assert dtype == "f64"
        _DeviceSteps.launches += 1
        ad, bd, fd = _to(a), _to(b), _to(f)
        na, nb = torch.empty_like(ad), torch.empty_like(bd)
        lbcube.reaction_step(fluid_lattice, scalar_lattice, ad, na, bd, nb, fd, omega_c, k, ref_a=ref_a, ref_b=ref_b)
        return na.cpu().numpy(), nb.cpu().numpy()


@pytest.fixture(autouse=True, params=[False, True], ids=["reference_order", "deviation_form"])
def _steps_on_the_device(gpu, monkeypatch, request):
    print(f"case reference suite on the GPU: {request.node.name}", flush=True)
    monkeypatch.setattr(TR, "O", _DeviceSteps())
    _DeviceSteps.launches = 0
    _DeviceSteps.fast = request.param
    yield
    assert _DeviceSteps.launches > 0, "the test took no step on the device"


from test_oracle_reference import (  # noqa: E402,F401  (collected here, run on the device)
    test_advected_shear_3d, test_d3q19_mrt_observed_lag, test_mass_conservation_d3q19,
    test_passive_scalar_advection_diffusion, test_reaction_kinetics_batch, test_segregated_reactive_mixing,
    test_shear_wave_decay, test_spatial_convergence_2d, test_spatial_convergence_3d, test_taylor_green_2d_total_energy,
    test_taylor_green_3d_velocity, test_temporal_convergence, test_tke_angled_shear_3d, test_tke_shear_wave,
    test_tke_taylor_green_2d)
