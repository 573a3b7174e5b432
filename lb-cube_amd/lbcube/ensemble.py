"""Batched ensembles of 2D reactive-mixing instances (lbc_ensemble_* of the C ABI).

An Ensemble owns two time levels of three packed buffers (fluid, A, B) in
torch CUDA memory; instance i holds Q population planes in each, padded to a
multiple of 64 cells apart (`fluid(i)` and `species(i)` are views of shape
(Q, ny_i, nx_i)).  One launch
advances every instance; `run(..., graph=True)` replays a captured hipGraph.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import STORAGE_NAMES, _stream, check, lib

_COLL = {"BGK": 0, "TRT": 1}
_STORE = {"f32": 0, "f64": 1, "f16": 2, "bf16": 3}


class Ensemble:
    def __init__(self, collision, storage, sizes, omega, omega_c, k_react, device="cuda", fast=False, ref_a=None,
                 ref_b=None):
        self.collision, self.storage = collision, storage
        self.sizes = [(int(nx), int(ny)) for nx, ny in sizes]
        n = len(self.sizes)
        nx = (ctypes.c_uint32 * n)(*[s[0] for s in self.sizes])
        ny = (ctypes.c_uint32 * n)(*[s[1] for s in self.sizes])
        om = (ctypes.c_double * n)(*map(float, omega))
        oc = (ctypes.c_double * n)(*map(float, omega_c))
        kk = (ctypes.c_double * n)(*map(float, k_react))
        handle = ctypes.c_void_p()
        check(lib().lbc_ensemble_create(_COLL[collision], _STORE[storage], n, nx, ny, om, oc, kk,
                                        ctypes.byref(handle)), "lbc_ensemble_create")
        self.handle = handle
        fe, se = ctypes.c_uint64(), ctypes.c_uint64()
        fo, so, ps = (ctypes.c_uint64 * n)(), (ctypes.c_uint64 * n)(), (ctypes.c_uint64 * n)()
        check(lib().lbc_ensemble_layout(handle, ctypes.byref(fe), ctypes.byref(se), fo, so, ps),
              "lbc_ensemble_layout")
        self.f_off, self.s_off, self.plane = list(fo), list(so), list(ps)
        dt = STORAGE_NAMES[storage]
        self.F = [torch.zeros(fe.value, dtype=dt, device=device) for _ in range(2)]
        self.A = [torch.zeros(se.value, dtype=dt, device=device) for _ in range(2)]
        self.B = [torch.zeros(se.value, dtype=dt, device=device) for _ in range(2)]
        self.level = 0
        self.cells = sum(x * y for x, y in self.sizes)
        self.set_fast(fast)
        self.set_references(ref_a, ref_b)

    def set_fast(self, fast):
        """Deviation-form arithmetic (LBC_FLAG_FAST) for every instance."""
        check(lib().lbc_ensemble_set_flags(self.handle, 1 if fast else 0), "lbc_ensemble_set_flags")
        self.fast = bool(fast)

    def set_references(self, ref_a=None, ref_b=None):
        """Per-instance reference concentrations of A and B (sequences of the
        instance count, or None for 0): f16/bf16 storage holds g - w ref."""
        n = len(self.sizes)
        self.ref_a = [0.0] * n if ref_a is None else [float(x) for x in ref_a]
        self.ref_b = [0.0] * n if ref_b is None else [float(x) for x in ref_b]
        ra, rb = (ctypes.c_double * n)(*self.ref_a), (ctypes.c_double * n)(*self.ref_b)
        check(lib().lbc_ensemble_set_references(self.handle, ra, rb), "lbc_ensemble_set_references")

    def _planes(self, buf, off, q, i):
        nx, ny = self.sizes[i]
        ps = self.plane[i]
        return buf[off:off + q * ps].view(q, ps)[:, :nx * ny].view(q, ny, nx)

    def fluid(self, i, level=None):
        """Instance i's fluid populations, a (9, ny, nx) view (planes padded apart)."""
        lv = self.level if level is None else level
        return self._planes(self.F[lv], self.f_off[i], 9, i)

    def species(self, i, level=None):
        lv = self.level if level is None else level
        return (self._planes(self.A[lv], self.s_off[i], 5, i), self._planes(self.B[lv], self.s_off[i], 5, i))

    def set_block(self, threads):
        """Work-items per block: 128, 256, 512 or 1024."""
        check(lib().lbc_ensemble_set_block(self.handle, threads), "lbc_ensemble_set_block")

    def _ptrs(self):
        c, n = self.level, 1 - self.level
        return (self.F[c].data_ptr(), self.F[n].data_ptr(), self.A[c].data_ptr(), self.A[n].data_ptr(),
                self.B[c].data_ptr(), self.B[n].data_ptr())

    def run(self, steps, graph=False, stream=None):
        f0, f1, a0, a1, b0, b1 = self._ptrs()
        check(lib().lbc_ensemble_run(self.handle, f0, f1, a0, a1, b0, b1, int(steps), int(graph), _stream(stream)),
              "lbc_ensemble_run")
        self.level = (self.level + steps) % 2

    def autotune(self, trial_steps=3, stream=None):
        f0, f1, a0, a1, b0, b1 = self._ptrs()
        chosen = ctypes.c_uint32()
        ms = (ctypes.c_float * 4)()
        check(lib().lbc_ensemble_autotune(self.handle, f0, f1, a0, a1, b0, b1, int(trial_steps),
                                          ctypes.byref(chosen), ms, _stream(stream)), "lbc_ensemble_autotune")
        # every candidate advances the state by 2 + 2 * trial_steps steps: an even count
        return chosen.value, {t: ms[i] for i, t in enumerate((128, 256, 512, 1024))}

    def close(self):
        if self.handle:
            lib().lbc_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def reactive_mixing_ensemble(count=512, seed=0x5EED1234, nmin=64, nmax=1024, re=(100.0, 2000.0), da=(0.1, 100.0),
                             u0=0.04):
    """The batched-ensemble configuration: edge N log-uniform in [nmin, nmax],
    Re uniform, Da log-uniform; nu = u0 N / Re, tau = 3 nu + 1/2, Sc = 1
    (tau_s = tau), k = Da u0 / N.  Returns per-instance (N, omega, omega_c, k)."""
    g = np.random.Generator(np.random.PCG64(seed))
    n = np.rint(np.exp(g.uniform(math.log(nmin), math.log(nmax), count))).astype(int)
    r = g.uniform(re[0], re[1], count)
    d = np.exp(g.uniform(math.log(da[0]), math.log(da[1]), count))
    nu = u0 * n / r
    tau = 3.0 * nu + 0.5
    k = d * u0 / n
    return n, 1.0 / tau, 1.0 / tau, k
