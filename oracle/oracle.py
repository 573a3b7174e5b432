"""THE ORACLE (test infrastructure only): ctypes view of oracle/liblbm_oracle.so.

The library is a CPU restatement of LB-Cube's step_cpu / step_scalar_cpu /
step_reaction_AB loops and statistics (include/lattice_core.hpp:146-474,
include/lattice_io.hpp:182-338); see oracle/lbm_oracle.cpp for the per-function
citations and how each part is pinned.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.

Arrays are numpy arrays in the SoA layout [Q][Z][Y][X] (x fastest), exactly the
layout of LatticeMemory (include/lattice_memory.hpp:67-272).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")

LATTICES = {"D2Q9": 0, "D3Q19": 1, "D3Q27": 2, "D2Q5": 3, "D3Q7": 4}
COLLISIONS = {"BGK": 0, "TRT": 1, "MRT": 2, "RLBM": 3}
# "f80": long double (x87 extended precision), the yardstick for comparing
# FP64 and FP32 evaluation orders; steps and runs accept it.
DTYPES = {"f32": 0, "f64": 1, "f16": 2, "bf16": 3, "f80": 4}
NP_STORAGE = {"f32": np.float32, "f64": np.float64, "f16": np.uint16, "bf16": np.uint16, "f80": np.longdouble}
Q_OF = {"D2Q9": 9, "D3Q19": 19, "D3Q27": 27, "D2Q5": 5, "D3Q7": 7}
D_OF = {"D2Q9": 2, "D3Q19": 3, "D3Q27": 3, "D2Q5": 2, "D3Q7": 3}

_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (g++ only, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        vp, ll, d, i = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_double, ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        sig = {
            "orc_fluid_step": [i, i, i, vp, vp, ll, ll, ll, d, i],
            "orc_fluid_run": [i, i, i, vp, vp, ll, ll, ll, d, i, i],
            "orc_scalar_step": [i, i, i, vp, vp, vp, ll, ll, ll, d, d, i, d],
            "orc_reaction_step": [i, i, i, vp, vp, vp, vp, vp, ll, ll, ll, d, d, i, d, d],
            "orc_reactive_run": [i, i, i, i, vp, vp, vp, vp, vp, vp, ll, ll, ll, d, d, d, i, i, d, d],
            "orc_collide_cell": [i, i, i, vp, d],
            "orc_macro_feq": [i, vp, vp, vp],
            "orc_mrt_forward": [i, vp, vp],
            "orc_mrt_inverse": [i, vp, vp],
            "orc_reaction_source": [i, vp, vp, d, vp, ll],
            "orc_reactive_stats": [i, i, vp, vp, ll, ll, ll, d, i, dp, d, d],
            "orc_fluid_diagnostics": [i, i, vp, ll, ll, ll, i, dp],
            "orc_macro_fields": [i, i, vp, ll, ll, ll, vp, vp],
            "orc_scalar_fields": [i, i, vp, ll, ll, ll, vp, d],
            "orc_max_speed": [i, i, vp, ll, ll, ll, dp],
            "orc_feq_field": [i, i, vp, vp, vp, ll],
            "orc_scalar_feq_field": [i, i, vp, vp, vp, ll, d],
            "orc_f16_roundtrip": [vp, vp, vp, ll, i],
            "orc_reactive_gradient": [i, i, vp, vp, vp, ll, ll, d, d, d, i, dp, vp, vp, vp, dp],
            "orc_reaction_grad": [i, vp, vp, d, vp, vp, ll],
            "orc_phi_slope": [i, vp, vp, vp, ll],
            "orc_reactive_adjoint_step": [i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, ll, ll, d, d, d, dp],
            "orc_reactive_gradient_3d": [i, i, i, vp, vp, vp, ll, ll, ll, d, d, d, i, dp, vp, vp, vp, dp],
            "orc_reactive_adjoint_step_3d": [i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, ll, ll, ll, d, d, d, dp],
            "orc_max_threads": [],
        }
        for name, args in sig.items():
            fn = getattr(_lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ValueError(f"oracle: unsupported combination in {what}")


def _dims(lattice, shape):
    nx, ny = shape[-1], shape[-2]
    nz = shape[-3] if D_OF[lattice] == 3 else 1
    return nx, ny, nz


def max_threads() -> int:
    return lib().orc_max_threads()


def fluid_step(lattice, collision, dtype, cur, omega, threads=1):
    """One pull-stream/collide step (lattice_core.hpp:146-216); returns next."""
    nxt = np.empty_like(cur)
    nx, ny, nz = _dims(lattice, cur.shape)
    _check(lib().orc_fluid_step(LATTICES[lattice], COLLISIONS[collision], DTYPES[dtype], _p(cur), _p(nxt),
                                nx, ny, nz, float(omega), threads), "fluid_step")
    return nxt


def fluid_run(lattice, collision, dtype, state, omega, steps, threads=1):
    """`steps` fluid steps; returns the final populations (input untouched)."""
    buf = np.array(state, copy=True)
    tmp = np.empty_like(buf)
    nx, ny, nz = _dims(lattice, buf.shape)
    _check(lib().orc_fluid_run(LATTICES[lattice], COLLISIONS[collision], DTYPES[dtype], _p(buf), _p(tmp),
                               nx, ny, nz, float(omega), int(steps), threads), "fluid_run")
    return buf


def scalar_step(fluid_lattice, scalar_lattice, dtype, cur_s, cur_f, omega_c, source=0.0, threads=1, ref=0.0):
    """step_scalar_cpu (lattice_core.hpp:238-337); returns the next scalar state.
    ref: the field's reference concentration (f16/bf16 storage holds g - w ref)."""
    nxt = np.empty_like(cur_s)
    nx, ny, nz = _dims(fluid_lattice, cur_s.shape)
    _check(lib().orc_scalar_step(LATTICES[fluid_lattice], LATTICES[scalar_lattice], DTYPES[dtype], _p(cur_s),
                                 _p(nxt), _p(cur_f), nx, ny, nz, float(omega_c), float(source), threads,
                                 float(ref)),
           "scalar_step")
    return nxt


def reaction_step(fluid_lattice, scalar_lattice, dtype, cur_a, cur_b, cur_f, omega_c, k_react, threads=1, ref_a=0.0,
                  ref_b=0.0):
    """step_reaction_AB (lattice_core.hpp:357-474); returns (next_a, next_b)."""
    na, nb = np.empty_like(cur_a), np.empty_like(cur_b)
    nx, ny, nz = _dims(fluid_lattice, cur_a.shape)
    _check(lib().orc_reaction_step(LATTICES[fluid_lattice], LATTICES[scalar_lattice], DTYPES[dtype], _p(cur_a),
                                   _p(na), _p(cur_b), _p(nb), _p(cur_f), nx, ny, nz, float(omega_c),
                                   float(k_react), threads, float(ref_a), float(ref_b)), "reaction_step")
    return na, nb


def reactive_run(fluid_lattice, collision, scalar_lattice, dtype, f, a, b, omega, omega_c, k_react, steps,
                 threads=1, ref_a=0.0, ref_b=0.0):
    """Fluid step then reaction step, `steps` times (src/simulate_reaction.cpp:548-554)."""
    f, a, b = (np.array(x, copy=True) for x in (f, a, b))
    ft, at, bt = np.empty_like(f), np.empty_like(a), np.empty_like(b)
    nx, ny, nz = _dims(fluid_lattice, f.shape)
    _check(lib().orc_reactive_run(LATTICES[fluid_lattice], COLLISIONS[collision], LATTICES[scalar_lattice],
                                  DTYPES[dtype], _p(f), _p(ft), _p(a), _p(at), _p(b), _p(bt), nx, ny, nz,
                                  float(omega), float(omega_c), float(k_react), int(steps), threads, float(ref_a),
                                  float(ref_b)), "reactive_run")
    return f, a, b


def collide_cell(lattice, collision, f, omega):
    """Collide one cell in place-free form; f is float32 or float64 of length Q."""
    out = np.array(f, copy=True)
    dt = 1 if out.dtype == np.float64 else 0
    _check(lib().orc_collide_cell(LATTICES[lattice], COLLISIONS[collision], dt, _p(out), float(omega)),
           "collide_cell")
    return out


def macro_feq(lattice, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    ru = np.zeros(1 + D_OF[lattice])
    eq = np.zeros(Q_OF[lattice])
    _check(lib().orc_macro_feq(LATTICES[lattice], _p(f), _p(ru), _p(eq)), "macro_feq")
    return ru[0], ru[1:], eq


def mrt_forward(lattice, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    m = np.zeros_like(f)
    _check(lib().orc_mrt_forward(LATTICES[lattice], _p(f), _p(m)), "mrt_forward")
    return m


def mrt_inverse(lattice, m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    f = np.zeros_like(m)
    _check(lib().orc_mrt_inverse(LATTICES[lattice], _p(m), _p(f)), "mrt_inverse")
    return f


def reaction_source(dtype, a, b, k):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros_like(a)
    _check(lib().orc_reaction_source(DTYPES[dtype], _p(a), _p(b), float(k), _p(out), a.size), "reaction_source")
    return out


STAT_NAMES = ("mean_A", "mean_B", "var_A", "var_B", "covariance", "segregation_intensity",
              "true_reaction_rate", "mixed_reaction_rate")


def reactive_stats(scalar_lattice, dtype, a, b, k_react, wide=False, ref_a=0.0, ref_b=0.0):
    """compute_reactive_stats (lattice_io.hpp:256-338); wide=True sums in long double."""
    out = (ctypes.c_double * 8)()
    nx, ny, nz = _dims(scalar_lattice, a.shape)
    _check(lib().orc_reactive_stats(LATTICES[scalar_lattice], DTYPES[dtype], _p(a), _p(b), nx, ny, nz,
                                    float(k_react), int(wide), out, float(ref_a),
                                    float(ref_b)), "reactive_stats")
    return np.array(out[:])


def fluid_diagnostics(lattice, dtype, f, wide=False):
    """log_diagnostics sums (lattice_io.hpp:182-234): mass, kinetic energy, max|u|."""
    out = (ctypes.c_double * 3)()
    nx, ny, nz = _dims(lattice, f.shape)
    _check(lib().orc_fluid_diagnostics(LATTICES[lattice], DTYPES[dtype], _p(f), nx, ny, nz, int(wide), out),
           "fluid_diagnostics")
    return np.array(out[:])


def reactive_adjoint_step(collision, dtype, f, a, b, nuf1, nua1, nub1, omega, omega_c, k, lattice="D2Q9"):
    """One step of the oracle's discrete adjoint (D2Q9 + 2 D2Q5, or D3Q19 /
    D3Q27 + 2 D3Q7) from arbitrary adjoint inputs at the collision cell;
    returns (nuf0, nua0, nub0, djdw)."""
    arrs = [np.ascontiguousarray(x, dtype=NP_STORAGE[dtype]) for x in (f, a, b, nuf1, nua1, nub1)]
    outs = [np.empty_like(arrs[0]), np.empty_like(arrs[1]), np.empty_like(arrs[2])]
    nx, ny, nz = _dims(lattice, arrs[0].shape)
    dw = (ctypes.c_double * 1)()
    if lattice == "D2Q9":
        _check(lib().orc_reactive_adjoint_step(COLLISIONS[collision], DTYPES[dtype], *(_p(x) for x in arrs),
                                               *(_p(x) for x in outs), nx, ny, float(omega), float(omega_c),
                                               float(k), dw), "reactive_adjoint_step")
    else:
        _check(lib().orc_reactive_adjoint_step_3d(LATTICES[lattice], COLLISIONS[collision], DTYPES[dtype],
                                                  *(_p(x) for x in arrs), *(_p(x) for x in outs), nx, ny, nz,
                                                  float(omega), float(omega_c), float(k), dw),
               "reactive_adjoint_step_3d")
    return outs[0], outs[1], outs[2], dw[0]


def macro_fields(lattice, dtype, f):
    """Per-cell (rho, u) of the VTK writers (lattice_io.hpp:386-451), as float64."""
    nx, ny, nz = _dims(lattice, f.shape)
    rho = np.empty(f.shape[1:], dtype=np.float64)
    u = np.empty((D_OF[lattice],) + tuple(f.shape[1:]), dtype=np.float64)
    _check(lib().orc_macro_fields(LATTICES[lattice], DTYPES[dtype], _p(f), nx, ny, nz, _p(rho), _p(u)),
           "macro_fields")
    return rho, u


def scalar_fields(scalar_lattice, dtype, g, ref=0.0):
    """Per-cell concentration (compute_concentration), as float64."""
    nx, ny, nz = _dims(scalar_lattice, g.shape)
    conc = np.empty(g.shape[1:], dtype=np.float64)
    _check(lib().orc_scalar_fields(LATTICES[scalar_lattice], DTYPES[dtype], _p(g), nx, ny, nz, _p(conc),
                                   float(ref)),
           "scalar_fields")
    return conc


def max_speed(lattice, dtype, f):
    """compute_max_velocity (src/turbulent_tgv_3d.cpp:149-173): max |u|, inf if any speed is not finite."""
    out = (ctypes.c_double * 1)()
    nx, ny, nz = _dims(lattice, f.shape)
    _check(lib().orc_max_speed(LATTICES[lattice], DTYPES[dtype], _p(f), nx, ny, nz, out), "max_speed")
    return out[0]


def feq_field(lattice, dtype, rho, u):
    """Populations at equilibrium for density rho[...] and velocity u[D, ...]."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty((Q_OF[lattice],) + rho.shape, dtype=NP_STORAGE[dtype])
    _check(lib().orc_feq_field(LATTICES[lattice], DTYPES[dtype], _p(rho), _p(u), _p(out), rho.size), "feq_field")
    return out


def scalar_feq_field(scalar_lattice, dtype, conc, u, ref=0.0):
    conc = np.ascontiguousarray(conc, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty((Q_OF[scalar_lattice],) + conc.shape, dtype=NP_STORAGE[dtype])
    _check(lib().orc_scalar_feq_field(LATTICES[scalar_lattice], DTYPES[dtype], _p(conc), _p(u), _p(out),
                                      conc.size, float(ref)), "scalar_feq_field")
    return out


def half_roundtrip(x, bf16=False):
    x = np.ascontiguousarray(x, dtype=np.float32)
    h = np.empty(x.shape, dtype=np.uint16)
    back = np.empty_like(x)
    lib().orc_f16_roundtrip(_p(x), _p(h), _p(back), x.size, int(bf16))
    return h, back


def decode(lattice, dtype, pops, ref=0.0):
    """Populations as float64 from a storage array: adds back w_i for f16/bf16
    fluid storage, w_i * ref for f16/bf16 scalar storage (ref: the field's
    reference concentration; the product formed in float32 as the kernels do)."""
    if dtype == "f80":
        return pops
    if dtype in ("f32", "f64"):
        return pops.astype(np.float64)
    if dtype == "f16":
        dev = pops.view(np.float16).astype(np.float32)
    else:
        dev = (pops.astype(np.uint32) << 16).view(np.float32)
    w = np.asarray(WEIGHTS[lattice], dtype=np.float32).reshape((-1,) + (1,) * (pops.ndim - 1))
    if lattice in ("D2Q5", "D3Q7"):
        w = (w * np.float32(ref)).astype(np.float32)
    return (w + dev).astype(np.float64)


WEIGHTS = {
    "D2Q9": [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4,
    "D3Q19": [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12,
    "D3Q27": [8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8,
    "D2Q5": [1 / 3] + [1 / 6] * 4,
    "D3Q7": [1 / 4] + [1 / 8] * 6,
}


def reactive_gradient(collision, dtype, f0, a0, b0, omega, omega_c, k, steps, lattice="D2Q9"):
    """J = <A>(0) - <A>(T) of the reactive loop (D2Q9 + 2 D2Q5, or D3Q19 /
    D3Q27 + 2 D3Q7) and its discrete adjoint.

    Returns (J, dJ/df0, dJ/da0, dJ/db0, dJ/domega_c).  Project-written CPU adjoint
    (test infrastructure): the reference has no adjoint."""
    f0, a0, b0 = (np.ascontiguousarray(x) for x in (f0, a0, b0))
    gf, ga, gb = np.empty_like(f0), np.empty_like(a0), np.empty_like(b0)
    J, dw = ctypes.c_double(), ctypes.c_double()
    nx, ny, nz = _dims(lattice, f0.shape)
    if lattice == "D2Q9":
        _check(lib().orc_reactive_gradient(COLLISIONS[collision], DTYPES[dtype], _p(f0), _p(a0), _p(b0), nx, ny,
                                           float(omega), float(omega_c), float(k), int(steps), ctypes.byref(J),
                                           _p(gf), _p(ga), _p(gb), ctypes.byref(dw)), "reactive_gradient")
    else:
        _check(lib().orc_reactive_gradient_3d(LATTICES[lattice], COLLISIONS[collision], DTYPES[dtype], _p(f0), _p(a0),
                                              _p(b0), nx, ny, nz, float(omega), float(omega_c), float(k),
                                              int(steps), ctypes.byref(J), _p(gf), _p(ga), _p(gb),
                                              ctypes.byref(dw)), "reactive_gradient_3d")
    return J.value, gf, ga, gb, dw.value


def reaction_grad(a, b, k, dtype="f64"):
    """dS/dA, dS/dB of the exact one-step A + B -> C increment, evaluated in
    `dtype` ("f64" or "f32"; inputs rounded to it), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    da, db = np.empty_like(a), np.empty_like(b)
    _check(lib().orc_reaction_grad(DTYPES[dtype], _p(a), _p(b), float(k), _p(da), _p(db), a.size), "reaction_grad")
    return da, db


def phi_slope(z, dtype="f64"):
    """phi(z) = expm1(z)/z and phi'(z) as the adjoint's reaction Jacobian forms them in `dtype`."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    phi, dphi = np.empty_like(z), np.empty_like(z)
    _check(lib().orc_phi_slope(DTYPES[dtype], _p(z), _p(phi), _p(dphi), z.size), "phi_slope")
    return phi, dphi
