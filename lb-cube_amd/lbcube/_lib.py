"""Loader for the gfx950 kernel library (liblbcube_hip.so, C ABI in include/lbcube.h).

There is no fallback: if the library is missing or cannot be loaded, every
call raises.  Build it with ``make -C lb-cube_amd`` (or
``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(PKG_DIR), "liblbcube_hip.so")

_lib = None

c_int, c_u32, c_u64, c_double, c_vp = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)

# name -> (argtypes, restype); mirrors include/lbcube.h
SIGNATURES = {
    "lbc_fluid_step": ([c_int, c_int, c_int, c_vp, c_vp, c_u64, c_u64, c_u64, c_double, c_u32, c_u32, c_u32, c_vp],
                       c_int),
    "lbc_fluid_step_ex": ([c_int, c_int, c_int, c_u32, c_vp, c_vp, c_u64, c_u64, c_u64, c_double, c_u32, c_u32,
                           c_u32, c_vp], c_int),
    "lbc_fluid_step_pitched": ([c_int, c_int, c_int, c_u32, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_double, c_u32,
                                c_u32, c_u32, c_vp], c_int),
    "lbc_pull_stream": ([c_int, c_int, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_plane_copy": ([c_int, c_int, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_fluid_step_strided": ([c_int, c_int, c_int, c_u32, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_u64, c_double,
                                c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_pull_stream_strided": ([c_int, c_int, c_int, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_u64, c_u32, c_u32,
                                 c_u32, c_vp], c_int),
    "lbc_relayout": ([c_int, c_int, c_vp, c_u64, c_u64, c_vp, c_u64, c_u64, c_u64, c_u64, c_u64, c_vp], c_int),
    "lbc_relayout_tiled": ([c_int, c_int, c_vp, c_u64, c_vp, c_u64, c_u64, c_u64, c_u64, c_vp], c_int),
    "lbc_fluid_step_tiled": ([c_int, c_int, c_int, c_u32, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_double, c_u32,
                              c_u32, c_u32, c_vp], c_int),
    "lbc_reactive_fused_step_tiled": ([c_int, c_int, c_int, c_int, c_u32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64,
                                       c_u64, c_u64, c_u64, c_double, c_double, c_double, c_double, c_double, c_u32,
                                       c_u32, c_u32, c_vp], c_int),
    "lbc_reactive_pull_stream": ([c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64,
                                  c_u64, c_u64, c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_reactive_stats_tiled": ([c_int, c_int, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_double, c_double, c_double,
                                  _dp, c_vp], c_int),
    "lbc_scalar_step": ([c_int, c_int, c_int, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64, c_double, c_double, c_double,
                         c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_reaction_ab_step": ([c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64, c_double,
                              c_double, c_double, c_double, c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_scalar_step_tiled": ([c_int, c_int, c_int, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64, c_double, c_double,
                               c_double, c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_reaction_ab_step_tiled": ([c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64, c_u64,
                                    c_double, c_double, c_double, c_double, c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_reactive_fused_step": ([c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64,
                                 c_double, c_double, c_double, c_double, c_double, c_u32, c_u32, c_u32, c_vp], c_int),
    "lbc_reactive_fused_step_ex": ([c_int, c_int, c_int, c_int, c_u32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64,
                                    c_u64, c_double, c_double, c_double, c_double, c_double, c_u32, c_u32, c_u32,
                                    c_vp], c_int),
    "lbc_fluid_diagnostics": ([c_int, c_int, c_vp, c_u64, c_u64, c_u64, _dp, c_vp], c_int),
    "lbc_reactive_stats": ([c_int, c_int, c_vp, c_vp, c_u64, c_u64, c_u64, c_double, c_double, c_double, _dp, c_vp],
                           c_int),
    "lbc_reactive_stats_noalloc": ([c_int, c_int, c_vp, c_vp, c_u64, c_u64, c_u64, c_double, c_double, c_double, _dp,
                                    c_vp], c_int),
    "lbc_max_speed": ([c_int, c_int, c_vp, c_u64, c_u64, c_u64, _dp, c_vp], c_int),
    "lbc_macro_fields": ([c_int, c_int, c_vp, c_u64, c_u64, c_u64, c_vp, c_vp, c_vp], c_int),
    "lbc_scalar_fields": ([c_int, c_int, c_vp, c_u64, c_u64, c_u64, c_double, c_vp, c_vp], c_int),
    "lbc_reactive_gradient_workspace": ([c_int, c_u64, c_u64, c_int, c_int, c_u32, c_u32,
                                         ctypes.POINTER(ctypes.c_uint64)], c_int),
    "lbc_reactive_gradient": ([c_int, c_int, c_u32, c_vp, c_vp, c_vp, c_u64, c_u64, c_int, c_int, c_double, c_double,
                               c_double, _dp, c_vp, c_vp, c_vp, _dp, c_vp, c_u64, c_u32, c_u32, c_vp], c_int),
    "lbc_reactive_gradient_workspace_3d": ([c_int, c_int, c_u64, c_u64, c_u64, c_int, c_int, c_u32, c_u32, c_u32,
                                            ctypes.POINTER(c_u64)], c_int),
    "lbc_reactive_gradient_3d": ([c_int, c_int, c_int, c_u32, c_vp, c_vp, c_vp, c_u64, c_u64, c_u64, c_int, c_int,
                                  c_double, c_double, c_double, _dp, c_vp, c_vp, c_vp, _dp, c_vp, c_u64, c_u32, c_u32,
                                  c_u32, c_vp], c_int),
    "lbc_ensemble_create": ([c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)], c_int),
    "lbc_ensemble_layout": ([c_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), c_vp, c_vp,
                             c_vp], c_int),
    "lbc_ensemble_set_block": ([c_vp, c_u32], c_int),
    "lbc_ensemble_set_flags": ([c_vp, c_u32], c_int),
    "lbc_ensemble_set_references": ([c_vp, c_vp, c_vp], c_int),
    "lbc_ensemble_step": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp], c_int),
    "lbc_ensemble_run": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp], c_int),
    "lbc_ensemble_autotune": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, _u32p, ctypes.POINTER(ctypes.c_float),
                               c_vp], c_int),
    "lbc_ensemble_destroy": ([c_vp], c_int),
    "lbc_stream_copy": ([c_vp, c_vp, c_u64, c_vp], c_int),
    "lbc_default_block": ([c_int, c_u64, c_u64, c_u64, _u32p], None),
    "lbc_status_string": ([c_int], ctypes.c_char_p),
    "lbc_abi_version": ([], c_int),
}


class LbcError(RuntimeError):
    """A C entry point returned a non-zero hipError_t."""

    def __init__(self, status: int, what: str):
        self.status = status
        msg = lib().lbc_status_string(status)
        super().__init__(f"{what} failed: hipError {status} ({msg.decode() if msg else '?'})")


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C lb-cube_amd` first")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        raise LbcError(status, what)
