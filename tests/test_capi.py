"""The C ABI library loads without a GPU, exports every entry point that
include/lbcube.h declares, and validates arguments in the reference's order
(src/lattice_cuda.cu:488-503) before touching the device."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

import lbcube
from lbcube import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_INVALID_VALUE, HIP_INVALID_CONFIGURATION, HIP_INVALID_DEVICE_POINTER = 1, 9, 17
FAKE_A, FAKE_B = 0x1000, 0x2000  # never dereferenced: validation fails first


def declared_entry_points():
    with open(os.path.join(ROOT, "include", "lbcube.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(lbc_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_entry_points()
    assert len(names) >= 9
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} missing from the Python binding"


@pytest.mark.parametrize("std", ["c99", "c11"])
def test_header_compiles_as_c(tmp_path, std):
    """include/lbcube.h is a C header: a C translation unit that includes it
    and calls an entry point compiles with -pedantic -Werror."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "uses_lbcube.c"
    src.write_text('#include "lbcube.h"\nint main(void) { return lbc_abi_version() == 2 ? 0 : 1; }\n')
    r = subprocess.run([cc, f"-std={std}", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_abi_version():
    assert lbcube.lib().lbc_abi_version() == 2


def fluid(ptr_a, ptr_b, nx=8, ny=8, nz=8, block=(16, 8, 1), lattice=1, coll=0, storage=0):
    return lbcube.lib().lbc_fluid_step(lattice, coll, storage, ptr_a, ptr_b, nx, ny, nz, 1.0, *block, None)


def test_validation_order_fluid():
    assert fluid(None, FAKE_B) == HIP_INVALID_DEVICE_POINTER
    assert fluid(FAKE_A, None, block=(0, 8, 1)) == HIP_INVALID_DEVICE_POINTER  # pointers first
    assert fluid(FAKE_A, FAKE_B, block=(0, 8, 1)) == HIP_INVALID_CONFIGURATION
    assert fluid(FAKE_A, FAKE_B, block=(16, 8, 0), nz=0) == HIP_INVALID_CONFIGURATION  # block before extents
    assert fluid(FAKE_A, FAKE_B, ny=0) == HIP_INVALID_VALUE
    assert fluid(FAKE_A, FAKE_B, block=(64, 32, 1)) == HIP_INVALID_CONFIGURATION  # > 1024 work-items
    assert fluid(FAKE_A, FAKE_A) == HIP_INVALID_VALUE  # ping-pong buffers must not alias
    # at most 2^32 - 2^20 cells, each extent below 2^31, at most 2^31 - 1 blocks
    assert fluid(FAKE_A, FAKE_B, nx=1 << 12, ny=1 << 10, nz=1 << 10) == HIP_INVALID_VALUE  # 2^32 cells
    assert fluid(FAKE_A, FAKE_B, nx=(1 << 12) - 1, ny=1 << 10, nz=(1 << 10) + 1) == HIP_INVALID_VALUE
    assert fluid(FAKE_A, FAKE_B, nx=1 << 31, ny=1, nz=1) == HIP_INVALID_VALUE
    assert fluid(FAKE_A, FAKE_B, nx=1 << 12, ny=1 << 10, nz=(1 << 10) - 1, block=(1, 1, 1)) == \
        HIP_INVALID_CONFIGURATION  # 2^32 - 2^22 cells in 1-cell blocks
    # extents whose 64-bit product wraps to 0 (2^31 * 2^17 * 2^16) are still rejected
    assert fluid(FAKE_A, FAKE_B, nx=1 << 31, ny=1 << 17, nz=1 << 16, block=(2, 4, 2)) == HIP_INVALID_VALUE


def test_gradient_3d_validation():
    """The 3D adjoint: D3Q19 / D3Q27 fluids only, BGK, TRT, MRT or RLBM, whole waves."""
    L = lbcube.lib()
    out = ctypes.c_uint64()
    assert L.lbc_reactive_gradient_workspace_3d(1, 0, 16, 16, 16, 8, 4, 256, 1, 1, ctypes.byref(out)) == 0
    small = out.value
    assert L.lbc_reactive_gradient_workspace_3d(2, 0, 16, 16, 16, 8, 4, 256, 1, 1, ctypes.byref(out)) == 0
    assert out.value > small  # D3Q27 states hold 41 planes against D3Q19's 33
    assert L.lbc_reactive_gradient_workspace_3d(1, 0, 16, 16, 16, 8, 4, 100, 1, 1, ctypes.byref(out)) == \
        HIP_INVALID_CONFIGURATION
    for fluid in (0, 3, 4):
        assert L.lbc_reactive_gradient_workspace_3d(fluid, 0, 16, 16, 16, 8, 4, 256, 1, 1, ctypes.byref(out)) == \
            HIP_INVALID_VALUE
    loss, dw = ctypes.c_double(), ctypes.c_double()

    def grad(fluid, coll, block=(256, 1, 1)):
        return L.lbc_reactive_gradient_3d(fluid, coll, 0, 0, FAKE_A, FAKE_B, 0x3000, 16, 16, 16, 8, 4, 1.5, 1.2, 0.1,
                                          ctypes.byref(loss), 0x4000, 0x5000, 0x6000, ctypes.byref(dw), 0x7000, 1 << 40,
                                          *block, None)

    assert grad(1, 4) == HIP_INVALID_VALUE  # unknown collisions
    assert grad(2, 7) == HIP_INVALID_VALUE
    assert grad(1, -1) == HIP_INVALID_VALUE
    assert grad(0, 0) == HIP_INVALID_VALUE  # a 2D fluid
    assert grad(1, 0, block=(96, 1, 1)) == HIP_INVALID_CONFIGURATION


def test_gradient_needs_whole_waves():
    out = ctypes.c_uint64()
    lib = lbcube.lib()
    assert lib.lbc_reactive_gradient_workspace(0, 64, 64, 8, 4, 256, 1, ctypes.byref(out)) == 0 and out.value > 0
    assert lib.lbc_reactive_gradient_workspace(0, 64, 64, 8, 4, 100, 1, ctypes.byref(out)) == HIP_INVALID_CONFIGURATION
    # the adjoint keeps at most 2^29 cells per instance
    assert lib.lbc_reactive_gradient_workspace(0, 1 << 15, 1 << 14, 8, 4, 256, 1, ctypes.byref(out)) == 0
    assert lib.lbc_reactive_gradient_workspace(0, 1 << 15, (1 << 14) + 1, 8, 4, 256, 1,
                                               ctypes.byref(out)) == HIP_INVALID_VALUE


def test_two_d_lattice_ignores_z():
    # D2Q9: the zero-block check runs before z is normalised (as in the
    # reference), then z extent and block z are forced to 1, so nz = 0 passes
    # validation; the aliasing check still fires afterwards.
    assert fluid(FAKE_A, FAKE_B, nz=3, block=(16, 8, 0), lattice=0) == HIP_INVALID_CONFIGURATION
    assert fluid(FAKE_A, FAKE_A, nz=0, block=(16, 8, 4), lattice=0) == HIP_INVALID_VALUE
    assert fluid(FAKE_A, FAKE_B, nz=0, block=(16, 8, 1), lattice=0, coll=9) == HIP_INVALID_VALUE


@pytest.mark.parametrize("bad", [(5, 0, 0), (1, 7, 0), (1, 0, 9)])
def test_unknown_codes_rejected(bad):
    lattice, coll, storage = bad
    assert fluid(FAKE_A, FAKE_B, lattice=lattice, coll=coll, storage=storage) == HIP_INVALID_VALUE


def fluid_ex(flags, lattice=1, coll=2, storage=2, block=(256, 1, 1)):
    return lbcube.lib().lbc_fluid_step_ex(lattice, coll, storage, flags, FAKE_A, FAKE_B, 8, 8, 8, 1.0, *block, None)


def test_mfma_flag_validation():
    """LBC_FLAG_MFMA only with LBC_FLAG_FAST, MRT, D3Q19 / D3Q27 and 16-bit
    storage, and only with blocks of whole 64-wide wavefronts."""
    fast, mfma = 1, 4
    assert fluid_ex(mfma) == HIP_INVALID_VALUE  # without the deviation form
    assert fluid_ex(fast | mfma, coll=1) == HIP_INVALID_VALUE  # TRT
    assert fluid_ex(fast | mfma, coll=3) == HIP_INVALID_VALUE  # RLBM
    assert fluid_ex(fast | mfma, lattice=0) == HIP_INVALID_VALUE  # D2Q9
    assert fluid_ex(fast | mfma, storage=0) == HIP_INVALID_VALUE  # FP32
    assert fluid_ex(fast | mfma, storage=1) == HIP_INVALID_VALUE  # FP64
    assert fluid_ex(fast | mfma, block=(100, 1, 1)) == HIP_INVALID_CONFIGURATION
    assert fluid_ex(fast | mfma, lattice=2, storage=3, block=(8, 4, 1)) == HIP_INVALID_CONFIGURATION  # 32
    assert fluid_ex(fast | 8) == HIP_INVALID_VALUE  # unknown bit
    assert fluid_ex(fast | 2) == HIP_INVALID_VALUE  # the adjoint's flag


def test_validation_scalar_and_reaction():
    L = lbcube.lib()
    assert L.lbc_scalar_step(0, 3, 0, FAKE_A, FAKE_B, None, 8, 8, 1, 1.0, 0.0, 0.0, 16, 8, 1, None) == \
        HIP_INVALID_DEVICE_POINTER
    assert L.lbc_scalar_step(0, 3, 0, FAKE_A, FAKE_B, 0x3000, 8, 8, 1, 1.0, 0.0, 0.0, 0, 8, 1, None) == \
        HIP_INVALID_CONFIGURATION
    assert L.lbc_scalar_step(1, 4, 0, FAKE_A, FAKE_B, 0x3000, 8, 8, 0, 1.0, 0.0, 0.0, 16, 8, 1, None) == HIP_INVALID_VALUE
    assert L.lbc_scalar_step(0, 4, 0, FAKE_A, FAKE_B, 0x3000, 8, 8, 1, 1.0, 0.0, 0.0, 16, 8, 1, None) == HIP_INVALID_VALUE
    assert L.lbc_reaction_ab_step(0, 3, 0, FAKE_A, FAKE_B, 0x3000, None, 0x5000, 8, 8, 1, 1.0, 0.1, 0.0, 0.0, 16, 8, 1,
                                  None) == HIP_INVALID_DEVICE_POINTER
    assert L.lbc_reaction_ab_step(0, 3, 0, FAKE_A, FAKE_A, 0x3000, 0x4000, 0x5000, 8, 8, 1, 1.0, 0.1, 0.0, 0.0, 16, 8, 1,
                                  None) == HIP_INVALID_VALUE
    assert L.lbc_reactive_fused_step(0, 0, 3, 0, FAKE_A, FAKE_B, 0x3000, 0x4000, 0x5000, 0x6000, 8, 0, 1, 1.0, 1.0,
                                     0.1, 0.0, 0.0, 16, 8, 1, None) == HIP_INVALID_VALUE


def test_default_block():
    assert lbcube.default_block("D3Q19", 256, 256, 256) == (256, 1, 1)
    assert lbcube.default_block("D2Q9", 64, 64, 1) == (256, 1, 1)
    assert lbcube.default_block("D3Q27", 384, 384, 384) == (256, 1, 1)


def test_validation_fields_and_guard():
    L = lbcube.lib()
    out = (ctypes.c_double * 1)()
    assert L.lbc_max_speed(1, 0, None, 8, 8, 8, out, None) == HIP_INVALID_DEVICE_POINTER
    assert L.lbc_max_speed(1, 0, FAKE_A, 8, 0, 8, out, None) == HIP_INVALID_VALUE
    assert L.lbc_max_speed(4, 0, FAKE_A, 8, 8, 8, out, None) == HIP_INVALID_VALUE  # scalar lattice
    assert L.lbc_macro_fields(1, 0, FAKE_A, 8, 8, 8, None, FAKE_B, None) == HIP_INVALID_DEVICE_POINTER
    assert L.lbc_macro_fields(1, 0, FAKE_A, 0, 8, 8, FAKE_B, 0x3000, None) == HIP_INVALID_VALUE
    assert L.lbc_macro_fields(3, 0, FAKE_A, 8, 8, 8, FAKE_B, 0x3000, None) == HIP_INVALID_VALUE
    assert L.lbc_scalar_fields(3, 0, None, 8, 8, 1, 0.0, FAKE_B, None) == HIP_INVALID_DEVICE_POINTER
    assert L.lbc_scalar_fields(0, 0, FAKE_A, 8, 8, 1, 0.0, FAKE_B, None) == HIP_INVALID_VALUE  # fluid lattice


def test_validation_noalloc_stats():
    """The no-allocation statistics entry point validates like lbc_reactive_stats
    before touching the device."""
    L = lbcube.lib()
    out = (ctypes.c_double * 8)()
    for fn in (L.lbc_reactive_stats, L.lbc_reactive_stats_noalloc):
        assert fn(3, 0, None, FAKE_B, 8, 8, 1, 0.1, 0.0, 0.0, out, None) == HIP_INVALID_DEVICE_POINTER
        assert fn(3, 0, FAKE_A, FAKE_B, 8, 8, 1, 0.1, 0.0, 0.0, None, None) == HIP_INVALID_DEVICE_POINTER
        assert fn(3, 0, FAKE_A, FAKE_B, 0, 8, 1, 0.1, 0.0, 0.0, out, None) == HIP_INVALID_VALUE
        assert fn(4, 0, FAKE_A, FAKE_B, 8, 8, 0, 0.1, 0.0, 0.0, out, None) == HIP_INVALID_VALUE
        assert fn(4, 0, FAKE_A, FAKE_B, 1 << 12, 1 << 10, 1 << 10, 0.1, 0.0, 0.0, out, None) == HIP_INVALID_VALUE



def test_validation_pitched_and_pull_stream():
    """The pitched step and the pull-stream step validate in the reference's
    order, then reject a plane pitch below the cell count."""
    L = lbcube.lib()
    for fn in (lambda *a: L.lbc_fluid_step_pitched(1, 1, 0, 0, *a[:2], *a[2:5], a[5], 1.0, *a[6:], None),
               lambda *a: L.lbc_pull_stream(1, 0, *a[:2], *a[2:5], a[5], *a[6:], None),
               lambda *a: L.lbc_plane_copy(1, 0, *a[:2], *a[2:5], a[5], *a[6:], None)):
        assert fn(None, FAKE_B, 8, 8, 8, 0, 256, 1, 1) == HIP_INVALID_DEVICE_POINTER
        assert fn(FAKE_A, FAKE_B, 8, 8, 8, 0, 0, 1, 1) == HIP_INVALID_CONFIGURATION
        assert fn(FAKE_A, FAKE_B, 8, 0, 8, 0, 256, 1, 1) == HIP_INVALID_VALUE
        assert fn(FAKE_A, FAKE_A, 8, 8, 8, 0, 256, 1, 1) == HIP_INVALID_VALUE
        assert fn(FAKE_A, FAKE_B, 8, 8, 8, 511, 256, 1, 1) == HIP_INVALID_VALUE  # pitch below 512 cells
        assert fn(FAKE_A, FAKE_B, 8, 8, 8, (1 << 51) + 1, 256, 1, 1) == HIP_INVALID_VALUE
    assert L.lbc_fluid_step_pitched(1, 1, 0, 2, FAKE_A, FAKE_B, 8, 8, 8, 0, 1.0, 256, 1, 1, None) == \
        HIP_INVALID_VALUE  # unknown flag
    assert L.lbc_pull_stream(4, 0, FAKE_A, FAKE_B, 8, 8, 8, 0, 256, 1, 1, None) == HIP_INVALID_VALUE  # scalar lattice
    assert L.lbc_plane_copy(4, 0, FAKE_A, FAKE_B, 8, 8, 8, 0, 256, 1, 1, None) == HIP_INVALID_VALUE
    assert L.lbc_plane_copy(1, 7, FAKE_A, FAKE_B, 8, 8, 8, 0, 256, 1, 1, None) == HIP_INVALID_VALUE  # storage


def test_validation_strided_layouts():
    """Strided layouts on an 8^3 D3Q19 domain (64-cell slabs, 512 cells): the
    reference's planes with a pitch below the cell count, slab-spaced planes
    closer than a slab, slabs closer than Q planes, per-plane offsets of 2^32
    elements and more, and a scalar lattice are rejected before any launch;
    so are aliasing buffers and unknown storage in lbc_relayout."""
    L = lbcube.lib()
    step = lambda a, b, plane, slab, lattice=1, nz=8: L.lbc_fluid_step_strided(  # noqa: E731
        lattice, 1, 0, 0, a, b, 8, 8, nz, plane, slab, 1.0, 256, 1, 1, None)
    pull = lambda a, b, plane, slab, lattice=1, nz=8: L.lbc_pull_stream_strided(  # noqa: E731
        lattice, 0, 1, a, b, 8, 8, nz, plane, slab, 256, 1, 1, None)
    relayout_src = lambda a, b, plane, slab, lattice=1, nz=8: L.lbc_relayout(  # noqa: E731
        lattice, 0, a, plane, slab, b, 0, 0, 8, 8, nz, None)
    relayout_dst = lambda a, b, plane, slab, lattice=1, nz=8: L.lbc_relayout(  # noqa: E731
        lattice, 0, a, 0, 0, b, plane, slab, 8, 8, nz, None)
    for fn in (step, pull, relayout_src, relayout_dst):
        assert fn(None, FAKE_B, 0, 0) == HIP_INVALID_DEVICE_POINTER
        assert fn(FAKE_A, FAKE_A, 0, 0) == HIP_INVALID_VALUE
        assert fn(FAKE_A, FAKE_B, 511, 64) == HIP_INVALID_VALUE  # reference planes, pitch below 512 cells
        assert fn(FAKE_A, FAKE_B, 63, 19 * 64) == HIP_INVALID_VALUE  # planes closer than a slab
        assert fn(FAKE_A, FAKE_B, 0, 19 * 64 - 1) == HIP_INVALID_VALUE  # slabs closer than 19 planes
        assert fn(FAKE_A, FAKE_B, 100, 19 * 100 - 1) == HIP_INVALID_VALUE
        assert fn(FAKE_A, FAKE_B, (1 << 51) + 1, 1 << 60) == HIP_INVALID_VALUE
        # (nz - 1) * slab_stride + 64 must stay below 2^32 elements
        assert fn(FAKE_A, FAKE_B, 0, (0xFFFFFFFF - 64) // 7 + 1) == HIP_INVALID_VALUE
        assert fn(FAKE_A, FAKE_B, 0, 0, lattice=4) == HIP_INVALID_VALUE  # scalar lattice
        assert fn(FAKE_A, FAKE_B, 0, 1 << 33, nz=1) == HIP_INVALID_VALUE  # slab gap of 2^32 or more
    assert L.lbc_relayout(1, 7, FAKE_A, 0, 0, FAKE_B, 0, 19 * 64, 8, 8, 8, None) == HIP_INVALID_VALUE  # storage


def test_validation_tiles():
    """Tiles: rows must divide ny (in 3D, rows = ny is the slab order);
    Q times the cells must stay below 2^32; scalar lattices are no fluid;
    checks come after the reference's pointer / block / extent order."""
    L = lbcube.lib()
    fluid = lambda lattice, nx, ny, nz, rows, a=FAKE_A: L.lbc_fluid_step_tiled(  # noqa: E731
        lattice, 1, 0, 0, a, FAKE_B, nx, ny, nz, rows, 1.0, 256, 1, 1, None)
    fused = lambda fl, sl, nx, ny, nz, rows: L.lbc_reactive_fused_step_tiled(  # noqa: E731
        fl, 1, sl, 0, 0, FAKE_A, FAKE_B, 0x3000, 0x4000, 0x5000, 0x6000, nx, ny, nz, rows, 1.0, 1.0, 0.1, 0.0, 0.0,
        256, 1, 1, None)
    out = (ctypes.c_double * 8)()
    stats = lambda sl, nx, ny, nz, rows: L.lbc_reactive_stats_tiled(  # noqa: E731
        sl, 0, FAKE_A, FAKE_B, nx, ny, nz, rows, 0.1, 0.0, 0.0, out, None)
    relayout = lambda lattice, nx, ny, nz, r0, r1: L.lbc_relayout_tiled(  # noqa: E731
        lattice, 0, FAKE_A, r0, FAKE_B, r1, nx, ny, nz, None)
    assert fluid(0, 16, 12, 1, 4, a=None) == HIP_INVALID_DEVICE_POINTER
    assert fluid(0, 16, 12, 1, 5) == HIP_INVALID_VALUE  # 5 does not divide 12
    assert fluid(0, 16, 12, 1, 24) == HIP_INVALID_VALUE
    assert fluid(1, 8, 8, 8, 3) == HIP_INVALID_VALUE  # 3 does not divide ny = 8 (3D)
    assert fluid(3, 16, 12, 1, 4) == HIP_INVALID_VALUE  # a scalar lattice
    assert fluid(1, 1 << 12, 1 << 10, 1 << 7, 1 << 10) == HIP_INVALID_VALUE  # 19 x 2^29 elements
    assert fluid(1, 610, 610, 610, 610) == HIP_INVALID_VALUE  # 19 x 610^3 > 2^32: slab order
    assert fluid(1, 610, 610, 610, 61) == HIP_INVALID_VALUE  # and tiles
    assert fused(0, 3, 16, 12, 1, 5) == HIP_INVALID_VALUE
    assert fused(1, 4, 8, 8, 8, 3) == HIP_INVALID_VALUE
    assert fused(3, 3, 16, 12, 1, 4) == HIP_INVALID_VALUE  # scalar lattice as the fluid
    # 9 x 21888^2 = 4,311,760,896 elements: one step past the 2D limit
    assert fused(0, 3, 21888, 21888, 1, 4) == HIP_INVALID_VALUE
    assert relayout(0, 21888, 21888, 1, 0, 4) == HIP_INVALID_VALUE
    scalar = lambda fl, sl, nx, ny, nz, rows, a=FAKE_A: L.lbc_scalar_step_tiled(  # noqa: E731
        fl, sl, 0, a, FAKE_B, 0x3000, nx, ny, nz, rows, 1.0, 0.0, 0.0, 256, 1, 1, None)
    react = lambda fl, sl, nx, ny, nz, rows: L.lbc_reaction_ab_step_tiled(  # noqa: E731
        fl, sl, 0, FAKE_A, FAKE_B, 0x3000, 0x4000, 0x5000, nx, ny, nz, rows, 1.0, 0.1, 0.0, 0.0, 256, 1, 1, None)
    assert scalar(0, 3, 16, 12, 1, 4, a=None) == HIP_INVALID_DEVICE_POINTER
    assert scalar(0, 3, 16, 12, 1, 5) == HIP_INVALID_VALUE  # 5 does not divide 12
    assert scalar(3, 3, 16, 12, 1, 4) == HIP_INVALID_VALUE  # a scalar lattice as the fluid
    assert scalar(1, 4, 8, 8, 8, 3) == HIP_INVALID_VALUE
    assert scalar(0, 3, 21888, 21888, 1, 4) == HIP_INVALID_VALUE  # 9 x 21888^2 elements
    assert react(0, 3, 16, 12, 1, 5) == HIP_INVALID_VALUE
    assert react(3, 3, 16, 12, 1, 4) == HIP_INVALID_VALUE
    assert react(1, 4, 610, 610, 610, 610) == HIP_INVALID_VALUE  # 19 x 610^3 elements
    assert stats(3, 16, 12, 1, 5) == HIP_INVALID_VALUE
    assert stats(4, 8, 8, 8, 3) == HIP_INVALID_VALUE
    assert stats(0, 16, 12, 1, 4) == HIP_INVALID_VALUE  # a fluid lattice
    assert relayout(0, 16, 12, 1, 0, 5) == HIP_INVALID_VALUE
    assert relayout(4, 8, 8, 8, 8, 3) == HIP_INVALID_VALUE
    assert relayout(7, 8, 8, 8, 0, 8) == HIP_INVALID_VALUE  # unknown lattice
    assert L.lbc_relayout_tiled(0, 0, FAKE_A, 0, FAKE_A, 4, 16, 12, 1, None) == HIP_INVALID_VALUE  # src == dst
    assert L.lbc_relayout_tiled(0, 9, FAKE_A, 0, FAKE_B, 4, 16, 12, 1, None) == HIP_INVALID_VALUE  # storage
    pull = lambda fl, sl, nx, ny, nz, rows, st=0, a=FAKE_A: L.lbc_reactive_pull_stream(  # noqa: E731
        fl, sl, st, 1, a, FAKE_B, 0x3000, 0x4000, 0x5000, 0x6000, nx, ny, nz, rows, 256, 1, 1, None)
    assert pull(0, 3, 16, 12, 1, 4, a=None) == HIP_INVALID_DEVICE_POINTER
    assert pull(0, 3, 16, 12, 1, 5) == HIP_INVALID_VALUE
    assert pull(1, 4, 8, 8, 8, 3) == HIP_INVALID_VALUE
    assert pull(3, 3, 16, 12, 1, 0) == HIP_INVALID_VALUE  # scalar lattice as the fluid
    assert pull(1, 3, 8, 8, 8, 0) == HIP_INVALID_VALUE  # mismatched pair
    assert pull(0, 3, 16, 12, 1, 4, st=9) == HIP_INVALID_VALUE  # storage
    assert L.lbc_fluid_step_strided(1, 1, 0, 8, FAKE_A, FAKE_B, 8, 8, 8, 0, 19 * 64, 1.0, 256, 1, 1, None) == \
        HIP_INVALID_VALUE  # unknown flag
