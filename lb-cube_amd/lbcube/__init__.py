"""Python view of the MI355X lattice Boltzmann kernels (liblbcube_hip.so).

Populations live in torch CUDA tensors laid out [Q][Z][Y][X] (x fastest; 2D:
[Q][Y][X]), the SoA layout of LB-Cube's LatticeMemory
(include/lattice_memory.hpp:67-272).  The storage format follows the tensor
dtype: float32 / float64 hold populations; float16 / bfloat16 hold
(f_i - w_i) for fluid lattices and g_i for scalar lattices (see
include/lbcube.h).  Every call is asynchronous on the current torch stream and
raises LbcError on a non-zero status; no call allocates device memory except
the statistics calls (one small per-device workspace, allocated once).
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import LbcError, check, lib  # noqa: F401

LATTICES = {"D2Q9": 0, "D3Q19": 1, "D3Q27": 2, "D2Q5": 3, "D3Q7": 4}
COLLISIONS = {"BGK": 0, "TRT": 1, "MRT": 2, "RLBM": 3}
Q = {"D2Q9": 9, "D3Q19": 19, "D3Q27": 27, "D2Q5": 5, "D3Q7": 7}
DIM = {"D2Q9": 2, "D3Q19": 3, "D3Q27": 3, "D2Q5": 2, "D3Q7": 3}
STORAGE = {torch.float32: 0, torch.float64: 1, torch.float16: 2, torch.bfloat16: 3}
STORAGE_NAMES = {"f32": torch.float32, "f64": torch.float64, "f16": torch.float16, "bf16": torch.bfloat16}
WEIGHTS = {
    "D2Q9": [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4,
    "D3Q19": [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12,
    "D3Q27": [8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8,
    "D2Q5": [1 / 3] + [1 / 6] * 4,
    "D3Q7": [1 / 4] + [1 / 8] * 6,
}
# Lattice velocities, direction order of include/lattice_traits.hpp:62-305.
VELOCITIES = {
    "D2Q9": [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)],
    "D3Q19": [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, 1, 0),
              (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 1), (0, -1, 1),
              (0, 1, -1), (0, -1, -1)],
    "D2Q5": [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1)],
    "D3Q7": [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)],
}
VELOCITIES["D3Q27"] = VELOCITIES["D3Q19"] + [(sx, sy, sz) for sz in (1, -1) for sy in (1, -1) for sx in (1, -1)]
SCALAR_PAIRS = {("D2Q9", "D2Q5"), ("D3Q19", "D3Q7"), ("D3Q27", "D3Q7")}


def _stream(stream):
    if stream is None:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


def _extents(lattice, t):
    if DIM[lattice] == 3:
        _, nz, ny, nx = t.shape
    else:
        _, ny, nx = t.shape
        nz = 1
    return nx, ny, nz


def _check_pop(lattice, t, name):
    if not (t.is_cuda and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous CUDA tensor")
    if t.dim() != DIM[lattice] + 1 or t.shape[0] != Q[lattice]:
        raise ValueError(f"{name} must have shape (Q={Q[lattice]}, ...) with {DIM[lattice]} spatial axes")
    if t.dtype not in STORAGE:
        raise ValueError(f"{name}: unsupported dtype {t.dtype}")


def _check_set(named):
    """Every tensor of one call must cover the same cells with the same storage
    format on the same device: the kernels take the extents from one tensor and
    the storage code from one dtype, so a mismatch would read or write past
    the end of a smaller allocation."""
    (n0, t0), rest = named[0], named[1:]
    for n, t in rest:
        if tuple(t.shape[1:]) != tuple(t0.shape[1:]):
            raise ValueError(f"{n} covers cells {tuple(t.shape[1:])}, {n0} covers {tuple(t0.shape[1:])}")
        if t.dtype != t0.dtype:
            raise ValueError(f"{n} has dtype {t.dtype}, {n0} has {t0.dtype}: one storage format per call")
        if t.device != t0.device:
            raise ValueError(f"{n} is on {t.device}, {n0} on {t0.device}")


def default_block(lattice, nx, ny, nz=1):
    out = (ctypes.c_uint32 * 3)()
    lib().lbc_default_block(LATTICES[lattice], nx, ny, nz, out)
    return tuple(out)


def _block(lattice, block, nx, ny, nz):
    return default_block(lattice, nx, ny, nz) if block is None else tuple(block)


FLAG_FAST = 1
FLAG_MFMA = 4


def _step_flags(fast, mfma):
    return (FLAG_FAST if fast else 0) | (FLAG_MFMA if mfma else 0)


def fluid_step(lattice, collision, cur, nxt, omega, block=None, stream=None, fast=False, mfma=False):
    """Fused pull-stream + collide: nxt <- step(cur). Replaces launch_step_gpu.

    fast=True selects the deviation-form FMA arithmetic for BGK, TRT and MRT
    (LBC_FLAG_FAST); the default evaluates the reference's operation order.
    mfma=True (with fast=True; D3Q19 / D3Q27 MRT, float16 / bfloat16 storage)
    runs that form's moment transforms on the matrix cores (LBC_FLAG_MFMA)."""
    _check_pop(lattice, cur, "cur")
    _check_pop(lattice, nxt, "nxt")
    _check_set([("cur", cur), ("nxt", nxt)])
    nx, ny, nz = _extents(lattice, cur)
    bx, by, bz = _block(lattice, block, nx, ny, nz)
    check(lib().lbc_fluid_step_ex(LATTICES[lattice], COLLISIONS[collision], STORAGE[cur.dtype],
                                  _step_flags(fast, mfma), cur.data_ptr(), nxt.data_ptr(), nx, ny, nz, float(omega),
                                  bx, by, bz, _stream(stream)), "lbc_fluid_step_ex")


def _check_flat(lattice, t, name, extents, pitch):
    if not (t.is_cuda and t.is_contiguous()) or t.dtype not in STORAGE:
        raise ValueError(f"{name} must be a contiguous CUDA tensor of a storage dtype")
    cells = extents[0] * extents[1] * extents[2]
    plane = pitch if pitch else cells
    if t.numel() < (Q[lattice] - 1) * plane + cells:
        raise ValueError(f"{name} holds {t.numel()} elements; {Q[lattice]} planes of pitch {plane} need more")


def fluid_step_pitched(lattice, collision, cur, nxt, extents, plane_pitch, omega, block=None, stream=None,
                       fast=False, mfma=False):
    """lbc_fluid_step_pitched on flat buffers: plane i at element i * plane_pitch
    (0: the contiguous layout); extents = (nx, ny, nz)."""
    nx, ny, nz = (int(v) for v in extents)
    if DIM[lattice] == 2:
        nz = 1
    for t, n in ((cur, "cur"), (nxt, "nxt")):
        _check_flat(lattice, t, n, (nx, ny, nz), plane_pitch)
    if cur.dtype != nxt.dtype or cur.device != nxt.device:
        raise ValueError("cur and nxt must share dtype and device")
    bx, by, bz = _block(lattice, block, nx, ny, nz)
    check(lib().lbc_fluid_step_pitched(LATTICES[lattice], COLLISIONS[collision], STORAGE[cur.dtype],
                                       _step_flags(fast, mfma), cur.data_ptr(), nxt.data_ptr(), nx, ny, nz,
                                       int(plane_pitch), float(omega), bx, by, bz, _stream(stream)),
          "lbc_fluid_step_pitched")


def slab_strides(lattice, extents):
    """(plane_stride, slab_stride) of the dense [Z][Q][Y][X] order: the Q
    population planes of each xy-slab side by side."""
    nx, ny = int(extents[0]), int(extents[1])
    return nx * ny, Q[lattice] * nx * ny


def _check_storage(t, name):
    if not (t.is_cuda and t.is_contiguous()) or t.dtype not in STORAGE:
        raise ValueError(f"{name} must be a contiguous CUDA tensor of a storage dtype")


def _check_strided(lattice, t, name, extents, plane_stride, slab_stride):
    _check_storage(t, name)
    nx, ny, nz = extents
    slab = nx * ny
    if slab_stride in (0, slab):
        plane, step = (plane_stride or slab * nz), slab
    else:
        plane, step = (plane_stride or slab), slab_stride
    need = (Q[lattice] - 1) * plane + (nz - 1) * step + slab
    if t.numel() < need:
        raise ValueError(f"{name} holds {t.numel()} elements; the layout needs {need}")


def _strided_args(lattice, cur, nxt, extents, plane_stride, slab_stride):
    nx, ny, nz = (int(v) for v in extents)
    if DIM[lattice] == 2:
        nz = 1
    for t, n in ((cur, "cur"), (nxt, "nxt")):
        _check_strided(lattice, t, n, (nx, ny, nz), int(plane_stride), int(slab_stride))
    if cur.dtype != nxt.dtype or cur.device != nxt.device:
        raise ValueError("cur and nxt must share dtype and device")
    return nx, ny, nz


def fluid_step_strided(lattice, collision, cur, nxt, extents, plane_stride, slab_stride, omega, block=None,
                       stream=None, fast=False, mfma=False):
    """lbc_fluid_step_strided on flat buffers: population i of cell (x, y, z)
    at element i * plane_stride + z * slab_stride + y * nx + x (slab_strides()
    gives the dense [Z][Q][Y][X] order); extents = (nx, ny, nz)."""
    nx, ny, nz = _strided_args(lattice, cur, nxt, extents, plane_stride, slab_stride)
    bx, by, bz = _block(lattice, block, nx, ny, nz)
    check(lib().lbc_fluid_step_strided(LATTICES[lattice], COLLISIONS[collision], STORAGE[cur.dtype],
                                       _step_flags(fast, mfma), cur.data_ptr(), nxt.data_ptr(), nx, ny, nz,
                                       int(plane_stride), int(slab_stride), float(omega), bx, by, bz,
                                       _stream(stream)), "lbc_fluid_step_strided")


def fluid_step_slabs(lattice, collision, cur, nxt, omega, block=None, stream=None, fast=False, mfma=False):
    """fluid_step on populations stored slab by slab: tensors of shape
    (nz, Q, ny, nx) (3D lattices)."""
    if DIM[lattice] != 3:
        raise ValueError("the slab order applies to 3D lattices")
    for t, n in ((cur, "cur"), (nxt, "nxt")):
        if t.dim() != 4 or t.shape[1] != Q[lattice]:
            raise ValueError(f"{n} must have shape (nz, {Q[lattice]}, ny, nx)")
    if cur.shape != nxt.shape:
        raise ValueError("cur and nxt must share a shape")
    ext = (cur.shape[3], cur.shape[2], cur.shape[0])
    fluid_step_strided(lattice, collision, cur, nxt, ext, *slab_strides(lattice, ext), omega, block=block,
                       stream=stream, fast=fast, mfma=mfma)


def relayout(lattice, src, dst, extents, src_strides=(0, 0), dst_strides=(0, 0), stream=None):
    """lbc_relayout: copy a population set between two strided layouts
    ((plane_stride, slab_stride) pairs; (0, 0) is the reference's planes)."""
    nx, ny, nz = (int(v) for v in extents)
    if DIM[lattice] == 2:
        nz = 1
    _check_strided(lattice, src, "src", (nx, ny, nz), *(int(v) for v in src_strides))
    _check_strided(lattice, dst, "dst", (nx, ny, nz), *(int(v) for v in dst_strides))
    if src.dtype != dst.dtype or src.device != dst.device:
        raise ValueError("src and dst must share dtype and device")
    check(lib().lbc_relayout(LATTICES[lattice], STORAGE[src.dtype], src.data_ptr(), int(src_strides[0]),
                             int(src_strides[1]), dst.data_ptr(), int(dst_strides[0]), int(dst_strides[1]), nx, ny,
                             nz, _stream(stream)), "lbc_relayout")


def to_slabs(lattice, f, stream=None):
    """Populations of shape (Q, nz, ny, nx) -> a new tensor of shape
    (nz, Q, ny, nx) in the slab order (lbc_relayout)."""
    _check_pop(lattice, f, "f")
    nx, ny, nz = _extents(lattice, f)
    out = torch.empty((nz, Q[lattice], ny, nx), dtype=f.dtype, device=f.device)
    relayout(lattice, f, out, (nx, ny, nz), (0, 0), slab_strides(lattice, (nx, ny)), stream=stream)
    return out


def from_slabs(lattice, h, stream=None):
    """The inverse of to_slabs: (nz, Q, ny, nx) -> (Q, nz, ny, nx)."""
    nz, q, ny, nx = h.shape
    out = torch.empty((q, nz, ny, nx), dtype=h.dtype, device=h.device)
    relayout(lattice, h, out, (nx, ny, nz), slab_strides(lattice, (nx, ny)), (0, 0), stream=stream)
    return out


def _tiled_extents(lattice, t, name):
    """(nx, ny, nz, rows) of a tiled population tensor: (tiles, Q, rows, nx)
    in 2D; in 3D (nz, Q, ny, nx) for the slab order (a tile is an xy-slab) or
    (nz, ny / rows, Q, rows, nx) for tiles of fewer rows."""
    _check_storage(t, name)
    if DIM[lattice] == 3 and t.dim() == 5 and t.shape[2] == Q[lattice]:
        nz, tiles, _, rows, nx = (int(v) for v in t.shape)
        return nx, tiles * rows, nz, rows
    if t.dim() != 4 or t.shape[1] != Q[lattice]:
        raise ValueError(f"{name} must have shape (tiles, {Q[lattice]}, rows, nx)"
                         + (f" or (nz, tiles, {Q[lattice]}, rows, nx)" if DIM[lattice] == 3 else ""))
    tiles, _, rows, nx = (int(v) for v in t.shape)
    if DIM[lattice] == 2:
        return nx, tiles * rows, 1, rows
    return nx, rows, tiles, rows


def to_tiles(lattice, f, rows=None, stream=None):
    """Populations of shape (Q, [nz,] ny, nx) -> a new tensor in tiles of
    `rows` rows (lbc_relayout_tiled), each tile holding the Q planes of its
    rows back to back: 2D (ny / rows, Q, rows, nx); 3D (nz, Q, ny, nx) for the
    slab order (rows = ny, the default) or (nz, ny / rows, Q, rows, nx)."""
    _check_pop(lattice, f, "f")
    nx, ny, nz = _extents(lattice, f)
    rows = ny if (rows is None and DIM[lattice] == 3) else int(rows)
    if rows <= 0 or ny % rows:
        raise ValueError(f"rows must divide ny = {ny}")
    if DIM[lattice] == 3:
        shape = (nz, Q[lattice], ny, nx) if rows == ny else (nz, ny // rows, Q[lattice], rows, nx)
    else:
        shape = (ny // rows, Q[lattice], rows, nx)
    out = torch.empty(shape, dtype=f.dtype, device=f.device)
    check(lib().lbc_relayout_tiled(LATTICES[lattice], STORAGE[f.dtype], f.data_ptr(), 0, out.data_ptr(), rows, nx,
                                   ny, nz, _stream(stream)), "lbc_relayout_tiled")
    return out


def from_tiles(lattice, t, stream=None):
    """The inverse of to_tiles: back to the reference's planes (Q, [nz,] ny, nx)."""
    nx, ny, nz, rows = _tiled_extents(lattice, t, "t")
    shape = (Q[lattice], ny, nx) if DIM[lattice] == 2 else (Q[lattice], nz, ny, nx)
    out = torch.empty(shape, dtype=t.dtype, device=t.device)
    check(lib().lbc_relayout_tiled(LATTICES[lattice], STORAGE[t.dtype], t.data_ptr(), rows, out.data_ptr(), 0, nx,
                                   ny, nz, _stream(stream)), "lbc_relayout_tiled")
    return out


def fluid_step_tiled(lattice, collision, cur, nxt, omega, block=None, stream=None, fast=False, mfma=False):
    """fluid_step on populations in tiles (to_tiles): lbc_fluid_step_tiled."""
    nx, ny, nz, rows = _tiled_extents(lattice, cur, "cur")
    if nxt.shape != cur.shape or nxt.dtype != cur.dtype:
        raise ValueError("cur and nxt must share shape and dtype")
    _tiled_extents(lattice, nxt, "nxt")
    bx, by, bz = _block(lattice, block, nx, ny, nz)
    check(lib().lbc_fluid_step_tiled(LATTICES[lattice], COLLISIONS[collision], STORAGE[cur.dtype],
                                     _step_flags(fast, mfma), cur.data_ptr(), nxt.data_ptr(), nx, ny, nz, rows,
                                     float(omega), bx, by, bz, _stream(stream)), "lbc_fluid_step_tiled")


def reactive_fused_step_tiled(fluid_lattice, collision, scalar_lattice, cur_f, nxt_f, cur_a, nxt_a, cur_b, nxt_b,
                              omega, omega_c, k_react, block=None, stream=None, fast=False, ref_a=0.0, ref_b=0.0):
    """reactive_fused_step on the three population sets in tiles of the same
    rows (to_tiles): lbc_reactive_fused_step_tiled."""
    nx, ny, nz, rows = _tiled_extents(scalar_lattice, cur_a, "cur_a")
    for t, n, lat in ((cur_f, "cur_f", fluid_lattice), (nxt_f, "nxt_f", fluid_lattice), (nxt_a, "nxt_a", scalar_lattice),
                      (cur_b, "cur_b", scalar_lattice), (nxt_b, "nxt_b", scalar_lattice)):
        if _tiled_extents(lat, t, n) != (nx, ny, nz, rows) or t.dtype != cur_a.dtype:
            raise ValueError(f"{n} must match cur_a's tiles and dtype")
    bx, by, bz = _block(scalar_lattice, block, nx, ny, nz)
    check(lib().lbc_reactive_fused_step_tiled(LATTICES[fluid_lattice], COLLISIONS[collision], LATTICES[scalar_lattice],
                                              STORAGE[cur_f.dtype], 1 if fast else 0, cur_f.data_ptr(),
                                              nxt_f.data_ptr(), cur_a.data_ptr(), nxt_a.data_ptr(), cur_b.data_ptr(),
                                              nxt_b.data_ptr(), nx, ny, nz, rows, float(omega), float(omega_c),
                                              float(k_react), float(ref_a), float(ref_b), bx, by, bz,
                                              _stream(stream)), "lbc_reactive_fused_step_tiled")


def _same_tiles(ref, ref_name, sets):
    """(nx, ny, nz, rows) of `ref`; every (tensor, name, lattice) of `sets`
    must have the same tiles and dtype."""
    ext = _tiled_extents(ref[1], ref[0], ref_name)
    for t, n, lat in sets:
        if _tiled_extents(lat, t, n) != ext or t.dtype != ref[0].dtype:
            raise ValueError(f"{n} must match {ref_name}'s tiles and dtype")
    return ext


def scalar_step_tiled(fluid_lattice, scalar_lattice, cur_s, nxt_s, cur_f, omega_c, source=0.0, block=None,
                      stream=None, ref=0.0):
    """scalar_step with the scalar and the fluid in tiles of the same rows
    (to_tiles): lbc_scalar_step_tiled."""
    nx, ny, nz, rows = _same_tiles((cur_s, scalar_lattice), "cur_s",
                                   ((nxt_s, "nxt_s", scalar_lattice), (cur_f, "cur_f", fluid_lattice)))
    if nxt_s.data_ptr() in (cur_s.data_ptr(), cur_f.data_ptr()):
        raise ValueError("nxt_s must not alias cur_s or cur_f")
    bx, by, bz = _block(scalar_lattice, block, nx, ny, nz)
    check(lib().lbc_scalar_step_tiled(LATTICES[fluid_lattice], LATTICES[scalar_lattice], STORAGE[cur_s.dtype],
                                      cur_s.data_ptr(), nxt_s.data_ptr(), cur_f.data_ptr(), nx, ny, nz, rows,
                                      float(omega_c), float(source), float(ref), bx, by, bz, _stream(stream)),
          "lbc_scalar_step_tiled")


def reaction_step_tiled(fluid_lattice, scalar_lattice, cur_a, nxt_a, cur_b, nxt_b, cur_f, omega_c, k_react,
                        block=None, stream=None, ref_a=0.0, ref_b=0.0):
    """reaction_step with both species and the fluid in tiles of the same
    rows (to_tiles): lbc_reaction_ab_step_tiled."""
    nx, ny, nz, rows = _same_tiles((cur_a, scalar_lattice), "cur_a",
                                   ((nxt_a, "nxt_a", scalar_lattice), (cur_b, "cur_b", scalar_lattice),
                                    (nxt_b, "nxt_b", scalar_lattice), (cur_f, "cur_f", fluid_lattice)))
    bx, by, bz = _block(scalar_lattice, block, nx, ny, nz)
    check(lib().lbc_reaction_ab_step_tiled(LATTICES[fluid_lattice], LATTICES[scalar_lattice], STORAGE[cur_a.dtype],
                                           cur_a.data_ptr(), nxt_a.data_ptr(), cur_b.data_ptr(), nxt_b.data_ptr(),
                                           cur_f.data_ptr(), nx, ny, nz, rows, float(omega_c), float(k_react),
                                           float(ref_a), float(ref_b), bx, by, bz, _stream(stream)),
          "lbc_reaction_ab_step_tiled")


def reactive_pull_stream(fluid_lattice, scalar_lattice, cur_f, nxt_f, cur_a, nxt_a, cur_b, nxt_b, tiled=False,
                         shift=True, block=None, stream=None):
    """lbc_reactive_pull_stream: the fused step's access pattern without
    arithmetic (the ceiling of reactive_fused_step on the same buffers), on
    the reference's planes or (tiled=True) on tensors from to_tiles."""
    if tiled:
        nx, ny, nz, rows = _tiled_extents(scalar_lattice, cur_a, "cur_a")
        sets = ((cur_f, "cur_f", fluid_lattice), (nxt_f, "nxt_f", fluid_lattice), (nxt_a, "nxt_a", scalar_lattice),
                (cur_b, "cur_b", scalar_lattice), (nxt_b, "nxt_b", scalar_lattice))
        for t, n, lat in sets:
            if _tiled_extents(lat, t, n) != (nx, ny, nz, rows) or t.dtype != cur_a.dtype:
                raise ValueError(f"{n} must match cur_a's tiles and dtype")
    else:
        rows = 0
        _check_pop(fluid_lattice, cur_f, "cur_f")
        _check_pop(fluid_lattice, nxt_f, "nxt_f")
        for t, n in ((cur_a, "cur_a"), (nxt_a, "nxt_a"), (cur_b, "cur_b"), (nxt_b, "nxt_b")):
            _check_pop(scalar_lattice, t, n)
        _check_set([("cur_f", cur_f), ("nxt_f", nxt_f), ("cur_a", cur_a), ("nxt_a", nxt_a), ("cur_b", cur_b),
                    ("nxt_b", nxt_b)])
        nx, ny, nz = _extents(scalar_lattice, cur_a)
    bx, by, bz = _block(scalar_lattice, block, nx, ny, nz)
    check(lib().lbc_reactive_pull_stream(LATTICES[fluid_lattice], LATTICES[scalar_lattice], STORAGE[cur_f.dtype],
                                         int(bool(shift)), cur_f.data_ptr(), nxt_f.data_ptr(), cur_a.data_ptr(),
                                         nxt_a.data_ptr(), cur_b.data_ptr(), nxt_b.data_ptr(), nx, ny, nz, rows, bx,
                                         by, bz, _stream(stream)), "lbc_reactive_pull_stream")


def reactive_stats_tiled(scalar_lattice, a, b, k_react, stream=None, ref_a=0.0, ref_b=0.0):
    """reactive_stats of species in tiles (to_tiles): lbc_reactive_stats_tiled;
    the same sums in the same order, so the same results bit for bit."""
    nx, ny, nz, rows = _tiled_extents(scalar_lattice, a, "a")
    if _tiled_extents(scalar_lattice, b, "b") != (nx, ny, nz, rows) or b.dtype != a.dtype:
        raise ValueError("b must match a's tiles and dtype")
    out = (ctypes.c_double * 8)()
    check(lib().lbc_reactive_stats_tiled(LATTICES[scalar_lattice], STORAGE[a.dtype], a.data_ptr(), b.data_ptr(), nx,
                                         ny, nz, rows, float(k_react), float(ref_a), float(ref_b), out,
                                         _stream(stream)), "lbc_reactive_stats_tiled")
    return dict(zip(STAT_NAMES, out))


def pull_stream_strided(lattice, cur, nxt, extents, plane_stride, slab_stride, block=None, stream=None, shift=True):
    """lbc_pull_stream_strided: the pull-streaming step (shift=True) or the
    unshifted plane copy on a strided layout."""
    nx, ny, nz = _strided_args(lattice, cur, nxt, extents, plane_stride, slab_stride)
    bx, by, bz = _block(lattice, block, nx, ny, nz)
    check(lib().lbc_pull_stream_strided(LATTICES[lattice], STORAGE[cur.dtype], int(bool(shift)), cur.data_ptr(),
                                        nxt.data_ptr(), nx, ny, nz, int(plane_stride), int(slab_stride), bx, by, bz,
                                        _stream(stream)), "lbc_pull_stream_strided")


def pull_stream(lattice, cur, nxt, extents=None, plane_pitch=0, block=None, stream=None, shift=True):
    """The pull-streaming step alone (lbc_pull_stream), or with shift=False
    the same plane streams copied in place (lbc_plane_copy).  Population
    tensors of shape (Q, ...) for the contiguous layout, or flat buffers with
    `extents` and `plane_pitch`."""
    if extents is None:
        _check_pop(lattice, cur, "cur")
        _check_pop(lattice, nxt, "nxt")
        _check_set([("cur", cur), ("nxt", nxt)])
        nx, ny, nz = _extents(lattice, cur)
    else:
        nx, ny, nz = (int(v) for v in extents)
        if DIM[lattice] == 2:
            nz = 1
        for t, n in ((cur, "cur"), (nxt, "nxt")):
            _check_flat(lattice, t, n, (nx, ny, nz), plane_pitch)
        if cur.dtype != nxt.dtype or cur.device != nxt.device:
            raise ValueError("cur and nxt must share dtype and device")
    bx, by, bz = _block(lattice, block, nx, ny, nz)
    name = "lbc_pull_stream" if shift else "lbc_plane_copy"
    check(getattr(lib(), name)(LATTICES[lattice], STORAGE[cur.dtype], cur.data_ptr(), nxt.data_ptr(), nx, ny, nz,
                               int(plane_pitch), bx, by, bz, _stream(stream)), name)


def scalar_step(fluid_lattice, scalar_lattice, cur_s, nxt_s, cur_f, omega_c, source=0.0, block=None, stream=None,
                ref=0.0):
    """Passive-scalar ADR step. Replaces launch_scalar_step_gpu.  `ref`: the
    field's reference concentration (f16/bf16 storage holds g - w ref)."""
    _check_pop(scalar_lattice, cur_s, "cur_s")
    _check_pop(scalar_lattice, nxt_s, "nxt_s")
    _check_pop(fluid_lattice, cur_f, "cur_f")
    _check_set([("cur_s", cur_s), ("nxt_s", nxt_s), ("cur_f", cur_f)])
    nx, ny, nz = _extents(scalar_lattice, cur_s)
    bx, by, bz = _block(scalar_lattice, block, nx, ny, nz)
    check(lib().lbc_scalar_step(LATTICES[fluid_lattice], LATTICES[scalar_lattice], STORAGE[cur_s.dtype],
                                cur_s.data_ptr(), nxt_s.data_ptr(), cur_f.data_ptr(), nx, ny, nz, float(omega_c),
                                float(source), float(ref), bx, by, bz, _stream(stream)), "lbc_scalar_step")


def reaction_step(fluid_lattice, scalar_lattice, cur_a, nxt_a, cur_b, nxt_b, cur_f, omega_c, k_react, block=None,
                  stream=None, ref_a=0.0, ref_b=0.0):
    """Fused A + B -> C ADR step. Replaces launch_reaction_AB_gpu.  ref_a/ref_b:
    the species' reference concentrations."""
    for t, n in ((cur_a, "cur_a"), (nxt_a, "nxt_a"), (cur_b, "cur_b"), (nxt_b, "nxt_b")):
        _check_pop(scalar_lattice, t, n)
    _check_pop(fluid_lattice, cur_f, "cur_f")
    _check_set([("cur_a", cur_a), ("nxt_a", nxt_a), ("cur_b", cur_b), ("nxt_b", nxt_b), ("cur_f", cur_f)])
    nx, ny, nz = _extents(scalar_lattice, cur_a)
    bx, by, bz = _block(scalar_lattice, block, nx, ny, nz)
    check(lib().lbc_reaction_ab_step(LATTICES[fluid_lattice], LATTICES[scalar_lattice], STORAGE[cur_a.dtype],
                                     cur_a.data_ptr(), nxt_a.data_ptr(), cur_b.data_ptr(), nxt_b.data_ptr(),
                                     cur_f.data_ptr(), nx, ny, nz, float(omega_c), float(k_react), float(ref_a),
                                     float(ref_b), bx, by, bz, _stream(stream)), "lbc_reaction_ab_step")


def reactive_fused_step(fluid_lattice, collision, scalar_lattice, cur_f, nxt_f, cur_a, nxt_a, cur_b, nxt_b, omega,
                        omega_c, k_react, block=None, stream=None, fast=False, ref_a=0.0, ref_b=0.0):
    """Fluid step + A + B -> C step in one pass (same results as the two calls);
    fast=True selects the deviation-form arithmetic (LBC_FLAG_FAST); ref_a/ref_b:
    the species' reference concentrations."""
    _check_pop(fluid_lattice, cur_f, "cur_f")
    _check_pop(fluid_lattice, nxt_f, "nxt_f")
    for t, n in ((cur_a, "cur_a"), (nxt_a, "nxt_a"), (cur_b, "cur_b"), (nxt_b, "nxt_b")):
        _check_pop(scalar_lattice, t, n)
    _check_set([("cur_f", cur_f), ("nxt_f", nxt_f), ("cur_a", cur_a), ("nxt_a", nxt_a), ("cur_b", cur_b),
                ("nxt_b", nxt_b)])
    nx, ny, nz = _extents(scalar_lattice, cur_a)
    bx, by, bz = _block(scalar_lattice, block, nx, ny, nz)
    check(lib().lbc_reactive_fused_step_ex(LATTICES[fluid_lattice], COLLISIONS[collision], LATTICES[scalar_lattice],
                                           STORAGE[cur_f.dtype], 1 if fast else 0, cur_f.data_ptr(), nxt_f.data_ptr(),
                                           cur_a.data_ptr(), nxt_a.data_ptr(), cur_b.data_ptr(), nxt_b.data_ptr(), nx,
                                           ny, nz, float(omega), float(omega_c), float(k_react), float(ref_a),
                                           float(ref_b), bx, by, bz, _stream(stream)), "lbc_reactive_fused_step_ex")


def fluid_diagnostics(lattice, f, stream=None):
    """(total mass, total kinetic energy, max |u|), double accumulation; synchronizes."""
    _check_pop(lattice, f, "f")
    nx, ny, nz = _extents(lattice, f)
    out = (ctypes.c_double * 3)()
    check(lib().lbc_fluid_diagnostics(LATTICES[lattice], STORAGE[f.dtype], f.data_ptr(), nx, ny, nz, out,
                                      _stream(stream)), "lbc_fluid_diagnostics")
    return tuple(out)


def max_speed(lattice, f, stream=None):
    """Stability guard (compute_max_velocity): max |u|, inf if any cell's speed
    is not finite; synchronizes."""
    _check_pop(lattice, f, "f")
    nx, ny, nz = _extents(lattice, f)
    out = (ctypes.c_double * 1)()
    check(lib().lbc_max_speed(LATTICES[lattice], STORAGE[f.dtype], f.data_ptr(), nx, ny, nz, out, _stream(stream)),
          "lbc_max_speed")
    return out[0]


def macro_fields(lattice, f, stream=None):
    """(rho, u): float64 device tensors of shape f.shape[1:] and (D,) + f.shape[1:]."""
    _check_pop(lattice, f, "f")
    nx, ny, nz = _extents(lattice, f)
    rho = torch.empty(f.shape[1:], dtype=torch.float64, device=f.device)
    u = torch.empty((len(VELOCITIES[lattice][0]),) + tuple(f.shape[1:]), dtype=torch.float64, device=f.device)
    check(lib().lbc_macro_fields(LATTICES[lattice], STORAGE[f.dtype], f.data_ptr(), nx, ny, nz, rho.data_ptr(),
                                 u.data_ptr(), _stream(stream)), "lbc_macro_fields")
    return rho, u


def scalar_fields(scalar_lattice, g, stream=None, ref=0.0):
    """Concentration per cell: a float64 device tensor of shape g.shape[1:]."""
    _check_pop(scalar_lattice, g, "g")
    nx, ny, nz = _extents(scalar_lattice, g)
    conc = torch.empty(g.shape[1:], dtype=torch.float64, device=g.device)
    check(lib().lbc_scalar_fields(LATTICES[scalar_lattice], STORAGE[g.dtype], g.data_ptr(), nx, ny, nz, float(ref),
                                  conc.data_ptr(), _stream(stream)), "lbc_scalar_fields")
    return conc


STAT_NAMES = ("mean_A", "mean_B", "var_A", "var_B", "covariance", "segregation_intensity", "true_reaction_rate",
              "mixed_reaction_rate")


def reactive_stats(scalar_lattice, a, b, k_react, stream=None, ref_a=0.0, ref_b=0.0, no_alloc=False):
    """compute_reactive_stats on the device; returns a dict; synchronizes.
    no_alloc=True calls lbc_reactive_stats_noalloc (partials in the library's
    static device array: no device allocation, same results)."""
    _check_pop(scalar_lattice, a, "a")
    _check_pop(scalar_lattice, b, "b")
    _check_set([("a", a), ("b", b)])
    nx, ny, nz = _extents(scalar_lattice, a)
    out = (ctypes.c_double * 8)()
    fn = lib().lbc_reactive_stats_noalloc if no_alloc else lib().lbc_reactive_stats
    check(fn(LATTICES[scalar_lattice], STORAGE[a.dtype], a.data_ptr(), b.data_ptr(), nx, ny, nz, float(k_react),
             float(ref_a), float(ref_b), out, _stream(stream)),
          "lbc_reactive_stats_noalloc" if no_alloc else "lbc_reactive_stats")
    return dict(zip(STAT_NAMES, out))


def weights_tensor(lattice, dtype=torch.float32, device="cuda"):
    return torch.tensor(WEIGHTS[lattice], dtype=torch.float64).to(dtype).to(device)


def _offsets(lattice, ref, device, ndim):
    """Per-population storage offsets in float32: w_i for a fluid lattice,
    w_i * ref for a scalar lattice (the products the kernels form)."""
    w = weights_tensor(lattice, torch.float32, device)
    if lattice in ("D2Q5", "D3Q7"):
        w = w * torch.tensor(float(ref), dtype=torch.float32, device=device)
    return w.view((-1,) + (1,) * (ndim - 1))


def encode(lattice, f, storage, ref=0.0):
    """Populations (float tensor) -> storage tensor: f16/bf16 hold f - w for a
    fluid lattice and g - w * ref for a scalar lattice (ref: the field's
    reference concentration)."""
    dt = STORAGE_NAMES[storage] if isinstance(storage, str) else storage
    if dt in (torch.float32, torch.float64):
        return f.to(dt).contiguous()
    return (f.to(torch.float32) - _offsets(lattice, ref, f.device, f.dim())).to(dt).contiguous()


def decode(lattice, t, ref=0.0):
    """Storage tensor -> float64 populations."""
    if t.dtype in (torch.float32, torch.float64):
        return t.to(torch.float64)
    return (_offsets(lattice, ref, t.device, t.dim()) + t.to(torch.float32)).to(torch.float64)


def stream_copy(src, dst, stream=None):
    """Streaming copy of src into dst (same byte size, multiple of 16 bytes)."""
    if not (src.is_cuda and dst.is_cuda and src.is_contiguous() and dst.is_contiguous()):
        raise ValueError("src and dst must be contiguous CUDA tensors")
    nbytes = src.numel() * src.element_size()
    if nbytes != dst.numel() * dst.element_size():
        raise ValueError("src and dst must have the same size in bytes")
    check(lib().lbc_stream_copy(src.data_ptr(), dst.data_ptr(), nbytes, _stream(stream)), "lbc_stream_copy")
