// Fused pull-stream + collide kernels for D2Q9 / D3Q19 / D3Q27 — the gfx950
// replacement of LB-Cube's kernel_step (src/lattice_cuda.cu:149-213) behind
// lbc_fluid_step (launch_step_gpu, src/lattice_cuda.cu:476-520).
//
// One work-item owns one destination cell; cells are numbered linearly (x
// fastest), so the 64 lanes of a wave cover 64 consecutive cells and each of
// the Q population planes is read and written with unit stride.  A population is pulled from x - c_i
// (periodic), collided in registers and stored once: Q loads + Q stores per
// cell, the whole algorithmic traffic.  Periodic wrap uses selects on 32-bit
// coordinates; the plane offset of population i is a wave-uniform 64-bit base
// (per-lane offsets 32-bit, or 64-bit for planes of more than 4 GiB).
#include <hip/hip_runtime.h>

#include "../../include/lbcube.h"
#include "lbc_fast.hpp"
#include "lbc_launch.hpp"
#include "lbc_mfma.hpp"
#include "lbc_physics.hpp"
#include "lbc_storage.hpp"

namespace lbc {

// FAST selects the deviation-form arithmetic (lbc_fast.hpp) behind
// LBC_FLAG_FAST: BGK, TRT and MRT on every lattice; RLBM always evaluates the
// reference's order.  WIDE: 64-bit per-lane addresses, for planes of more than
// 4 GiB.
// `plane`: elements from one population plane to the next (the cell count for
// the reference's contiguous [Q][Z][Y][X] buffers; larger for pitched ones).
template <class L, int CT, int ST, bool FAST, int MAXT, bool WIDE>
__global__ __launch_bounds__(MAXT) void fluid_step_kernel(const typename Storage<ST>::S* __restrict__ cur,
                                                          typename Storage<ST>::S* __restrict__ nxt, Cells c,
                                                          size_t plane, typename Storage<ST>::R omega,
                                                          typename Storage<ST>::R omega_minus) {
    using St = Storage<ST>;
    using R = typename St::R;
    // tiled 2D grids keep the dispatch order (as the fused reactive kernel)
    const uint32_t block =
        (renumber_blocks<L::D, typename Storage<ST>::S>() && c.tiles == 1) ? xcd_block() : blockIdx.x;
    const uint32_t dst = block * blockDim.x + threadIdx.x;
    if (dst >= c.count) return;
    const CellCoord q = cell_coord(dst, c);
    const Pull pull = locate<L::D>(dst, q, c, c.zgap);
    const uint32_t cell = pull.dst;
    R f[L::Q];
    if constexpr (FAST) {
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            const uint32_t src = pull.src(L::cx[i], L::cy[i], L::cz[i]);
            f[i] = St::load_dev(at<WIDE>(plane_ptr(cur, i, plane), src), R(L::w[i]));
        }
        collide_dev<L, CT, R>(f, omega, omega_minus, [&](int i, R v) {
            St::store_dev(at<WIDE>(plane_ptr_opaque(nxt, i, plane), cell), v, R(L::w[i]));
        });
    } else {
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            const uint32_t src = pull.src(L::cx[i], L::cy[i], L::cz[i]);
            const R off = kOffsetStorage<ST> ? R(L::w[i]) : R(0);
            f[i] = St::load(at<WIDE>(plane_ptr(cur, i, plane), src), off);
        }
        collide_cell<L, CT, R>(f, omega, omega_minus, [&](int i, R v) {
            St::store(at<WIDE>(plane_ptr_opaque(nxt, i, plane), cell), v, kOffsetStorage<ST> ? R(L::w[i]) : R(0));
        });
    }
}

// LBC_FLAG_MFMA: the deviation-form MRT of D3Q19 / D3Q27 with both moment
// transforms on the matrix cores (lbc_mfma.hpp), FP16 / BF16 storage.  The
// launcher guarantees blocks of whole waves; a wave that starts past the
// domain's end leaves at once, and the lanes of a partial last wave take part
// in the products with zero populations and store nothing.
template <int ST> __device__ __forceinline__ float decode16(uint32_t bits) {
    if constexpr (ST == kF16) return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(bits)));
    else return __builtin_bit_cast(float, bits << 16);
}

template <class L, int ST, int MAXT, bool WIDE>
__global__ __launch_bounds__(MAXT) void fluid_step_mfma_kernel(const uint16_t* __restrict__ cur,
                                                               uint16_t* __restrict__ nxt, Cells c, size_t plane,
                                                               float omega) {
    using St = Storage<ST>;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t dst = blockIdx.x * blockDim.x + threadIdx.x;
    if (dst - lane >= c.count) return;  // the whole wave lies past the domain
    const bool live = dst < c.count;
    uint32_t hw[16];  // the cell's stored deviations in pairs (population 2v in the low half), zero-padded
#pragma unroll
    for (int v = 0; v < 16; ++v) hw[v] = 0u;
    uint32_t off = 0u;  // the cell's element offset in its planes
    if (live) {
        const CellCoord q = cell_coord(dst, c);
        const Pull pull = locate<L::D>(dst, q, c, c.zgap);
        off = pull.dst;
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            const uint32_t raw = get(at<WIDE>(plane_ptr(cur, i, plane), pull.src(L::cx[i], L::cy[i], L::cz[i])));
            hw[i / 2] |= raw << (16 * (i % 2));
        }
    }
    const MfmaOperands& T = mfma_operands<L, ST>();
    uint32_t a[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) a[v] = T.a[0][lane][v];
    float m[32];
    mfma_product<ST>(a, hw, m);  // m = M h
    float p[32];
    if constexpr (L::Q == 19) mfma_relax_d3q19<ST>(m, p, omega);
    else mfma_relax_d3q27<ST>(m, p, omega);
    uint32_t pw[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) pw[v] = pack2<ST>(p[2 * v], p[2 * v + 1]);
#pragma unroll
    for (int v = 0; v < 8; ++v) a[v] = T.a[1][lane][v];
    float out[32];
    mfma_product<ST>(a, pw, out);  // out = M^T p'
    if (!live) return;
    constexpr float inv_scale = 1.0f / kMfmaScale<ST>;
    const uint32_t cell = fresh_lane_value(off);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const float h = decode16<ST>((hw[i / 2] >> (16 * (i % 2))) & 0xFFFFu);
        St::store_dev(at<WIDE>(plane_ptr_opaque(nxt, i, plane), cell), fmadd(-inv_scale, out[i], h), 0.0f);
    }
}

// The pull-streaming step alone: next[i][x] = cur[i][x - c_i] for every
// population, with the fluid kernel's addressing, plane order, store form and
// launch shape, and no collision (the identity collision of BGK at omega = 0).
// Its rate is the bandwidth ceiling of the collide-stream kernels' access
// pattern: the same 2 Q streams with the x-shifted pulls, without arithmetic.
// SHIFT = false: the same 2 Q plane streams read without the shifts
// (next[i][x] = cur[i][x]), the ceiling of the plane streams themselves.
template <class L, int ST, int MAXT, bool WIDE, bool SHIFT = true>
__global__ __launch_bounds__(MAXT) void pull_stream_kernel(const typename Storage<ST>::S* __restrict__ cur,
                                                           typename Storage<ST>::S* __restrict__ nxt, Cells c,
                                                           size_t plane) {
    const uint32_t block =
        (renumber_blocks<L::D, typename Storage<ST>::S>() && c.tiles == 1) ? xcd_block() : blockIdx.x;
    const uint32_t dst = block * blockDim.x + threadIdx.x;
    if (dst >= c.count) return;
    const CellCoord q = cell_coord(dst, c);
    const Pull pull = locate<L::D>(dst, q, c, c.zgap);
    const uint32_t cell = pull.dst;
    typename Storage<ST>::S v[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
        v[i] = get(at<WIDE>(plane_ptr(cur, i, plane), SHIFT ? pull.src(L::cx[i], L::cy[i], L::cz[i]) : cell));
#pragma unroll
    for (int i = 0; i < L::Q; ++i) put(at<WIDE>(plane_ptr_opaque(nxt, i, plane), cell), v[i]);
}

// Kernels that run faster with residency capped at kCappedWavesPerCu (measured
// on MI355X): 3D lattices with FP32 storage or FP16 in the deviation form
// (the reference-order D3Q27 MRT included since its constant divisions made it
// memory-bound: +0.9% at 256^3, +2.7% at 384^3; RLBM in FP32: D3Q19 38.8-38.9k
// MLUPS against 37.2-37.3k uncapped at 256^3, D3Q27 26.8k against 26.1-26.3k).
// Left uncapped: BF16 (its software rounding makes the step VALU-heavier:
// 71.0k against 74.4k MLUPS capped), FP64 (unchanged by the cap) and 16-bit
// storage in the reference order (RLBM has no other form).
template <class L, int CT, int ST, bool FAST> constexpr bool capped_residency() {
    if (L::D != 3 || ST == kF64 || ST == kBF16) return false;
    if (kOffsetStorage<ST> && !FAST) return false;
    return true;
}

// Resident waves per CU where the residency is capped: kCappedWavesPerCu,
// except the D3Q19 step in FP16 with the deviation-form arithmetic, which
// keeps more waves busy per byte moved: at 256^3, 24 waves per CU gave TRT
// 74.5-75.4k MLUPS against 70.4-74.6k at 16 (81.5-81.9k against 79.4-79.8k
// with the planes stored slab by slab) and MRT 75.1-75.8k against 73.8-74.0k
// (81.3-81.8k against 77.2-77.6k); 28 gave the same, 32 less.  D3Q27 in FP16
// (50.7-50.8k at 24 against 51.4-51.7k) and every FP32 step stay at 16 (D3Q27
// FP32 TRT / MRT / RLBM at 256^3 and 384^3: 12 waves lost about 7%, 20 were
// within 1% and 24 lost 1-2% against 16).
template <class L, int CT, int ST, bool FAST> constexpr uint32_t capped_waves() {
    return (L::Q == 19 && ST == kF16 && FAST) ? 24u : kCappedWavesPerCu;
}

template <class L, int CT, int ST, bool FAST>
hipError_t launch_fluid(const void* cur, void* nxt, const LaunchShape& s, uint64_t plane, double omega,
                        hipStream_t stream) {
    using St = Storage<ST>;
    using R = typename St::R;
    const R w = static_cast<R>(omega);
    const R wm = CT == kTRT ? trt_omega_minus<R>(w) : R(0);
    const uint32_t lds =
        capped_residency<L, CT, ST, FAST>() ? capped_lds(s.threads(), capped_waves<L, CT, ST, FAST>()) : 0;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, s.grid, s.block, lds, stream, static_cast<const typename St::S*>(cur),
                           static_cast<typename St::S*>(nxt), s.cells, static_cast<size_t>(plane), w, wm);
    };
    const bool wide = wide_planes<typename St::S>(s);
    if (s.threads() <= 256)
        wide ? go(fluid_step_kernel<L, CT, ST, FAST, 256, true>) : go(fluid_step_kernel<L, CT, ST, FAST, 256, false>);
    else
        wide ? go(fluid_step_kernel<L, CT, ST, FAST, 1024, true>) : go(fluid_step_kernel<L, CT, ST, FAST, 1024, false>);
    return hipGetLastError();
}

template <class L, int CT, bool FAST>
hipError_t dispatch_storage(int storage, const void* cur, void* nxt, const LaunchShape& s, uint64_t plane,
                            double omega, hipStream_t st) {
    switch (storage) {
        case kF32: return launch_fluid<L, CT, kF32, FAST>(cur, nxt, s, plane, omega, st);
        case kF64: return launch_fluid<L, CT, kF64, FAST>(cur, nxt, s, plane, omega, st);
        case kF16: return launch_fluid<L, CT, kF16, FAST>(cur, nxt, s, plane, omega, st);
        case kBF16: return launch_fluid<L, CT, kBF16, FAST>(cur, nxt, s, plane, omega, st);
        default: return hipErrorInvalidValue;
    }
}

template <class L>
hipError_t dispatch_collision(int collision, int storage, bool fast, const void* cur, void* nxt,
                              const LaunchShape& s, uint64_t plane, double omega, hipStream_t st) {
    switch (collision) {
        case kBGK:
            return fast ? dispatch_storage<L, kBGK, true>(storage, cur, nxt, s, plane, omega, st)
                        : dispatch_storage<L, kBGK, false>(storage, cur, nxt, s, plane, omega, st);
        case kTRT:
            return fast ? dispatch_storage<L, kTRT, true>(storage, cur, nxt, s, plane, omega, st)
                        : dispatch_storage<L, kTRT, false>(storage, cur, nxt, s, plane, omega, st);
        case kMRT:
            return fast ? dispatch_storage<L, kMRT, true>(storage, cur, nxt, s, plane, omega, st)
                        : dispatch_storage<L, kMRT, false>(storage, cur, nxt, s, plane, omega, st);
        case kRLBM: return dispatch_storage<L, kRLBM, false>(storage, cur, nxt, s, plane, omega, st);
        default: return hipErrorInvalidValue;
    }
}

template <class L, int ST>
hipError_t launch_fluid_mfma(const void* cur, void* nxt, const LaunchShape& s, uint64_t plane, double omega,
                             hipStream_t stream) {
    const float w = static_cast<float>(omega);
    const uint32_t lds = capped_lds(s.threads());
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, s.grid, s.block, lds, stream, static_cast<const uint16_t*>(cur),
                           static_cast<uint16_t*>(nxt), s.cells, static_cast<size_t>(plane), w);
    };
    const bool wide = wide_planes<uint16_t>(s);
    if (s.threads() <= 256)
        wide ? go(fluid_step_mfma_kernel<L, ST, 256, true>) : go(fluid_step_mfma_kernel<L, ST, 256, false>);
    else
        wide ? go(fluid_step_mfma_kernel<L, ST, 1024, true>) : go(fluid_step_mfma_kernel<L, ST, 1024, false>);
    return hipGetLastError();
}

// LBC_FLAG_MFMA: with LBC_FLAG_FAST, MRT, D3Q19 or D3Q27 and FP16 or BF16
// storage only; blocks of whole waves.
hipError_t dispatch_mfma(int lattice, int collision, int storage, uint32_t flags, const void* cur, void* nxt,
                         const LaunchShape& s, uint64_t plane, double omega, hipStream_t st) {
    if (!(flags & LBC_FLAG_FAST) || collision != kMRT) return hipErrorInvalidValue;
    if (lattice != kD3Q19 && lattice != kD3Q27) return hipErrorInvalidValue;
    if (storage != kF16 && storage != kBF16) return hipErrorInvalidValue;
    if (s.threads() % 64 != 0) return hipErrorInvalidConfiguration;
    if (lattice == kD3Q19)
        return storage == kF16 ? launch_fluid_mfma<D3Q19, kF16>(cur, nxt, s, plane, omega, st)
                               : launch_fluid_mfma<D3Q19, kBF16>(cur, nxt, s, plane, omega, st);
    return storage == kF16 ? launch_fluid_mfma<D3Q27, kF16>(cur, nxt, s, plane, omega, st)
                           : launch_fluid_mfma<D3Q27, kBF16>(cur, nxt, s, plane, omega, st);
}

// The pull-streaming step shares the residency of the TRT step it bounds.
template <class L, int ST, bool SHIFT>
hipError_t launch_stream(const void* cur, void* nxt, const LaunchShape& s, uint64_t plane, hipStream_t stream) {
    using S = typename Storage<ST>::S;
    constexpr bool kFast = kOffsetStorage<ST>;
    const uint32_t lds =
        capped_residency<L, kTRT, ST, kFast>() ? capped_lds(s.threads(), capped_waves<L, kTRT, ST, kFast>()) : 0;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, s.grid, s.block, lds, stream, static_cast<const S*>(cur), static_cast<S*>(nxt),
                           s.cells, static_cast<size_t>(plane));
    };
    const bool wide = wide_planes<S>(s);
    if (s.threads() <= 256)
        wide ? go(pull_stream_kernel<L, ST, 256, true, SHIFT>) : go(pull_stream_kernel<L, ST, 256, false, SHIFT>);
    else
        wide ? go(pull_stream_kernel<L, ST, 1024, true, SHIFT>) : go(pull_stream_kernel<L, ST, 1024, false, SHIFT>);
    return hipGetLastError();
}

template <class L, bool SHIFT>
hipError_t stream_storage(int storage, const void* cur, void* nxt, const LaunchShape& s, uint64_t plane,
                          hipStream_t st) {
    switch (storage) {
        case kF32: return launch_stream<L, kF32, SHIFT>(cur, nxt, s, plane, st);
        case kF64: return launch_stream<L, kF64, SHIFT>(cur, nxt, s, plane, st);
        case kF16: return launch_stream<L, kF16, SHIFT>(cur, nxt, s, plane, st);
        case kBF16: return launch_stream<L, kBF16, SHIFT>(cur, nxt, s, plane, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t plane_layout(int lattice, LaunchShape& s, uint64_t plane_stride, uint64_t slab_stride, uint64_t& plane);

template <bool SHIFT>
hipError_t stream_entry(int lattice, int storage, const void* current, void* next, uint64_t nx, uint64_t ny,
                        uint64_t nz, uint64_t pitch, uint64_t slab_stride, uint32_t block_x, uint32_t block_y,
                        uint32_t block_z, void* stream) {
    if (current == nullptr || next == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(lattice), nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (current == next) return hipErrorInvalidValue;
    uint64_t plane = 0;
    e = plane_layout(lattice, s, pitch, slab_stride, plane);
    if (e != hipSuccess) return e;
    const hipStream_t st = as_stream(stream);
    switch (lattice) {
        case kD2Q9: return stream_storage<D2Q9, SHIFT>(storage, current, next, s, plane, st);
        case kD3Q19: return stream_storage<D3Q19, SHIFT>(storage, current, next, s, plane, st);
        case kD3Q27: return stream_storage<D3Q27, SHIFT>(storage, current, next, s, plane, st);
        default: return hipErrorInvalidValue;
    }
}

// Plane layout of a call, in elements.  Population i of cell (x, y, z) lies
// at i * plane_stride + z * slab_stride + y * nx + x.
//  - slab_stride 0 or nx ny: the reference's planes, [Q][Z][Y][X];
//    plane_stride 0 (the cell count) or at least the cell count (pitched).
//  - a larger slab_stride: the planes stored slab by slab, [Z][Q][Y][X] when
//    dense (plane_stride 0, i.e. nx ny, and slab_stride Q nx ny);
//    plane_stride at least nx ny and slab_stride at least Q plane_stride, so
//    no two (population, slab) pairs share an element.
// A lane's offset within its plane, (nz - 1) slab_stride + nx ny - 1, must stay
// below 2^32 elements; plane_stride at most 2^51 (27 planes of 8-byte values
// stay below 2^59 bytes).
hipError_t plane_layout(int lattice, LaunchShape& s, uint64_t plane_stride, uint64_t slab_stride, uint64_t& plane) {
    const uint64_t cells = s.cells.count;
    const uint64_t slab = static_cast<uint64_t>(s.cells.nx) * s.cells.ny;
    const uint64_t q = lattice == kD2Q9 ? 9 : lattice == kD3Q19 ? 19 : lattice == kD3Q27 ? 27 : 0;
    if (q == 0) return hipErrorInvalidValue;
    if (slab_stride == 0 || slab_stride == slab) {
        plane = plane_stride == 0 ? cells : plane_stride;
        if (plane < cells || plane > (1ull << 51)) return hipErrorInvalidValue;
        s.cells.zgap = 0;
        return hipSuccess;
    }
    plane = plane_stride == 0 ? slab : plane_stride;
    if (plane < slab || plane > (1ull << 51) || slab_stride / q < plane) return hipErrorInvalidValue;
    if (slab_stride - slab >= (1ull << 32)) return hipErrorInvalidValue;
    if (s.cells.nz > 1 && slab_stride > (0xFFFFFFFFull - slab) / (s.cells.nz - 1)) return hipErrorInvalidValue;
    s.cells.zgap = static_cast<uint32_t>(slab_stride - slab);
    return hipSuccess;
}

// One side of a layout conversion: population i of cell (x, y, z) at element
// i * plane + (z ny + y) nx + x + t * zgap, with t = (z ny + y) / rows (the
// tile of Cells; rows = ny for the strided layouts, whose tiles are slabs).
struct LayoutSide {
    uint64_t plane, zgap;
    FastDiv by_rows;
};

// Copies every population of every cell from one layout to another (e.g. the
// reference's [Q][Z][Y][X] planes to the slab order [Z][Q][Y][X], or to 2D
// tiles of rows, and back): one work-item per cell, x fastest, so both sides
// are read and written in runs of consecutive elements.  64-bit offsets: a
// layout conversion, not a step.
template <class T>
__global__ __launch_bounds__(256) void relayout_kernel(const T* __restrict__ src, T* __restrict__ dst, Cells c,
                                                       int q, LayoutSide from, LayoutSide to) {
    const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= c.count) return;
    const CellCoord p = cell_coord(cell, c);
    const uint32_t row = static_cast<uint32_t>(p.z) * c.ny + static_cast<uint32_t>(p.y);
    const uint64_t ts = fast_div(row, from.by_rows), td = fast_div(row, to.by_rows);
    const uint64_t a = cell + ts * from.zgap, b = cell + td * to.zgap;
    for (int i = 0; i < q; ++i)
        dst[static_cast<uint64_t>(i) * to.plane + b] = src[static_cast<uint64_t>(i) * from.plane + a];
}

hipError_t launch_relayout(int storage, const void* src, void* dst, const LaunchShape& s, int q, LayoutSide from,
                           LayoutSide to, hipStream_t st) {
    auto go = [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL(relayout_kernel<T>, s.grid, s.block, 0, st, static_cast<const T*>(src),
                           static_cast<T*>(dst), s.cells, q, from, to);
        return hipGetLastError();
    };
    switch (storage) {
        case kF32: return go(uint32_t{});
        case kF64: return go(uint64_t{});
        case kF16:
        case kBF16: return go(uint16_t{});
        default: return hipErrorInvalidValue;
    }
}

// A layout side from strides (lbc_relayout) or from rows per tile
// (lbc_relayout_tiled: rows = 0 is the reference's planes).
hipError_t strided_side(int lattice, const LaunchShape& shape, uint64_t plane_stride, uint64_t slab_stride,
                        LayoutSide& side) {
    LaunchShape s = shape;
    uint64_t plane = 0;
    const hipError_t e = plane_layout(lattice, s, plane_stride, slab_stride, plane);
    side = LayoutSide{plane, s.cells.zgap, s.cells.by_rows};
    return e;
}

hipError_t tiled_side(int lattice, const LaunchShape& shape, uint64_t rows, LayoutSide& side) {
    LaunchShape s = shape;
    uint64_t plane = 0;
    const int q = q_of(lattice);
    if (q == 0) return hipErrorInvalidValue;
    const hipError_t e = tile_layout(is_2d(lattice), rows, q, s, plane);
    side = LayoutSide{plane, s.cells.zgap, s.cells.by_rows};
    return e;
}

template <class SideFn>
hipError_t relayout_entry(int lattice, int storage, const void* src, void* dst, uint64_t nx, uint64_t ny,
                          uint64_t nz, void* stream, SideFn sides) {
    if (src == nullptr || dst == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(lattice), nx, ny, nz, 256, 1, 1, s);
    if (e != hipSuccess) return e;
    if (src == dst) return hipErrorInvalidValue;
    LayoutSide from{}, to{};
    e = sides(s, from, to);
    if (e != hipSuccess) return e;
    return launch_relayout(storage, src, dst, s, q_of(lattice), from, to, as_stream(stream));
}

// tiled: dense tiles of `tile_rows` rows (tile_layout); otherwise the strided
// layout (pitch, slab_stride) of plane_layout.
hipError_t fluid_entry(int lattice, int collision, int storage, uint32_t flags, const void* current, void* next,
                       uint64_t nx, uint64_t ny, uint64_t nz, uint64_t pitch, uint64_t slab_stride, double omega,
                       uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream, bool tiled = false,
                       uint64_t tile_rows = 0) {
    if (current == nullptr || next == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(lattice), nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (current == next) return hipErrorInvalidValue;
    if (flags & ~static_cast<uint32_t>(LBC_FLAG_FAST | LBC_FLAG_MFMA)) return hipErrorInvalidValue;
    uint64_t plane = 0;
    if (tiled) {
        if (q_of(lattice) == 0 || lattice == kD2Q5 || lattice == kD3Q7) return hipErrorInvalidValue;
        e = tile_layout(is_2d(lattice), tile_rows, q_of(lattice), s, plane);
    } else {
        e = plane_layout(lattice, s, pitch, slab_stride, plane);
    }
    if (e != hipSuccess) return e;
    const bool fast = (flags & LBC_FLAG_FAST) != 0;
    const hipStream_t st = as_stream(stream);
    if (flags & LBC_FLAG_MFMA) return dispatch_mfma(lattice, collision, storage, flags, current, next, s, plane, omega, st);
    switch (lattice) {
        case kD2Q9: return dispatch_collision<D2Q9>(collision, storage, fast, current, next, s, plane, omega, st);
        case kD3Q19: return dispatch_collision<D3Q19>(collision, storage, fast, current, next, s, plane, omega, st);
        case kD3Q27: return dispatch_collision<D3Q27>(collision, storage, fast, current, next, s, plane, omega, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lbc

extern "C" int lbc_fluid_step_ex(int lattice, int collision, int storage, uint32_t flags, const void* current,
                                 void* next, uint64_t nx, uint64_t ny, uint64_t nz, double omega, uint32_t block_x,
                                 uint32_t block_y, uint32_t block_z, void* stream) {
    return lbc::fluid_entry(lattice, collision, storage, flags, current, next, nx, ny, nz, 0, 0, omega, block_x,
                            block_y, block_z, stream);
}

extern "C" int lbc_fluid_step_pitched(int lattice, int collision, int storage, uint32_t flags, const void* current,
                                      void* next, uint64_t nx, uint64_t ny, uint64_t nz, uint64_t plane_pitch,
                                      double omega, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                                      void* stream) {
    return lbc::fluid_entry(lattice, collision, storage, flags, current, next, nx, ny, nz, plane_pitch, 0, omega,
                            block_x, block_y, block_z, stream);
}

extern "C" int lbc_fluid_step_strided(int lattice, int collision, int storage, uint32_t flags, const void* current,
                                      void* next, uint64_t nx, uint64_t ny, uint64_t nz, uint64_t plane_stride,
                                      uint64_t slab_stride, double omega, uint32_t block_x, uint32_t block_y,
                                      uint32_t block_z, void* stream) {
    return lbc::fluid_entry(lattice, collision, storage, flags, current, next, nx, ny, nz, plane_stride, slab_stride,
                            omega, block_x, block_y, block_z, stream);
}

extern "C" int lbc_pull_stream(int lattice, int storage, const void* current, void* next, uint64_t nx, uint64_t ny,
                               uint64_t nz, uint64_t plane_pitch, uint32_t block_x, uint32_t block_y,
                               uint32_t block_z, void* stream) {
    return lbc::stream_entry<true>(lattice, storage, current, next, nx, ny, nz, plane_pitch, 0, block_x, block_y,
                                   block_z, stream);
}

extern "C" int lbc_pull_stream_strided(int lattice, int storage, int shift, const void* current, void* next,
                                       uint64_t nx, uint64_t ny, uint64_t nz, uint64_t plane_stride,
                                       uint64_t slab_stride, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                                       void* stream) {
    return shift ? lbc::stream_entry<true>(lattice, storage, current, next, nx, ny, nz, plane_stride, slab_stride,
                                           block_x, block_y, block_z, stream)
                 : lbc::stream_entry<false>(lattice, storage, current, next, nx, ny, nz, plane_stride, slab_stride,
                                            block_x, block_y, block_z, stream);
}

extern "C" int lbc_plane_copy(int lattice, int storage, const void* current, void* next, uint64_t nx, uint64_t ny,
                              uint64_t nz, uint64_t plane_pitch, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                              void* stream) {
    return lbc::stream_entry<false>(lattice, storage, current, next, nx, ny, nz, plane_pitch, 0, block_x, block_y,
                                    block_z, stream);
}

extern "C" int lbc_fluid_step(int lattice, int collision, int storage, const void* current, void* next, uint64_t nx,
                              uint64_t ny, uint64_t nz, double omega, uint32_t block_x, uint32_t block_y,
                              uint32_t block_z, void* stream) {
    return lbc_fluid_step_ex(lattice, collision, storage, 0u, current, next, nx, ny, nz, omega, block_x, block_y,
                             block_z, stream);
}

extern "C" int lbc_relayout(int lattice, int storage, const void* src, uint64_t src_plane_stride,
                            uint64_t src_slab_stride, void* dst, uint64_t dst_plane_stride, uint64_t dst_slab_stride,
                            uint64_t nx, uint64_t ny, uint64_t nz, void* stream) {
    return lbc::relayout_entry(lattice, storage, src, dst, nx, ny, nz, stream,
                               [&](const lbc::LaunchShape& s, lbc::LayoutSide& from, lbc::LayoutSide& to) {
                                   hipError_t e = lbc::strided_side(lattice, s, src_plane_stride, src_slab_stride, from);
                                   if (e != hipSuccess) return e;
                                   return lbc::strided_side(lattice, s, dst_plane_stride, dst_slab_stride, to);
                               });
}

extern "C" int lbc_relayout_tiled(int lattice, int storage, const void* src, uint64_t src_tile_rows, void* dst,
                                  uint64_t dst_tile_rows, uint64_t nx, uint64_t ny, uint64_t nz, void* stream) {
    return lbc::relayout_entry(lattice, storage, src, dst, nx, ny, nz, stream,
                               [&](const lbc::LaunchShape& s, lbc::LayoutSide& from, lbc::LayoutSide& to) {
                                   hipError_t e = lbc::tiled_side(lattice, s, src_tile_rows, from);
                                   if (e != hipSuccess) return e;
                                   return lbc::tiled_side(lattice, s, dst_tile_rows, to);
                               });
}

extern "C" int lbc_fluid_step_tiled(int lattice, int collision, int storage, uint32_t flags, const void* current,
                                    void* next, uint64_t nx, uint64_t ny, uint64_t nz, uint64_t tile_rows,
                                    double omega, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                                    void* stream) {
    return lbc::fluid_entry(lattice, collision, storage, flags, current, next, nx, ny, nz, 0, 0, omega, block_x,
                            block_y, block_z, stream, true, tile_rows);
}
