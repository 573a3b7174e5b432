// Advection-diffusion(-reaction) kernels for D2Q5 / D3Q7 scalars carried by a
// D2Q9 / D3Q19 / D3Q27 fluid — the gfx950 replacements of LB-Cube's
// kernel_scalar_step (src/lattice_cuda.cu:234-327) and kernel_reaction_AB
// (src/lattice_cuda.cu:349-457), plus a single-pass kernel that fuses the fluid
// step with the reaction step.
//
// Same mapping as the fluid kernel: one work-item per cell, cells numbered
// linearly along the lanes of a wave64, unit-stride loads and stores of every
// population plane; WIDE instantiations (64-bit per-lane addresses) for planes
// of more than 4 GiB.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "../../include/lbcube.h"
#include "lbc_launch.hpp"
#include "lbc_physics.hpp"
#include "lbc_reactive.hpp"
#include "lbc_storage.hpp"

namespace lbc {

// Scalar step (REACT = false: one species with a uniform source) and the
// A + B -> C step (REACT = true: two species, source from the exact one-step
// reaction increment).  The advecting velocity is the fluid's at the
// destination cell, unshifted (include/lattice_core.hpp:259-265, :381-387).
// TILED: the buffers stored in tiles (tile_layout), the fluid's spaced by
// c.zgap and the species' by zgap_s, each `plane` cells per population and
// tile; tiled grids keep the dispatch order of their blocks, as the fused
// step's do.
template <class F, class S, int ST, bool REACT, int MAXT, bool WIDE, bool TILED>
__global__ __launch_bounds__(MAXT) void scalar_step_kernel(
    const typename Storage<ST>::S* __restrict__ cur_a, typename Storage<ST>::S* __restrict__ nxt_a,
    const typename Storage<ST>::S* __restrict__ cur_b, typename Storage<ST>::S* __restrict__ nxt_b,
    const typename Storage<ST>::S* __restrict__ cur_f, Cells c, uint32_t plane, uint32_t zgap_s,
    typename Storage<ST>::R omega_c, typename Storage<ST>::R param, typename Storage<ST>::R ref_a,
    typename Storage<ST>::R ref_b) {
    using St = Storage<ST>;
    using R = typename St::R;
    const uint32_t block = (!TILED && renumber_blocks<S::D, typename Storage<ST>::S>()) ? xcd_block() : blockIdx.x;
    const uint32_t cell = block * blockDim.x + threadIdx.x;
    if (cell >= c.count) return;
    const CellCoord q = cell_coord(cell, c);
    const int x = q.x, y = q.y, z = q.z;
    const int nx = static_cast<int>(c.nx), ny = static_cast<int>(c.ny), nz = static_cast<int>(c.nz);
    const size_t pl = TILED ? plane : c.count;
    // the fluid's element of this cell in its plane
    const uint32_t fcell = TILED ? locate<S::D>(cell, q, c, c.zgap).dst : cell;
    R fl[F::Q];
#pragma unroll
    for (int i = 0; i < F::Q; ++i)
        fl[i] = St::load(at<WIDE>(plane_ptr(cur_f, i, pl), fcell), kOffsetStorage<ST> ? R(F::w[i]) : R(0));
    const Macro<R> m = macro_state<F, R>(fl);
    const Pull pull = TILED ? locate<S::D>(cell, q, c, zgap_s) : Pull(cell, x, y, z, nx, ny, nz);
    R ga[S::Q], gb[S::Q];
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        const uint32_t src = pull.src(S::cx[i], S::cy[i], S::cz[i]);
        ga[i] = St::load(at<WIDE>(plane_ptr(cur_a, i, pl), src), R(S::w[i]) * ref_a);
        if constexpr (REACT) gb[i] = St::load(at<WIDE>(plane_ptr(cur_b, i, pl), src), R(S::w[i]) * ref_b);
    }
    if constexpr (REACT) {
        const R ca = concentration<S, R>(ga);
        const R cb = concentration<S, R>(gb);
        const R src = reaction_source<R>(ca, cb, param);
        collide_scalar<S, R>(ga, m.ux, m.uy, m.uz, omega_c, src);
        collide_scalar<S, R>(gb, m.ux, m.uy, m.uz, omega_c, src);
    } else {
        collide_scalar<S, R>(ga, m.ux, m.uy, m.uz, omega_c, param);
    }
    const uint32_t scell = TILED ? pull.dst : cell;
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        St::store(at<WIDE>(plane_ptr(nxt_a, i, pl), scell), ga[i], R(S::w[i]) * ref_a);
        if constexpr (REACT) St::store(at<WIDE>(plane_ptr(nxt_b, i, pl), scell), gb[i], R(S::w[i]) * ref_b);
    }
}

// TILED: the three buffers stored in tiles (tile_layout), whose gaps differ
// with the lattices' Q, so the fluid and the species take pulls of their own;
// otherwise one pull serves all three.  Tiled 2D grids keep the dispatch
// order of their blocks (measured at 2048^2, tiles of 4 rows: FP32 39.0k
// MLUPS against 37.2k with the XCD renumbering, FP16 deviation form 76.0k
// against 69.5k), while the reference's planes gain from the renumbering.
template <class F, int CT, class S, int ST, bool FAST, int MAXT, bool WIDE, bool TILED>
__global__ __launch_bounds__(MAXT) void reactive_fused_kernel(
    const typename Storage<ST>::S* __restrict__ cur_f, typename Storage<ST>::S* __restrict__ nxt_f,
    const typename Storage<ST>::S* __restrict__ cur_a, typename Storage<ST>::S* __restrict__ nxt_a,
    const typename Storage<ST>::S* __restrict__ cur_b, typename Storage<ST>::S* __restrict__ nxt_b, Cells c,
    uint32_t plane, uint32_t zgap_s, typename Storage<ST>::R omega, typename Storage<ST>::R omega_minus,
    typename Storage<ST>::R omega_c, typename Storage<ST>::R k_react, typename Storage<ST>::R ref_a,
    typename Storage<ST>::R ref_b) {
    const uint32_t block =
        (!TILED && renumber_blocks<F::D, typename Storage<ST>::S>()) ? xcd_block() : blockIdx.x;
    const uint32_t p = block * blockDim.x + threadIdx.x;
    if (p >= c.count) return;
    const CellCoord q = cell_coord(p, c);
    Pull pf(p, q.x, q.y, q.z, static_cast<int>(c.nx), static_cast<int>(c.ny), static_cast<int>(c.nz));
    Pull ps = pf;
    if constexpr (TILED) {  // the fluid's tiles are spaced by c.zgap, the species' by zgap_s
        pf = locate<F::D>(p, q, c, c.zgap);
        ps = locate<F::D>(p, q, c, zgap_s);
    }
    // the reference's planes: the cell count, a 32-bit value (as the address
    // arithmetic of the untiled kernels has always seen it)
    const size_t pl = TILED ? plane : c.count;
    if constexpr (FAST)
        reactive_cell_fast<F, CT, S, ST, WIDE>(cur_f, nxt_f, cur_a, nxt_a, cur_b, nxt_b, pf, ps, pl, omega,
                                               omega_minus, omega_c, k_react, ref_a, ref_b);
    else
        reactive_cell<F, CT, S, ST, WIDE>(cur_f, nxt_f, cur_a, nxt_a, cur_b, nxt_b, pf, ps, pl, omega, omega_minus,
                                          omega_c, k_react, ref_a, ref_b);
}

// The fused step's access pattern alone: every fluid and species population
// pulled from x - c_i (SHIFT) or read in place, and stored, with the fused
// kernel's addressing, block order, launch shape and residency and no
// arithmetic.  Its rate is the ceiling of the fused step's memory pattern on
// the same buffers (the reference's planes or tiles).
template <class F, class S, int ST, int MAXT, bool WIDE, bool TILED, bool SHIFT>
__global__ __launch_bounds__(MAXT) void reactive_pull_kernel(
    const typename Storage<ST>::S* __restrict__ cur_f, typename Storage<ST>::S* __restrict__ nxt_f,
    const typename Storage<ST>::S* __restrict__ cur_a, typename Storage<ST>::S* __restrict__ nxt_a,
    const typename Storage<ST>::S* __restrict__ cur_b, typename Storage<ST>::S* __restrict__ nxt_b, Cells c,
    uint32_t plane, uint32_t zgap_s) {
    using T = typename Storage<ST>::S;
    const uint32_t block = (!TILED && renumber_blocks<F::D, T>()) ? xcd_block() : blockIdx.x;
    const uint32_t p = block * blockDim.x + threadIdx.x;
    if (p >= c.count) return;
    const CellCoord q = cell_coord(p, c);
    Pull pf(p, q.x, q.y, q.z, static_cast<int>(c.nx), static_cast<int>(c.ny), static_cast<int>(c.nz));
    Pull ps = pf;
    if constexpr (TILED) {
        pf = locate<F::D>(p, q, c, c.zgap);
        ps = locate<F::D>(p, q, c, zgap_s);
    }
    const size_t pl = TILED ? plane : c.count;
    T v[F::Q];
#pragma unroll
    for (int i = 0; i < F::Q; ++i)
        v[i] = get(at<WIDE>(plane_ptr(cur_f, i, pl), SHIFT ? pf.src(F::cx[i], F::cy[i], F::cz[i]) : pf.dst));
#pragma unroll
    for (int i = 0; i < F::Q; ++i) put(at<WIDE>(plane_ptr(nxt_f, i, pl), pf.dst), v[i]);
    T va[S::Q], vb[S::Q];
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        const uint32_t src = SHIFT ? ps.src(S::cx[i], S::cy[i], S::cz[i]) : ps.dst;
        va[i] = get(at<WIDE>(plane_ptr(cur_a, i, pl), src));
        vb[i] = get(at<WIDE>(plane_ptr(cur_b, i, pl), src));
    }
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        put(at<WIDE>(plane_ptr(nxt_a, i, pl), ps.dst), va[i]);
        put(at<WIDE>(plane_ptr(nxt_b, i, pl), ps.dst), vb[i]);
    }
}

// Residency cap (kCappedWavesPerCu, lbc_launch.hpp), measured on the 256^3
// TGV with two D3Q7 species.  FP32: scalar step 41.2k -> 45.0k MLUPS,
// reaction step 30.9k -> 31.8k, fused TRT step 21.0k -> 21.3k (both forms).
// FP16: the fused step gains 3-4.5% (39.2k -> 40.4k, deviation form 39.3k ->
// 41.1k), the scalar step loses 7%, the reaction step gains 1.4% (left
// uncapped).  BF16 loses up to 15% and FP64 is unchanged, so both stay
// uncapped, as do 2D grids (the fused 2048^2 step lost 1-5%).
template <int D, int ST, bool FUSED> constexpr bool capped_residency_adr() {
    return D == 3 && (ST == kF32 || (FUSED && ST == kF16));
}

template <class F, class S, int ST, bool REACT>
hipError_t launch_scalar(const void* ca, void* na, const void* cb, void* nb, const void* cf, const LaunchShape& s,
                         double omega_c, double param, double ref_a, double ref_b, hipStream_t stream) {
    using St = Storage<ST>;
    using T = typename St::S;
    using R = typename St::R;
    // `s` carries the tiles of the call (tile_layout, with the fluid's gap);
    // one 2D tile holding every row is the reference's planes
    const bool tiled = s.cells.zgap != 0 && (S::D == 3 || s.cells.tiles > 1);
    const uint32_t plane = tiled ? s.cells.zgap / (F::Q - 1) : s.cells.count;
    const uint32_t zgap_s = tiled ? (S::Q - 1) * plane : 0u;
    const uint32_t lds = capped_residency_adr<S::D, ST, false>() ? capped_lds(s.threads()) : 0;
    auto args = [&](auto kern) {
        hipLaunchKernelGGL(kern, s.grid, s.block, lds, stream, static_cast<const T*>(ca), static_cast<T*>(na),
                           static_cast<const T*>(cb), static_cast<T*>(nb), static_cast<const T*>(cf), s.cells, plane,
                           zgap_s, static_cast<R>(omega_c), static_cast<R>(param), static_cast<R>(ref_a),
                           static_cast<R>(ref_b));
    };
    const bool wide = wide_planes<T>(s);
    auto pick = [&](auto tl) {
        constexpr bool TL = decltype(tl)::value;
        if (s.threads() <= 256)
            wide ? args(scalar_step_kernel<F, S, ST, REACT, 256, true, TL>)
                 : args(scalar_step_kernel<F, S, ST, REACT, 256, false, TL>);
        else
            wide ? args(scalar_step_kernel<F, S, ST, REACT, 1024, true, TL>)
                 : args(scalar_step_kernel<F, S, ST, REACT, 1024, false, TL>);
    };
    tiled ? pick(std::true_type{}) : pick(std::false_type{});
    return hipGetLastError();
}

template <class F, class S, bool REACT>
hipError_t scalar_by_storage(int storage, const void* ca, void* na, const void* cb, void* nb, const void* cf,
                             const LaunchShape& s, double omega_c, double param, double ra, double rb,
                             hipStream_t st) {
    switch (storage) {
        case kF32: return launch_scalar<F, S, kF32, REACT>(ca, na, cb, nb, cf, s, omega_c, param, ra, rb, st);
        case kF64: return launch_scalar<F, S, kF64, REACT>(ca, na, cb, nb, cf, s, omega_c, param, ra, rb, st);
        case kF16: return launch_scalar<F, S, kF16, REACT>(ca, na, cb, nb, cf, s, omega_c, param, ra, rb, st);
        case kBF16: return launch_scalar<F, S, kBF16, REACT>(ca, na, cb, nb, cf, s, omega_c, param, ra, rb, st);
        default: return hipErrorInvalidValue;
    }
}

template <bool REACT>
hipError_t scalar_dispatch(int fl, int sl, int storage, const void* ca, void* na, const void* cb, void* nb,
                           const void* cf, const LaunchShape& s, double omega_c, double param, double ra, double rb,
                           hipStream_t st) {
    if (fl == kD2Q9 && sl == kD2Q5)
        return scalar_by_storage<D2Q9, D2Q5, REACT>(storage, ca, na, cb, nb, cf, s, omega_c, param, ra, rb, st);
    if (fl == kD3Q19 && sl == kD3Q7)
        return scalar_by_storage<D3Q19, D3Q7, REACT>(storage, ca, na, cb, nb, cf, s, omega_c, param, ra, rb, st);
    if (fl == kD3Q27 && sl == kD3Q7)
        return scalar_by_storage<D3Q27, D3Q7, REACT>(storage, ca, na, cb, nb, cf, s, omega_c, param, ra, rb, st);
    return hipErrorInvalidValue;
}

// `s` carries the tiles of the call (tile_layout) with the fluid's gap; the
// species' gap follows from their Q.
template <class F, int CT, class S, int ST, bool FAST>
hipError_t launch_fused(const void* cf, void* nf, const void* ca, void* na, const void* cb, void* nb,
                        const LaunchShape& s, double omega, double omega_c, double k, double ra, double rb,
                        hipStream_t stream) {
    // one 2D tile holding every row is the reference's planes
    const bool tiled = s.cells.zgap != 0 && (F::D == 3 || s.cells.tiles > 1);
    const uint32_t plane = tiled ? s.cells.zgap / (F::Q - 1) : s.cells.count;  // a tile's cells, or all cells
    const uint32_t zgap_s = tiled ? (S::Q - 1) * plane : 0u;
    using St = Storage<ST>;
    using T = typename St::S;
    using R = typename St::R;
    const R w = static_cast<R>(omega);
    const R wm = CT == kTRT ? trt_omega_minus<R>(w) : R(0);
    const uint32_t lds = capped_residency_adr<F::D, ST, true>() ? capped_lds(s.threads()) : 0;
    auto args = [&](auto kern) {
        hipLaunchKernelGGL(kern, s.grid, s.block, lds, stream, static_cast<const T*>(cf), static_cast<T*>(nf),
                           static_cast<const T*>(ca), static_cast<T*>(na), static_cast<const T*>(cb),
                           static_cast<T*>(nb), s.cells, plane, zgap_s, w, wm, static_cast<R>(omega_c),
                           static_cast<R>(k), static_cast<R>(ra), static_cast<R>(rb));
    };
    const bool wide = wide_planes<T>(s);
    auto pick = [&](auto tiled) {
        constexpr bool TL = decltype(tiled)::value;
        if (s.threads() <= 256)
            wide ? args(reactive_fused_kernel<F, CT, S, ST, FAST, 256, true, TL>)
                 : args(reactive_fused_kernel<F, CT, S, ST, FAST, 256, false, TL>);
        else
            wide ? args(reactive_fused_kernel<F, CT, S, ST, FAST, 1024, true, TL>)
                 : args(reactive_fused_kernel<F, CT, S, ST, FAST, 1024, false, TL>);
    };
    tiled ? pick(std::true_type{}) : pick(std::false_type{});
    return hipGetLastError();
}

template <class F, class S, int CT, bool FAST>
hipError_t fused_by_storage(int storage, const void* cf, void* nf, const void* ca, void* na, const void* cb,
                            void* nb, const LaunchShape& s, double omega, double omega_c, double k, double ra,
                            double rb, hipStream_t st) {
    switch (storage) {
        case kF32: return launch_fused<F, CT, S, kF32, FAST>(cf, nf, ca, na, cb, nb, s, omega, omega_c, k, ra, rb, st);
        case kF64: return launch_fused<F, CT, S, kF64, FAST>(cf, nf, ca, na, cb, nb, s, omega, omega_c, k, ra, rb, st);
        case kF16: return launch_fused<F, CT, S, kF16, FAST>(cf, nf, ca, na, cb, nb, s, omega, omega_c, k, ra, rb, st);
        case kBF16:
            return launch_fused<F, CT, S, kBF16, FAST>(cf, nf, ca, na, cb, nb, s, omega, omega_c, k, ra, rb, st);
        default: return hipErrorInvalidValue;
    }
}

template <class F, class S>
hipError_t fused_by_collision(int collision, int storage, bool fast, const void* cf, void* nf, const void* ca,
                              void* na, const void* cb, void* nb, const LaunchShape& s, double omega, double omega_c,
                              double k, double ra, double rb, hipStream_t st) {
#define LBC_FUSED(CT, FAST) fused_by_storage<F, S, CT, FAST>(storage, cf, nf, ca, na, cb, nb, s, omega, omega_c, k, ra, rb, st)
    switch (collision) {
        case kBGK: return fast ? LBC_FUSED(kBGK, true) : LBC_FUSED(kBGK, false);
        case kTRT: return fast ? LBC_FUSED(kTRT, true) : LBC_FUSED(kTRT, false);
        case kMRT: return fast ? LBC_FUSED(kMRT, true) : LBC_FUSED(kMRT, false);
        case kRLBM: return LBC_FUSED(kRLBM, false);
        default: return hipErrorInvalidValue;
    }
#undef LBC_FUSED
}

template <class F, class S, int ST, bool SHIFT>
hipError_t launch_reactive_pull(const void* cf, void* nf, const void* ca, void* na, const void* cb, void* nb,
                                const LaunchShape& s, hipStream_t stream) {
    using T = typename Storage<ST>::S;
    const bool tiled = s.cells.zgap != 0 && (F::D == 3 || s.cells.tiles > 1);
    const uint32_t plane = tiled ? s.cells.zgap / (F::Q - 1) : s.cells.count;
    const uint32_t zgap_s = tiled ? (S::Q - 1) * plane : 0u;
    const uint32_t lds = capped_residency_adr<F::D, ST, true>() ? capped_lds(s.threads()) : 0;
    auto args = [&](auto kern) {
        hipLaunchKernelGGL(kern, s.grid, s.block, lds, stream, static_cast<const T*>(cf), static_cast<T*>(nf),
                           static_cast<const T*>(ca), static_cast<T*>(na), static_cast<const T*>(cb),
                           static_cast<T*>(nb), s.cells, plane, zgap_s);
    };
    const bool wide = wide_planes<T>(s);
    auto pick = [&](auto tl) {
        constexpr bool TL = decltype(tl)::value;
        if (s.threads() <= 256)
            wide ? args(reactive_pull_kernel<F, S, ST, 256, true, TL, SHIFT>)
                 : args(reactive_pull_kernel<F, S, ST, 256, false, TL, SHIFT>);
        else
            wide ? args(reactive_pull_kernel<F, S, ST, 1024, true, TL, SHIFT>)
                 : args(reactive_pull_kernel<F, S, ST, 1024, false, TL, SHIFT>);
    };
    tiled ? pick(std::true_type{}) : pick(std::false_type{});
    return hipGetLastError();
}

template <class F, class S, bool SHIFT>
hipError_t reactive_pull_by_storage(int storage, const void* cf, void* nf, const void* ca, void* na, const void* cb,
                                    void* nb, const LaunchShape& s, hipStream_t st) {
    switch (storage) {
        case kF32: return launch_reactive_pull<F, S, kF32, SHIFT>(cf, nf, ca, na, cb, nb, s, st);
        case kF64: return launch_reactive_pull<F, S, kF64, SHIFT>(cf, nf, ca, na, cb, nb, s, st);
        case kF16: return launch_reactive_pull<F, S, kF16, SHIFT>(cf, nf, ca, na, cb, nb, s, st);
        case kBF16: return launch_reactive_pull<F, S, kBF16, SHIFT>(cf, nf, ca, na, cb, nb, s, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lbc

namespace lbc {

// The tiles of a call's buffers (tile_rows 0: the reference's planes), with
// the fluid's gap; the fluid lattice must be a fluid.
hipError_t scalar_tiles(int fluid_lattice, int scalar_lattice, uint64_t tile_rows, LaunchShape& s) {
    if (tile_rows == 0) return hipSuccess;
    const int qf = q_of(fluid_lattice);
    if (qf == 0 || fluid_lattice == kD2Q5 || fluid_lattice == kD3Q7) return hipErrorInvalidValue;
    uint64_t plane = 0;
    return tile_layout(is_2d(scalar_lattice), tile_rows, qf, s, plane);
}

hipError_t scalar_entry(int fluid_lattice, int scalar_lattice, int storage, const void* current_scalar,
                        void* next_scalar, const void* current_fluid, uint64_t nx, uint64_t ny, uint64_t nz,
                        uint64_t tile_rows, double omega_c, double source_term, double reference, uint32_t block_x,
                        uint32_t block_y, uint32_t block_z, void* stream) {
    if (current_scalar == nullptr || next_scalar == nullptr || current_fluid == nullptr)
        return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(scalar_lattice), nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (current_scalar == next_scalar || next_scalar == current_fluid) return hipErrorInvalidValue;
    if ((e = scalar_tiles(fluid_lattice, scalar_lattice, tile_rows, s)) != hipSuccess) return e;
    return scalar_dispatch<false>(fluid_lattice, scalar_lattice, storage, current_scalar, next_scalar, nullptr,
                                  nullptr, current_fluid, s, omega_c, source_term, reference, 0.0, as_stream(stream));
}

hipError_t reaction_entry(int fluid_lattice, int scalar_lattice, int storage, const void* current_a, void* next_a,
                          const void* current_b, void* next_b, const void* current_fluid, uint64_t nx, uint64_t ny,
                          uint64_t nz, uint64_t tile_rows, double omega_c, double k_react, double reference_a,
                          double reference_b, uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream) {
    if (current_a == nullptr || next_a == nullptr || current_b == nullptr || next_b == nullptr ||
        current_fluid == nullptr)
        return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(scalar_lattice), nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (current_a == next_a || current_b == next_b || next_a == next_b) return hipErrorInvalidValue;
    if ((e = scalar_tiles(fluid_lattice, scalar_lattice, tile_rows, s)) != hipSuccess) return e;
    return scalar_dispatch<true>(fluid_lattice, scalar_lattice, storage, current_a, next_a, current_b, next_b,
                                 current_fluid, s, omega_c, k_react, reference_a, reference_b, as_stream(stream));
}

}  // namespace lbc

extern "C" int lbc_scalar_step(int fluid_lattice, int scalar_lattice, int storage, const void* current_scalar,
                               void* next_scalar, const void* current_fluid, uint64_t nx, uint64_t ny, uint64_t nz,
                               double omega_c, double source_term, double reference, uint32_t block_x,
                               uint32_t block_y, uint32_t block_z, void* stream) {
    return lbc::scalar_entry(fluid_lattice, scalar_lattice, storage, current_scalar, next_scalar, current_fluid, nx,
                             ny, nz, 0, omega_c, source_term, reference, block_x, block_y, block_z, stream);
}

extern "C" int lbc_scalar_step_tiled(int fluid_lattice, int scalar_lattice, int storage, const void* current_scalar,
                                     void* next_scalar, const void* current_fluid, uint64_t nx, uint64_t ny,
                                     uint64_t nz, uint64_t tile_rows, double omega_c, double source_term,
                                     double reference, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                                     void* stream) {
    return lbc::scalar_entry(fluid_lattice, scalar_lattice, storage, current_scalar, next_scalar, current_fluid, nx,
                             ny, nz, tile_rows, omega_c, source_term, reference, block_x, block_y, block_z, stream);
}

extern "C" int lbc_reaction_ab_step(int fluid_lattice, int scalar_lattice, int storage, const void* current_a,
                                    void* next_a, const void* current_b, void* next_b, const void* current_fluid,
                                    uint64_t nx, uint64_t ny, uint64_t nz, double omega_c, double k_react,
                                    double reference_a, double reference_b, uint32_t block_x, uint32_t block_y,
                                    uint32_t block_z, void* stream) {
    return lbc::reaction_entry(fluid_lattice, scalar_lattice, storage, current_a, next_a, current_b, next_b,
                               current_fluid, nx, ny, nz, 0, omega_c, k_react, reference_a, reference_b, block_x,
                               block_y, block_z, stream);
}

extern "C" int lbc_reaction_ab_step_tiled(int fluid_lattice, int scalar_lattice, int storage, const void* current_a,
                                          void* next_a, const void* current_b, void* next_b,
                                          const void* current_fluid, uint64_t nx, uint64_t ny, uint64_t nz,
                                          uint64_t tile_rows, double omega_c, double k_react, double reference_a,
                                          double reference_b, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                                          void* stream) {
    return lbc::reaction_entry(fluid_lattice, scalar_lattice, storage, current_a, next_a, current_b, next_b,
                               current_fluid, nx, ny, nz, tile_rows, omega_c, k_react, reference_a, reference_b,
                               block_x, block_y, block_z, stream);
}

namespace lbc {

// The fused step in the reference's planes (tile_rows = 0) or in dense tiles
// of tile_rows rows (tile_layout) for all three buffers.
hipError_t fused_entry(int fluid_lattice, int collision, int scalar_lattice, int storage, uint32_t flags,
                       const void* current_fluid, void* next_fluid, const void* current_a, void* next_a,
                       const void* current_b, void* next_b, uint64_t nx, uint64_t ny, uint64_t nz,
                       uint64_t tile_rows, double omega, double omega_c, double k_react, double reference_a,
                       double reference_b, uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream) {
    if (current_fluid == nullptr || next_fluid == nullptr || current_a == nullptr || next_a == nullptr ||
        current_b == nullptr || next_b == nullptr)
        return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(scalar_lattice), nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (current_fluid == next_fluid || current_a == next_a || current_b == next_b || next_a == next_b)
        return hipErrorInvalidValue;
    if (flags & ~static_cast<uint32_t>(LBC_FLAG_FAST)) return hipErrorInvalidValue;
    if (tile_rows != 0) {
        const int qf = q_of(fluid_lattice);
        if (qf == 0 || fluid_lattice == kD2Q5 || fluid_lattice == kD3Q7) return hipErrorInvalidValue;
        uint64_t plane = 0;
        e = tile_layout(is_2d(scalar_lattice), tile_rows, qf, s, plane);
        if (e != hipSuccess) return e;
    }
    const bool fast = (flags & LBC_FLAG_FAST) != 0;
    const hipStream_t st = as_stream(stream);
    if (fluid_lattice == kD2Q9 && scalar_lattice == kD2Q5)
        return fused_by_collision<D2Q9, D2Q5>(collision, storage, fast, current_fluid, next_fluid, current_a, next_a,
                                              current_b, next_b, s, omega, omega_c, k_react, reference_a,
                                              reference_b, st);
    if (fluid_lattice == kD3Q19 && scalar_lattice == kD3Q7)
        return fused_by_collision<D3Q19, D3Q7>(collision, storage, fast, current_fluid, next_fluid, current_a,
                                               next_a, current_b, next_b, s, omega, omega_c, k_react, reference_a,
                                              reference_b, st);
    if (fluid_lattice == kD3Q27 && scalar_lattice == kD3Q7)
        return fused_by_collision<D3Q27, D3Q7>(collision, storage, fast, current_fluid, next_fluid, current_a,
                                               next_a, current_b, next_b, s, omega, omega_c, k_react, reference_a,
                                              reference_b, st);
    return hipErrorInvalidValue;
}

}  // namespace lbc

extern "C" int lbc_reactive_fused_step_ex(int fluid_lattice, int collision, int scalar_lattice, int storage,
                                          uint32_t flags, const void* current_fluid, void* next_fluid,
                                          const void* current_a, void* next_a, const void* current_b, void* next_b,
                                          uint64_t nx, uint64_t ny, uint64_t nz, double omega, double omega_c,
                                          double k_react, double reference_a, double reference_b, uint32_t block_x,
                                          uint32_t block_y, uint32_t block_z, void* stream) {
    return lbc::fused_entry(fluid_lattice, collision, scalar_lattice, storage, flags, current_fluid, next_fluid,
                            current_a, next_a, current_b, next_b, nx, ny, nz, 0, omega, omega_c, k_react,
                            reference_a, reference_b, block_x, block_y, block_z, stream);
}

extern "C" int lbc_reactive_fused_step_tiled(int fluid_lattice, int collision, int scalar_lattice, int storage,
                                             uint32_t flags, const void* current_fluid, void* next_fluid,
                                             const void* current_a, void* next_a, const void* current_b,
                                             void* next_b, uint64_t nx, uint64_t ny, uint64_t nz, uint64_t tile_rows,
                                             double omega, double omega_c, double k_react, double reference_a,
                                             double reference_b, uint32_t block_x, uint32_t block_y,
                                             uint32_t block_z, void* stream) {
    return lbc::fused_entry(fluid_lattice, collision, scalar_lattice, storage, flags, current_fluid, next_fluid,
                            current_a, next_a, current_b, next_b, nx, ny, nz, tile_rows, omega, omega_c, k_react,
                            reference_a, reference_b, block_x, block_y, block_z, stream);
}

extern "C" int lbc_reactive_fused_step(int fluid_lattice, int collision, int scalar_lattice, int storage,
                                       const void* current_fluid, void* next_fluid, const void* current_a,
                                       void* next_a, const void* current_b, void* next_b, uint64_t nx, uint64_t ny,
                                       uint64_t nz, double omega, double omega_c, double k_react,
                                       double reference_a, double reference_b, uint32_t block_x, uint32_t block_y,
                                       uint32_t block_z, void* stream) {
    return lbc_reactive_fused_step_ex(fluid_lattice, collision, scalar_lattice, storage, 0u, current_fluid,
                                      next_fluid, current_a, next_a, current_b, next_b, nx, ny, nz, omega, omega_c,
                                      k_react, reference_a, reference_b, block_x, block_y, block_z, stream);
}

extern "C" int lbc_reactive_pull_stream(int fluid_lattice, int scalar_lattice, int storage, int shift,
                                        const void* current_fluid, void* next_fluid, const void* current_a,
                                        void* next_a, const void* current_b, void* next_b, uint64_t nx, uint64_t ny,
                                        uint64_t nz, uint64_t tile_rows, uint32_t block_x, uint32_t block_y,
                                        uint32_t block_z, void* stream) {
    using namespace lbc;
    if (current_fluid == nullptr || next_fluid == nullptr || current_a == nullptr || next_a == nullptr ||
        current_b == nullptr || next_b == nullptr)
        return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(scalar_lattice), nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (current_fluid == next_fluid || current_a == next_a || current_b == next_b || next_a == next_b)
        return hipErrorInvalidValue;
    const int qf = q_of(fluid_lattice);
    if (qf == 0 || fluid_lattice == kD2Q5 || fluid_lattice == kD3Q7) return hipErrorInvalidValue;
    uint64_t plane = 0;
    e = tile_layout(is_2d(scalar_lattice), tile_rows, qf, s, plane);
    if (e != hipSuccess) return e;
    const hipStream_t st = as_stream(stream);
    auto go = [&](auto f, auto sc) -> hipError_t {
        using F = decltype(f);
        using S = decltype(sc);
        return shift ? reactive_pull_by_storage<F, S, true>(storage, current_fluid, next_fluid, current_a, next_a,
                                                           current_b, next_b, s, st)
                     : reactive_pull_by_storage<F, S, false>(storage, current_fluid, next_fluid, current_a, next_a,
                                                            current_b, next_b, s, st);
    };
    if (fluid_lattice == kD2Q9 && scalar_lattice == kD2Q5) return go(D2Q9{}, D2Q5{});
    if (fluid_lattice == kD3Q19 && scalar_lattice == kD3Q7) return go(D3Q19{}, D3Q7{});
    if (fluid_lattice == kD3Q27 && scalar_lattice == kD3Q7) return go(D3Q27{}, D3Q7{});
    return hipErrorInvalidValue;
}
