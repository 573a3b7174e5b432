// lattice_hip.hpp — C++ launcher API of the MI355X kernels, mirroring LB-Cube's
// include/lattice_cuda.cuh (names, template parameters, argument order and
// defaults).  It defines no lattice types: the launchers take any lattice
// traits type shaped like the reference's (include/lattice_traits.hpp:62-305:
// static D, Q, cs2 and, when present, the direction table `c`), so a reference
// driver keeps its own lattice_traits.hpp and LatticeMemory and switches by
// including this header instead of lattice_cuda.cuh and renaming cuda* runtime
// calls to hip*.  Programs without the reference's headers take the same
// traits from lbcube_lattices.hpp.  Every launcher is an inline forwarder to
// the C ABI of lbcube.h; the caller compiles no device code.
//
//   launch_step_gpu<L, Real, CT>(cur, next, nx, ny, nz, omega, block)
//       lattice_cuda.cuh:179-190, src/lattice_cuda.cu:476-520
//   launch_step_gpu<L, Real, CT>(mem, cur, next, omega, block)
//       lattice_cuda.cuh:310-319, src/lattice_cuda.cu:680-698
//   launch_scalar_step_gpu<F, S, Real>(cur_s, next_s, cur_f, nx, ny, nz, omega_c, source, block)
//       lattice_cuda.cuh:209-222, src/lattice_cuda.cu:539-589
//   launch_scalar_step_gpu<F, S, Real>(scalar_mem, fluid_mem, cur_s, next_s, cur_f, omega_c, source, block)
//       lattice_cuda.cuh:243-255, src/lattice_cuda.cu:716-747
//   launch_reaction_AB_gpu<F, S, Real>(cur_a, next_a, cur_b, next_b, cur_f, nx, ny, nz, omega_c, k, block)
//       lattice_cuda.cuh:276-291, src/lattice_cuda.cu:610-666
//
// The memory overloads accept any class template Memory<Lattice, Real> with
// x_extent(), y_extent() and z_extent() (the reference's LatticeMemory,
// include/lattice_memory.hpp:67-272, or a container of the caller's); as in the
// reference, Lattice and Real are deduced from it or given explicitly, and only
// the extents are read.
//
// Differences: launchers return hipError_t; the default block is
// default_cuda_block_size() = {256, 1, 1} (four wave64s) instead of {16, 8, 1};
// the kernels cover cells linearly, so any block gives the same results and
// only its size matters for speed; an optional trailing hipStream_t selects
// the stream (the reference always uses the default stream).  A traits type
// whose (D, Q) is not one of the five lattices, whose cs2 differs, or whose
// direction table is not in the library's plane order is rejected at compile
// time (the planes of [Q][Z][Y][X] follow that order).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <concepts>
#include <cstddef>
#include <type_traits>

#include "lbcube.h"

namespace lbm {

#ifndef LB_CUBE_COLLISION_TYPE_DEFINED
#define LB_CUBE_COLLISION_TYPE_DEFINED
// CollisionType, include/lattice_core.hpp:33-42 (same enumerator order; the
// reference guards its two copies with this macro).
enum class CollisionType { BGK, TRT, MRT, RLBM };
#endif

namespace hip_detail {

// The kernel library's direction order per lattice (the order of
// include/lattice_traits.hpp:62-305), flattened [q * D + d].
inline constexpr std::array<int, 18> d2q9_c{0, 0, 1, 0, 0, 1, -1, 0, 0, -1, 1, 1, -1, 1, -1, -1, 1, -1};
inline constexpr std::array<int, 57> d3q19_c{
    0, 0, 0,  1, 0, 0,  -1, 0, 0,  0, 1, 0,  0, -1, 0,  0, 0, 1,  0, 0, -1,
    1, 1, 0,  -1, 1, 0,  1, -1, 0,  -1, -1, 0,
    1, 0, 1,  -1, 0, 1,  1, 0, -1,  -1, 0, -1,
    0, 1, 1,  0, -1, 1,  0, 1, -1,  0, -1, -1};
inline constexpr std::array<int, 81> d3q27_c{
    0, 0, 0,  1, 0, 0,  -1, 0, 0,  0, 1, 0,  0, -1, 0,  0, 0, 1,  0, 0, -1,
    1, 1, 0,  -1, 1, 0,  1, -1, 0,  -1, -1, 0,
    1, 0, 1,  -1, 0, 1,  1, 0, -1,  -1, 0, -1,
    0, 1, 1,  0, -1, 1,  0, 1, -1,  0, -1, -1,
    1, 1, 1,  -1, 1, 1,  1, -1, 1,  -1, -1, 1,  1, 1, -1,  -1, 1, -1,  1, -1, -1,  -1, -1, -1};
inline constexpr std::array<int, 10> d2q5_c{0, 0, 1, 0, 0, 1, -1, 0, 0, -1};
inline constexpr std::array<int, 21> d3q7_c{0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};

// ABI code of the lattice with D dimensions and Q populations, or -1.  Scalar
// and fluid lattices differ in (D, Q), so the pair names the table.
constexpr int code_of(int d, int q) noexcept {
    if (d == 2 && q == 9) return LBC_D2Q9;
    if (d == 3 && q == 19) return LBC_D3Q19;
    if (d == 3 && q == 27) return LBC_D3Q27;
    if (d == 2 && q == 5) return LBC_D2Q5;
    if (d == 3 && q == 7) return LBC_D3Q7;
    return -1;
}

constexpr double cs2_of(int code) noexcept { return code == LBC_D3Q7 ? 1.0 / 4.0 : 1.0 / 3.0; }

template <std::size_t N> constexpr bool same_table(const auto& c, const std::array<int, N>& mine) {
    if (c.size() != N) return false;
    for (std::size_t k = 0; k < N; ++k)
        if (static_cast<int>(c[k]) != mine[k]) return false;
    return true;
}

template <class L> constexpr bool directions_match() {
    if constexpr (!requires { L::c; }) {
        return true;  // a traits type without a table: (D, Q) and cs2 decide
    } else {
        constexpr int code = code_of(L::D, L::Q);
        if constexpr (code == LBC_D2Q9) return same_table(L::c, d2q9_c);
        else if constexpr (code == LBC_D3Q19) return same_table(L::c, d3q19_c);
        else if constexpr (code == LBC_D3Q27) return same_table(L::c, d3q27_c);
        else if constexpr (code == LBC_D2Q5) return same_table(L::c, d2q5_c);
        else return same_table(L::c, d3q7_c);
    }
}

template <class M>
concept HasExtents = requires(const M& m) {
    { m.x_extent() } -> std::convertible_to<std::size_t>;
    { m.y_extent() } -> std::convertible_to<std::size_t>;
    { m.z_extent() } -> std::convertible_to<std::size_t>;
};

}  // namespace hip_detail

// A lattice traits type the kernel library implements: static D and Q of one
// of D2Q9, D3Q19, D3Q27, D2Q5, D3Q7, the matching cs2 and, when a direction
// table `c` is present, the library's direction order.
template <class L>
concept HipLattice = requires {
    { L::D } -> std::convertible_to<int>;
    { L::Q } -> std::convertible_to<int>;
    { L::cs2 } -> std::convertible_to<double>;
} && (hip_detail::code_of(L::D, L::Q) >= 0) && (static_cast<double>(L::cs2) == hip_detail::cs2_of(
         hip_detail::code_of(L::D, L::Q))) && hip_detail::directions_match<L>();

// ABI code (enum lbc_lattice of lbcube.h) of a lattice traits type.
template <HipLattice L> constexpr int lattice_code() noexcept { return hip_detail::code_of(L::D, L::Q); }

template <class Real> constexpr int storage_code() noexcept {
    static_assert(std::is_same_v<Real, float> || std::is_same_v<Real, double>, "Real must be float or double");
    return std::is_same_v<Real, float> ? LBC_F32 : LBC_F64;
}

inline dim3 default_cuda_block_size() noexcept { return dim3{256, 1, 1}; }

template <HipLattice Lattice, std::floating_point Real, CollisionType CT = CollisionType::BGK>
hipError_t launch_step_gpu(const Real* current_populations, Real* next_populations, std::size_t x_extent,
                           std::size_t y_extent, std::size_t z_extent, Real omega,
                           dim3 block = default_cuda_block_size(), hipStream_t stream = nullptr) {
    return static_cast<hipError_t>(lbc_fluid_step(lattice_code<Lattice>(), static_cast<int>(CT), storage_code<Real>(),
                                                  current_populations, next_populations, x_extent, y_extent,
                                                  z_extent, static_cast<double>(omega), block.x, block.y, block.z,
                                                  stream));
}

template <HipLattice Lattice, std::floating_point Real, CollisionType CT = CollisionType::BGK,
          template <class, class> class Memory>
    requires hip_detail::HasExtents<Memory<Lattice, Real>>
hipError_t launch_step_gpu(const Memory<Lattice, Real>& mem, const Real* current_populations, Real* next_populations,
                           Real omega, dim3 block = default_cuda_block_size(), hipStream_t stream = nullptr) {
    return launch_step_gpu<Lattice, Real, CT>(current_populations, next_populations, mem.x_extent(), mem.y_extent(),
                                              mem.z_extent(), omega, block, stream);
}

template <HipLattice FluidLattice, HipLattice ScalarLattice, std::floating_point Real>
hipError_t launch_scalar_step_gpu(const Real* current_scalar, Real* next_scalar, const Real* current_fluid,
                                  std::size_t x_extent, std::size_t y_extent, std::size_t z_extent, Real omega_c,
                                  Real source_term = Real{}, dim3 block = default_cuda_block_size(),
                                  hipStream_t stream = nullptr) {
    static_assert(FluidLattice::D == ScalarLattice::D, "fluid and scalar lattices must have the same dimension");
    return static_cast<hipError_t>(lbc_scalar_step(lattice_code<FluidLattice>(), lattice_code<ScalarLattice>(),
                                                   storage_code<Real>(), current_scalar, next_scalar, current_fluid,
                                                   x_extent, y_extent, z_extent, static_cast<double>(omega_c),
                                                   static_cast<double>(source_term), 0.0, block.x, block.y,
                                                   block.z, stream));
}

template <HipLattice FluidLattice, HipLattice ScalarLattice, std::floating_point Real,
          template <class, class> class Memory>
    requires hip_detail::HasExtents<Memory<ScalarLattice, Real>> && hip_detail::HasExtents<Memory<FluidLattice, Real>>
hipError_t launch_scalar_step_gpu(const Memory<ScalarLattice, Real>& scalar_mem,
                                  const Memory<FluidLattice, Real>& fluid_mem, const Real* current_scalar,
                                  Real* next_scalar, const Real* current_fluid, Real omega_c,
                                  Real source_term = Real{}, dim3 block = default_cuda_block_size(),
                                  hipStream_t stream = nullptr) {
    // Extent mismatch check of src/lattice_cuda.cu:731-735.
    if (scalar_mem.x_extent() != fluid_mem.x_extent() || scalar_mem.y_extent() != fluid_mem.y_extent() ||
        scalar_mem.z_extent() != fluid_mem.z_extent())
        return hipErrorInvalidValue;
    return launch_scalar_step_gpu<FluidLattice, ScalarLattice, Real>(
        current_scalar, next_scalar, current_fluid, scalar_mem.x_extent(), scalar_mem.y_extent(),
        scalar_mem.z_extent(), omega_c, source_term, block, stream);
}

template <HipLattice FluidLattice, HipLattice ScalarLattice, std::floating_point Real>
hipError_t launch_reaction_AB_gpu(const Real* current_a, Real* next_a, const Real* current_b, Real* next_b,
                                  const Real* current_fluid, std::size_t x_extent, std::size_t y_extent,
                                  std::size_t z_extent, Real omega_c, Real k_react,
                                  dim3 block = default_cuda_block_size(), hipStream_t stream = nullptr) {
    static_assert(FluidLattice::D == ScalarLattice::D, "fluid and scalar lattices must have the same dimension");
    return static_cast<hipError_t>(lbc_reaction_ab_step(lattice_code<FluidLattice>(), lattice_code<ScalarLattice>(),
                                                        storage_code<Real>(), current_a, next_a, current_b, next_b,
                                                        current_fluid, x_extent, y_extent, z_extent,
                                                        static_cast<double>(omega_c), static_cast<double>(k_react),
                                                        0.0, 0.0, block.x, block.y, block.z, stream));
}

}  // namespace lbm
