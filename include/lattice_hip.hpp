// lattice_hip.hpp — C++ launcher API of the MI355X kernels, mirroring LB-Cube's
// include/lattice_cuda.cuh (names, template parameters, argument order and
// defaults), so the reference's drivers switch by changing the include and the
// runtime prefix (cuda* -> hip*).  Every launcher is an inline forwarder to the
// C ABI of lbcube.h; no device code is compiled by the caller.
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
// Differences: launchers return hipError_t; the default block is
// default_cuda_block_size() = {256, 1, 1} (four wave64s) instead of {16, 8, 1};
// the kernels cover cells linearly, so any block gives the same results and
// only its size matters for speed; an optional trailing hipStream_t selects
// the stream (the reference always uses the default stream).
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <concepts>
#include <cstddef>
#include <type_traits>
#include <vector>

#include "lbcube.h"

namespace lbm {

#ifndef LB_CUBE_COLLISION_TYPE_DEFINED
#define LB_CUBE_COLLISION_TYPE_DEFINED
// CollisionType, include/lattice_core.hpp:33-42 (same enumerator order).
enum class CollisionType { BGK, TRT, MRT, RLBM };
#endif

// Lattice tags with the reference's names (include/lattice_traits.hpp:62-305);
// tables live in the kernel library, the tags carry D, Q and the ABI code.
struct D2Q9 { static constexpr int D = 2, Q = 9, abi = LBC_D2Q9; static constexpr double cs2 = 1.0 / 3.0; };
struct D3Q19 { static constexpr int D = 3, Q = 19, abi = LBC_D3Q19; static constexpr double cs2 = 1.0 / 3.0; };
struct D3Q27 { static constexpr int D = 3, Q = 27, abi = LBC_D3Q27; static constexpr double cs2 = 1.0 / 3.0; };
struct D2Q5 { static constexpr int D = 2, Q = 5, abi = LBC_D2Q5; static constexpr double cs2 = 1.0 / 3.0; };
struct D3Q7 { static constexpr int D = 3, Q = 7, abi = LBC_D3Q7; static constexpr double cs2 = 1.0 / 4.0; };

template <class L>
concept IsLatticeModel = requires {
    { L::D } -> std::convertible_to<int>;
    { L::Q } -> std::convertible_to<int>;
    { L::abi } -> std::convertible_to<int>;
};

template <class Real> constexpr int storage_code() {
    static_assert(std::is_same_v<Real, float> || std::is_same_v<Real, double>, "Real must be float or double");
    return std::is_same_v<Real, float> ? LBC_F32 : LBC_F64;
}

// Host SoA ping-pong container with the layout contract of LatticeMemory
// (include/lattice_memory.hpp:67-272): two flat buffers [Q][Z][Y][X], x fastest.
// The launchers read only its extents.
template <IsLatticeModel L, std::floating_point Real> class LatticeMemory {
public:
    LatticeMemory(std::size_t x, std::size_t y) requires(L::D == 2)
        : x_(x), y_(y), z_(1), pop_0(L::Q * x * y), pop_1(L::Q * x * y) {}
    LatticeMemory(std::size_t x, std::size_t y, std::size_t z) requires(L::D == 3)
        : x_(x), y_(y), z_(z), pop_0(L::Q * x * y * z), pop_1(L::Q * x * y * z) {}
    std::size_t x_extent() const noexcept { return x_; }
    std::size_t y_extent() const noexcept { return y_; }
    std::size_t z_extent() const noexcept { return z_; }
    std::size_t population_count() const noexcept { return pop_0.size(); }
    Real* current() noexcept { return cur_ == 0 ? pop_0.data() : pop_1.data(); }
    Real* next() noexcept { return cur_ == 0 ? pop_1.data() : pop_0.data(); }
    const Real* current() const noexcept { return cur_ == 0 ? pop_0.data() : pop_1.data(); }
    // Element (i, z, y, x) of the current buffer.
    Real& at(int i, std::size_t z, std::size_t y, std::size_t x) noexcept {
        return current()[((static_cast<std::size_t>(i) * z_ + z) * y_ + y) * x_ + x];
    }
    void swap_buffers() noexcept { cur_ = 1 - cur_; }

private:
    std::size_t x_, y_, z_;
    std::vector<Real> pop_0, pop_1;
    int cur_ = 0;
};

inline dim3 default_cuda_block_size() noexcept { return dim3{256, 1, 1}; }

template <IsLatticeModel Lattice, std::floating_point Real, CollisionType CT = CollisionType::BGK>
hipError_t launch_step_gpu(const Real* current_populations, Real* next_populations, std::size_t x_extent,
                           std::size_t y_extent, std::size_t z_extent, Real omega,
                           dim3 block = default_cuda_block_size(), hipStream_t stream = nullptr) {
    return static_cast<hipError_t>(lbc_fluid_step(Lattice::abi, static_cast<int>(CT), storage_code<Real>(),
                                                  current_populations, next_populations, x_extent, y_extent,
                                                  z_extent, static_cast<double>(omega), block.x, block.y, block.z,
                                                  stream));
}

template <IsLatticeModel Lattice, std::floating_point Real, CollisionType CT = CollisionType::BGK>
hipError_t launch_step_gpu(const LatticeMemory<Lattice, Real>& mem, const Real* current_populations,
                           Real* next_populations, Real omega, dim3 block = default_cuda_block_size(),
                           hipStream_t stream = nullptr) {
    return launch_step_gpu<Lattice, Real, CT>(current_populations, next_populations, mem.x_extent(), mem.y_extent(),
                                              mem.z_extent(), omega, block, stream);
}

template <IsLatticeModel FluidLattice, IsLatticeModel ScalarLattice, std::floating_point Real>
hipError_t launch_scalar_step_gpu(const Real* current_scalar, Real* next_scalar, const Real* current_fluid,
                                  std::size_t x_extent, std::size_t y_extent, std::size_t z_extent, Real omega_c,
                                  Real source_term = Real{}, dim3 block = default_cuda_block_size(),
                                  hipStream_t stream = nullptr) {
    static_assert(FluidLattice::D == ScalarLattice::D);
    return static_cast<hipError_t>(lbc_scalar_step(FluidLattice::abi, ScalarLattice::abi, storage_code<Real>(),
                                                   current_scalar, next_scalar, current_fluid, x_extent, y_extent,
                                                   z_extent, static_cast<double>(omega_c),
                                                   static_cast<double>(source_term), 0.0, block.x, block.y,
                                                   block.z, stream));
}

template <IsLatticeModel FluidLattice, IsLatticeModel ScalarLattice, std::floating_point Real>
hipError_t launch_scalar_step_gpu(const LatticeMemory<ScalarLattice, Real>& scalar_mem,
                                  const LatticeMemory<FluidLattice, Real>& fluid_mem, const Real* current_scalar,
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

template <IsLatticeModel FluidLattice, IsLatticeModel ScalarLattice, std::floating_point Real>
hipError_t launch_reaction_AB_gpu(const Real* current_a, Real* next_a, const Real* current_b, Real* next_b,
                                  const Real* current_fluid, std::size_t x_extent, std::size_t y_extent,
                                  std::size_t z_extent, Real omega_c, Real k_react,
                                  dim3 block = default_cuda_block_size(), hipStream_t stream = nullptr) {
    static_assert(FluidLattice::D == ScalarLattice::D);
    return static_cast<hipError_t>(lbc_reaction_ab_step(FluidLattice::abi, ScalarLattice::abi, storage_code<Real>(),
                                                        current_a, next_a, current_b, next_b, current_fluid,
                                                        x_extent, y_extent, z_extent, static_cast<double>(omega_c),
                                                        static_cast<double>(k_react), 0.0, 0.0, block.x, block.y,
                                                        block.z, stream));
}

}  // namespace lbm
