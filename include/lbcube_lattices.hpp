// lbcube_lattices.hpp — lattice traits shaped like LB-Cube's
// include/lattice_traits.hpp:62-305 (static D, Q, cs2, weights, directions /
// c and opposite, in the same direction order), for programs that use
// lattice_hip.hpp without the reference's headers (the drivers of this
// repository).  A program that includes the reference's lattice_traits.hpp
// uses its types instead and must not include this header: both define
// lbm::D2Q9, lbm::D3Q19, lbm::D3Q27, lbm::D2Q5 and lbm::D3Q7.
//
// The tables are the lattices' definitions (weights of the standard D2Q9,
// D3Q19, D3Q27 quadratures; D2Q5 with cs2 = 1/3 and D3Q7 with cs2 = 1/4), and
// the direction order is the plane order of the kernels' [Q][Z][Y][X] layout.
#pragma once

#include <array>
#include <cstddef>
#include <vector>

#include "lattice_hip.hpp"

namespace lbm {

namespace lattice_detail {
// opposite[i] = the j with c_j = -c_i.
template <int D, int Q> constexpr std::array<int, Q> opposites(const std::array<int, Q * D>& c) {
    std::array<int, Q> opp{};
    for (int i = 0; i < Q; ++i) {
        for (int j = 0; j < Q; ++j) {
            bool neg = true;
            for (int d = 0; d < D; ++d) neg = neg && (c[j * D + d] == -c[i * D + d]);
            if (neg) opp[i] = j;
        }
    }
    return opp;
}
}  // namespace lattice_detail

struct D2Q9 {
    static constexpr int D = 2;
    static constexpr int Q = 9;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr std::array<double, Q> weights{4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr std::array<int, Q * D> directions = hip_detail::d2q9_c;
    static constexpr auto c = directions;
    static constexpr std::array<int, Q> opposite = lattice_detail::opposites<D, Q>(directions);
};

struct D3Q19 {
    static constexpr int D = 3;
    static constexpr int Q = 19;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr std::array<double, Q> weights{
        1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr std::array<int, Q * D> directions = hip_detail::d3q19_c;
    static constexpr auto c = directions;
    static constexpr std::array<int, Q> opposite = lattice_detail::opposites<D, Q>(directions);
};

struct D3Q27 {
    static constexpr int D = 3;
    static constexpr int Q = 27;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr std::array<double, Q> weights{
        8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
        1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
        1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
        1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
    static constexpr std::array<int, Q * D> directions = hip_detail::d3q27_c;
    static constexpr auto c = directions;
    static constexpr std::array<int, Q> opposite = lattice_detail::opposites<D, Q>(directions);
};

struct D2Q5 {
    static constexpr int D = 2;
    static constexpr int Q = 5;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr std::array<double, Q> weights{1.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
    static constexpr std::array<int, Q * D> directions = hip_detail::d2q5_c;
    static constexpr auto c = directions;
    static constexpr std::array<int, Q> opposite = lattice_detail::opposites<D, Q>(directions);
};

struct D3Q7 {
    static constexpr int D = 3;
    static constexpr int Q = 7;
    static constexpr double cs2 = 1.0 / 4.0;
    static constexpr std::array<double, Q> weights{1.0 / 4.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0,
                                                   1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0};
    static constexpr std::array<int, Q * D> directions = hip_detail::d3q7_c;
    static constexpr auto c = directions;
    static constexpr std::array<int, Q> opposite = lattice_detail::opposites<D, Q>(directions);
};

static_assert(HipLattice<D2Q9> && HipLattice<D3Q19> && HipLattice<D3Q27> && HipLattice<D2Q5> && HipLattice<D3Q7>);

// Host SoA ping-pong container with the layout contract of the reference's
// LatticeMemory (include/lattice_memory.hpp:67-272): two flat buffers
// [Q][Z][Y][X], x fastest; the memory overloads of lattice_hip.hpp read its
// extents.  Element access by (i, z, y, x) replaces the reference's mdspan views.
template <HipLattice L, class Real> class HostLattice {
public:
    HostLattice(std::size_t x, std::size_t y) requires(L::D == 2)
        : x_(x), y_(y), z_(1), pop_(2 * static_cast<std::size_t>(L::Q) * x * y) {}
    HostLattice(std::size_t x, std::size_t y, std::size_t z) requires(L::D == 3)
        : x_(x), y_(y), z_(z), pop_(2 * static_cast<std::size_t>(L::Q) * x * y * z) {}
    std::size_t x_extent() const noexcept { return x_; }
    std::size_t y_extent() const noexcept { return y_; }
    std::size_t z_extent() const noexcept { return z_; }
    std::size_t population_count() const noexcept { return pop_.size() / 2; }
    Real* current() noexcept { return pop_.data() + cur_ * population_count(); }
    const Real* current() const noexcept { return pop_.data() + cur_ * population_count(); }
    Real* next() noexcept { return pop_.data() + (1 - cur_) * population_count(); }
    Real& operator()(int i, std::size_t z, std::size_t y, std::size_t x) noexcept {
        return current()[((static_cast<std::size_t>(i) * z_ + z) * y_ + y) * x_ + x];
    }
    void swap_buffers() noexcept { cur_ = 1 - cur_; }

private:
    std::size_t x_, y_, z_;
    std::vector<Real> pop_;
    std::size_t cur_ = 0;
};

}  // namespace lbm
