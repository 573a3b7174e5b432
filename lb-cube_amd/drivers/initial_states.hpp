// Initial states of the drivers, on the host, in the reference's operation
// order so that a driver started without -load begins from the populations the
// reference's executable would build:
//   equilibrium         compute_equilibrium, include/lattice_physics.hpp:267-290
//   scalar_equilibrium  compute_scalar_equilibrium, include/lattice_physics.hpp:643-659
//   taylor_green_3d     initialize_taylor_green_state_3d, src/main.cpp:371-411
//   taylor_green_2d     taylor_green_macro, src/simulate_reaction.cpp:173-192 and
//                       src/compare_operators.cpp:64-81
// Dot products are taken left to right (c_x u_x + c_y u_y) + c_z u_z.
#pragma once

#include <cmath>
#include <cstddef>
#include <numbers>

#include "lbcube_lattices.hpp"

namespace lbc_init {

template <lbm::HipLattice L, class Real> Real equilibrium(int i, Real rho, const Real* u) {
    Real cu = static_cast<Real>(L::c[i * L::D]) * u[0];
    Real uu = u[0] * u[0];
    for (int d = 1; d < L::D; ++d) {
        cu += static_cast<Real>(L::c[i * L::D + d]) * u[d];
        uu += u[d] * u[d];
    }
    const Real weight = static_cast<Real>(L::weights[i]);
    const Real inv_cs2 = static_cast<Real>(1.0) / static_cast<Real>(L::cs2);
    return weight * rho *
           (static_cast<Real>(1.0) + inv_cs2 * cu + static_cast<Real>(0.5) * inv_cs2 * inv_cs2 * cu * cu -
            static_cast<Real>(0.5) * inv_cs2 * uu);
}

template <lbm::HipLattice S, class Real> Real scalar_equilibrium(int i, Real conc, const Real* u) {
    Real cu{};
    for (int d = 0; d < S::D; ++d) cu += static_cast<Real>(S::c[i * S::D + d]) * u[d];
    return static_cast<Real>(S::weights[i]) * conc * (static_cast<Real>(1.0) + cu / static_cast<Real>(S::cs2));
}

struct Macro {
    double rho;
    double u[3];
};

// The 3D Taylor-Green vortex with its density perturbation (rho = 1 + ...).
inline Macro taylor_green_3d(std::size_t x, std::size_t y, std::size_t z, std::size_t nx, std::size_t ny,
                             std::size_t nz, double u0) {
    const double cs2 = 1.0 / 3.0;
    const double kx = 2.0 * std::numbers::pi_v<double> / static_cast<double>(nx);
    const double ky = 2.0 * std::numbers::pi_v<double> / static_cast<double>(ny);
    const double kz = 2.0 * std::numbers::pi_v<double> / static_cast<double>(nz);
    const double px = kx * static_cast<double>(x), py = ky * static_cast<double>(y), pz = kz * static_cast<double>(z);
    const double perturbation = (u0 * u0) / (16.0 * cs2) * (std::cos(2.0 * px) + std::cos(2.0 * py)) *
                                (std::cos(2.0 * pz) + 2.0);
    return {1.0 + perturbation,
            {u0 * std::sin(px) * std::cos(py) * std::cos(pz), -u0 * std::cos(px) * std::sin(py) * std::cos(pz), 0.0}};
}

// The 2D Taylor-Green vortex of the reactive drivers (rho = 1 - ...).
inline Macro taylor_green_2d(std::size_t x, std::size_t y, std::size_t nx, std::size_t ny, double u0) {
    const double kx = 2.0 * std::numbers::pi_v<double> / static_cast<double>(nx);
    const double ky = 2.0 * std::numbers::pi_v<double> / static_cast<double>(ny);
    const double px = kx * static_cast<double>(x), py = ky * static_cast<double>(y);
    const double perturbation =
        (u0 * u0) / (4.0 * static_cast<double>(lbm::D2Q9::cs2)) * (std::cos(2.0 * px) + std::cos(2.0 * py));
    return {1.0 - perturbation, {-u0 * std::cos(px) * std::sin(py), u0 * std::sin(px) * std::cos(py), 0.0}};
}

}  // namespace lbc_init
