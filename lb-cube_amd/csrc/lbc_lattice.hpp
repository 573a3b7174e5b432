// Lattice descriptors for the gfx950 kernels.
//
// Direction order, weights, velocities and opposite maps are those of
// LB-Cube's include/lattice_traits.hpp:62-305: the direction order fixes the
// order of the SoA population planes and of the MRT bases, so it has to be the
// same.  Velocities are stored as three components for every lattice (z = 0 in
// 2D) so one kernel template serves 2D and 3D.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbc {

enum LatticeId : int { kD2Q9 = 0, kD3Q19 = 1, kD3Q27 = 2, kD2Q5 = 3, kD3Q7 = 4 };
enum CollisionId : int { kBGK = 0, kTRT = 1, kMRT = 2, kRLBM = 3 };
// Storage formats: kF32/kF64 hold populations as they are; kF16/kBF16 hold
// (f_i - w_i) for fluid populations and g_i for scalar populations, rounded to
// nearest-even, with arithmetic in float.
enum StorageId : int { kF32 = 0, kF64 = 1, kF16 = 2, kBF16 = 3 };

struct D2Q9 {
    static constexpr int id = kD2Q9, D = 2, Q = 9;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr int cx[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int cy[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
};

struct D3Q19 {
    static constexpr int id = kD3Q19, D = 3, Q = 19;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {1.0 / 3.0,
                                    1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
};

struct D3Q27 {
    static constexpr int id = kD3Q27, D = 3, Q = 27;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {8.0 / 27.0,
                                    2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
                                    1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
                                    1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
                                    1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
                                    1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
    static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                                  1, -1, 1, -1, 1, -1, 1, -1};
    static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1,
                                  1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1,
                                  1, 1, 1, 1, -1, -1, -1, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                                   26, 25, 24, 23, 22, 21, 20, 19};
};

struct D2Q5 {
    static constexpr int id = kD2Q5, D = 2, Q = 5;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {1.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
    static constexpr int cx[Q] = {0, 1, 0, -1, 0};
    static constexpr int cy[Q] = {0, 0, 1, 0, -1};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0};
    static constexpr int opp[Q] = {0, 3, 4, 1, 2};
};

struct D3Q7 {
    static constexpr int id = kD3Q7, D = 3, Q = 7;
    static constexpr double cs2 = 1.0 / 4.0;
    static constexpr double w[Q] = {1.0 / 4.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0};
    static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0};
    static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5};
};

// Compile-time check of the opposite map: an involution with c[opp] = -c.
template <class L> constexpr bool opposites_valid() {
    for (int i = 0; i < L::Q; ++i) {
        const int j = L::opp[i];
        if (j < 0 || j >= L::Q || L::opp[j] != i) return false;
        if (L::cx[i] + L::cx[j] != 0 || L::cy[i] + L::cy[j] != 0 || L::cz[i] + L::cz[j] != 0) return false;
    }
    return true;
}
static_assert(opposites_valid<D2Q9>() && opposites_valid<D3Q19>() && opposites_valid<D3Q27>() &&
              opposites_valid<D2Q5>() && opposites_valid<D3Q7>());

}  // namespace lbc
