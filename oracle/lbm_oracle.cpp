// ============================================================================
// THE ORACLE — test infrastructure only.
//
// A CPU restatement of LB-Cube's fused pull-stream/collide loop, passive-scalar
// loop, fused A+B->C reaction loop and host statistics, written to follow the
// reference's operation order so that the HIP kernels can be checked against
// it bit-for-bit where both sides evaluate the same IEEE operations.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, and only as the checker or the timed CPU baseline.  Nothing in
// the product (lb-cube_amd/) links, loads or calls it.
//
// Reference files restated (paths relative to the LB-Cube source tree):
//   include/lattice_traits.hpp   lattice tables, lines 62-305
//   include/lattice_physics.hpp  macro/feq/BGK/TRT/RLBM/scalar/reaction, 116-743
//   include/lattice_mrt.hpp      D2Q9/D3Q19/D3Q27 MRT, 38-1305
//   include/lattice_core.hpp     step_cpu 146-216, step_scalar_cpu 238-337,
//                                step_reaction_AB 357-474, pull index 70-78
//   include/lattice_io.hpp       log_diagnostics 182-234,
//                                compute_reactive_stats 256-338
//
// Pinning (what checks each part; the checks live in tests/test_oracle_*.py):
//   tables, pull index, opposite map ... involution / c[opp]=-c / omega=0
//                                        permutation KAT with unique ids
//   macro + feq ........................ isotropy sums (sum feq = rho,
//                                        sum feq c = rho u), rest fixed point
//   BGK/TRT/MRT/RLBM ................... the reference's analytic decay tests
//                                        (tests/test_validation.cpp) restated,
//                                        TRT(w-=w+) == BGK, D2Q9/D3Q27 MRT with
//                                        all rates = w == BGK, conservation
//   RLBM reconstruction ................ Hermite moments of orders 0-3 of the
//                                        post-collision state against the
//                                        recursive a3 = u Pi + permutations
//                                        (Malaspinas 2015), D3Q19 pairs, and
//                                        degree <= 3 of f_neq / w
//   MRT transforms ..................... the reference's literal identity
//                                        vectors (test_validation.cpp:1532-1582)
//                                        and M M^-1 = I by brute force
//   scalar ADR ......................... advected decaying sine closed form
//                                        (test_validation.cpp:1590-1646), in 2D
//                                        on D2Q5 and in 3D on D3Q7 (cs2 = 1/4);
//                                        equilibrium moments C, C u, C cs2 I
//                                        through a one-cell step
//   reaction source .................... batch closed form C0/(1+C0 k t),
//                                        high-precision ODE integration
//   statistics ......................... closed forms (uniform, complementary
//                                        half planes) and long double sums
//   discrete adjoint (2D and 3D, ....... central differences of this forward,
//   BGK/TRT/MRT/RLBM; no counterpart     the one-step identity <L1, J v> =
//   in the reference)                    <J^T L1, v> against an extended-
//                                        precision tangent, and dJ/d omega_c
//                                        over 200-400 steps against extended-
//                                        precision differences
//   Eigen's summation grouping in dot() is not observable here; dot products
//   are taken left to right (c_x u_x + c_y u_y) + c_z u_z.
// ============================================================================

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using i64 = long long;

// ---------------------------------------------------------------------------
// Lattice tables — include/lattice_traits.hpp:62-305 (same direction order).
// ---------------------------------------------------------------------------
struct D2Q9 {
    static constexpr int D = 2, Q = 9;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr int c[Q * D] = {0, 0, 1, 0, 0, 1, -1, 0, 0, -1, 1, 1, -1, 1, -1, -1, 1, -1};
    static constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
};
struct D3Q19 {
    static constexpr int D = 3, Q = 19;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {1.0 / 3.0,
                                    1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    static constexpr int c[Q * D] = {0, 0, 0,   1, 0, 0,   -1, 0, 0,  0, 1, 0,   0, -1, 0,
                                     0, 0, 1,   0, 0, -1,  1, 1, 0,   -1, 1, 0,  1, -1, 0,
                                     -1, -1, 0, 1, 0, 1,   -1, 0, 1,  1, 0, -1,  -1, 0, -1,
                                     0, 1, 1,   0, -1, 1,  0, 1, -1,  0, -1, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
};
struct D3Q27 {
    static constexpr int D = 3, Q = 27;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {8.0 / 27.0,
                                    2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
                                    1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
                                    1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
                                    1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
                                    1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
    static constexpr int c[Q * D] = {0, 0, 0,   1, 0, 0,   -1, 0, 0,  0, 1, 0,   0, -1, 0,
                                     0, 0, 1,   0, 0, -1,  1, 1, 0,   -1, 1, 0,  1, -1, 0,
                                     -1, -1, 0, 1, 0, 1,   -1, 0, 1,  1, 0, -1,  -1, 0, -1,
                                     0, 1, 1,   0, -1, 1,  0, 1, -1,  0, -1, -1,
                                     1, 1, 1,   -1, 1, 1,  1, -1, 1,  -1, -1, 1,
                                     1, 1, -1,  -1, 1, -1, 1, -1, -1, -1, -1, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                                   26, 25, 24, 23, 22, 21, 20, 19};
};
struct D2Q5 {
    static constexpr int D = 2, Q = 5;
    static constexpr double cs2 = 1.0 / 3.0;
    static constexpr double w[Q] = {1.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
    static constexpr int c[Q * D] = {0, 0, 1, 0, 0, 1, -1, 0, 0, -1};
    static constexpr int opp[Q] = {0, 3, 4, 1, 2};
};
struct D3Q7 {
    static constexpr int D = 3, Q = 7;
    static constexpr double cs2 = 1.0 / 4.0;  // lattice_traits.hpp:279
    static constexpr double w[Q] = {1.0 / 4.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0, 1.0 / 8.0};
    static constexpr int c[Q * D] = {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1};
    static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5};
};

enum Lat { L_D2Q9 = 0, L_D3Q19 = 1, L_D3Q27 = 2, L_D2Q5 = 3, L_D3Q7 = 4 };
enum Coll { C_BGK = 0, C_TRT = 1, C_MRT = 2, C_RLBM = 3 };
// T_F80: x87 extended precision (long double, 64-bit significand), test-only:
// the same algorithm with rounding errors ~2^11 times smaller than FP64, the
// yardstick for comparing two FP64 (or FP32) evaluation orders.
enum Dt { T_F32 = 0, T_F64 = 1, T_F16 = 2, T_BF16 = 3, T_F80 = 4 };

// ---------------------------------------------------------------------------
// Storage codecs.  F32/F64 store populations as-is (the reference's
// LatticeMemory<L, Real>).  F16/BF16 are the project's reduced-precision
// storage: the stored value is (f - offset) rounded to nearest-even, with
// offset = w_i for fluid populations and 0 for scalar populations; arithmetic
// runs in float.  Conversions are IEEE 754 round-to-nearest-even, written
// from the format definitions.
// ---------------------------------------------------------------------------
inline uint16_t f32_to_f16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t a = x & 0x7fffffffu;
    if (a >= 0x7f800000u) return static_cast<uint16_t>(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
    if (a >= 0x477ff000u) return static_cast<uint16_t>(sign | 0x7c00u);  // rounds to +-inf
    if (a < 0x38800000u) {                                               // half subnormal range
        if (a < 0x33000000u) return static_cast<uint16_t>(sign);
        const uint32_t e = a >> 23;
        const uint32_t m = (a & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126u - e;
        uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return static_cast<uint16_t>(sign | q);
    }
    const uint32_t r = a - (112u << 23);
    uint32_t q = r >> 13;
    const uint32_t rem = r & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
    return static_cast<uint16_t>(sign | q);
}
inline float f16_to_f32(uint16_t h) {
    const uint32_t sign = (static_cast<uint32_t>(h) & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu;
    const uint32_t m = h & 0x3ffu;
    float out;
    if (e == 0) {
        out = std::ldexp(static_cast<float>(m), -24);
        return sign ? -out : out;
    }
    uint32_t bits;
    if (e == 31) bits = sign | 0x7f800000u | (m << 13);
    else bits = sign | ((e + 112u) << 23) | (m << 13);
    std::memcpy(&out, &bits, 4);
    return out;
}
inline uint16_t f32_to_bf16(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((x >> 16) | 0x40u);
    x += 0x7fffu + ((x >> 16) & 1u);
    return static_cast<uint16_t>(x >> 16);
}
inline float bf16_to_f32(uint16_t h) {
    const uint32_t bits = static_cast<uint32_t>(h) << 16;
    float out;
    std::memcpy(&out, &bits, 4);
    return out;
}

template <int DT> struct Codec;
template <> struct Codec<T_F32> {
    using S = float; using R = float;
    static R load(S v, R) { return v; }
    static S store(R f, R) { return f; }
};
template <> struct Codec<T_F64> {
    using S = double; using R = double;
    static R load(S v, R) { return v; }
    static S store(R f, R) { return f; }
};
template <> struct Codec<T_F80> {
    using S = long double; using R = long double;
    static R load(S v, R) { return v; }
    static S store(R f, R) { return f; }
};
template <> struct Codec<T_F16> {
    using S = uint16_t; using R = float;
    static R load(S v, R off) { return off + f16_to_f32(v); }
    static S store(R f, R off) { return f32_to_f16(f - off); }
};
template <> struct Codec<T_BF16> {
    using S = uint16_t; using R = float;
    static R load(S v, R off) { return off + bf16_to_f32(v); }
    static S store(R f, R off) { return f32_to_bf16(f - off); }
};

// ---------------------------------------------------------------------------
// Stateless per-cell physics — include/lattice_physics.hpp.
// ---------------------------------------------------------------------------
template <class L, class R> struct Macro {
    R rho{};
    R u[3]{};
};

// compute_macro_state, lattice_physics.hpp:227-249
template <class L, class R> Macro<L, R> macro_state(const R* f) {
    R density{};
    R mom[3]{};
    for (int i = 0; i < L::Q; ++i) {
        density += f[i];
        for (int d = 0; d < L::D; ++d) mom[d] += f[i] * static_cast<R>(L::c[i * L::D + d]);
    }
    Macro<L, R> m;
    m.rho = density;
    for (int d = 0; d < L::D; ++d) m.u[d] = mom[d] / density;
    return m;
}

// Left-to-right dot product of a lattice direction with a velocity.
template <class L, class R> R dir_dot(int i, const R* u) {
    R s = static_cast<R>(L::c[i * L::D]) * u[0];
    for (int d = 1; d < L::D; ++d) s += static_cast<R>(L::c[i * L::D + d]) * u[d];
    return s;
}
template <class L, class R> R vel_dot(const R* u) {
    R s = u[0] * u[0];
    for (int d = 1; d < L::D; ++d) s += u[d] * u[d];
    return s;
}

// compute_equilibrium, lattice_physics.hpp:267-290
template <class L, class R> R feq(int i, const Macro<L, R>& m) {
    const R weight = static_cast<R>(L::w[i]);
    const R cs2 = static_cast<R>(L::cs2);
    const R inv_cs2 = static_cast<R>(1.0) / cs2;
    const R cu = dir_dot<L, R>(i, m.u);
    const R uu = vel_dot<L, R>(m.u);
    return weight * m.rho *
           (static_cast<R>(1.0) + inv_cs2 * cu + static_cast<R>(0.5) * inv_cs2 * inv_cs2 * cu * cu -
            static_cast<R>(0.5) * inv_cs2 * uu);
}

// collide_bgk, lattice_physics.hpp:307-318
template <class L, class R> void collide_bgk(R* f, R omega) {
    const Macro<L, R> m = macro_state<L, R>(f);
    for (int i = 0; i < L::Q; ++i) f[i] += omega * (feq<L, R>(i, m) - f[i]);
}

// compute_omega_minus, lattice_physics.hpp:336-344 (Lambda = 1/4)
template <class R> R omega_minus(R omega_plus) {
    const R magic = R{1} / R{4};
    const R tau_plus = R{1} / omega_plus;
    const R tau_minus = magic / (tau_plus - R{1} / R{2}) + R{1} / R{2};
    return R{1} / tau_minus;
}

// collide_trt, lattice_physics.hpp:366-398
template <class L, class R> void collide_trt(R* f, const Macro<L, R>& m, R wp, R wm) {
    R eq[L::Q], post[L::Q];
    for (int i = 0; i < L::Q; ++i) eq[i] = feq<L, R>(i, m);
    for (int i = 0; i < L::Q; ++i) {
        const int j = L::opp[i];
        const R even = R{0.5} * (f[i] + f[j]);
        const R eq_even = R{0.5} * (eq[i] + eq[j]);
        const R odd = R{0.5} * (f[i] - f[j]);
        const R eq_odd = R{0.5} * (eq[i] - eq[j]);
        post[i] = f[i] - wp * (even - eq_even) - wm * (odd - eq_odd);
    }
    for (int i = 0; i < L::Q; ++i) f[i] = post[i];
}

// ---------------------------------------------------------------------------
// RLBM — lattice_physics.hpp:116-211 (Pi_neq) and 422-605 (reconstruction).
// ---------------------------------------------------------------------------
template <class L, class R> struct Pi { R xx{}, yy{}, zz{}, xy{}, xz{}, yz{}; };

template <class L, class R> Pi<L, R> pi_neq(const R* f, const R* eq) {
    Pi<L, R> p;
    for (int i = 0; i < L::Q; ++i) {
        const R fn = f[i] - eq[i];
        const R cx = static_cast<R>(L::c[i * L::D]);
        const R cy = static_cast<R>(L::c[i * L::D + 1]);
        if constexpr (L::D == 2) {
            p.xx += fn * cx * cx;
            p.yy += fn * cy * cy;
            p.xy += fn * cx * cy;
        } else {
            const R cz = static_cast<R>(L::c[i * L::D + 2]);
            p.xx += fn * cx * cx;
            p.yy += fn * cy * cy;
            p.zz += fn * cz * cz;
            p.xy += fn * cx * cy;
            p.xz += fn * cx * cz;
            p.yz += fn * cy * cz;
        }
    }
    return p;
}

template <class L, class R> void collide_rlbm(R* f, const Macro<L, R>& m, R omega) {
    R eq[L::Q];
    for (int i = 0; i < L::Q; ++i) eq[i] = feq<L, R>(i, m);
    const Pi<L, R> p = pi_neq<L, R>(f, eq);
    const R cs2 = static_cast<R>(L::cs2);
    const R pref = R{9} / R{2};
    const R ux = m.u[0], uy = m.u[1];
    if constexpr (L::Q == 9) {
        const R a_xxx = R{3} * ux * p.xx;
        const R a_yyy = R{3} * uy * p.yy;
        const R a_xxy = R{2} * ux * p.xy + uy * p.xx;
        const R a_xyy = R{2} * uy * p.xy + ux * p.yy;
        for (int i = 0; i < L::Q; ++i) {
            const R cx = static_cast<R>(L::c[i * 2]);
            const R cy = static_cast<R>(L::c[i * 2 + 1]);
            const R q_xx = cx * cx - cs2;
            const R q_yy = cy * cy - cs2;
            const R q_xy = cx * cy;
            const R second = q_xx * p.xx + q_yy * p.yy + R{2} * q_xy * p.xy;
            const R h_xxx = cx * cx * cx - R{3} * cs2 * cx;
            const R h_yyy = cy * cy * cy - R{3} * cs2 * cy;
            const R h_xxy = cx * cx * cy - cs2 * cy;
            const R h_xyy = cx * cy * cy - cs2 * cx;
            const R third = a_xxx * h_xxx + a_yyy * h_yyy + R{3} * (a_xxy * h_xxy + a_xyy * h_xyy);
            const R rec = static_cast<R>(L::w[i]) * pref * (second + third);
            f[i] = eq[i] + (R{1} - omega) * rec;
        }
    } else if constexpr (L::Q == 19) {
        const R uz = m.u[2];
        const R a_xxy = R{2} * ux * p.xy + uy * p.xx;
        const R a_xxz = R{2} * ux * p.xz + uz * p.xx;
        const R a_xyy = R{2} * uy * p.xy + ux * p.yy;
        const R a_yyz = R{2} * uy * p.yz + uz * p.yy;
        const R a_xzz = R{2} * uz * p.xz + ux * p.zz;
        const R a_yzz = R{2} * uz * p.yz + uy * p.zz;
        for (int i = 0; i < L::Q; ++i) {
            const R cx = static_cast<R>(L::c[i * 3]);
            const R cy = static_cast<R>(L::c[i * 3 + 1]);
            const R cz = static_cast<R>(L::c[i * 3 + 2]);
            const R q_xx = cx * cx - cs2;
            const R q_yy = cy * cy - cs2;
            const R q_zz = cz * cz - cs2;
            const R q_xy = cx * cy;
            const R q_xz = cx * cz;
            const R q_yz = cy * cz;
            const R second = q_xx * p.xx + q_yy * p.yy + q_zz * p.zz +
                             R{2} * (q_xy * p.xy + q_xz * p.xz + q_yz * p.yz);
            const R h_xxy = cx * cx * cy - cs2 * cy;
            const R h_xxz = cx * cx * cz - cs2 * cz;
            const R h_xyy = cx * cy * cy - cs2 * cx;
            const R h_yyz = cy * cy * cz - cs2 * cz;
            const R h_xzz = cx * cz * cz - cs2 * cx;
            const R h_yzz = cy * cz * cz - cs2 * cy;
            // paired third-order projection, lattice_physics.hpp:523-529
            const R third = R{3} * (h_xxy + h_yzz) * (a_xxy + a_yzz) + (h_xxy - h_yzz) * (a_xxy - a_yzz) +
                            R{3} * (h_xzz + h_xyy) * (a_xzz + a_xyy) + (h_xzz - h_xyy) * (a_xzz - a_xyy) +
                            R{3} * (h_yyz + h_xxz) * (a_yyz + a_xxz) + (h_yyz - h_xxz) * (a_yyz - a_xxz);
            const R rec = static_cast<R>(L::w[i]) * pref * (second + third);
            f[i] = eq[i] + (R{1} - omega) * rec;
        }
    } else {
        const R uz = m.u[2];
        const R a_xxx = R{3} * ux * p.xx;
        const R a_yyy = R{3} * uy * p.yy;
        const R a_zzz = R{3} * uz * p.zz;
        const R a_xxy = R{2} * ux * p.xy + uy * p.xx;
        const R a_xxz = R{2} * ux * p.xz + uz * p.xx;
        const R a_xyy = ux * p.yy + R{2} * uy * p.xy;
        const R a_yyz = R{2} * uy * p.yz + uz * p.yy;
        const R a_xzz = ux * p.zz + R{2} * uz * p.xz;
        const R a_yzz = uy * p.zz + R{2} * uz * p.yz;
        const R a_xyz = ux * p.yz + uy * p.xz + uz * p.xy;
        for (int i = 0; i < L::Q; ++i) {
            const R cx = static_cast<R>(L::c[i * 3]);
            const R cy = static_cast<R>(L::c[i * 3 + 1]);
            const R cz = static_cast<R>(L::c[i * 3 + 2]);
            const R q_xx = cx * cx - cs2;
            const R q_yy = cy * cy - cs2;
            const R q_zz = cz * cz - cs2;
            const R q_xy = cx * cy;
            const R q_xz = cx * cz;
            const R q_yz = cy * cz;
            const R second = q_xx * p.xx + q_yy * p.yy + q_zz * p.zz +
                             R{2} * (q_xy * p.xy + q_xz * p.xz + q_yz * p.yz);
            const R h_xxx = cx * cx * cx - R{3} * cs2 * cx;
            const R h_yyy = cy * cy * cy - R{3} * cs2 * cy;
            const R h_zzz = cz * cz * cz - R{3} * cs2 * cz;
            const R h_xxy = cx * cx * cy - cs2 * cy;
            const R h_xxz = cx * cx * cz - cs2 * cz;
            const R h_xyy = cx * cy * cy - cs2 * cx;
            const R h_yyz = cy * cy * cz - cs2 * cz;
            const R h_xzz = cx * cz * cz - cs2 * cx;
            const R h_yzz = cy * cz * cz - cs2 * cy;
            const R h_xyz = cx * cy * cz;
            const R third = a_xxx * h_xxx + a_yyy * h_yyy + a_zzz * h_zzz +
                            R{3} * (a_xxy * h_xxy + a_xxz * h_xxz + a_xyy * h_xyy + a_yyz * h_yyz +
                                    a_xzz * h_xzz + a_yzz * h_yzz) +
                            R{6} * a_xyz * h_xyz;
            const R rec = static_cast<R>(L::w[i]) * pref * (second + third);
            f[i] = eq[i] + (R{1} - omega) * rec;
        }
    }
}

// ---------------------------------------------------------------------------
// MRT — include/lattice_mrt.hpp.
// The inverse transforms of D2Q9 and D3Q19 are M^T D^-1 (D = squared row
// norms) summed in the reference's term order: natural order for D2Q9
// (lattice_mrt.hpp:470-534), and m0..m9, m14, m10, m15, m11..m13, m16..m18 for
// D3Q19 (lattice_mrt.hpp:674-892).  Each coefficient is the correctly rounded
// quotient M[r][i] / D[r], which equals the reference's reduced constants
// (e.g. -30/2394 = -5/399).
// ---------------------------------------------------------------------------
// Moment basis rows as integer polynomials of the velocity.
static constexpr int MD2Q9[9][9] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1},
    {-4, -1, -1, -1, -1, 2, 2, 2, 2},
    {4, -2, -2, -2, -2, 1, 1, 1, 1},
    {0, 1, 0, -1, 0, 1, -1, -1, 1},
    {0, -2, 0, 2, 0, 1, -1, -1, 1},
    {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1},
    {0, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 1, -1, 1, -1}};

// D3Q19 rows from the moment expressions in lattice_mrt.hpp:617-661.
static int MD3Q19[19][19];
static bool md3q19_ready = false;
static void build_md3q19() {
    if (md3q19_ready) return;
    for (int i = 0; i < 19; ++i) {
        const int cx = D3Q19::c[i * 3], cy = D3Q19::c[i * 3 + 1], cz = D3Q19::c[i * 3 + 2];
        const int c2 = cx * cx + cy * cy + cz * cz;
        MD3Q19[0][i] = 1;
        MD3Q19[1][i] = 19 * c2 - 30;
        MD3Q19[2][i] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
        MD3Q19[3][i] = cx;
        MD3Q19[4][i] = (5 * c2 - 9) * cx;
        MD3Q19[5][i] = cy;
        MD3Q19[6][i] = (5 * c2 - 9) * cy;
        MD3Q19[7][i] = cz;
        MD3Q19[8][i] = (5 * c2 - 9) * cz;
        MD3Q19[9][i] = 3 * cx * cx - c2;
        MD3Q19[10][i] = cy * cy - cz * cz;
        MD3Q19[11][i] = cx * cy;
        MD3Q19[12][i] = cy * cz;
        MD3Q19[13][i] = cx * cz;
        MD3Q19[14][i] = (3 * c2 - 5) * (3 * cx * cx - c2);
        MD3Q19[15][i] = (3 * c2 - 5) * (cy * cy - cz * cz);
        MD3Q19[16][i] = (cy * cy - cz * cz) * cx;
        MD3Q19[17][i] = (cz * cz - cx * cx) * cy;
        MD3Q19[18][i] = (cx * cx - cy * cy) * cz;
    }
    md3q19_ready = true;
}
static constexpr int D3Q19_INV_ORDER[19] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 14, 10, 15, 11, 12, 13, 16, 17, 18};

template <class R, int Q>
void inverse_by_rows(const int (*M)[Q], const int* order, const R* m, R* f) {
    int norm[Q];
    for (int r = 0; r < Q; ++r) {
        norm[r] = 0;
        for (int i = 0; i < Q; ++i) norm[r] += M[r][i] * M[r][i];
    }
    for (int i = 0; i < Q; ++i) {
        bool first = true;
        R acc{};
        for (int k = 0; k < Q; ++k) {
            const int r = order[k];
            if (M[r][i] == 0) continue;
            const R coef = static_cast<R>(M[r][i]) / static_cast<R>(norm[r]);
            if (first) { acc = coef * m[r]; first = false; }
            else acc = acc + coef * m[r];
        }
        f[i] = acc;
    }
}

// compute_moments_d2q9, lattice_mrt.hpp:439-455
template <class R> void moments_d2q9(const R* f, R* m) {
    m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    m[1] = -R{4} * f[0] - f[1] - f[2] - f[3] - f[4] + R{2} * (f[5] + f[6] + f[7] + f[8]);
    m[2] = R{4} * f[0] - R{2} * (f[1] + f[2] + f[3] + f[4]) + f[5] + f[6] + f[7] + f[8];
    m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    m[4] = -R{2} * f[1] + R{2} * f[3] + f[5] - f[6] - f[7] + f[8];
    m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    m[6] = -R{2} * f[2] + R{2} * f[4] + f[5] + f[6] - f[7] - f[8];
    m[7] = f[1] - f[2] + f[3] - f[4];
    m[8] = f[5] - f[6] + f[7] - f[8];
}
static constexpr int NATURAL9[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
template <class R> void populations_d2q9(const R* m, R* f) { inverse_by_rows<R, 9>(MD2Q9, NATURAL9, m, f); }

// compute_moments_d3q19, lattice_mrt.hpp:617-661
template <class R> void moments_d3q19(const R* f, R* m) {
    m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] +
           f[14] + f[15] + f[16] + f[17] + f[18];
    m[1] = -R{30} * f[0] - R{11} * (f[1] + f[2] + f[3] + f[4] + f[5] + f[6]) +
           R{8} * (f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18]);
    m[2] = R{12} * f[0] - R{4} * (f[1] + f[2] + f[3] + f[4] + f[5] + f[6]) + f[7] + f[8] + f[9] + f[10] + f[11] +
           f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    m[3] = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
    m[4] = -R{4} * f[1] + R{4} * f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
    m[5] = f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18];
    m[6] = -R{4} * f[3] + R{4} * f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18];
    m[7] = f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18];
    m[8] = -R{4} * f[5] + R{4} * f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18];
    m[9] = R{2} * (f[1] + f[2]) - f[3] - f[4] - f[5] - f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] +
           f[14] - R{2} * (f[15] + f[16] + f[17] + f[18]);
    m[10] = f[3] + f[4] - f[5] - f[6] + f[7] + f[8] + f[9] + f[10] - f[11] - f[12] - f[13] - f[14];
    m[11] = f[7] - f[8] - f[9] + f[10];
    m[12] = f[15] - f[16] - f[17] + f[18];
    m[13] = f[11] - f[12] - f[13] + f[14];
    m[14] = -R{4} * (f[1] + f[2]) + R{2} * (f[3] + f[4] + f[5] + f[6]) + f[7] + f[8] + f[9] + f[10] + f[11] +
            f[12] + f[13] + f[14] - R{2} * (f[15] + f[16] + f[17] + f[18]);
    m[15] = -R{2} * (f[3] + f[4]) + R{2} * (f[5] + f[6]) + f[7] + f[8] + f[9] + f[10] - f[11] - f[12] - f[13] - f[14];
    m[16] = f[7] - f[8] + f[9] - f[10] - f[11] + f[12] - f[13] + f[14];
    m[17] = -f[7] - f[8] + f[9] + f[10] + f[15] - f[16] + f[17] - f[18];
    m[18] = f[11] + f[12] - f[13] - f[14] - f[15] - f[16] + f[17] + f[18];
}
template <class R> void populations_d3q19(const R* m, R* f) {
    build_md3q19();
    inverse_by_rows<R, 19>(MD3Q19, D3Q19_INV_ORDER, m, f);
}

// collide_mrt_d2q9, lattice_mrt.hpp:548-603 (rates :38-65)
template <class R> void collide_mrt_d2q9(R* f, const Macro<D2Q9, R>& mac, R s_nu) {
    R m[9];
    moments_d2q9<R>(f, m);
    const R rho = mac.rho, ux = mac.u[0], uy = mac.u[1];
    const R jx = rho * ux, jy = rho * uy;
    const R usq = ux * ux + uy * uy;
    const R eq[9] = {rho, -R{2} * rho + R{3} * rho * usq, rho - R{3} * rho * usq, jx, -jx, jy, -jy,
                     rho * (ux * ux - uy * uy), rho * ux * uy};
    const R s_e = R{1.63}, s_eps = R{1.14}, s_q = R{1.92};
    m[1] -= s_e * (m[1] - eq[1]);
    m[2] -= s_eps * (m[2] - eq[2]);
    m[4] -= s_q * (m[4] - eq[4]);
    m[6] -= s_q * (m[6] - eq[6]);
    m[7] -= s_nu * (m[7] - eq[7]);
    m[8] -= s_nu * (m[8] - eq[8]);
    populations_d2q9<R>(m, f);
}

// collide_mrt_d3q19, lattice_mrt.hpp:908-1009; ghost rates fixed to 1 (:980-1001)
template <class R> void collide_mrt_d3q19(R* f, const Macro<D3Q19, R>& mac, R omega) {
    R m[19];
    moments_d3q19<R>(f, m);
    const R rho = mac.rho, ux = mac.u[0], uy = mac.u[1], uz = mac.u[2];
    const R jx = rho * ux, jy = rho * uy, jz = rho * uz;
    const R ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const R usq = ux2 + uy2 + uz2;
    const R m9 = R{3} * rho * (R{2} * ux2 - uy2 - uz2);
    const R m10 = rho * (uy2 - uz2);
    const R eq[19] = {rho,
                      -R{11} * rho + R{19} * rho * usq,
                      R{3} * rho - (R{11} / R{2}) * rho * usq,
                      jx, -(R{2} / R{3}) * jx,
                      jy, -(R{2} / R{3}) * jy,
                      jz, -(R{2} / R{3}) * jz,
                      m9, m10, rho * ux * uy, rho * uy * uz, rho * ux * uz,
                      R{0}, R{0}, R{0}, R{0}, R{0}};
    const R rates[19] = {R{0}, R{1}, R{1}, R{0}, R{1}, R{0}, R{1}, R{0}, R{1}, omega,
                         omega, omega, omega, omega, R{1}, R{1}, R{1}, R{1}, R{1}};
    for (int i = 0; i < 19; ++i) m[i] -= rates[i] * (m[i] - eq[i]);
    populations_d3q19<R>(m, f);
}

// D3Q27 separable tensor-product transforms, lattice_mrt.hpp:206-426, 1023-1225.
template <class R> R axis_basis(int order, int c) {
    const R v = static_cast<R>(c);
    if (order == 0) return R{1};
    if (order == 1) return v;
    return R{3} * v * v - R{2};
}
template <class R> R axis_norm(int order) { return order == 0 ? R{3} : (order == 1 ? R{2} : R{6}); }
inline int tix(int a, int b, int c) { return (a * 3 + b) * 3 + c; }

template <class R> void moments_d3q27(const R* f, R* mo) {
    R p[27], xt[27], xyt[27], raw[27];
    for (int i = 0; i < 27; ++i)
        p[tix(D3Q27::c[i * 3] + 1, D3Q27::c[i * 3 + 1] + 1, D3Q27::c[i * 3 + 2] + 1)] = f[i];
    for (int ax = 0; ax < 3; ++ax)
        for (int iy = 0; iy < 3; ++iy)
            for (int iz = 0; iz < 3; ++iz) {
                R v{};
                for (int ix = 0; ix < 3; ++ix) v += axis_basis<R>(ax, ix - 1) * p[tix(ix, iy, iz)];
                xt[tix(ax, iy, iz)] = v;
            }
    for (int ax = 0; ax < 3; ++ax)
        for (int ay = 0; ay < 3; ++ay)
            for (int iz = 0; iz < 3; ++iz) {
                R v{};
                for (int iy = 0; iy < 3; ++iy) v += axis_basis<R>(ay, iy - 1) * xt[tix(ax, iy, iz)];
                xyt[tix(ax, ay, iz)] = v;
            }
    for (int ax = 0; ax < 3; ++ax)
        for (int ay = 0; ay < 3; ++ay)
            for (int az = 0; az < 3; ++az) {
                R v{};
                for (int iz = 0; iz < 3; ++iz) v += axis_basis<R>(az, iz - 1) * xyt[tix(ax, ay, iz)];
                raw[tix(ax, ay, az)] = v;
            }
    const R r200 = raw[tix(2, 0, 0)], r020 = raw[tix(0, 2, 0)], r002 = raw[tix(0, 0, 2)];
    mo[0] = raw[tix(0, 0, 0)];
    mo[1] = raw[tix(1, 0, 0)];
    mo[2] = raw[tix(0, 1, 0)];
    mo[3] = raw[tix(0, 0, 1)];
    mo[4] = r200 + r020 + r002;
    mo[5] = R{2} * r200 - r020 - r002;
    mo[6] = r020 - r002;
    static constexpr int rest[21][3] = {{1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1},
                                        {1, 0, 2}, {0, 1, 2}, {1, 1, 1}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1},
                                        {1, 2, 1}, {1, 1, 2}, {2, 2, 1}, {2, 1, 2}, {1, 2, 2}, {2, 2, 2}, {0, 0, 0}};
    for (int k = 0; k < 20; ++k) mo[7 + k] = raw[tix(rest[k][0], rest[k][1], rest[k][2])];
}

template <class R> void populations_d3q27(const R* mo, R* f) {
    R raw[27], zi[27], yzi[27], pt[27];
    raw[tix(0, 0, 0)] = mo[0];
    raw[tix(1, 0, 0)] = mo[1];
    raw[tix(0, 1, 0)] = mo[2];
    raw[tix(0, 0, 1)] = mo[3];
    raw[tix(2, 0, 0)] = (mo[4] + mo[5]) / R{3};
    raw[tix(0, 2, 0)] = (R{2} * mo[4] - mo[5] + R{3} * mo[6]) / R{6};
    raw[tix(0, 0, 2)] = (R{2} * mo[4] - mo[5] - R{3} * mo[6]) / R{6};
    static constexpr int rest[20][3] = {{1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {0, 2, 1},
                                        {1, 0, 2}, {0, 1, 2}, {1, 1, 1}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1},
                                        {1, 2, 1}, {1, 1, 2}, {2, 2, 1}, {2, 1, 2}, {1, 2, 2}, {2, 2, 2}};
    for (int k = 0; k < 20; ++k) raw[tix(rest[k][0], rest[k][1], rest[k][2])] = mo[7 + k];
    for (int ax = 0; ax < 3; ++ax)
        for (int ay = 0; ay < 3; ++ay)
            for (int iz = 0; iz < 3; ++iz) {
                R v{};
                for (int az = 0; az < 3; ++az)
                    v += raw[tix(ax, ay, az)] * axis_basis<R>(az, iz - 1) / axis_norm<R>(az);
                zi[tix(ax, ay, iz)] = v;
            }
    for (int ax = 0; ax < 3; ++ax)
        for (int iy = 0; iy < 3; ++iy)
            for (int iz = 0; iz < 3; ++iz) {
                R v{};
                for (int ay = 0; ay < 3; ++ay)
                    v += zi[tix(ax, ay, iz)] * axis_basis<R>(ay, iy - 1) / axis_norm<R>(ay);
                yzi[tix(ax, iy, iz)] = v;
            }
    for (int ix = 0; ix < 3; ++ix)
        for (int iy = 0; iy < 3; ++iy)
            for (int iz = 0; iz < 3; ++iz) {
                R v{};
                for (int ax = 0; ax < 3; ++ax)
                    v += yzi[tix(ax, iy, iz)] * axis_basis<R>(ax, ix - 1) / axis_norm<R>(ax);
                pt[tix(ix, iy, iz)] = v;
            }
    for (int i = 0; i < 27; ++i)
        f[i] = pt[tix(D3Q27::c[i * 3] + 1, D3Q27::c[i * 3 + 1] + 1, D3Q27::c[i * 3 + 2] + 1)];
}

// collide_mrt_d3q27, lattice_mrt.hpp:1242-1305 (rates :161-204)
template <class R> void collide_mrt_d3q27(R* f, const Macro<D3Q27, R>& mac, R s_nu) {
    R m[27], feqp[27], eq[27];
    moments_d3q27<R>(f, m);
    const R rho = mac.rho, ux = mac.u[0], uy = mac.u[1], uz = mac.u[2];
    const R usq = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < 27; ++i) {
        const R cx = static_cast<R>(D3Q27::c[i * 3]);
        const R cy = static_cast<R>(D3Q27::c[i * 3 + 1]);
        const R cz = static_cast<R>(D3Q27::c[i * 3 + 2]);
        const R cu = cx * ux + cy * uy + cz * uz;
        feqp[i] = static_cast<R>(D3Q27::w[i]) * rho * (R{1} + R{3} * cu + R{4.5} * cu * cu - R{1.5} * usq);
    }
    moments_d3q27<R>(feqp, eq);
    const R s_b = R{1.19}, s_e = R{1.1}, s_eps = R{1.2}, s_q = R{1.2};
    const R rates[27] = {R{0}, R{0}, R{0}, R{0}, s_b, s_nu, s_nu, s_nu, s_nu, s_nu, s_e, s_e, s_e, s_e,
                         s_e, s_e, s_eps, s_q, s_q, s_q, s_q, s_q, s_q, s_q, s_q, s_q, s_q};
    for (int i = 0; i < 27; ++i) m[i] -= rates[i] * (m[i] - eq[i]);
    populations_d3q27<R>(m, f);
}

// detail::collide_cell, lattice_core.hpp:90-126
template <class L, int CT, class R> void collide_cell(R* f, R omega) {
    if constexpr (CT == C_BGK) {
        collide_bgk<L, R>(f, omega);
    } else if constexpr (CT == C_TRT) {
        const Macro<L, R> m = macro_state<L, R>(f);
        collide_trt<L, R>(f, m, omega, omega_minus<R>(omega));
    } else if constexpr (CT == C_MRT) {
        const Macro<L, R> m = macro_state<L, R>(f);
        if constexpr (L::Q == 9) collide_mrt_d2q9<R>(f, m, omega);
        else if constexpr (L::Q == 19) collide_mrt_d3q19<R>(f, m, omega);
        else collide_mrt_d3q27<R>(f, m, omega);
    } else {
        const Macro<L, R> m = macro_state<L, R>(f);
        collide_rlbm<L, R>(f, m, omega);
    }
}

// Scalar physics, lattice_physics.hpp:618-694
template <class S, class R> R concentration(const R* g) {
    R c{};
    for (int i = 0; i < S::Q; ++i) c += g[i];
    return c;
}
template <class S, class R> R scalar_feq(int i, R conc, const R* u) {
    R cu{};
    for (int d = 0; d < S::D; ++d) cu += static_cast<R>(S::c[i * S::D + d]) * u[d];
    const R weight = static_cast<R>(S::w[i]);
    const R cs2 = static_cast<R>(S::cs2);
    return weight * conc * (static_cast<R>(1.0) + cu / cs2);
}
template <class S, class R> void collide_scalar(R* g, const R* u, R omega_c, R source) {
    const R conc = concentration<S, R>(g);
    for (int i = 0; i < S::Q; ++i) {
        const R weight = static_cast<R>(S::w[i]);
        const R eq = scalar_feq<S, R>(i, conc, u);
        g[i] += omega_c * (eq - g[i]) + weight * source;
    }
}

// compute_reaction_ab_source, lattice_physics.hpp:712-743
template <class R> R reaction_source(R a, R b, R k) {
    const R abs_a = a < R{} ? -a : a;
    const R abs_b = b < R{} ? -b : b;
    const R diff = b - a;
    const R abs_diff = diff < R{} ? -diff : diff;
    const R tol = static_cast<R>(64.0) * std::numeric_limits<R>::epsilon() * (static_cast<R>(1.0) + abs_a + abs_b);
    R after{};
    if (abs_diff <= tol) {
        after = a / (static_cast<R>(1.0) + k * a);
    } else {
        using std::exp;
        const R att = exp(-k * diff);
        after = a * diff * att / (b - a * att);
    }
    return after - a;
}

// periodic_pull_index, lattice_core.hpp:70-78
inline i64 pull(i64 x, int c, i64 n) { return (x - c + n) % n; }

struct Grid {
    i64 nx, ny, nz;
    i64 cells() const { return nx * ny * nz; }
};

// step_cpu, lattice_core.hpp:146-216 (no swap: the caller swaps buffers).
template <class L, int CT, int DT>
void fluid_step(const void* curv, void* nxtv, Grid g, double omega_d, int threads) {
    using C = Codec<DT>;
    using S = typename C::S;
    using R = typename C::R;
    const S* cur = static_cast<const S*>(curv);
    S* nxt = static_cast<S*>(nxtv);
    const R omega = static_cast<R>(omega_d);
    const i64 N = g.cells();
    const bool off = (DT == T_F16 || DT == T_BF16);
    (void)threads;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (i64 zy = 0; zy < g.nz * g.ny; ++zy) {
        const i64 z = zy / g.ny, y = zy % g.ny;
        for (i64 x = 0; x < g.nx; ++x) {
            R f[L::Q];
            for (int i = 0; i < L::Q; ++i) {
                const i64 sx = pull(x, L::c[i * L::D], g.nx);
                const i64 sy = pull(y, L::c[i * L::D + 1], g.ny);
                const i64 sz = L::D == 3 ? pull(z, L::c[i * L::D + 2], g.nz) : 0;
                const R o = off ? static_cast<R>(L::w[i]) : R{0};
                f[i] = C::load(cur[i * N + (sz * g.ny + sy) * g.nx + sx], o);
            }
            collide_cell<L, CT, R>(f, omega);
            for (int i = 0; i < L::Q; ++i) {
                const R o = off ? static_cast<R>(L::w[i]) : R{0};
                nxt[i * N + (z * g.ny + y) * g.nx + x] = C::store(f[i], o);
            }
        }
    }
}

// step_scalar_cpu (lattice_core.hpp:238-337) and step_reaction_AB
// (lattice_core.hpp:357-474), one cell at a time; `reaction` selects the
// fused A+B->C form.  The fluid velocity comes from the fluid buffer at the
// destination cell, unshifted.
// Storage offset of scalar population i of a field with reference
// concentration ref: F16/BF16 hold g_i - w_i ref (formed in the arithmetic
// type, as the kernels form it); FP32/FP64 codecs ignore it.
template <class S, class R> R scalar_offset(int i, R ref) { return static_cast<R>(S::w[i]) * ref; }

template <class F, class S, int DT>
void scalar_like_step(const void* cur_av, void* nxt_av, const void* cur_bv, void* nxt_bv, const void* cur_fv, Grid g,
                      double omega_cd, double param, bool reaction, int threads, double ref_a = 0.0,
                      double ref_b = 0.0) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    const St* ca = static_cast<const St*>(cur_av);
    St* na = static_cast<St*>(nxt_av);
    const St* cb = static_cast<const St*>(cur_bv);
    St* nb = static_cast<St*>(nxt_bv);
    const St* cf = static_cast<const St*>(cur_fv);
    const R omega_c = static_cast<R>(omega_cd);
    const R prm = static_cast<R>(param);
    const R ra = static_cast<R>(ref_a), rb = static_cast<R>(ref_b);
    const i64 N = g.cells();
    const bool off = (DT == T_F16 || DT == T_BF16);
    (void)threads;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (i64 zy = 0; zy < g.nz * g.ny; ++zy) {
        const i64 z = zy / g.ny, y = zy % g.ny;
        for (i64 x = 0; x < g.nx; ++x) {
            const i64 cell = (z * g.ny + y) * g.nx + x;
            R fl[F::Q];
            for (int i = 0; i < F::Q; ++i)
                fl[i] = C::load(cf[i * N + cell], off ? static_cast<R>(F::w[i]) : R{0});
            const Macro<F, R> m = macro_state<F, R>(fl);
            R ga[S::Q], gb[S::Q];
            for (int i = 0; i < S::Q; ++i) {
                const i64 sx = pull(x, S::c[i * S::D], g.nx);
                const i64 sy = pull(y, S::c[i * S::D + 1], g.ny);
                const i64 sz = S::D == 3 ? pull(z, S::c[i * S::D + 2], g.nz) : 0;
                const i64 src = i * N + (sz * g.ny + sy) * g.nx + sx;
                ga[i] = C::load(ca[src], scalar_offset<S, R>(i, ra));
                if (reaction) gb[i] = C::load(cb[src], scalar_offset<S, R>(i, rb));
            }
            if (reaction) {
                const R conc_a = concentration<S, R>(ga);
                const R conc_b = concentration<S, R>(gb);
                const R src = reaction_source<R>(conc_a, conc_b, prm);
                collide_scalar<S, R>(ga, m.u, omega_c, src);
                collide_scalar<S, R>(gb, m.u, omega_c, src);
            } else {
                collide_scalar<S, R>(ga, m.u, omega_c, prm);
            }
            for (int i = 0; i < S::Q; ++i) {
                na[i * N + cell] = C::store(ga[i], scalar_offset<S, R>(i, ra));
                if (reaction) nb[i * N + cell] = C::store(gb[i], scalar_offset<S, R>(i, rb));
            }
        }
    }
}

// compute_reactive_stats, lattice_io.hpp:256-338 (accumulation in Real as the
// reference does) plus a long double accumulation used as the brute-force sum.
template <class S, int DT>
void reactive_stats(const void* av, const void* bv, Grid g, double k_d, double* out, bool wide, double ref_a = 0.0,
                    double ref_b = 0.0) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    const R ra = static_cast<R>(ref_a), rb = static_cast<R>(ref_b);
    const St* a = static_cast<const St*>(av);
    const St* b = static_cast<const St*>(bv);
    const i64 N = g.cells();
    if (!wide) {
        R sa{}, sb{}, saa{}, sbb{}, sab{};
        for (i64 cell = 0; cell < N; ++cell) {
            R ca{}, cb{};
            for (int i = 0; i < S::Q; ++i) ca += C::load(a[i * N + cell], scalar_offset<S, R>(i, ra));
            for (int i = 0; i < S::Q; ++i) cb += C::load(b[i * N + cell], scalar_offset<S, R>(i, rb));
            sa += ca; sb += cb; saa += ca * ca; sbb += cb * cb; sab += ca * cb;
        }
        const R k = static_cast<R>(k_d);
        const R inv = R{1} / static_cast<R>(N);
        const R ma = sa * inv, mb = sb * inv, maa = saa * inv, mbb = sbb * inv, mab = sab * inv;
        const R mixed = k * ma * mb;
        const R cov = mab - ma * mb;
        const R mp = ma * mb;
        const R seg = std::abs(mp) > std::numeric_limits<R>::epsilon() ? cov / mp : R{};
        const R res[8] = {ma, mb, maa - ma * ma, mbb - mb * mb, cov, seg, k * mab, mixed};
        for (int j = 0; j < 8; ++j) out[j] = static_cast<double>(res[j]);
    } else {
        long double sa = 0, sb = 0, saa = 0, sbb = 0, sab = 0;
        for (i64 cell = 0; cell < N; ++cell) {
            R ca{}, cb{};
            for (int i = 0; i < S::Q; ++i) ca += C::load(a[i * N + cell], scalar_offset<S, R>(i, ra));
            for (int i = 0; i < S::Q; ++i) cb += C::load(b[i * N + cell], scalar_offset<S, R>(i, rb));
            const long double la = ca, lb = cb;
            sa += la; sb += lb; saa += la * la; sbb += lb * lb; sab += la * lb;
        }
        const long double k = k_d;
        const long double inv = 1.0L / static_cast<long double>(N);
        const long double ma = sa * inv, mb = sb * inv, maa = saa * inv, mbb = sbb * inv, mab = sab * inv;
        const long double cov = mab - ma * mb;
        const long double mp = ma * mb;
        const long double seg = std::fabs(mp) > std::numeric_limits<R>::epsilon() ? cov / mp : 0.0L;
        const long double res[8] = {ma, mb, maa - ma * ma, mbb - mb * mb, cov, seg, k * mab, k * ma * mb};
        for (int j = 0; j < 8; ++j) out[j] = static_cast<double>(res[j]);
    }
}

// log_diagnostics, lattice_io.hpp:182-234: total mass, total kinetic energy
// (sum of rho |u|^2 / 2) and max |u|.  `wide` accumulates in long double.
template <class L, int DT> void fluid_diagnostics(const void* fv, Grid g, double* out, bool wide) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    const St* f = static_cast<const St*>(fv);
    const i64 N = g.cells();
    const bool off = (DT == T_F16 || DT == T_BF16);
    R mass{}, ke{}, umax{};
    long double lmass = 0, lke = 0;
    for (i64 cell = 0; cell < N; ++cell) {
        R loc[L::Q];
        for (int i = 0; i < L::Q; ++i) loc[i] = C::load(f[i * N + cell], off ? static_cast<R>(L::w[i]) : R{0});
        const Macro<L, R> m = macro_state<L, R>(loc);
        const R sp2 = vel_dot<L, R>(m.u);
        if (wide) {
            lmass += m.rho;
            lke += 0.5L * static_cast<long double>(m.rho) * static_cast<long double>(sp2);
        } else {
            mass += m.rho;
            ke += static_cast<R>(0.5) * m.rho * sp2;
        }
        umax = std::max(umax, std::sqrt(sp2));
    }
    out[0] = wide ? static_cast<double>(lmass) : static_cast<double>(mass);
    out[1] = wide ? static_cast<double>(lke) : static_cast<double>(ke);
    out[2] = static_cast<double>(umax);
}

// Per-cell fields printed by the VTK writers (macro_at of write_vtk,
// include/lattice_io.hpp:386-451; concentration_at of write_reaction_vtk,
// src/simulate_reaction.cpp:253-333): compute_macro_state / compute_concentration
// in the storage's arithmetic type, widened to double.
template <class L, int DT> void macro_fields(const void* fv, Grid g, double* rho, double* u) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    const St* f = static_cast<const St*>(fv);
    const i64 N = g.cells();
    const bool off = (DT == T_F16 || DT == T_BF16);
    for (i64 cell = 0; cell < N; ++cell) {
        R loc[L::Q];
        for (int i = 0; i < L::Q; ++i) loc[i] = C::load(f[i * N + cell], off ? static_cast<R>(L::w[i]) : R{0});
        const Macro<L, R> m = macro_state<L, R>(loc);
        rho[cell] = static_cast<double>(m.rho);
        for (int d = 0; d < L::D; ++d) u[d * N + cell] = static_cast<double>(m.u[d]);
    }
}

template <class S, int DT> void scalar_fields(const void* gv, Grid g, double* conc, double ref) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    const R r = static_cast<R>(ref);
    const St* p = static_cast<const St*>(gv);
    const i64 N = g.cells();
    for (i64 cell = 0; cell < N; ++cell) {
        R loc[S::Q];
        for (int i = 0; i < S::Q; ++i) loc[i] = C::load(p[i * N + cell], scalar_offset<S, R>(i, r));
        conc[cell] = static_cast<double>(concentration<S, R>(loc));
    }
}

// compute_max_velocity, src/turbulent_tgv_3d.cpp:149-173: the largest |u|
// (sqrt of the squared norm in the arithmetic type), infinity as soon as one
// cell's speed is not finite.
template <class L, int DT> double max_speed(const void* fv, Grid g) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    const St* f = static_cast<const St*>(fv);
    const i64 N = g.cells();
    const bool off = (DT == T_F16 || DT == T_BF16);
    R best{};
    for (i64 cell = 0; cell < N; ++cell) {
        R loc[L::Q];
        for (int i = 0; i < L::Q; ++i) loc[i] = C::load(f[i * N + cell], off ? static_cast<R>(L::w[i]) : R{0});
        const Macro<L, R> m = macro_state<L, R>(loc);
        const R speed = std::sqrt(vel_dot<L, R>(m.u));
        if (!std::isfinite(speed)) return std::numeric_limits<double>::infinity();
        best = std::max(best, speed);
    }
    return static_cast<double>(best);
}

// Equilibrium fields from macroscopic fields (initialisers of
// src/main.cpp:368-411 and src/simulate_reaction.cpp:197-232 go through
// compute_equilibrium / compute_scalar_equilibrium per cell).
template <class L, int DT> void feq_field(const double* rho, const double* u, void* outv, i64 N) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    St* out = static_cast<St*>(outv);
    const bool off = (DT == T_F16 || DT == T_BF16);
    for (i64 cell = 0; cell < N; ++cell) {
        Macro<L, R> m;
        m.rho = static_cast<R>(rho[cell]);
        for (int d = 0; d < L::D; ++d) m.u[d] = static_cast<R>(u[d * N + cell]);
        for (int i = 0; i < L::Q; ++i)
            out[i * N + cell] = C::store(feq<L, R>(i, m), off ? static_cast<R>(L::w[i]) : R{0});
    }
}
template <class S, int DT> void scalar_feq_field(const double* conc, const double* u, void* outv, i64 N, double ref) {
    using C = Codec<DT>;
    using St = typename C::S;
    using R = typename C::R;
    const R r = static_cast<R>(ref);
    St* out = static_cast<St*>(outv);
    for (i64 cell = 0; cell < N; ++cell) {
        R uu[3];
        for (int d = 0; d < S::D; ++d) uu[d] = static_cast<R>(u[d * N + cell]);
        for (int i = 0; i < S::Q; ++i)
            out[i * N + cell] = C::store(scalar_feq<S, R>(i, static_cast<R>(conc[cell]), uu), scalar_offset<S, R>(i, r));
    }
}


// ===========================================================================
// Discrete adjoint of the reactive loop, 2D (D2Q9 + 2 D2Q5) and 3D (D3Q19 /
// D3Q27 + 2 D3Q7) (project-written; no counterpart in the reference).
// Forward step n: f_{n+1} = stream+collide(f_n) (BGK, TRT, MRT or RLBM),
// then (a_{n+1}, b_{n+1}) = A+B->C step of (a_n, b_n) advected by u(f_{n+1})
// at the same cell (the order of src/simulate_reaction.cpp:548-554).
// Adjoint buffers are kept "at the collision cell": nu[i][x] holds the
// derivative with respect to the population i gathered by cell x, so the
// adjoint of the stored state is Lambda[i][y] = nu[i][y + c_i] (reverse pull).
// Pinned by central finite differences of the forward and the one-step
// dot-product identity (tests).
// ===========================================================================

// Derivatives of the exact one-step A+B->C map.  With D = B - A (invariant)
// and z = k D, A' = A / (1 + k A phi(z) + expm1(z)), phi(z) = expm1(z) / z; the
// form has no cancellation at A = B.  Returns dS/dA and dS/dB of S = A' - A.
// phi(z) = expm1(z)/z = sum_n z^n/(n+1)! and phi'(z) = sum_n n z^(n-1)/(n+1)!.
// The closed form phi' = (z e^z - expm1(z))/z^2 cancels near 0 (its numerator
// is ~z^2/2), so below |z| = 1 (float) or 0.1 (double) both come from the
// series through z^10, whose first omitted term is below one ulp there;
// above, phi' = (z + expm1(z)(z - 1))/z^2 (at most ~5 ulps lost, at z = -1).
template <class R> void phi_and_slope(R z, R& phi, R& dphi, R& em1) {
    const R limit = sizeof(R) == 4 ? R(1) : R(0.1);
    if (std::fabs(z) < limit) {
        // phi: 1/(m+1)! for z^m, m = 0..10; phi': n/(n+1)! = 1/(n! (n+1)/n) for z^(n-1), n = 1..11
        static constexpr double phi_den[11] = {1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800, 39916800};
        static constexpr double slope_den[11] = {2, 3, 8, 30, 144, 840, 5760, 45360, 403200, 3991680, 43545600};
        phi = R(0);
        dphi = R(0);
        for (int m = 10; m >= 0; --m) phi = phi * z + R(1) / static_cast<R>(phi_den[m]);
        for (int m = 10; m >= 0; --m) dphi = dphi * z + R(1) / static_cast<R>(slope_den[m]);
        em1 = z * phi;
    } else {
        em1 = std::expm1(z);
        phi = em1 / z;
        dphi = (z + em1 * (z - R(1))) / (z * z);
    }
}

template <class R> void reaction_grad(R a, R b, R k, R& dsa, R& dsb) {
    const R d = b - a;
    const R z = k * d;
    R phi, dphi, em1;
    phi_and_slope(z, phi, dphi, em1);
    // A' = A / (1 + s), s = k A phi + expm1(z).  dA'/dA - 1 would subtract 1
    // from a quotient near 1; with e^z - phi = z phi' and z + k A = k B:
    //   dS/dA = (k^2 A B phi' - k B phi - s^2) / (1 + s)^2
    //   dS/dB = -k A (k A phi' + e^z) / (1 + s)^2
    const R e = R(1) + em1;
    const R s = k * a * phi + em1;
    const R den = R(1) + s;
    dsa = (k * k * a * b * dphi - k * b * phi - s * s) / (den * den);
    dsb = -a * k * (k * a * dphi + e) / (den * den);
}

// Rows of the D3Q19 moment basis (M[r][i] = moments of the unit population
// e_i) and their norms sum_i M[r][i]^2, for the MRT adjoint.
struct D3q19Rows {
    double m[19][19];
    double norm[19];
};
template <class R> D3q19Rows d3q19_rows() {
    D3q19Rows b{};
    for (int i = 0; i < 19; ++i) {
        R e[19] = {}, mo[19];
        e[i] = R(1);
        moments_d3q19<R>(e, mo);
        for (int r = 0; r < 19; ++r) b.m[r][i] = static_cast<double>(mo[r]);
    }
    for (int r = 0; r < 19; ++r) {
        b.norm[r] = 0;
        for (int i = 0; i < 19; ++i) b.norm[r] += b.m[r][i] * b.m[r][i];
    }
    return b;
}

// The fluid lattice F (D2Q9, D3Q19, D3Q27; all cs2 = 1/3) and the species
// lattice S (D2Q5, D3Q7) of one reactive pair; BGK, TRT, MRT and RLBM.
template <class F, class S, class R, int CT>
void reactive_adjoint_step(const R* f, const R* a, const R* b, const R* nuf1, const R* nua1, const R* nub1, R* nuf0,
                           R* nua0, R* nub0, Grid g, R omega, R omega_c, R k, long double* djdw) {
    static_assert(F::D == S::D, "a fluid and a species lattice of one dimension");
    constexpr int D = F::D, QF = F::Q, QS = S::Q;
    const i64 N = g.cells();
    const R wm = CT == C_TRT ? omega_minus<R>(omega) : R(0);
    const R cs2 = static_cast<R>(S::cs2);
    long double acc = 0;
    for (i64 z = 0; z < g.nz; ++z)
    for (i64 y = 0; y < g.ny; ++y) {
        for (i64 x = 0; x < g.nx; ++x) {
            const i64 cell = (z * g.ny + y) * g.nx + x;
            // source of a population with velocity c (sign: +1 the pull, -1 the reverse pull)
            auto at = [&](const int* c, int sign) {
                const i64 sx = pull(x, sign * c[0], g.nx), sy = pull(y, sign * c[1], g.ny);
                const i64 sz = D == 3 ? pull(z, sign * c[2], g.nz) : 0;
                return (sz * g.ny + sy) * g.nx + sx;
            };
            R fl[QF], fp[QF], lf[QF];
            for (int i = 0; i < QF; ++i) {
                fl[i] = f[i * N + at(&F::c[i * D], 1)];
                // reverse pull: the adjoint of the population stored at this cell
                lf[i] = nuf1[i * N + at(&F::c[i * D], -1)];
                fp[i] = fl[i];
            }
            collide_cell<F, CT, R>(fp, omega);
            const Macro<F, R> mp = macro_state<F, R>(fp);  // what the scalar step sees
            R ga[QS], gb[QS], la[QS], lb[QS];
            for (int j = 0; j < QS; ++j) {
                const i64 src = j * N + at(&S::c[j * D], 1);
                const i64 rsrc = j * N + at(&S::c[j * D], -1);
                ga[j] = a[src];
                gb[j] = b[src];
                la[j] = nua1[rsrc];
                lb[j] = nub1[rsrc];
            }
            const R ca = concentration<S, R>(ga), cb = concentration<S, R>(gb);
            R dsa, dsb;
            reaction_grad<R>(ca, cb, k, dsa, dsb);
            R ea = 0, eb = 0, wa = 0, wb = 0, pa[3] = {0, 0, 0}, pb[3] = {0, 0, 0};
            for (int j = 0; j < QS; ++j) {
                const R w = static_cast<R>(S::w[j]);
                R cu = 0;
                for (int d = 0; d < D; ++d) cu += static_cast<R>(S::c[j * D + d]) * mp.u[d];
                const R e = R(1) + cu / cs2;
                ea += la[j] * w * e;
                eb += lb[j] * w * e;
                wa += la[j] * w;
                wb += lb[j] * w;
                for (int d = 0; d < D; ++d) {
                    pa[d] += la[j] * w * static_cast<R>(S::c[j * D + d]);
                    pb[d] += lb[j] * w * static_cast<R>(S::c[j * D + d]);
                }
                acc += static_cast<long double>(la[j]) * (scalar_feq<S, R>(j, ca, mp.u) - ga[j]) +
                       static_cast<long double>(lb[j]) * (scalar_feq<S, R>(j, cb, mp.u) - gb[j]);
            }
            const R sw = wa + wb;
            for (int j = 0; j < QS; ++j) {
                nua0[j * N + cell] = (R(1) - omega_c) * la[j] + omega_c * ea + sw * dsa;
                nub0[j * N + cell] = (R(1) - omega_c) * lb[j] + omega_c * eb + sw * dsb;
            }
            R lu[3];
            for (int d = 0; d < D; ++d) lu[d] = omega_c / cs2 * (ca * pa[d] + cb * pb[d]);
            // velocity coupling: d u' / d fp_i = (c_i - u') / rho'
            for (int i = 0; i < QF; ++i) {
                R t = 0;
                for (int d = 0; d < D; ++d) t += lu[d] * (static_cast<R>(F::c[i * D + d]) - mp.u[d]);
                lf[i] += t / mp.rho;
            }
            // transposed collision Jacobian at the pre-collision state:
            // d feq_q / d f_i = w_q [P_q + G_q . (c_i - u)], P_q = feq_q / (w_q rho),
            // G_q = dP_q / du = 3 c_q + 9 (c_q . u) c_q - 3 u (cs2 = 1/3)
            const Macro<F, R> m = macro_state<F, R>(fl);
            R uu = 0;
            for (int d = 0; d < D; ++d) uu += m.u[d] * m.u[d];
            auto moments = [&](const R* lam, R& al, R* bd) {
                al = 0;
                for (int d = 0; d < D; ++d) bd[d] = 0;
                for (int q = 0; q < QF; ++q) {
                    const R w = static_cast<R>(F::w[q]);
                    R cu = 0;
                    for (int d = 0; d < D; ++d) cu += static_cast<R>(F::c[q * D + d]) * m.u[d];
                    const R P = R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu;
                    al += lam[q] * w * P;
                    for (int d = 0; d < D; ++d) {
                        const R cd = static_cast<R>(F::c[q * D + d]);
                        bd[d] += lam[q] * w * (R(3) * cd + R(9) * cu * cd - R(3) * m.u[d]);
                    }
                }
            };
            auto dev = [&](int i, const R* bd) {
                R t = 0;
                for (int d = 0; d < D; ++d) t += bd[d] * (static_cast<R>(F::c[i * D + d]) - m.u[d]);
                return t;
            };
            if constexpr (CT == C_MRT && F::Q == 19) {
                // collide_mrt_d3q19: f' = M^-1 [(I - S) M f + S m_eq(rho, j)]
                // with an orthogonal basis and diagonal S, so M^-1 (I - S) M is
                // symmetric and J^T lam = M^-1 (I - S) M lam + (d m_eq/d f)^T S mu,
                // mu = D^-1 M lam; d m_eq / d f_i = d/drho + c_i . d/dj of the
                // equilibria written in (rho, j): rho u_a u_b = j_a j_b / rho.
                static const D3q19Rows B = d3q19_rows<R>();
                const R rates[19] = {R{0}, R{1}, R{1}, R{0}, R{1}, R{0}, R{1}, R{0}, R{1}, omega,
                                     omega, omega, omega, omega, R{1}, R{1}, R{1}, R{1}, R{1}};
                R mo[19], keep[19], back[19], ne[19];
                moments_d3q19<R>(lf, mo);
                for (int r = 0; r < 19; ++r) {
                    keep[r] = (R(1) - rates[r]) * mo[r];
                    ne[r] = rates[r] * mo[r] / static_cast<R>(B.norm[r]);
                }
                populations_d3q19<R>(keep, back);
                const R ux = m.u[0], uy = m.u[1], uz = m.u[2];
                const R grho = ne[1] * (R(-11) - R(19) * uu) + ne[2] * (R(3) + R(5.5) * uu) -
                               ne[9] * R(3) * (R(2) * ux * ux - uy * uy - uz * uz) - ne[10] * (uy * uy - uz * uz) -
                               ne[11] * ux * uy - ne[12] * uy * uz - ne[13] * ux * uz;
                const R gj[3] = {R(38) * ux * ne[1] - R(11) * ux * ne[2] - R(2) / R(3) * ne[4] + R(12) * ux * ne[9] +
                                     uy * ne[11] + uz * ne[13],
                                 R(38) * uy * ne[1] - R(11) * uy * ne[2] - R(2) / R(3) * ne[6] - R(6) * uy * ne[9] +
                                     R(2) * uy * ne[10] + ux * ne[11] + uz * ne[12],
                                 R(38) * uz * ne[1] - R(11) * uz * ne[2] - R(2) / R(3) * ne[8] - R(6) * uz * ne[9] -
                                     R(2) * uz * ne[10] + uy * ne[12] + ux * ne[13]};
                for (int i = 0; i < 19; ++i) {
                    R t = grho;
                    for (int d = 0; d < 3; ++d) t += static_cast<R>(F::c[i * 3 + d]) * gj[d];
                    nuf0[i * N + cell] = back[i] + t;
                }
            } else if constexpr (CT == C_MRT && F::Q == 27) {
                // collide_mrt_d3q27: f' = f - M^-1 S M (f - feq), M^-1 S M symmetric
                // (orthogonal basis, diagonal S), so J^T lam = lam - w + (d feq/d f)^T w
                // with w = M^-1 S M lam.
                const R s_b = R{1.19}, s_e = R{1.1}, s_eps = R{1.2}, s_q = R{1.2};
                const R rates[27] = {R{0}, R{0}, R{0}, R{0}, s_b, omega, omega, omega, omega, omega, s_e, s_e, s_e,
                                     s_e, s_e, s_e, s_eps, s_q, s_q, s_q, s_q, s_q, s_q, s_q, s_q, s_q, s_q};
                R mo[27], w[27];
                moments_d3q27<R>(lf, mo);
                for (int r = 0; r < 27; ++r) mo[r] *= rates[r];
                populations_d3q27<R>(mo, w);
                R al, bd[3];
                moments(w, al, bd);
                for (int i = 0; i < 27; ++i) nuf0[i * N + cell] = lf[i] - w[i] + al + dev(i, bd);
            } else if constexpr (CT == C_MRT) {
                static_assert(std::is_same<F, D2Q9>::value, "the MRT adjoint covers D2Q9, D3Q19 and D3Q27");
                // f' = M^-1 [(I - S) M f + S m_eq(rho, j)] with M^-1 = M^T D^-1
                // (collide_mrt_d2q9, lattice_mrt.hpp:548-603), so
                // J^T lam = M^T (I - S) mu + (d m_eq / d f)^T S mu, mu = D^-1 M lam,
                // and d/df_i = d/drho + c_ix d/djx + c_iy d/djy.
                static const int rows[9][9] = {{1, 1, 1, 1, 1, 1, 1, 1, 1},    {-4, -1, -1, -1, -1, 2, 2, 2, 2},
                                               {4, -2, -2, -2, -2, 1, 1, 1, 1}, {0, 1, 0, -1, 0, 1, -1, -1, 1},
                                               {0, -2, 0, 2, 0, 1, -1, -1, 1},  {0, 0, 1, 0, -1, 1, 1, -1, -1},
                                               {0, 0, -2, 0, 2, 1, 1, -1, -1},  {0, 1, -1, 1, -1, 0, 0, 0, 0},
                                               {0, 0, 0, 0, 0, 1, -1, 1, -1}};
                const R sr[9] = {R(0), R{1.63}, R{1.14}, R(0), R{1.92}, R(0), R{1.92}, omega, omega};
                R mu[9], nm[9], ne[9];
                for (int r = 0; r < 9; ++r) {
                    R acc_r = 0, norm = 0;
                    for (int i = 0; i < 9; ++i) {
                        acc_r += R(rows[r][i]) * lf[i];
                        norm += R(rows[r][i] * rows[r][i]);
                    }
                    mu[r] = acc_r / norm;
                    nm[r] = (R(1) - sr[r]) * mu[r];
                    ne[r] = sr[r] * mu[r];
                }
                const R ux = m.u[0], uy = m.u[1];
                const R grho = ne[1] * (R(-2) - R(3) * uu) + ne[2] * (R(1) + R(3) * uu) - ne[7] * (ux * ux - uy * uy) -
                               ne[8] * ux * uy;
                const R gjx = R(6) * ux * (ne[1] - ne[2]) - ne[4] + R(2) * ux * ne[7] + uy * ne[8];
                const R gjy = R(6) * uy * (ne[1] - ne[2]) - ne[6] - R(2) * uy * ne[7] + ux * ne[8];
                for (int i = 0; i < 9; ++i) {
                    R t = 0;
                    for (int r = 0; r < 9; ++r) t += R(rows[r][i]) * nm[r];
                    const R cx = static_cast<R>(F::c[2 * i]), cy = static_cast<R>(F::c[2 * i + 1]);
                    nuf0[i * N + cell] = t + grho + cx * gjx + cy * gjy;
                }
            } else if constexpr (CT == C_BGK) {
                R al, bd[3];
                moments(lf, al, bd);
                for (int i = 0; i < QF; ++i) nuf0[i * N + cell] = (R(1) - omega) * lf[i] + omega * (al + dev(i, bd));
            } else if constexpr (CT == C_RLBM) {
                // collide_rlbm (lattice_physics.hpp:422-605): f' = feq(rho, u) +
                // (1 - omega) w (9/2) [second(Pi) + third(u, Pi)] with
                // Pi_ab = sum c_a c_b (f - feq) = M2_ab - rho cs2 delta_ab - j_a j_b / rho
                // (fourth-order isotropy of D2Q9 / D3Q19 / D3Q27) and the
                // third-order coefficients a_k bilinear in u and Pi, so f'
                // depends on f only through rho, j and M2.  With psi = (1 - omega)
                // w (9/2) lam, g_P = d(psi . [second + third]) / d Pi_P over the
                // independent components P (xx, yy, zz, xy, xz, yz) and g_u the
                // same with respect to u:
                //   J^T lam = (d feq/d f)^T lam + sum_P g_P [(c - u)_a (c - u)_b
                //             - cs2 delta_ab] + g_u . (c - u) / rho.
                const R c3 = R(1) / R(3);
                const R u3[3] = {m.u[0], m.u[1], D == 3 ? m.u[2] : R(0)};
                auto cc = [&](int i, int d) { return d < D ? static_cast<R>(F::c[i * D + d]) : R(0); };
                // P index of the pair (a, b): 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz
                auto pidx = [](int a, int b) {
                    if (a > b) std::swap(a, b);
                    return a == b ? a : (a == 0 ? (b == 1 ? 3 : 4) : 5);
                };
                R pi[6] = {0, 0, 0, 0, 0, 0};
                for (int q = 0; q < QF; ++q) {
                    R cu = 0;
                    for (int d = 0; d < D; ++d) cu += cc(q, d) * m.u[d];
                    const R eq = static_cast<R>(F::w[q]) * m.rho * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu);
                    const R fn = fl[q] - eq;
                    for (int a = 0; a < D; ++a)
                        for (int b2 = a; b2 < D; ++b2) pi[pidx(a, b2)] += cc(q, a) * cc(q, b2) * fn;
                }
                R g[6] = {0, 0, 0, 0, 0, 0}, gu[3] = {0, 0, 0};
                // H_k = sum psi T_k with the forward's third-order combination
                R Hk[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz
                for (int i = 0; i < QF; ++i) {
                    const R psi = (R(1) - omega) * static_cast<R>(F::w[i]) * R(4.5) * lf[i];
                    const R cx = cc(i, 0), cy = cc(i, 1), cz = cc(i, 2);
                    g[0] += psi * (cx * cx - c3);
                    g[1] += psi * (cy * cy - c3);
                    if (D == 3) g[2] += psi * (cz * cz - c3);
                    g[3] += R(2) * psi * cx * cy;
                    g[4] += R(2) * psi * cx * cz;
                    g[5] += R(2) * psi * cy * cz;
                    const R hxxx = cx * cx * cx - R(3) * c3 * cx, hyyy = cy * cy * cy - R(3) * c3 * cy;
                    const R hzzz = cz * cz * cz - R(3) * c3 * cz;
                    const R hxxy = cx * cx * cy - c3 * cy, hxxz = cx * cx * cz - c3 * cz;
                    const R hxyy = cx * cy * cy - c3 * cx, hyyz = cy * cy * cz - c3 * cz;
                    const R hxzz = cx * cz * cz - c3 * cx, hyzz = cy * cz * cz - c3 * cy, hxyz = cx * cy * cz;
                    if constexpr (QF == 9) {
                        Hk[0] += psi * hxxx;
                        Hk[1] += psi * hyyy;
                        Hk[3] += psi * R(3) * hxxy;
                        Hk[5] += psi * R(3) * hxyy;
                    } else if constexpr (QF == 19) {  // the paired projection of the forward
                        Hk[3] += psi * (R(4) * hxxy + R(2) * hyzz);
                        Hk[8] += psi * (R(2) * hxxy + R(4) * hyzz);
                        Hk[7] += psi * (R(4) * hxzz + R(2) * hxyy);
                        Hk[5] += psi * (R(2) * hxzz + R(4) * hxyy);
                        Hk[6] += psi * (R(4) * hyyz + R(2) * hxxz);
                        Hk[4] += psi * (R(2) * hyyz + R(4) * hxxz);
                    } else {
                        Hk[0] += psi * hxxx;
                        Hk[1] += psi * hyyy;
                        Hk[2] += psi * hzzz;
                        Hk[3] += psi * R(3) * hxxy;
                        Hk[4] += psi * R(3) * hxxz;
                        Hk[5] += psi * R(3) * hxyy;
                        Hk[6] += psi * R(3) * hyyz;
                        Hk[7] += psi * R(3) * hxzz;
                        Hk[8] += psi * R(3) * hyzz;
                        Hk[9] += psi * R(6) * hxyz;
                    }
                }
                // a_k holds coef u_c Pi_ab: d/dPi_ab += H coef u_c, d/du_c += H coef Pi_ab
                auto bil = [&](R H, int uc, int a, int b2, R coef) {
                    g[pidx(a, b2)] += H * coef * u3[uc];
                    gu[uc] += H * coef * pi[pidx(a, b2)];
                };
                if (QF != 19) {
                    bil(Hk[0], 0, 0, 0, R(3));  // a_xxx = 3 ux pxx
                    bil(Hk[1], 1, 1, 1, R(3));  // a_yyy = 3 uy pyy
                }
                if (QF == 27) bil(Hk[2], 2, 2, 2, R(3));
                bil(Hk[3], 0, 0, 1, R(2));  // a_xxy = 2 ux pxy + uy pxx
                bil(Hk[3], 1, 0, 0, R(1));
                bil(Hk[5], 1, 0, 1, R(2));  // a_xyy = 2 uy pxy + ux pyy
                bil(Hk[5], 0, 1, 1, R(1));
                if (D == 3) {
                    bil(Hk[4], 0, 0, 2, R(2));  // a_xxz = 2 ux pxz + uz pxx
                    bil(Hk[4], 2, 0, 0, R(1));
                    bil(Hk[6], 1, 1, 2, R(2));  // a_yyz = 2 uy pyz + uz pyy
                    bil(Hk[6], 2, 1, 1, R(1));
                    bil(Hk[7], 2, 0, 2, R(2));  // a_xzz = 2 uz pxz + ux pzz
                    bil(Hk[7], 0, 2, 2, R(1));
                    bil(Hk[8], 2, 1, 2, R(2));  // a_yzz = 2 uz pyz + uy pzz
                    bil(Hk[8], 1, 2, 2, R(1));
                }
                if (QF == 27) {  // a_xyz = ux pyz + uy pxz + uz pxy
                    bil(Hk[9], 0, 1, 2, R(1));
                    bil(Hk[9], 1, 0, 2, R(1));
                    bil(Hk[9], 2, 0, 1, R(1));
                }
                R al, bd[3];
                moments(lf, al, bd);
                for (int i = 0; i < QF; ++i) {
                    R dc[3];
                    for (int d = 0; d < 3; ++d) dc[d] = cc(i, d) - u3[d];
                    R t = al + dev(i, bd);
                    for (int a = 0; a < D; ++a)
                        for (int b2 = a; b2 < D; ++b2) t += g[pidx(a, b2)] * (dc[a] * dc[b2] - (a == b2 ? c3 : R(0)));
                    R tu = 0;
                    for (int d = 0; d < D; ++d) tu += gu[d] * dc[d];
                    nuf0[i * N + cell] = t + tu / m.rho;
                }
            } else {
                R lp[QF], lm[QF];
                for (int i = 0; i < QF; ++i) {
                    lp[i] = R(0.5) * (lf[i] + lf[F::opp[i]]);
                    lm[i] = R(0.5) * (lf[i] - lf[F::opp[i]]);
                }
                R ap, bp[3], am, bm[3];
                moments(lp, ap, bp);
                moments(lm, am, bm);
                for (int i = 0; i < QF; ++i)
                    nuf0[i * N + cell] = lf[i] - omega * lp[i] - wm * lm[i] + omega * (ap + dev(i, bp)) +
                                         wm * (am + dev(i, bm));
            }
        }
    }
    *djdw += acc;
}

template <class S, class R> R mean_conc(const R* g, i64 N) {
    long double s = 0;
    for (i64 j = 0; j < S::Q * N; ++j) s += g[j];
    return static_cast<R>(s / N);
}

// J = <C>(T) = <A>(0) - <A>(T).  Runs the forward (all states kept), then the
// adjoint; returns J, dJ/d(f0, a0, b0) (as adjoints of the stored initial
// populations) and dJ/d(omega_c).
template <class F, class S, class R, int CT>
void reactive_gradient(const R* f0, const R* a0, const R* b0, Grid g, R omega, R omega_c, R k, int steps, R* J,
                       R* gf, R* ga, R* gb, R* djdwc) {
    constexpr int D = F::D;
    const i64 N = g.cells();
    const size_t nf = F::Q * N, ns = S::Q * N;
    std::vector<R> Fs((steps + 1) * nf), A((steps + 1) * ns), B((steps + 1) * ns);
    std::copy(f0, f0 + nf, Fs.begin());
    std::copy(a0, a0 + ns, A.begin());
    std::copy(b0, b0 + ns, B.begin());
    constexpr int DT = std::is_same<R, double>::value ? T_F64 : T_F32;
    for (int n = 0; n < steps; ++n) {
        fluid_step<F, CT, DT>(&Fs[n * nf], &Fs[(n + 1) * nf], g, omega, 1);
        scalar_like_step<F, S, DT>(&A[n * ns], &A[(n + 1) * ns], &B[n * ns], &B[(n + 1) * ns], &Fs[(n + 1) * nf], g,
                                   omega_c, k, true, 1);
    }
    *J = mean_conc<S, R>(&A[0], N) - mean_conc<S, R>(&A[steps * ns], N);
    std::vector<R> nf1(nf, R(0)), na1(ns, R(-1) / static_cast<R>(N)), nb1(ns, R(0));
    std::vector<R> nf0(nf), na0(ns), nb0(ns);
    long double dj = 0;
    for (int n = steps - 1; n >= 0; --n) {
        reactive_adjoint_step<F, S, R, CT>(&Fs[n * nf], &A[n * ns], &B[n * ns], nf1.data(), na1.data(), nb1.data(),
                                           nf0.data(), na0.data(), nb0.data(), g, omega, omega_c, k, &dj);
        nf1.swap(nf0);
        na1.swap(na0);
        nb1.swap(nb0);
    }
    for (i64 z = 0; z < g.nz; ++z)
        for (i64 y = 0; y < g.ny; ++y)
            for (i64 x = 0; x < g.nx; ++x) {
                const i64 cell = (z * g.ny + y) * g.nx + x;
                auto rev = [&](const int* c) {
                    const i64 sz = D == 3 ? pull(z, -c[2], g.nz) : 0;
                    return (sz * g.ny + pull(y, -c[1], g.ny)) * g.nx + pull(x, -c[0], g.nx);
                };
                for (int i = 0; i < F::Q; ++i) gf[i * N + cell] = nf1[i * N + rev(&F::c[i * D])];
                for (int j = 0; j < S::Q; ++j) {
                    const i64 r = j * N + rev(&S::c[j * D]);
                    ga[j * N + cell] = na1[r] + R(1) / static_cast<R>(N);
                    gb[j * N + cell] = nb1[r];
                }
            }
    *djdwc = static_cast<R>(dj);
}

// Entry helpers of the adjoint: the collisions the adjoint covers (BGK, TRT,
// MRT, RLBM) for a fluid/species pair, dispatched at run time.
template <class F, class S>
int gradient_entry(int ct, int dt, const void* f0, const void* a0, const void* b0, Grid g, double omega,
                   double omega_c, double k, int steps, double* J, void* gf, void* ga, void* gb, double* djdwc) {
    auto run = [&](auto tag, auto ctc) -> int {
        using R = decltype(tag);
        constexpr int CT = decltype(ctc)::value;
        R j, w;
        reactive_gradient<F, S, R, CT>(static_cast<const R*>(f0), static_cast<const R*>(a0),
                                       static_cast<const R*>(b0), g, static_cast<R>(omega), static_cast<R>(omega_c),
                                       static_cast<R>(k), steps, &j, static_cast<R*>(gf), static_cast<R*>(ga),
                                       static_cast<R*>(gb), &w);
        *J = j;
        *djdwc = w;
        return 0;
    };
    auto by_ct = [&](auto tag) -> int {
        if (ct == C_BGK) return run(tag, std::integral_constant<int, C_BGK>{});
        if (ct == C_TRT) return run(tag, std::integral_constant<int, C_TRT>{});
        if (ct == C_MRT) return run(tag, std::integral_constant<int, C_MRT>{});
        if (ct == C_RLBM) return run(tag, std::integral_constant<int, C_RLBM>{});
        return -1;
    };
    if (dt == T_F64) return by_ct(double{});
    if (dt == T_F32) return by_ct(float{});
    return -1;
}

template <class F, class S>
int adjoint_step_entry(int ct, int dt, const void* f, const void* a, const void* b, const void* nuf1,
                       const void* nua1, const void* nub1, void* nuf0, void* nua0, void* nub0, Grid g, double omega,
                       double omega_c, double k, double* djdw) {
    long double acc = 0;
    auto run = [&](auto tag, auto ctc) -> int {
        using R = decltype(tag);
        constexpr int CT = decltype(ctc)::value;
        reactive_adjoint_step<F, S, R, CT>(
            static_cast<const R*>(f), static_cast<const R*>(a), static_cast<const R*>(b),
            static_cast<const R*>(nuf1), static_cast<const R*>(nua1), static_cast<const R*>(nub1),
            static_cast<R*>(nuf0), static_cast<R*>(nua0), static_cast<R*>(nub0), g, static_cast<R>(omega),
            static_cast<R>(omega_c), static_cast<R>(k), &acc);
        return 0;
    };
    auto by_ct = [&](auto tag) -> int {
        if (ct == C_BGK) return run(tag, std::integral_constant<int, C_BGK>{});
        if (ct == C_TRT) return run(tag, std::integral_constant<int, C_TRT>{});
        if (ct == C_MRT) return run(tag, std::integral_constant<int, C_MRT>{});
        if (ct == C_RLBM) return run(tag, std::integral_constant<int, C_RLBM>{});
        return -1;
    };
    int rc;
    if (dt == T_F64) rc = by_ct(double{});
    else if (dt == T_F80) rc = by_ct((long double){});
    else return -1;
    *djdw = static_cast<double>(acc);
    return rc;
}

}  // namespace orc

// ============================================================================
// C entry points (ctypes).  Return 0 on success, -1 for an unsupported
// combination.
// ============================================================================
using namespace orc;

#define ORC_DT_SWITCH(DTV, ...)                      \
    switch (DTV) {                                    \
        case T_F32: { constexpr int DT = T_F32; __VA_ARGS__; break; }   \
        case T_F64: { constexpr int DT = T_F64; __VA_ARGS__; break; }   \
        case T_F16: { constexpr int DT = T_F16; __VA_ARGS__; break; }   \
        case T_BF16: { constexpr int DT = T_BF16; __VA_ARGS__; break; } \
        case T_F80: { constexpr int DT = T_F80; __VA_ARGS__; break; }   \
        default: return -1;                           \
    }

static size_t elem_size(int dt) { return dt == T_F80 ? sizeof(long double) : dt == T_F64 ? 8 : (dt == T_F32 ? 4 : 2); }

#define ORC_CT_SWITCH(CTV, ...)                      \
    switch (CTV) {                                    \
        case C_BGK: { constexpr int CT = C_BGK; __VA_ARGS__; break; }   \
        case C_TRT: { constexpr int CT = C_TRT; __VA_ARGS__; break; }   \
        case C_MRT: { constexpr int CT = C_MRT; __VA_ARGS__; break; }   \
        case C_RLBM: { constexpr int CT = C_RLBM; __VA_ARGS__; break; } \
        default: return -1;                           \
    }

#define ORC_FLUID_SWITCH(LV, ...)                    \
    switch (LV) {                                     \
        case L_D2Q9: { using L = D2Q9; __VA_ARGS__; break; }  \
        case L_D3Q19: { using L = D3Q19; __VA_ARGS__; break; } \
        case L_D3Q27: { using L = D3Q27; __VA_ARGS__; break; } \
        default: return -1;                           \
    }

static Grid make_grid(int lat, long long nx, long long ny, long long nz) {
    Grid g{nx, ny, nz};
    if (lat == L_D2Q9 || lat == L_D2Q5) g.nz = 1;
    return g;
}

extern "C" {

int orc_fluid_step(int lat, int ct, int dt, const void* cur, void* nxt, long long nx, long long ny, long long nz,
                   double omega, int threads) {
    const Grid g = make_grid(lat, nx, ny, nz);
    ORC_FLUID_SWITCH(lat, ORC_CT_SWITCH(ct, ORC_DT_SWITCH(dt, (fluid_step<L, CT, DT>(cur, nxt, g, omega, threads)))));
    return 0;
}

// Runs `steps` fluid steps ping-ponging between buf and tmp; the final state is
// copied back into buf.
int orc_fluid_run(int lat, int ct, int dt, void* buf, void* tmp, long long nx, long long ny, long long nz,
                  double omega, int steps, int threads) {
    const Grid g = make_grid(lat, nx, ny, nz);
    int q = lat == L_D2Q9 ? 9 : (lat == L_D3Q19 ? 19 : 27);
    const size_t es = elem_size(dt);
    void* a = buf;
    void* b = tmp;
    for (int s = 0; s < steps; ++s) {
        const int rc = orc_fluid_step(lat, ct, dt, a, b, nx, ny, nz, omega, threads);
        if (rc) return rc;
        std::swap(a, b);
    }
    if (a != buf) std::memcpy(buf, a, es * q * g.cells());
    return 0;
}

static int scalar_dispatch(int flat, int slat, int dt, const void* ca, void* na, const void* cb, void* nb,
                           const void* cf, long long nx, long long ny, long long nz, double omega_c, double param,
                           bool reaction, int threads, double ra, double rb) {
    const Grid g = make_grid(flat, nx, ny, nz);
    if (flat == L_D2Q9 && slat == L_D2Q5) {
        ORC_DT_SWITCH(dt, (scalar_like_step<D2Q9, D2Q5, DT>(ca, na, cb, nb, cf, g, omega_c, param, reaction, threads,
                                                            ra, rb)));
    } else if (flat == L_D3Q19 && slat == L_D3Q7) {
        ORC_DT_SWITCH(dt, (scalar_like_step<D3Q19, D3Q7, DT>(ca, na, cb, nb, cf, g, omega_c, param, reaction, threads,
                                                             ra, rb)));
    } else if (flat == L_D3Q27 && slat == L_D3Q7) {
        ORC_DT_SWITCH(dt, (scalar_like_step<D3Q27, D3Q7, DT>(ca, na, cb, nb, cf, g, omega_c, param, reaction, threads,
                                                             ra, rb)));
    } else {
        return -1;
    }
    return 0;
}

int orc_scalar_step(int flat, int slat, int dt, const void* cur_s, void* nxt_s, const void* cur_f, long long nx,
                    long long ny, long long nz, double omega_c, double source, int threads, double ref) {
    return scalar_dispatch(flat, slat, dt, cur_s, nxt_s, nullptr, nullptr, cur_f, nx, ny, nz, omega_c, source, false,
                           threads, ref, 0.0);
}

int orc_reaction_step(int flat, int slat, int dt, const void* cur_a, void* nxt_a, const void* cur_b, void* nxt_b,
                      const void* cur_f, long long nx, long long ny, long long nz, double omega_c, double k_react,
                      int threads, double ref_a, double ref_b) {
    return scalar_dispatch(flat, slat, dt, cur_a, nxt_a, cur_b, nxt_b, cur_f, nx, ny, nz, omega_c, k_react, true,
                           threads, ref_a, ref_b);
}

// The reactive loop of src/simulate_reaction.cpp:548-554: a fluid step, then
// the reaction step reading the freshly written fluid state.  Final states are
// copied back into f, a and b.
int orc_reactive_run(int flat, int ct, int slat, int dt, void* f, void* ft, void* a, void* at, void* b, void* bt,
                     long long nx, long long ny, long long nz, double omega, double omega_c, double k_react,
                     int steps, int threads, double ref_a, double ref_b) {
    const Grid g = make_grid(flat, nx, ny, nz);
    const int qf = flat == L_D2Q9 ? 9 : (flat == L_D3Q19 ? 19 : 27);
    const int qs = slat == L_D2Q5 ? 5 : 7;
    const size_t es = elem_size(dt);
    void *f0 = f, *f1 = ft, *a0 = a, *a1 = at, *b0 = b, *b1 = bt;
    for (int s = 0; s < steps; ++s) {
        int rc = orc_fluid_step(flat, ct, dt, f0, f1, nx, ny, nz, omega, threads);
        if (rc) return rc;
        std::swap(f0, f1);
        rc = orc_reaction_step(flat, slat, dt, a0, a1, b0, b1, f0, nx, ny, nz, omega_c, k_react, threads, ref_a, ref_b);
        if (rc) return rc;
        std::swap(a0, a1);
        std::swap(b0, b1);
    }
    if (f0 != f) std::memcpy(f, f0, es * qf * g.cells());
    if (a0 != a) std::memcpy(a, a0, es * qs * g.cells());
    if (b0 != b) std::memcpy(b, b0, es * qs * g.cells());
    return 0;
}

// Single-cell physics probes (dt: 0 = float, 1 = double; arrays in that type).
int orc_collide_cell(int lat, int ct, int dt, void* f, double omega) {
    if (dt == T_F32) {
        ORC_FLUID_SWITCH(lat, ORC_CT_SWITCH(ct, (collide_cell<L, CT, float>(static_cast<float*>(f), (float)omega))));
    } else if (dt == T_F64) {
        ORC_FLUID_SWITCH(lat, ORC_CT_SWITCH(ct, (collide_cell<L, CT, double>(static_cast<double*>(f), omega))));
    } else {
        return -1;
    }
    return 0;
}

int orc_macro_feq(int lat, const double* f, double* rho_u, double* feq_out) {
    ORC_FLUID_SWITCH(lat, {
        const Macro<L, double> m = macro_state<L, double>(f);
        rho_u[0] = m.rho;
        for (int d = 0; d < L::D; ++d) rho_u[1 + d] = m.u[d];
        for (int i = 0; i < L::Q; ++i) feq_out[i] = feq<L, double>(i, m);
    });
    return 0;
}

int orc_mrt_forward(int lat, const double* f, double* m) {
    if (lat == L_D2Q9) moments_d2q9<double>(f, m);
    else if (lat == L_D3Q19) moments_d3q19<double>(f, m);
    else if (lat == L_D3Q27) moments_d3q27<double>(f, m);
    else return -1;
    return 0;
}

int orc_mrt_inverse(int lat, const double* m, double* f) {
    if (lat == L_D2Q9) populations_d2q9<double>(m, f);
    else if (lat == L_D3Q19) populations_d3q19<double>(m, f);
    else if (lat == L_D3Q27) populations_d3q27<double>(m, f);
    else return -1;
    return 0;
}

int orc_reaction_source(int dt, const double* a, const double* b, double k, double* out, long long n) {
    for (long long j = 0; j < n; ++j) {
        if (dt == T_F64) out[j] = reaction_source<double>(a[j], b[j], k);
        else if (dt == T_F32) out[j] = reaction_source<float>((float)a[j], (float)b[j], (float)k);
        else return -1;
    }
    return 0;
}

int orc_reactive_stats(int slat, int dt, const void* a, const void* b, long long nx, long long ny, long long nz,
                       double k_react, int wide, double* out, double ref_a, double ref_b) {
    const Grid g = make_grid(slat, nx, ny, nz);
    if (slat == L_D2Q5) {
        ORC_DT_SWITCH(dt, (reactive_stats<D2Q5, DT>(a, b, g, k_react, out, wide != 0, ref_a, ref_b)));
    } else if (slat == L_D3Q7) {
        ORC_DT_SWITCH(dt, (reactive_stats<D3Q7, DT>(a, b, g, k_react, out, wide != 0, ref_a, ref_b)));
    } else {
        return -1;
    }
    return 0;
}

int orc_fluid_diagnostics(int lat, int dt, const void* f, long long nx, long long ny, long long nz, int wide,
                          double* out) {
    const Grid g = make_grid(lat, nx, ny, nz);
    ORC_FLUID_SWITCH(lat, ORC_DT_SWITCH(dt, (fluid_diagnostics<L, DT>(f, g, out, wide != 0))));
    return 0;
}

int orc_macro_fields(int lat, int dt, const void* f, long long nx, long long ny, long long nz, double* rho,
                     double* u) {
    const Grid g = make_grid(lat, nx, ny, nz);
    ORC_FLUID_SWITCH(lat, ORC_DT_SWITCH(dt, (macro_fields<L, DT>(f, g, rho, u))));
    return 0;
}

int orc_scalar_fields(int slat, int dt, const void* p, long long nx, long long ny, long long nz, double* conc,
                      double ref) {
    const Grid g = make_grid(slat, nx, ny, nz);
    if (slat == L_D2Q5) {
        ORC_DT_SWITCH(dt, (scalar_fields<D2Q5, DT>(p, g, conc, ref)));
    } else if (slat == L_D3Q7) {
        ORC_DT_SWITCH(dt, (scalar_fields<D3Q7, DT>(p, g, conc, ref)));
    } else {
        return -1;
    }
    return 0;
}

int orc_max_speed(int lat, int dt, const void* f, long long nx, long long ny, long long nz, double* out) {
    const Grid g = make_grid(lat, nx, ny, nz);
    ORC_FLUID_SWITCH(lat, ORC_DT_SWITCH(dt, (*out = max_speed<L, DT>(f, g))));
    return 0;
}

int orc_feq_field(int lat, int dt, const double* rho, const double* u, void* out, long long ncells) {
    ORC_FLUID_SWITCH(lat, ORC_DT_SWITCH(dt, (feq_field<L, DT>(rho, u, out, ncells))));
    return 0;
}

int orc_scalar_feq_field(int slat, int dt, const double* conc, const double* u, void* out, long long ncells,
                         double ref) {
    if (slat == L_D2Q5) {
        ORC_DT_SWITCH(dt, (scalar_feq_field<D2Q5, DT>(conc, u, out, ncells, ref)));
    } else if (slat == L_D3Q7) {
        ORC_DT_SWITCH(dt, (scalar_feq_field<D3Q7, DT>(conc, u, out, ncells, ref)));
    } else {
        return -1;
    }
    return 0;
}

int orc_f16_roundtrip(const float* in, uint16_t* h, float* back, long long n, int bf) {
    for (long long j = 0; j < n; ++j) {
        h[j] = bf ? f32_to_bf16(in[j]) : f32_to_f16(in[j]);
        back[j] = bf ? bf16_to_f32(h[j]) : f16_to_f32(h[j]);
    }
    return 0;
}

// Reactive loop gradient (D2Q9 BGK/TRT/MRT + two D2Q5 species); dt 0 = float, 1 = double.
int orc_reactive_gradient(int ct, int dt, const void* f0, const void* a0, const void* b0, long long nx, long long ny,
                          double omega, double omega_c, double k, int steps, double* J, void* gf, void* ga, void* gb,
                          double* djdwc) {
    return gradient_entry<D2Q9, D2Q5>(ct, dt, f0, a0, b0, Grid{nx, ny, 1}, omega, omega_c, k, steps, J, gf, ga, gb,
                                      djdwc);
}

// The same for the 3D pairs (D3Q19 or D3Q27 fluid + two D3Q7 species; BGK, TRT, MRT).
int orc_reactive_gradient_3d(int fluid, int ct, int dt, const void* f0, const void* a0, const void* b0,
                             long long nx, long long ny, long long nz, double omega, double omega_c, double k,
                             int steps, double* J, void* gf, void* ga, void* gb, double* djdwc) {
    const Grid g{nx, ny, nz};
    if (fluid == L_D3Q19)
        return gradient_entry<D3Q19, D3Q7>(ct, dt, f0, a0, b0, g, omega, omega_c, k, steps, J, gf, ga, gb, djdwc);
    if (fluid == L_D3Q27)
        return gradient_entry<D3Q27, D3Q7>(ct, dt, f0, a0, b0, g, omega, omega_c, k, steps, J, gf, ga, gb, djdwc);
    return -1;
}

// One step of the discrete adjoint with arbitrary adjoint inputs (collision-cell
// convention) for the dot-product test; djdw receives <nu1, dStep/d omega_c>.
int orc_reactive_adjoint_step(int ct, int dt, const void* f, const void* a, const void* b, const void* nuf1,
                              const void* nua1, const void* nub1, void* nuf0, void* nua0, void* nub0, long long nx,
                              long long ny, double omega, double omega_c, double k, double* djdw) {
    return adjoint_step_entry<D2Q9, D2Q5>(ct, dt, f, a, b, nuf1, nua1, nub1, nuf0, nua0, nub0, Grid{nx, ny, 1},
                                          omega, omega_c, k, djdw);
}

int orc_reactive_adjoint_step_3d(int fluid, int ct, int dt, const void* f, const void* a, const void* b,
                                 const void* nuf1, const void* nua1, const void* nub1, void* nuf0, void* nua0,
                                 void* nub0, long long nx, long long ny, long long nz, double omega, double omega_c,
                                 double k, double* djdw) {
    const Grid g{nx, ny, nz};
    if (fluid == L_D3Q19)
        return adjoint_step_entry<D3Q19, D3Q7>(ct, dt, f, a, b, nuf1, nua1, nub1, nuf0, nua0, nub0, g, omega,
                                               omega_c, k, djdw);
    if (fluid == L_D3Q27)
        return adjoint_step_entry<D3Q27, D3Q7>(ct, dt, f, a, b, nuf1, nua1, nub1, nuf0, nua0, nub0, g, omega,
                                               omega_c, k, djdw);
    return -1;
}

int orc_phi_slope(int dt, const double* z, double* phi, double* dphi, long long n) {
    for (long long j = 0; j < n; ++j) {
        if (dt == T_F64) {
            double em1;
            phi_and_slope<double>(z[j], phi[j], dphi[j], em1);
        } else if (dt == T_F32) {
            float p, dp, em1;
            phi_and_slope<float>(static_cast<float>(z[j]), p, dp, em1);
            phi[j] = p;
            dphi[j] = dp;
        } else {
            return 1;
        }
    }
    return 0;
}

int orc_reaction_grad(int dt, const double* a, const double* b, double k, double* dsa, double* dsb, long long n) {
    if (dt == T_F64) {
        for (long long j = 0; j < n; ++j) reaction_grad<double>(a[j], b[j], k, dsa[j], dsb[j]);
        return 0;
    }
    if (dt != T_F32) return 1;
    for (long long j = 0; j < n; ++j) {
        float sa, sb;
        reaction_grad<float>(static_cast<float>(a[j]), static_cast<float>(b[j]), static_cast<float>(k), sa, sb);
        dsa[j] = sa;
        dsb[j] = sb;
    }
    return 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
