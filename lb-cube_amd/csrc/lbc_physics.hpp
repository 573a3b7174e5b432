// Per-cell lattice Boltzmann physics for the gfx950 kernels.
//
// Every formula follows LB-Cube's include/lattice_physics.hpp and
// include/lattice_mrt.hpp term for term, in the same evaluation order, and the
// kernels are compiled with -ffp-contract=off: with the same inputs a cell
// update therefore rounds exactly like the reference's CPU loop
// (include/lattice_core.hpp:146-474) wherever both sides evaluate the same IEEE
// operations.  The one freedom taken is to skip terms that multiply a lattice
// velocity component of 0 and to fold products by +-1 into additions and
// subtractions; both leave every value bit-identical (a running sum that starts
// at +0 never becomes -0, and the sign of a zero velocity projection does not
// reach any result).
#pragma once

#include "lbc_constdiv.hpp"
#include "lbc_lattice.hpp"

namespace lbc {

template <class R> struct Macro {
    R rho, ux, uy, uz;
};

// compute_macro_state, lattice_physics.hpp:227-249
template <class L, class R> __device__ __forceinline__ Macro<R> macro_state(const R* f) {
    R rho = R(0), mx = R(0), my = R(0), mz = R(0);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        rho += f[i];
        if (L::cx[i] > 0) mx += f[i];
        if (L::cx[i] < 0) mx -= f[i];
        if (L::cy[i] > 0) my += f[i];
        if (L::cy[i] < 0) my -= f[i];
        if (L::cz[i] > 0) mz += f[i];
        if (L::cz[i] < 0) mz -= f[i];
    }
    Macro<R> m;
    m.rho = rho;
    m.ux = mx / rho;
    m.uy = my / rho;
    m.uz = L::D == 3 ? mz / rho : R(0);
    return m;
}

// c_i . u, summed x then y then z (lattice_physics.hpp:282).
template <class L, class R> __device__ __forceinline__ R dir_dot(int i, R ux, R uy, R uz) {
    R s = R(0);
    bool started = false;
    if (L::cx[i] != 0) { s = L::cx[i] > 0 ? ux : -ux; started = true; }
    if (L::cy[i] != 0) { s = started ? (L::cy[i] > 0 ? s + uy : s - uy) : (L::cy[i] > 0 ? uy : -uy); started = true; }
    if (L::cz[i] != 0) { s = started ? (L::cz[i] > 0 ? s + uz : s - uz) : (L::cz[i] > 0 ? uz : -uz); }
    return s;
}

template <class L, class R> __device__ __forceinline__ R vel_sq(const Macro<R>& m) {
    return L::D == 3 ? m.ux * m.ux + m.uy * m.uy + m.uz * m.uz : m.ux * m.ux + m.uy * m.uy;
}

// compute_equilibrium, lattice_physics.hpp:267-290
template <class L, class R> __device__ __forceinline__ R feq(int i, const Macro<R>& m, R uu) {
    constexpr R inv_cs2 = R(1.0) / R(L::cs2);
    const R cu = dir_dot<L, R>(i, m.ux, m.uy, m.uz);
    return R(L::w[i]) * m.rho *
           (R(1.0) + inv_cs2 * cu + R(0.5) * inv_cs2 * inv_cs2 * cu * cu - R(0.5) * inv_cs2 * uu);
}

// A collision hands each post-collision value to a sink as soon as it is
// final.  The kernels store it there, so a wave's stores leave spread over
// its collision instead of in one burst at the end (measured: 1-2% more
// bandwidth); the values, and so the results, are the same.  The default sink
// discards.
struct NoSink {
    template <class R> __device__ __forceinline__ void operator()(int, R) const {}
};

// collide_bgk, lattice_physics.hpp:307-318
template <class L, class R, class Sink = NoSink>
__device__ __forceinline__ void collide_bgk(R* f, R omega, Sink sink = {}) {
    const Macro<R> m = macro_state<L, R>(f);
    const R uu = vel_sq<L, R>(m);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        f[i] += omega * (feq<L, R>(i, m, uu) - f[i]);
        sink(i, f[i]);
    }
}

// compute_omega_minus, lattice_physics.hpp:336-344, magic parameter 1/4.
template <class R> __host__ __device__ inline R trt_omega_minus(R omega_plus) {
    const R magic = R{1} / R{4};
    const R tau_plus = R{1} / omega_plus;
    const R tau_minus = magic / (tau_plus - R{1} / R{2}) + R{1} / R{2};
    return R{1} / tau_minus;
}

// collide_trt, lattice_physics.hpp:366-398
template <class L, class R, class Sink = NoSink>
__device__ __forceinline__ void collide_trt(R* f, R wp, R wm, Sink sink = {}) {
    const Macro<R> m = macro_state<L, R>(f);
    const R uu = vel_sq<L, R>(m);
    R eq[L::Q], post[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) eq[i] = feq<L, R>(i, m, uu);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const int j = L::opp[i];
        const R even = R(0.5) * (f[i] + f[j]);
        const R eq_even = R(0.5) * (eq[i] + eq[j]);
        const R odd = R(0.5) * (f[i] - f[j]);
        const R eq_odd = R(0.5) * (eq[i] - eq[j]);
        post[i] = f[i] - wp * (even - eq_even) - wm * (odd - eq_odd);
        sink(i, post[i]);
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i) f[i] = post[i];
}

// ---------------------------------------------------------------------------
// Regularized collision, lattice_physics.hpp:116-211 and 422-605.
// ---------------------------------------------------------------------------
template <class L, class R, class Sink = NoSink>
__device__ __forceinline__ void collide_rlbm(R* f, R omega, Sink sink = {}) {
    const Macro<R> m = macro_state<L, R>(f);
    const R uu = vel_sq<L, R>(m);
    R eq[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) eq[i] = feq<L, R>(i, m, uu);
    R pxx = R(0), pyy = R(0), pzz = R(0), pxy = R(0), pxz = R(0), pyz = R(0);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const R fn = f[i] - eq[i];
        if (L::cx[i] != 0) pxx += fn;
        if (L::cy[i] != 0) pyy += fn;
        if (L::cz[i] != 0) pzz += fn;
        if (L::cx[i] * L::cy[i] > 0) pxy += fn;
        if (L::cx[i] * L::cy[i] < 0) pxy -= fn;
        if (L::cx[i] * L::cz[i] > 0) pxz += fn;
        if (L::cx[i] * L::cz[i] < 0) pxz -= fn;
        if (L::cy[i] * L::cz[i] > 0) pyz += fn;
        if (L::cy[i] * L::cz[i] < 0) pyz -= fn;
    }
    const R cs2 = R(L::cs2);
    const R pref = R{9} / R{2};
    const R ux = m.ux, uy = m.uy, uz = m.uz;
    if constexpr (L::Q == 9) {
        const R a_xxx = R{3} * ux * pxx;
        const R a_yyy = R{3} * uy * pyy;
        const R a_xxy = R{2} * ux * pxy + uy * pxx;
        const R a_xyy = R{2} * uy * pxy + ux * pyy;
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            const R cx = R(L::cx[i]), cy = R(L::cy[i]);
            const R q_xx = cx * cx - cs2;
            const R q_yy = cy * cy - cs2;
            const R q_xy = cx * cy;
            const R second = q_xx * pxx + q_yy * pyy + R{2} * q_xy * pxy;
            const R h_xxx = cx * cx * cx - R{3} * cs2 * cx;
            const R h_yyy = cy * cy * cy - R{3} * cs2 * cy;
            const R h_xxy = cx * cx * cy - cs2 * cy;
            const R h_xyy = cx * cy * cy - cs2 * cx;
            const R third = a_xxx * h_xxx + a_yyy * h_yyy + R{3} * (a_xxy * h_xxy + a_xyy * h_xyy);
            const R rec = R(L::w[i]) * pref * (second + third);
            f[i] = eq[i] + (R{1} - omega) * rec;
            sink(i, f[i]);
        }
    } else if constexpr (L::Q == 19) {
        const R a_xxy = R{2} * ux * pxy + uy * pxx;
        const R a_xxz = R{2} * ux * pxz + uz * pxx;
        const R a_xyy = R{2} * uy * pxy + ux * pyy;
        const R a_yyz = R{2} * uy * pyz + uz * pyy;
        const R a_xzz = R{2} * uz * pxz + ux * pzz;
        const R a_yzz = R{2} * uz * pyz + uy * pzz;
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            const R cx = R(L::cx[i]), cy = R(L::cy[i]), cz = R(L::cz[i]);
            const R q_xx = cx * cx - cs2;
            const R q_yy = cy * cy - cs2;
            const R q_zz = cz * cz - cs2;
            const R q_xy = cx * cy;
            const R q_xz = cx * cz;
            const R q_yz = cy * cz;
            const R second = q_xx * pxx + q_yy * pyy + q_zz * pzz + R{2} * (q_xy * pxy + q_xz * pxz + q_yz * pyz);
            const R h_xxy = cx * cx * cy - cs2 * cy;
            const R h_xxz = cx * cx * cz - cs2 * cz;
            const R h_xyy = cx * cy * cy - cs2 * cx;
            const R h_yyz = cy * cy * cz - cs2 * cz;
            const R h_xzz = cx * cz * cz - cs2 * cx;
            const R h_yzz = cy * cz * cz - cs2 * cy;
            const R third = R{3} * (h_xxy + h_yzz) * (a_xxy + a_yzz) + (h_xxy - h_yzz) * (a_xxy - a_yzz) +
                            R{3} * (h_xzz + h_xyy) * (a_xzz + a_xyy) + (h_xzz - h_xyy) * (a_xzz - a_xyy) +
                            R{3} * (h_yyz + h_xxz) * (a_yyz + a_xxz) + (h_yyz - h_xxz) * (a_yyz - a_xxz);
            const R rec = R(L::w[i]) * pref * (second + third);
            f[i] = eq[i] + (R{1} - omega) * rec;
            sink(i, f[i]);
        }
    } else {
        const R a_xxx = R{3} * ux * pxx;
        const R a_yyy = R{3} * uy * pyy;
        const R a_zzz = R{3} * uz * pzz;
        const R a_xxy = R{2} * ux * pxy + uy * pxx;
        const R a_xxz = R{2} * ux * pxz + uz * pxx;
        const R a_xyy = ux * pyy + R{2} * uy * pxy;
        const R a_yyz = R{2} * uy * pyz + uz * pyy;
        const R a_xzz = ux * pzz + R{2} * uz * pxz;
        const R a_yzz = uy * pzz + R{2} * uz * pyz;
        const R a_xyz = ux * pyz + uy * pxz + uz * pxy;
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            const R cx = R(L::cx[i]), cy = R(L::cy[i]), cz = R(L::cz[i]);
            const R q_xx = cx * cx - cs2;
            const R q_yy = cy * cy - cs2;
            const R q_zz = cz * cz - cs2;
            const R q_xy = cx * cy;
            const R q_xz = cx * cz;
            const R q_yz = cy * cz;
            const R second = q_xx * pxx + q_yy * pyy + q_zz * pzz + R{2} * (q_xy * pxy + q_xz * pxz + q_yz * pyz);
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
                            R{3} * (a_xxy * h_xxy + a_xxz * h_xxz + a_xyy * h_xyy + a_yyz * h_yyz + a_xzz * h_xzz +
                                    a_yzz * h_yzz) +
                            R{6} * a_xyz * h_xyz;
            const R rec = R(L::w[i]) * pref * (second + third);
            f[i] = eq[i] + (R{1} - omega) * rec;
            sink(i, f[i]);
        }
    }
}

// ---------------------------------------------------------------------------
// MRT.  Forward transforms are the moment expressions of lattice_mrt.hpp
// (D2Q9 :439-455, D3Q19 :617-661).  The inverse transforms equal M^T D^-1 with
// D the squared row norms (lattice_mrt.hpp:470-534, :674-892); each population
// is summed in the reference's term order with coefficients M[r][i] / D[r]
// rounded once, which are the reference's reduced fractions.
// ---------------------------------------------------------------------------
template <int Q> struct Basis {
    int m[Q][Q];
    int norm[Q];
    int order[Q];
};

constexpr Basis<9> make_d2q9_basis() {
    Basis<9> b{};
    const int rows[9][9] = {{1, 1, 1, 1, 1, 1, 1, 1, 1},       {-4, -1, -1, -1, -1, 2, 2, 2, 2},
                            {4, -2, -2, -2, -2, 1, 1, 1, 1},    {0, 1, 0, -1, 0, 1, -1, -1, 1},
                            {0, -2, 0, 2, 0, 1, -1, -1, 1},     {0, 0, 1, 0, -1, 1, 1, -1, -1},
                            {0, 0, -2, 0, 2, 1, 1, -1, -1},     {0, 1, -1, 1, -1, 0, 0, 0, 0},
                            {0, 0, 0, 0, 0, 1, -1, 1, -1}};
    for (int r = 0; r < 9; ++r) {
        b.norm[r] = 0;
        b.order[r] = r;
        for (int i = 0; i < 9; ++i) {
            b.m[r][i] = rows[r][i];
            b.norm[r] += rows[r][i] * rows[r][i];
        }
    }
    return b;
}

constexpr Basis<19> make_d3q19_basis() {
    Basis<19> b{};
    for (int i = 0; i < 19; ++i) {
        const int cx = D3Q19::cx[i], cy = D3Q19::cy[i], cz = D3Q19::cz[i];
        const int c2 = cx * cx + cy * cy + cz * cz;
        const int p = 3 * cx * cx - c2;
        const int q = cy * cy - cz * cz;
        b.m[0][i] = 1;
        b.m[1][i] = 19 * c2 - 30;
        b.m[2][i] = (21 * c2 * c2 - 53 * c2 + 24) / 2;
        b.m[3][i] = cx;
        b.m[4][i] = (5 * c2 - 9) * cx;
        b.m[5][i] = cy;
        b.m[6][i] = (5 * c2 - 9) * cy;
        b.m[7][i] = cz;
        b.m[8][i] = (5 * c2 - 9) * cz;
        b.m[9][i] = p;
        b.m[10][i] = q;
        b.m[11][i] = cx * cy;
        b.m[12][i] = cy * cz;
        b.m[13][i] = cx * cz;
        b.m[14][i] = (3 * c2 - 5) * p;
        b.m[15][i] = (3 * c2 - 5) * q;
        b.m[16][i] = (cy * cy - cz * cz) * cx;
        b.m[17][i] = (cz * cz - cx * cx) * cy;
        b.m[18][i] = (cx * cx - cy * cy) * cz;
    }
    const int order[19] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 14, 10, 15, 11, 12, 13, 16, 17, 18};
    for (int r = 0; r < 19; ++r) {
        b.order[r] = order[r];
        b.norm[r] = 0;
        for (int i = 0; i < 19; ++i) b.norm[r] += b.m[r][i] * b.m[r][i];
    }
    return b;
}

constexpr Basis<9> kD2Q9Basis = make_d2q9_basis();
constexpr Basis<19> kD3Q19Basis = make_d3q19_basis();

template <class R, int Q, class Sink = NoSink>
__device__ __forceinline__ void mrt_inverse(const Basis<Q>& B, const R* m, R* f, Sink sink = {}) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        R acc = R(0);
        bool first = true;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const int r = B.order[k];
            if (B.m[r][i] != 0) {
                const R coef = R(B.m[r][i]) / R(B.norm[r]);
                acc = first ? coef * m[r] : acc + coef * m[r];
                first = false;
            }
        }
        f[i] = acc;
        sink(i, f[i]);
    }
}

// compute_moments_d2q9, lattice_mrt.hpp:439-455
template <class R> __device__ __forceinline__ void mrt_forward_d2q9(const R* f, R* m) {
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

// compute_moments_d3q19, lattice_mrt.hpp:617-661
template <class R> __device__ __forceinline__ void mrt_forward_d3q19(const R* f, R* m) {
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

// collide_mrt_d2q9, lattice_mrt.hpp:548-603; rates s_e 1.63, s_eps 1.14, s_q 1.92 (:38-65)
template <class R, class Sink = NoSink>
__device__ __forceinline__ void collide_mrt_d2q9(R* f, R s_nu, Sink sink = {}) {
    const Macro<R> mac = macro_state<D2Q9, R>(f);
    R m[9];
    mrt_forward_d2q9<R>(f, m);
    const R rho = mac.rho, ux = mac.ux, uy = mac.uy;
    const R jx = rho * ux, jy = rho * uy;
    const R usq = ux * ux + uy * uy;
    const R eq1 = -R{2} * rho + R{3} * rho * usq;
    const R eq2 = rho - R{3} * rho * usq;
    const R eq4 = -jx, eq6 = -jy;
    const R eq7 = rho * (ux * ux - uy * uy);
    const R eq8 = rho * ux * uy;
    m[1] -= R{1.63} * (m[1] - eq1);
    m[2] -= R{1.14} * (m[2] - eq2);
    m[4] -= R{1.92} * (m[4] - eq4);
    m[6] -= R{1.92} * (m[6] - eq6);
    m[7] -= s_nu * (m[7] - eq7);
    m[8] -= s_nu * (m[8] - eq8);
    mrt_inverse<R, 9>(kD2Q9Basis, m, f, sink);
}

// collide_mrt_d3q19, lattice_mrt.hpp:908-1009: only the five shear moments
// relax with s_nu; every other non-conserved moment relaxes with rate 1, and
// the ghost equilibria (moments 14-18) are zero.
template <class R, class Sink = NoSink>
__device__ __forceinline__ void collide_mrt_d3q19(R* f, R omega, Sink sink = {}) {
    const Macro<R> mac = macro_state<D3Q19, R>(f);
    R m[19];
    mrt_forward_d3q19<R>(f, m);
    const R rho = mac.rho, ux = mac.ux, uy = mac.uy, uz = mac.uz;
    const R jx = rho * ux, jy = rho * uy, jz = rho * uz;
    const R ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const R usq = ux2 + uy2 + uz2;
    R eq[19];
    eq[0] = rho;
    eq[1] = -R{11} * rho + R{19} * rho * usq;
    eq[2] = R{3} * rho - (R{11} / R{2}) * rho * usq;
    eq[3] = jx;
    eq[4] = -(R{2} / R{3}) * jx;
    eq[5] = jy;
    eq[6] = -(R{2} / R{3}) * jy;
    eq[7] = jz;
    eq[8] = -(R{2} / R{3}) * jz;
    eq[9] = R{3} * rho * (R{2} * ux2 - uy2 - uz2);
    eq[10] = rho * (uy2 - uz2);
    eq[11] = rho * ux * uy;
    eq[12] = rho * uy * uz;
    eq[13] = rho * ux * uz;
    eq[14] = eq[15] = eq[16] = eq[17] = eq[18] = R{0};
    // rate 0 for rho and j (moments 0, 3, 5, 7): m - 0 * (m - eq) == m.
#pragma unroll
    for (int r = 0; r < 19; ++r) {
        if (r == 0 || r == 3 || r == 5 || r == 7) continue;
        if (r >= 9 && r <= 13) m[r] -= omega * (m[r] - eq[r]);
        else m[r] -= R{1} * (m[r] - eq[r]);
    }
    mrt_inverse<R, 19>(kD3Q19Basis, m, f, sink);
}

// D3Q27 separable tensor-product MRT, lattice_mrt.hpp:206-426, 1023-1305.
__host__ __device__ constexpr int t27(int a, int b, int c) { return (a * 3 + b) * 3 + c; }
// Raw tensor moments carried by MRT rows 7..26, lattice_mrt.hpp:1092-1114.
struct Rest27 {
    int a[20], b[20], c[20];
};
constexpr Rest27 kRest27 = {{1, 0, 1, 2, 2, 1, 0, 1, 0, 1, 2, 2, 0, 2, 1, 1, 2, 2, 1, 2},
                            {1, 1, 0, 1, 0, 2, 2, 0, 1, 1, 2, 0, 2, 1, 2, 1, 2, 1, 2, 2},
                            {0, 1, 1, 0, 1, 0, 1, 2, 2, 1, 0, 2, 2, 1, 1, 2, 1, 2, 2, 2}};

// One separable pass of the forward transform over a line of three values at
// c = -1, 0, +1 (lattice_mrt.hpp:1038-1087): the sums P0 = a + b + c,
// P1 = -a + 0 b + c and P2 = a - 2 b + c in the reference's left-to-right
// order.  The zero term of P1 is left out, and a - 2 b is one fused
// multiply-add: 2 b is exact, so fma(-2, b, a) rounds like a + (-2 b).  The
// reference's sums start at +0; starting at the first term instead changes
// only the sign of a sum whose every term is -0 (the values are identical).
template <class R> __device__ __forceinline__ void forward_line(R a, R b, R c, R& p0, R& p1, R& p2) {
    p0 = (a + b) + c;
    p1 = c - a;
    if constexpr (sizeof(R) == 4) p2 = __builtin_fmaf(R(-2), b, a) + c;
    else p2 = __builtin_fma(R(-2), b, a) + c;
}

template <class R> __device__ __forceinline__ void mrt_forward_d3q27(const R* f, R* mo) {
    R p[27], xt[27], xyt[27], raw[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) p[t27(D3Q27::cx[i] + 1, D3Q27::cy[i] + 1, D3Q27::cz[i] + 1)] = f[i];
#pragma unroll
    for (int iy = 0; iy < 3; ++iy)
#pragma unroll
        for (int iz = 0; iz < 3; ++iz)
            forward_line<R>(p[t27(0, iy, iz)], p[t27(1, iy, iz)], p[t27(2, iy, iz)], xt[t27(0, iy, iz)],
                            xt[t27(1, iy, iz)], xt[t27(2, iy, iz)]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int iz = 0; iz < 3; ++iz)
            forward_line<R>(xt[t27(ax, 0, iz)], xt[t27(ax, 1, iz)], xt[t27(ax, 2, iz)], xyt[t27(ax, 0, iz)],
                            xyt[t27(ax, 1, iz)], xyt[t27(ax, 2, iz)]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int ay = 0; ay < 3; ++ay)
            forward_line<R>(xyt[t27(ax, ay, 0)], xyt[t27(ax, ay, 1)], xyt[t27(ax, ay, 2)], raw[t27(ax, ay, 0)],
                            raw[t27(ax, ay, 1)], raw[t27(ax, ay, 2)]);
    const R r200 = raw[t27(2, 0, 0)], r020 = raw[t27(0, 2, 0)], r002 = raw[t27(0, 0, 2)];
    mo[0] = raw[t27(0, 0, 0)];
    mo[1] = raw[t27(1, 0, 0)];
    mo[2] = raw[t27(0, 1, 0)];
    mo[3] = raw[t27(0, 0, 1)];
    mo[4] = r200 + r020 + r002;
    mo[5] = R{2} * r200 - r020 - r002;
    mo[6] = r020 - r002;
#pragma unroll
    for (int k = 0; k < 20; ++k) mo[7 + k] = raw[t27(kRest27.a[k], kRest27.b[k], kRest27.c[k])];
}

// One separable pass of the inverse over a line of three moments (orders 0,
// 1, 2) to the values at c = -1, 0, +1 (lattice_mrt.hpp:1165-1205): the
// reference's term raw * P(c) / norm, norms 3, 2, 6, summed over the orders
// left to right.  Each quotient is formed once and equals the reference's IEEE
// quotient bit for bit: v0 / 3 and v2 / 3 by the correctly rounded constant
// division of lbc_constdiv.hpp, v1 / 2 exactly, (-2 v2) / 6 = -(v2 / 3) (the
// same real number, so the same rounding), and v2 / 6 as (v2 / 3) / 2, which
// is exact while v2 / 6 is a normal number; below that (|v2| < 3 * 2^-125 in
// float, 3 * 2^-1021 in double) the IEEE division is taken.  The P1 term at
// c = 0 is a zero and is left out; the sums start at their first term (the
// sign of an all -0 sum aside, the same values as starting at +0).
template <class R> __device__ __forceinline__ void inverse_line(R v0, R v1, R v2, R& m1, R& z0, R& p1) {
    const R t0 = div_const<R, 3>(v0);
    const R h1 = v1 / R{2};
    const R t2 = div_const<R, 3>(v2);
    constexpr R tiny = sizeof(R) == 4 ? R(0x1p-123) : R(0x1p-1019);  // >= 3 * 2^-125 resp. 3 * 2^-1021
    R s2 = t2 / R{2};
    if (__builtin_fabs(v2) < tiny) [[unlikely]]
        s2 = v2 / R{6};
    m1 = (t0 - h1) + s2;
    z0 = t0 - t2;
    p1 = (t0 + h1) + s2;
}

template <class R, class Sink = NoSink>
__device__ __forceinline__ void mrt_inverse_d3q27(const R* mo, R* f, Sink sink = {}) {
    R raw[27], zi[27], yzi[27], pt[27];
    raw[t27(0, 0, 0)] = mo[0];
    raw[t27(1, 0, 0)] = mo[1];
    raw[t27(0, 1, 0)] = mo[2];
    raw[t27(0, 0, 1)] = mo[3];
    raw[t27(2, 0, 0)] = div_const<R, 3>(mo[4] + mo[5]);
    raw[t27(0, 2, 0)] = div_const<R, 6>(R{2} * mo[4] - mo[5] + R{3} * mo[6]);
    raw[t27(0, 0, 2)] = div_const<R, 6>(R{2} * mo[4] - mo[5] - R{3} * mo[6]);
#pragma unroll
    for (int k = 0; k < 20; ++k) raw[t27(kRest27.a[k], kRest27.b[k], kRest27.c[k])] = mo[7 + k];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int ay = 0; ay < 3; ++ay)
            inverse_line<R>(raw[t27(ax, ay, 0)], raw[t27(ax, ay, 1)], raw[t27(ax, ay, 2)], zi[t27(ax, ay, 0)],
                            zi[t27(ax, ay, 1)], zi[t27(ax, ay, 2)]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int iz = 0; iz < 3; ++iz)
            inverse_line<R>(zi[t27(ax, 0, iz)], zi[t27(ax, 1, iz)], zi[t27(ax, 2, iz)], yzi[t27(ax, 0, iz)],
                            yzi[t27(ax, 1, iz)], yzi[t27(ax, 2, iz)]);
#pragma unroll
    for (int iy = 0; iy < 3; ++iy)
#pragma unroll
        for (int iz = 0; iz < 3; ++iz)
            inverse_line<R>(yzi[t27(0, iy, iz)], yzi[t27(1, iy, iz)], yzi[t27(2, iy, iz)], pt[t27(0, iy, iz)],
                            pt[t27(1, iy, iz)], pt[t27(2, iy, iz)]);
#pragma unroll
    for (int i = 0; i < 27; ++i) {
        f[i] = pt[t27(D3Q27::cx[i] + 1, D3Q27::cy[i] + 1, D3Q27::cz[i] + 1)];
        sink(i, f[i]);
    }
}

// collide_mrt_d3q27, lattice_mrt.hpp:1242-1305; rates s_b 1.19, s_e 1.1,
// s_eps 1.2, s_q 1.2 (:161-204); equilibrium moments = M feq.
template <class R, class Sink = NoSink>
__device__ __forceinline__ void collide_mrt_d3q27(R* f, R s_nu, Sink sink = {}) {
    const Macro<R> mac = macro_state<D3Q27, R>(f);
    R m[27], fe[27], eq[27];
    mrt_forward_d3q27<R>(f, m);
    // compute_equilibrium_moments_d3q27 (lattice_mrt.hpp:1242-1268): the
    // second-order feq of each population, then the forward transform
    const R usq = vel_sq<D3Q27, R>(mac);
#pragma unroll
    for (int i = 0; i < 27; ++i) fe[i] = feq<D3Q27, R>(i, mac, usq);
    mrt_forward_d3q27<R>(fe, eq);
    const R s_b = R{1.19}, s_e = R{1.1}, s_eps = R{1.2}, s_q = R{1.2};
#pragma unroll
    for (int r = 4; r < 27; ++r) {
        const R s = r == 4 ? s_b : (r <= 9 ? s_nu : (r <= 15 ? s_e : (r == 16 ? s_eps : s_q)));
        m[r] -= s * (m[r] - eq[r]);
    }
    mrt_inverse_d3q27<R>(m, f, sink);
}

// detail::collide_cell, lattice_core.hpp:90-126
template <class L, int CT, class R, class Sink = NoSink>
__device__ __forceinline__ void collide_cell(R* f, R omega, R omega_minus, Sink sink = {}) {
    if constexpr (CT == kBGK) {
        collide_bgk<L, R>(f, omega, sink);
    } else if constexpr (CT == kTRT) {
        collide_trt<L, R>(f, omega, omega_minus, sink);
    } else if constexpr (CT == kMRT) {
        if constexpr (L::Q == 9) collide_mrt_d2q9<R>(f, omega, sink);
        else if constexpr (L::Q == 19) collide_mrt_d3q19<R>(f, omega, sink);
        else collide_mrt_d3q27<R>(f, omega, sink);
    } else {
        collide_rlbm<L, R>(f, omega, sink);
    }
}

// ---------------------------------------------------------------------------
// Scalar ADR, lattice_physics.hpp:618-694, and the A+B->C source :712-743.
// ---------------------------------------------------------------------------
template <class S, class R> __device__ __forceinline__ R concentration(const R* g) {
    R c{};
#pragma unroll
    for (int i = 0; i < S::Q; ++i) c += g[i];
    return c;
}

template <class S, class R> __device__ __forceinline__ R scalar_feq(int i, R conc, R ux, R uy, R uz) {
    const R cu = dir_dot<S, R>(i, ux, uy, uz);
    return R(S::w[i]) * conc * (R(1.0) + cu / R(S::cs2));
}

template <class S, class R>
__device__ __forceinline__ void collide_scalar(R* g, R ux, R uy, R uz, R omega_c, R source) {
    const R conc = concentration<S, R>(g);
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        const R eq = scalar_feq<S, R>(i, conc, ux, uy, uz);
        g[i] += omega_c * (eq - g[i]) + R(S::w[i]) * source;
    }
}

template <class R> struct Eps;
template <> struct Eps<float> { static constexpr float v = 1.1920928955078125e-07f; };
template <> struct Eps<double> { static constexpr double v = 2.220446049250313080847e-16; };

template <class R> __device__ __forceinline__ R reaction_source(R a, R b, R k) {
    const R abs_a = a < R{} ? -a : a;
    const R abs_b = b < R{} ? -b : b;
    const R diff = b - a;
    const R abs_diff = diff < R{} ? -diff : diff;
    const R tol = R(64.0) * Eps<R>::v * (R(1.0) + abs_a + abs_b);
    R after;
    if (abs_diff <= tol) {
        after = a / (R(1.0) + k * a);
    } else {
        const R att = exp(-k * diff);
        after = a * diff * att / (b - a * att);
    }
    return after - a;
}

}  // namespace lbc
