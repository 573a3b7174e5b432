// Discrete adjoint of the reactive loop: a D2Q9, D3Q19 or D3Q27 fluid (BGK,
// TRT, MRT or RLBM) plus two D2Q5 / D3Q7 species reacting A + B -> C
// (the loop of src/simulate_reaction.cpp:516-546, run by
// lbc_reactive_fused_step).  No counterpart in the reference.
//
// Forward step n: f_{n+1} = collide(pull f_n); (a, b)_{n+1} = reaction step of
// pulled (a, b)_n advected by u(f_{n+1}) at the same cell.  The adjoint keeps
// its state "at the collision cell": nu[i][x] is the derivative with respect
// to the population i that cell x gathered, so the adjoint of a stored
// population is Lambda[i][y] = nu[i][y + c_i], a pull along -c_i (reverse
// streaming).  One adjoint step per cell:
//   1. reverse-pull nu_f, nu_a, nu_b of step n+1 (the adjoints of the stored
//      f_{n+1}, a_{n+1}, b_{n+1} at this cell);
//   2. pull f_n and take its density and velocity, which the collision
//      conserves: u' of the forward reaction step up to rounding;
//   3. reaction step transpose: scalar BGK + source Jacobians, the reaction
//      Jacobian of the exact one-step map, and the velocity coupling into the
//      fluid adjoint through du'/df' = (c - u') / rho';
//   4. fluid collision transpose at the pre-collision state (BGK, TRT with
//      its even/odd split, MRT through the moment basis, or RLBM through the
//      non-equilibrium stress it reconstructs from), written to nu_f of step n.
// Long runs use checkpoints every `interval` steps and recompute each
// segment forward before walking it backward (lbc_reactive_gradient).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>
#include <utility>

#include "../../include/lbcube.h"
#include "lbc_fast.hpp"
#include "lbc_launch.hpp"
#include "lbc_physics.hpp"
#include "lbc_storage.hpp"

namespace lbc {

__device__ __forceinline__ float expm1_dev(float z) { return expm1f(z); }
__device__ __forceinline__ double expm1_dev(double z) { return expm1(z); }

// phi(z) = expm1(z) / z and phi'(z) = (z e^z - expm1(z)) / z^2 without the
// cancellation of that quotient near z = 0 (the numerator is ~z^2/2, formed
// from terms of size 1): below |z| = 1 (float) or 0.1 (double) both come from
// their Taylor series through z^10 (phi'(z) = sum_n n z^(n-1) / (n+1)!; the
// first omitted term is below one ulp at the switch), above it from expm1,
// where z + expm1(z) (z - 1) loses at most ~5 ulps (at z = -1).
template <class R> __device__ __forceinline__ void phi_and_slope(R z, R& phi, R& dphi, R& em1) {
    constexpr R limit = sizeof(R) == 4 ? R(1) : R(0.1);
    if (fabs(z) < limit) {
        phi = R(1) + z * (R(1) / R(2) + z * (R(1) / R(6) + z * (R(1) / R(24) + z * (R(1) / R(120) + z * (
                  R(1) / R(720) + z * (R(1) / R(5040) + z * (R(1) / R(40320) + z * (R(1) / R(362880) + z * (
                  R(1) / R(3628800) + z * (R(1) / R(39916800)))))))))));
        dphi = R(1) / R(2) + z * (R(1) / R(3) + z * (R(1) / R(8) + z * (R(1) / R(30) + z * (R(1) / R(144) + z * (
                   R(1) / R(840) + z * (R(1) / R(5760) + z * (R(1) / R(45360) + z * (R(1) / R(403200) + z * (
                   R(1) / R(3991680) + z * (R(1) / R(43545600)))))))))));
        em1 = z * phi;
    } else {
        em1 = expm1_dev(z);
        phi = em1 / z;
        dphi = (z + em1 * (z - R(1))) / (z * z);
    }
}

// dS/dA, dS/dB of S = A' - A for the exact one-step map of dA/dt = dB/dt =
// -k A B.  With D = B - A and z = k D: A' = A / (1 + s), s = k A phi(z) +
// expm1(z), phi(z) = expm1(z) / z, which has no cancellation at A = B.
// Differentiating, with e^z - phi = z phi' and z + k A = k B:
//   dS/dA = (k^2 A B phi' - k B phi - s^2) / (1 + s)^2
//   dS/dB = -k A (k A phi' + e^z) / (1 + s)^2,
// both free of the cancellation that dA'/dA - 1 (a quotient near 1 minus 1)
// would bring for small k.
template <class R> __device__ __forceinline__ void reaction_grad(R a, R b, R k, R& dsa, R& dsb) {
    const R d = b - a;
    const R z = k * d;
    R phi, dphi, em1;
    phi_and_slope(z, phi, dphi, em1);
    const R e = R(1) + em1;
    const R s = k * a * phi + em1;
    const R den = R(1) + s;
    const R inv2 = R(1) / (den * den);
    dsa = (k * k * a * b * dphi - k * b * phi - s * s) * inv2;
    dsb = -a * k * (k * a * dphi + e) * inv2;
}

// alpha = sum lam_q w_q P_q, beta = sum lam_q w_q grad_u P_q: the pieces of the
// transposed equilibrium Jacobian (d feq_q / d f_i = w_q [P_q + G_q . (c_i - u)],
// P_q = 1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u, G_q = 3 c + 9 (c.u) c - 3 u; cs2 =
// 1/3 for every fluid lattice).  sum_q w_q P_q = 1 exactly, but not with the
// weights rounded to binary (their FP32 sum is 1 + 7e-9 for D2Q9), which would
// make a constant adjoint grow by that much per step: the rest direction takes
// w_0 P_0 = 1 - sum_{q>0} w_q P_q.
template <class F, class R>
__device__ __forceinline__ void feq_transpose(const R* lam, R ux, R uy, R uz, R& al, R& bx, R& by, R& bz) {
    const R uu = F::D == 3 ? ux * ux + uy * uy + uz * uz : ux * ux + uy * uy;
    al = bx = by = bz = R(0);
    R wp_moving = R(0);
#pragma unroll
    for (int q = 1; q < F::Q; ++q) {
        const R w = R(F::w[q]);
        const R cx = R(F::cx[q]), cy = R(F::cy[q]), cz = R(F::cz[q]);
        const R cu = F::D == 3 ? cx * ux + cy * uy + cz * uz : cx * ux + cy * uy;
        const R wP = w * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu);
        wp_moving += wP;
        al += lam[q] * wP;
        bx += lam[q] * w * (R(3) * cx + R(9) * cu * cx - R(3) * ux);
        by += lam[q] * w * (R(3) * cy + R(9) * cu * cy - R(3) * uy);
        if constexpr (F::D == 3) bz += lam[q] * w * (R(3) * cz + R(9) * cu * cz - R(3) * uz);
    }
    al += lam[0] * (R(1) - wp_moving);
    bx += lam[0] * R(F::w[0]) * (-R(3) * ux);
    by += lam[0] * R(F::w[0]) * (-R(3) * uy);
    if constexpr (F::D == 3) bz += lam[0] * R(F::w[0]) * (-R(3) * uz);
}

// out = J^T lam of the regularized collision (collide_rlbm,
// lattice_physics.hpp:422-605) at the pre-collision populations f:
// f' = feq(rho, u) + (1 - omega) w (9/2) [second(Pi) + third(u, Pi)], where
// Pi = sum c c (f - feq) = M2 - rho cs2 I - j j / rho (the second moments of
// the second-order equilibrium are exact on D2Q9, D3Q19 and D3Q27) and the
// third-order coefficients a_k are bilinear in u and Pi.  With psi = (1 -
// omega) w (9/2) lam, g_P the derivative of psi . [second + third] with
// respect to the independent components of Pi and g_u that with respect to u:
//   J^T lam = (d feq / d f)^T lam + sum_P g_P [(c - u)_a (c - u)_b - cs2 delta_ab]
//             + g_u . (c - u) / rho,
// since d Pi_ab / d f_i = (c - u)_a (c - u)_b - cs2 delta_ab and d u / d f_i =
// (c_i - u) / rho.  H_k = sum psi T_k carries the forward's combination of the
// third-order Hermite polynomials h; D3Q19's paired projection 3 (h1 + h2)
// (a1 + a2) + (h1 - h2)(a1 - a2) is (4 h1 + 2 h2) a1 + (2 h1 + 4 h2) a2.
template <class F, class R>
__device__ __forceinline__ void rlbm_transpose(const R* f, const R* lam, const Macro<R>& m, R omega, R* out) {
    constexpr bool D3 = F::D == 3;
    const R cs2 = R(F::cs2);
    const R ux = m.ux, uy = m.uy, uz = m.uz;
    const R uu = vel_sq<F, R>(m);
    R pxx = R(0), pyy = R(0), pzz = R(0), pxy = R(0), pxz = R(0), pyz = R(0);
#pragma unroll
    for (int i = 0; i < F::Q; ++i) {
        const R fn = f[i] - feq<F, R>(i, m, uu);
        if (F::cx[i] != 0) pxx += fn;
        if (F::cy[i] != 0) pyy += fn;
        if (F::cz[i] != 0) pzz += fn;
        if (F::cx[i] * F::cy[i] > 0) pxy += fn;
        if (F::cx[i] * F::cy[i] < 0) pxy -= fn;
        if (F::cx[i] * F::cz[i] > 0) pxz += fn;
        if (F::cx[i] * F::cz[i] < 0) pxz -= fn;
        if (F::cy[i] * F::cz[i] > 0) pyz += fn;
        if (F::cy[i] * F::cz[i] < 0) pyz -= fn;
    }
    R gxx = R(0), gyy = R(0), gzz = R(0), gxy = R(0), gxz = R(0), gyz = R(0);
    R Hxxx = R(0), Hyyy = R(0), Hzzz = R(0), Hxxy = R(0), Hxxz = R(0), Hxyy = R(0), Hyyz = R(0), Hxzz = R(0),
      Hyzz = R(0), Hxyz = R(0);
    // coefficients below fold to constants once the loop is unrolled; the
    // zero ones are skipped rather than multiplied
    auto add = [](R& sum, R coef, R x) {
        if (coef != R(0)) sum += coef * x;
    };
    const R keep = (R(1) - omega) * R(4.5);
#pragma unroll
    for (int i = 0; i < F::Q; ++i) {
        const R psi = keep * R(F::w[i]) * lam[i];
        const R cx = R(F::cx[i]), cy = R(F::cy[i]), cz = R(F::cz[i]);
        add(gxx, cx * cx - cs2, psi);
        add(gyy, cy * cy - cs2, psi);
        if constexpr (D3) add(gzz, cz * cz - cs2, psi);
        add(gxy, R(2) * cx * cy, psi);
        if constexpr (D3) {
            add(gxz, R(2) * cx * cz, psi);
            add(gyz, R(2) * cy * cz, psi);
        }
        const R h_xxx = cx * cx * cx - R(3) * cs2 * cx, h_yyy = cy * cy * cy - R(3) * cs2 * cy;
        const R h_zzz = cz * cz * cz - R(3) * cs2 * cz;
        const R h_xxy = cx * cx * cy - cs2 * cy, h_xxz = cx * cx * cz - cs2 * cz;
        const R h_xyy = cx * cy * cy - cs2 * cx, h_yyz = cy * cy * cz - cs2 * cz;
        const R h_xzz = cx * cz * cz - cs2 * cx, h_yzz = cy * cz * cz - cs2 * cy, h_xyz = cx * cy * cz;
        if constexpr (F::Q == 9) {
            add(Hxxx, h_xxx, psi);
            add(Hyyy, h_yyy, psi);
            add(Hxxy, R(3) * h_xxy, psi);
            add(Hxyy, R(3) * h_xyy, psi);
        } else if constexpr (F::Q == 19) {
            add(Hxxy, R(4) * h_xxy + R(2) * h_yzz, psi);
            add(Hyzz, R(2) * h_xxy + R(4) * h_yzz, psi);
            add(Hxzz, R(4) * h_xzz + R(2) * h_xyy, psi);
            add(Hxyy, R(2) * h_xzz + R(4) * h_xyy, psi);
            add(Hyyz, R(4) * h_yyz + R(2) * h_xxz, psi);
            add(Hxxz, R(2) * h_yyz + R(4) * h_xxz, psi);
        } else {
            add(Hxxx, h_xxx, psi);
            add(Hyyy, h_yyy, psi);
            add(Hzzz, h_zzz, psi);
            add(Hxxy, R(3) * h_xxy, psi);
            add(Hxxz, R(3) * h_xxz, psi);
            add(Hxyy, R(3) * h_xyy, psi);
            add(Hyyz, R(3) * h_yyz, psi);
            add(Hxzz, R(3) * h_xzz, psi);
            add(Hyzz, R(3) * h_yzz, psi);
            add(Hxyz, R(6) * h_xyz, psi);
        }
    }
    // a_k = sum coef u_c Pi_ab: d/d Pi_ab gains H_k coef u_c, d/d u_c gains
    // H_k coef Pi_ab
    R gux = R(0), guy = R(0), guz = R(0);
    if constexpr (F::Q != 19) {
        gxx += R(3) * ux * Hxxx;  // a_xxx = 3 ux pxx
        gux += R(3) * pxx * Hxxx;
        gyy += R(3) * uy * Hyyy;  // a_yyy = 3 uy pyy
        guy += R(3) * pyy * Hyyy;
    }
    gxy += R(2) * ux * Hxxy;  // a_xxy = 2 ux pxy + uy pxx
    gux += R(2) * pxy * Hxxy;
    gxx += uy * Hxxy;
    guy += pxx * Hxxy;
    gxy += R(2) * uy * Hxyy;  // a_xyy = 2 uy pxy + ux pyy
    guy += R(2) * pxy * Hxyy;
    gyy += ux * Hxyy;
    gux += pyy * Hxyy;
    if constexpr (D3) {
        gxz += R(2) * ux * Hxxz;  // a_xxz = 2 ux pxz + uz pxx
        gux += R(2) * pxz * Hxxz;
        gxx += uz * Hxxz;
        guz += pxx * Hxxz;
        gyz += R(2) * uy * Hyyz;  // a_yyz = 2 uy pyz + uz pyy
        guy += R(2) * pyz * Hyyz;
        gyy += uz * Hyyz;
        guz += pyy * Hyyz;
        gxz += R(2) * uz * Hxzz;  // a_xzz = 2 uz pxz + ux pzz
        guz += R(2) * pxz * Hxzz;
        gzz += ux * Hxzz;
        gux += pzz * Hxzz;
        gyz += R(2) * uz * Hyzz;  // a_yzz = 2 uz pyz + uy pzz
        guz += R(2) * pyz * Hyzz;
        gzz += uy * Hyzz;
        guy += pzz * Hyzz;
    }
    if constexpr (F::Q == 27) {
        gzz += R(3) * uz * Hzzz;  // a_zzz = 3 uz pzz
        guz += R(3) * pzz * Hzzz;
        gyz += ux * Hxyz;  // a_xyz = ux pyz + uy pxz + uz pxy
        gux += pyz * Hxyz;
        gxz += uy * Hxyz;
        guy += pxz * Hxyz;
        gxy += uz * Hxyz;
        guz += pxy * Hxyz;
    }
    R al, bx, by, bz;
    feq_transpose<F, R>(lam, ux, uy, uz, al, bx, by, bz);
    const R irho = R(1) / m.rho;
    bx += gux * irho;
    by += guy * irho;
    if constexpr (D3) bz += guz * irho;
#pragma unroll
    for (int i = 0; i < F::Q; ++i) {
        const R dx = R(F::cx[i]) - ux, dy = R(F::cy[i]) - uy;
        R t = al + bx * dx + by * dy + gxx * (dx * dx - cs2) + gyy * (dy * dy - cs2) + gxy * (dx * dy);
        if constexpr (D3) {
            const R dz = R(F::cz[i]) - uz;
            t += bz * dz + gzz * (dz * dz - cs2) + gxz * (dx * dz) + gyz * (dy * dz);
        }
        out[i] = t;
    }
}

// w = M^-1 S M lam for the D3Q27 MRT (collide_mrt_d3q27: f' = f - M^-1 S M
// (f - feq)), by the separable transforms of its deviation form: three 1D
// passes of {1, c, 3c^2 - 2}, the rates divided by the row norms (the three
// diagonal second moments relaxed in the reference's combinations X + Y + Z,
// 2X - Y - Z, Y - Z, a symmetric 3 x 3 block), three inverse passes.  The
// map is symmetric (orthogonal rows, rates diagonal in the reference's
// basis), so it is also its own transpose.
template <class R> __device__ __forceinline__ void mrt27_operator(const R* lam, R* w, R s_nu) {
    using L = D3Q27;
    R p[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) p[g27(L::cx[i] + 1, L::cy[i] + 1, L::cz[i] + 1)] = lam[i];
    auto fwd = [](R& lo, R& mid, R& hi) {
        const R s = lo + hi;
        const R m0 = s + mid, m1 = hi - lo, m2 = s - R(2) * mid;
        lo = m0;
        mid = m1;
        hi = m2;
    };
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) fwd(p[g27(0, j, k)], p[g27(1, j, k)], p[g27(2, j, k)]);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) fwd(p[g27(j, 0, k)], p[g27(j, 1, k)], p[g27(j, 2, k)]);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) fwd(p[g27(j, k, 0)], p[g27(j, k, 1)], p[g27(j, k, 2)]);
    constexpr double n[3] = {3.0, 2.0, 6.0};
    const R s_b = R(1.19), s_e = R(1.1), s_eps = R(1.2), s_q = R(1.2);
    {
        const R X = p[g27(2, 0, 0)], Y = p[g27(0, 2, 0)], Z = p[g27(0, 0, 2)];
        const R dm4 = s_b * (X + Y + Z), dm5 = s_nu * (R(2) * X - Y - Z), dm6 = s_nu * (Y - Z);
        constexpr double k = 1.0 / 54.0;
        p[g27(2, 0, 0)] = (dm4 + dm5) * R(k / 3.0);
        p[g27(0, 2, 0)] = (R(3) * dm6 + R(2) * dm4 - dm5) * R(k / 6.0);
        p[g27(0, 0, 2)] = (R(-3) * dm6 + R(2) * dm4 - dm5) * R(k / 6.0);
    }
    p[g27(0, 0, 0)] = p[g27(1, 0, 0)] = p[g27(0, 1, 0)] = p[g27(0, 0, 1)] = R(0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int order = a + b + c;
                const int nonzero = (a > 0) + (b > 0) + (c > 0);
                if (order <= 1 || (order == 2 && nonzero == 1)) continue;
                R rate;
                if (order == 2) rate = s_nu;
                else if (order == 3 && nonzero == 2) rate = s_e;
                else if (order == 3) rate = s_eps;
                else rate = s_q;
                p[g27(a, b, c)] *= rate * R(1.0 / (n[a] * n[b] * n[c]));
            }
    auto inv = [](R& lo, R& mid, R& hi) {
        const R m0 = lo, m1 = mid, m2 = hi;
        const R t = m0 + m2;
        lo = t - m1;
        mid = m0 - R(2) * m2;
        hi = t + m1;
    };
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) inv(p[g27(j, k, 0)], p[g27(j, k, 1)], p[g27(j, k, 2)]);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) inv(p[g27(j, 0, k)], p[g27(j, 1, k)], p[g27(j, 2, k)]);
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) inv(p[g27(0, j, k)], p[g27(1, j, k)], p[g27(2, j, k)]);
#pragma unroll
    for (int i = 0; i < 27; ++i) w[i] = p[g27(L::cx[i] + 1, L::cy[i] + 1, L::cz[i] + 1)];
}

// F, S: the fluid and species lattices; RF: the forward states' type (the
// trajectory); R: the adjoint's state and arithmetic (R = double with RF =
// float: the FP64 adjoint of an FP32 run).
template <class F, class S, class RF, class R, int CT, int MAXT>
__global__ __launch_bounds__(MAXT) void reactive_adjoint_kernel(
    const RF* __restrict__ f, const RF* __restrict__ a, const RF* __restrict__ b, const R* __restrict__ nuf1,
    const R* __restrict__ nua1, const R* __restrict__ nub1, R* __restrict__ nuf0, R* __restrict__ nua0,
    R* __restrict__ nub0, Cells c, R omega, R omega_minus, R omega_c, R k, double* __restrict__ djdw) {
    static_assert(F::D == S::D, "a fluid and a species lattice of one dimension");

    constexpr int QF = F::Q, QS = S::Q;
    // XCD-aware block order as in the forward kernels (the reverse pulls share
    // block-edge lines the same way); each block keeps its own dJ/d omega_c
    // slot, indexed by the dispatch id, so a slot always sums the same cells
    const uint32_t block = renumber_blocks<F::D, R>() ? xcd_block() : blockIdx.x;
    const uint32_t cell = block * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (cell < c.count) {
        const CellCoord q = cell_coord(cell, c);
        const int nx = static_cast<int>(c.nx), ny = static_cast<int>(c.ny), nz = static_cast<int>(c.nz);
        const size_t plane = c.count;
        const Pull pull(cell, q.x, q.y, q.z, nx, ny, nz);
        R fl[QF], lf[QF];
#pragma unroll
        for (int i = 0; i < QF; ++i) {
            const uint32_t src = pull.src(F::cx[i], F::cy[i], F::cz[i]);
            const uint32_t rev = pull.src(-F::cx[i], -F::cy[i], -F::cz[i]);
            fl[i] = static_cast<R>(plane_ptr(f, i, plane)[src]);
            lf[i] = plane_ptr(nuf1, i, plane)[rev];
        }
        // BGK, TRT, MRT and RLBM conserve rho and j, so the velocity the forward
        // scalar step read from the post-collision state equals that of the
        // pulled state up to rounding: the linearisation point is taken from
        // the pulled populations, without re-colliding them.
        const Macro<R> m = macro_state<F, R>(fl);
        const Macro<R>& mp = m;
        R ga[QS], gb[QS], la[QS], lb[QS];
#pragma unroll
        for (int j = 0; j < QS; ++j) {
            const uint32_t src = pull.src(S::cx[j], S::cy[j], S::cz[j]);
            const uint32_t rev = pull.src(-S::cx[j], -S::cy[j], -S::cz[j]);
            ga[j] = static_cast<R>(plane_ptr(a, j, plane)[src]);
            gb[j] = static_cast<R>(plane_ptr(b, j, plane)[src]);
            la[j] = plane_ptr(nua1, j, plane)[rev];
            lb[j] = plane_ptr(nub1, j, plane)[rev];
        }
        const R ca = concentration<S, R>(ga), cb = concentration<S, R>(gb);
        R dsa, dsb;
        reaction_grad<R>(ca, cb, k, dsa, dsb);
        static_assert(1.0 / S::cs2 == 3.0 || 1.0 / S::cs2 == 4.0, "1 / cs2 of D2Q5 (3) or D3Q7 (4)");
        const R inv_cs2 = R(1.0 / S::cs2);
        // Rest direction by remainder (w_0 e_0 = 1 - sum_{j>0} w_j e_j, and
        // likewise w_0), so a constant adjoint is carried without the 3e-8
        // per-step growth of the species weights rounded to FP32.
        R ea = 0, eb = 0, wa = 0, wb = 0, pax = 0, pay = 0, paz = 0, pbx = 0, pby = 0, pbz = 0, we_moving = 0,
          w_moving = 0;
#pragma unroll
        for (int j = 1; j < QS; ++j) {
            const R w = R(S::w[j]);
            const R cx = R(S::cx[j]), cy = R(S::cy[j]), cz = R(S::cz[j]);
            const R cu = S::D == 3 ? cx * mp.ux + cy * mp.uy + cz * mp.uz : cx * mp.ux + cy * mp.uy;
            const R we = w * (R(1) + cu * inv_cs2);
            we_moving += we;
            w_moving += w;
            ea += la[j] * we;
            eb += lb[j] * we;
            wa += la[j] * w;
            wb += lb[j] * w;
            pax += la[j] * w * cx;
            pay += la[j] * w * cy;
            pbx += lb[j] * w * cx;
            pby += lb[j] * w * cy;
            if constexpr (S::D == 3) {
                paz += la[j] * w * cz;
                pbz += lb[j] * w * cz;
            }
            acc += static_cast<double>(la[j] * (ca * we - ga[j]) + lb[j] * (cb * we - gb[j]));
        }
        {
            const R we0 = R(1) - we_moving, w0 = R(1) - w_moving;
            ea += la[0] * we0;
            eb += lb[0] * we0;
            wa += la[0] * w0;
            wb += lb[0] * w0;
            acc += static_cast<double>(la[0] * (ca * we0 - ga[0]) + lb[0] * (cb * we0 - gb[0]));
        }
        const R sw = wa + wb;
#pragma unroll
        for (int j = 0; j < QS; ++j) {
            put(plane_ptr(nua0, j, plane) + cell, (R(1) - omega_c) * la[j] + omega_c * ea + sw * dsa);
            put(plane_ptr(nub0, j, plane) + cell, (R(1) - omega_c) * lb[j] + omega_c * eb + sw * dsb);
        }
        const R lux = omega_c * inv_cs2 * (ca * pax + cb * pbx);
        const R luy = omega_c * inv_cs2 * (ca * pay + cb * pby);
        const R luz = omega_c * inv_cs2 * (ca * paz + cb * pbz);
        const R irho = R(1) / mp.rho;
#pragma unroll
        for (int i = 0; i < QF; ++i) {
            R t = lux * (R(F::cx[i]) - mp.ux) + luy * (R(F::cy[i]) - mp.uy);
            if constexpr (F::D == 3) t += luz * (R(F::cz[i]) - mp.uz);
            lf[i] += t * irho;
        }
        // d feq / d f applied as (c_i - u) . b
        auto dev = [&](int i, R bx, R by, R bz) {
            R t = bx * (R(F::cx[i]) - m.ux) + by * (R(F::cy[i]) - m.uy);
            if constexpr (F::D == 3) t += bz * (R(F::cz[i]) - m.uz);
            return t;
        };
        if constexpr (CT == kMRT && QF == 19) {
            // collide_mrt_d3q19: J^T lam = M^T (I - S) mu + (d m_eq/d f)^T S mu,
            // mu = D^-1 M lam, with the reference's equilibria written in (rho,
            // j) (rho u_a u_b = j_a j_b / rho): rows 1, 2 (energy), 4/6/8
            // (-2/3 j), 9-13 (shear) depend on (rho, j); ghosts are zero.
            constexpr const Basis<19>& B = kD3Q19Basis;
            const R rates[19] = {R(0), R(1), R(1), R(0), R(1), R(0), R(1), R(0), R(1), omega,
                                 omega, omega, omega, omega, R(1), R(1), R(1), R(1), R(1)};
            R nm[19], ne[19];
#pragma unroll
            for (int r = 0; r < 19; ++r) {
                R acc_r = R(0);
#pragma unroll
                for (int i = 0; i < 19; ++i)
                    if (B.m[r][i] != 0) acc_r += R(B.m[r][i]) * lf[i];
                const R mu = acc_r / R(B.norm[r]);
                ne[r] = rates[r] * mu;
                nm[r] = mu - ne[r];
            }
            const R ux = m.ux, uy = m.uy, uz = m.uz, uu = ux * ux + uy * uy + uz * uz;
            const R grho = ne[1] * (R(-11) - R(19) * uu) + ne[2] * (R(3) + R(5.5) * uu) -
                           ne[9] * R(3) * (R(2) * ux * ux - uy * uy - uz * uz) - ne[10] * (uy * uy - uz * uz) -
                           ne[11] * ux * uy - ne[12] * uy * uz - ne[13] * ux * uz;
            const R two_thirds = R(2) / R(3);
            const R gjx = R(38) * ux * ne[1] - R(11) * ux * ne[2] - two_thirds * ne[4] + R(12) * ux * ne[9] +
                          uy * ne[11] + uz * ne[13];
            const R gjy = R(38) * uy * ne[1] - R(11) * uy * ne[2] - two_thirds * ne[6] - R(6) * uy * ne[9] +
                          R(2) * uy * ne[10] + ux * ne[11] + uz * ne[12];
            const R gjz = R(38) * uz * ne[1] - R(11) * uz * ne[2] - two_thirds * ne[8] - R(6) * uz * ne[9] -
                          R(2) * uz * ne[10] + uy * ne[12] + ux * ne[13];
#pragma unroll
            for (int i = 0; i < 19; ++i) {
                R t = R(0);
#pragma unroll
                for (int r = 0; r < 19; ++r)
                    if (B.m[r][i] != 0) t += R(B.m[r][i]) * nm[r];
                put(plane_ptr(nuf0, i, plane) + cell, t + grho + R(F::cx[i]) * gjx + R(F::cy[i]) * gjy +
                                                          R(F::cz[i]) * gjz);
            }
        } else if constexpr (CT == kMRT && QF == 27) {
            // collide_mrt_d3q27: f' = f - M^-1 S M (f - feq) with M^-1 S M
            // symmetric, so J^T lam = lam - w + (d feq/d f)^T w, w = M^-1 S M lam.
            R w[27];
            mrt27_operator<R>(lf, w, omega);
            R al, bx, by, bz;
            feq_transpose<F, R>(w, m.ux, m.uy, m.uz, al, bx, by, bz);
#pragma unroll
            for (int i = 0; i < 27; ++i)
                put(plane_ptr(nuf0, i, plane) + cell, lf[i] - w[i] + al + dev(i, bx, by, bz));
        } else if constexpr (CT == kMRT) {
            // f' = M^-1 [(I - S) M f + S m_eq(rho, j)], M^-1 = M^T D^-1
            // (collide_mrt_d2q9): J^T lam = M^T (I - S) mu + (d m_eq/d f)^T S mu
            // with mu = D^-1 M lam and d/df_i = d/drho + c_ix d/djx + c_iy d/djy.
            const R s[9] = {R(0), R(1.63), R(1.14), R(0), R(1.92), R(0), R(1.92), omega, omega};
            R mu[9], nm[9], ne[9];
            mrt_forward_d2q9<R>(lf, mu);
#pragma unroll
            for (int r = 0; r < 9; ++r) {
                mu[r] /= R(kD2Q9Basis.norm[r]);
                nm[r] = (R(1) - s[r]) * mu[r];
                ne[r] = s[r] * mu[r];
            }
            const R ux = m.ux, uy = m.uy, uu = ux * ux + uy * uy;
            const R grho = ne[1] * (R(-2) - R(3) * uu) + ne[2] * (R(1) + R(3) * uu) - ne[7] * (ux * ux - uy * uy) -
                           ne[8] * ux * uy;
            const R gjx = R(6) * ux * (ne[1] - ne[2]) - ne[4] + R(2) * ux * ne[7] + uy * ne[8];
            const R gjy = R(6) * uy * (ne[1] - ne[2]) - ne[6] - R(2) * uy * ne[7] + ux * ne[8];
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                R t = R(0);
#pragma unroll
                for (int r = 0; r < 9; ++r)
                    if (kD2Q9Basis.m[r][i] != 0) t += R(kD2Q9Basis.m[r][i]) * nm[r];
                put(plane_ptr(nuf0, i, plane) + cell,
                    t + grho + R(D2Q9::cx[i]) * gjx + R(D2Q9::cy[i]) * gjy);
            }
        } else if constexpr (CT == kRLBM) {
            R out[QF];
            rlbm_transpose<F, R>(fl, lf, m, omega, out);
#pragma unroll
            for (int i = 0; i < QF; ++i) put(plane_ptr(nuf0, i, plane) + cell, out[i]);
        } else if constexpr (CT == kBGK) {
            R al, bx, by, bz;
            feq_transpose<F, R>(lf, m.ux, m.uy, m.uz, al, bx, by, bz);
#pragma unroll
            for (int i = 0; i < QF; ++i)
                put(plane_ptr(nuf0, i, plane) + cell, (R(1) - omega) * lf[i] + omega * (al + dev(i, bx, by, bz)));
        } else {
            R lp[QF], lm[QF];
#pragma unroll
            for (int i = 0; i < QF; ++i) {
                lp[i] = R(0.5) * (lf[i] + lf[F::opp[i]]);
                lm[i] = R(0.5) * (lf[i] - lf[F::opp[i]]);
            }
            R ap, bxp, byp, bzp, am, bxm, bym, bzm;
            feq_transpose<F, R>(lp, m.ux, m.uy, m.uz, ap, bxp, byp, bzp);
            feq_transpose<F, R>(lm, m.ux, m.uy, m.uz, am, bxm, bym, bzm);
#pragma unroll
            for (int i = 0; i < QF; ++i)
                put(plane_ptr(nuf0, i, plane) + cell, lf[i] - omega * lp[i] - omega_minus * lm[i] +
                                                          omega * (ap + dev(i, bxp, byp, bzp)) +
                                                          omega_minus * (am + dev(i, bxm, bym, bzm)));
        }
    }
    // d J / d omega_c: per-block partial added to this block's own slot, so
    // successive launches accumulate without atomics (deterministic).
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ double wsum[MAXT / 64];
    const int t = threadIdx.x;
    if ((t & 63) == 0) wsum[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
        double s = 0.0;
        const int nw = (blockDim.x + 63) / 64;
        for (int w = 0; w < nw; ++w) s += wsum[w];
        djdw[blockIdx.x] += s;
    }
}

// Lambda[i][y] = nu[i][y + c_i] (+ shift): the adjoint of stored populations,
// rounded to the gradient's type RO.
template <class R, class RO, class L>
__global__ __launch_bounds__(1024) void reverse_pull_kernel(const R* __restrict__ nu, RO* __restrict__ lam, Cells c,
                                                            R shift) {
    const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= c.count) return;
    const CellCoord q = cell_coord(cell, c);
    const int nx = static_cast<int>(c.nx), ny = static_cast<int>(c.ny), nz = static_cast<int>(c.nz);
    const size_t plane = c.count;
    const Pull pull(cell, q.x, q.y, q.z, nx, ny, nz);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const uint32_t rev = pull.src(-L::cx[i], -L::cy[i], -L::cz[i]);
        lam[i * plane + cell] = static_cast<RO>(plane_ptr(nu, i, plane)[rev] + shift);
    }
}

template <class R> __global__ __launch_bounds__(256) void fill_kernel(R* p, size_t n, R v) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256) p[i] = v;
}

template <class R> __global__ __launch_bounds__(256) void zero_double_kernel(double* p, size_t n) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256) p[i] = 0.0;
}

// Sum of n values into out[0] (double), one block, fixed order.
template <class R> __global__ __launch_bounds__(256) void sum_kernel(const R* p, size_t n, double* out) {
    double s = 0.0;
    for (size_t i = threadIdx.x; i < n; i += 256) s += static_cast<double>(p[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (w[0] + w[1]) + (w[2] + w[3]);
}

// Per-block partial sums of a field, for large fields (grid-stride).
template <class R> __global__ __launch_bounds__(256) void partial_sum_kernel(const R* p, size_t n, double* part) {
    double s = 0.0;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256)
        s += static_cast<double>(p[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
}

struct GradLayout {
    size_t cells, state_elems, es, nseg, seg_slots, blocks;
    size_t off_ckpt, off_seg, off_nu, off_part, off_red, total;
};

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// planes: population planes of one state (Q of the fluid + 2 Q of the species).
inline GradLayout grad_layout(size_t es, size_t planes, uint64_t cells, int steps, int interval, dim3 grid) {
    GradLayout L{};
    L.cells = cells;
    L.state_elems = planes * L.cells;
    L.es = es;
    L.nseg = (static_cast<size_t>(steps) + interval - 1) / interval;
    L.seg_slots = static_cast<size_t>(interval);
    L.blocks = static_cast<size_t>(grid.x) * grid.y;
    size_t off = 0;
    L.off_ckpt = off;
    off = align_up(off + L.nseg * L.state_elems * es);  // states at the segment starts
    L.off_seg = off;
    off = align_up(off + 2 * L.seg_slots * L.state_elems * es);  // two recomputed segments
    L.off_nu = off;
    off = align_up(off + 2 * L.state_elems * sizeof(double));  // adjoint ping-pong (FP64 or FP32 elements)
    L.off_part = off;
    off = align_up(off + L.blocks * sizeof(double));
    L.off_red = off;
    off = align_up(off + 1024 * sizeof(double));
    L.total = off;
    return L;
}

// One state of the loop in a workspace slot: the fluid's Q planes, then each
// species' Q planes.
template <class F, class S, class R> struct State {
    R *f, *a, *b;
    State(R* base, size_t cells) : f(base), a(base + F::Q * cells), b(base + (F::Q + S::Q) * cells) {}
};

// The second stream of the reverse sweep and the events that order it
// against the caller's stream; destroyed when the sweep ends or fails.
struct SweepStreams {
    hipStream_t side = nullptr;
    hipEvent_t forward_done = nullptr, recomputed[2] = {nullptr, nullptr}, swept[2] = {nullptr, nullptr};
    hipError_t create() {
        hipError_t e = hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
        for (hipEvent_t* ev : {&forward_done, &recomputed[0], &recomputed[1], &swept[0], &swept[1]})
            if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
        return e;
    }
    ~SweepStreams() {
        if (side != nullptr) (void)hipStreamSynchronize(side);
        for (hipEvent_t ev : {forward_done, recomputed[0], recomputed[1], swept[0], swept[1]})
            if (ev != nullptr) (void)hipEventDestroy(ev);
        if (side != nullptr) (void)hipStreamDestroy(side);
    }
};

template <class F, class S> constexpr size_t kStatePlanes = F::Q + 2 * S::Q;

// F, S: the fluid and species lattices; R: the forward's storage and
// arithmetic; RA: the adjoint's (R, or double for the FP64 adjoint of an FP32
// forward).
template <class F, class S, class R, class RA, int CT>
hipError_t run_gradient(uint32_t flags, const R* f0, const R* a0, const R* b0, const LaunchShape& s, int steps,
                        int interval,
                        double omega_d, double omega_c_d, double k_d, double* J_out, R* gf, R* ga, R* gb,
                        double* djdwc_out, char* ws, hipStream_t st) {
    using St = State<F, S, R>;
    using StA = State<F, S, RA>;
    const int storage = std::is_same<R, double>::value ? kF64 : kF32;
    const GradLayout L = grad_layout(sizeof(R), kStatePlanes<F, S>, s.cells.count, steps, interval, s.grid);
    const size_t cells = L.cells, se = L.state_elems;
    R* ckpt = reinterpret_cast<R*>(ws + L.off_ckpt);
    R* seg = reinterpret_cast<R*>(ws + L.off_seg);
    RA* nu0 = reinterpret_cast<RA*>(ws + L.off_nu);
    RA* nu1 = nu0 + se;
    double* part = reinterpret_cast<double*>(ws + L.off_part);
    double* red = reinterpret_cast<double*>(ws + L.off_red);
    // the forward's parameters as it rounds them, widened for the adjoint
    const RA omega = static_cast<RA>(static_cast<R>(omega_d)), omega_c = static_cast<RA>(static_cast<R>(omega_c_d));
    const RA k = static_cast<RA>(static_cast<R>(k_d));
    const RA wm = CT == kTRT ? static_cast<RA>(trt_omega_minus<R>(static_cast<R>(omega_d))) : RA(0);
    const uint32_t forward_flags = flags & static_cast<uint32_t>(LBC_FLAG_FAST);
    hipError_t e;
    constexpr size_t QF = F::Q, QS = S::Q;
    auto step_on = [&](hipStream_t on, St in, St out) {
        return static_cast<hipError_t>(lbc_reactive_fused_step_ex(F::id, CT, S::id, storage, forward_flags, in.f,
                                                                  out.f, in.a, out.a, in.b, out.b, s.nx, s.ny, s.nz,
                                                                  omega_d, omega_c_d, k_d, 0.0, 0.0, s.block.x,
                                                                  s.block.y, s.block.z, on));
    };
    auto step = [&](St in, St out) { return step_on(st, in, out); };
    // forward with checkpoints at the segment starts
    if ((e = hipMemcpyAsync(ckpt, f0, QF * cells * sizeof(R), hipMemcpyDeviceToDevice, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(ckpt + QF * cells, a0, QS * cells * sizeof(R), hipMemcpyDeviceToDevice, st)) !=
        hipSuccess)
        return e;
    if ((e = hipMemcpyAsync(ckpt + (QF + QS) * cells, b0, QS * cells * sizeof(R), hipMemcpyDeviceToDevice, st)) !=
        hipSuccess)
        return e;
    R* last = nullptr;
    for (size_t sg = 0; sg < L.nseg; ++sg) {
        const int len = static_cast<int>(std::min<size_t>(interval, steps - sg * interval));
        St cur(ckpt + sg * se, cells);
        for (int t = 0; t < len; ++t) {
            const bool to_ckpt = (t == len - 1) && (sg + 1 < L.nseg);
            R* dst = to_ckpt ? ckpt + (sg + 1) * se : seg + ((t & 1) ? se : 0);
            St out(dst, cells);
            if ((e = step(cur, out)) != hipSuccess) return e;
            cur = out;
            last = dst;
        }
    }
    // J = <A>(0) - <A>(T)
    // (fixed-order two-level sums: 512 per-block partials in red[512..1023], then one block)
    hipLaunchKernelGGL((partial_sum_kernel<R>), dim3(512), dim3(256), 0, st, a0, QS * cells, red + 512);
    hipLaunchKernelGGL((sum_kernel<double>), dim3(1), dim3(256), 0, st, red + 512, 512, red);
    hipLaunchKernelGGL((partial_sum_kernel<R>), dim3(512), dim3(256), 0, st, last + QF * cells, QS * cells,
                       red + 512);
    hipLaunchKernelGGL((sum_kernel<double>), dim3(1), dim3(256), 0, st, red + 512, 512, red + 1);
    // The recomputes may overwrite the forward's scratch states from here on.
    // Segments shorter than kOverlapSteps stay on `stream`: two cross-stream
    // waits per segment cost more than the overlap gains (with interval 1 the
    // step rose from 24 to 34 us at 512^2).
    constexpr int kOverlapSteps = 8;
    const bool overlap = interval >= kOverlapSteps && L.nseg > 1;
    SweepStreams ss;
    if (overlap) {
        if ((e = ss.create()) != hipSuccess) return e;
        if ((e = hipEventRecord(ss.forward_done, st)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(ss.side, ss.forward_done, 0)) != hipSuccess) return e;
    }
    const hipStream_t rs = overlap ? ss.side : st;
    // adjoint initial condition: dJ/da_T = -1/N everywhere
    const RA inv_n = RA(1) / static_cast<RA>(cells);
    const unsigned fb = static_cast<unsigned>(std::min<size_t>(4096, (se + 255) / 256));
    hipLaunchKernelGGL((fill_kernel<RA>), dim3(fb), dim3(256), 0, st, nu1, QF * cells, RA(0));
    hipLaunchKernelGGL((fill_kernel<RA>), dim3(fb), dim3(256), 0, st, nu1 + QF * cells, QS * cells, -inv_n);
    hipLaunchKernelGGL((fill_kernel<RA>), dim3(fb), dim3(256), 0, st, nu1 + (QF + QS) * cells, QS * cells, RA(0));
    hipLaunchKernelGGL((zero_double_kernel<R>), dim3(4), dim3(256), 0, st, part, L.blocks);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // Reverse sweep.  Segment sgp's states 1..len-1 are recomputed into
    // segment buffer sgp % 2 (slots 1..len-1; state 0 is the checkpoint
    // itself, and with interval 1 nothing is recomputed) on the side stream,
    // while the caller's stream walks segment sgp + 1 backward from the other
    // buffer: the two are independent, and at a few hundred thousand cells
    // neither kernel fills the GPU.  Events order each walk after its
    // recompute and each recompute after the walk that last read its buffer
    // (without overlap everything runs in order on `stream`).
    auto seg_len = [&](size_t sgp) { return static_cast<int>(std::min<size_t>(interval, steps - sgp * interval)); };
    auto state_in = [&](size_t sgp, int t) {
        return t == 0 ? ckpt + sgp * se : seg + ((sgp & 1) * L.seg_slots + static_cast<size_t>(t)) * se;
    };
    auto recompute = [&](size_t sgp) {
        hipError_t err = hipSuccess;
        for (int t = 0; t + 1 < seg_len(sgp) && err == hipSuccess; ++t)
            err = step_on(rs, St(state_in(sgp, t), cells), St(state_in(sgp, t + 1), cells));
        return err == hipSuccess && overlap ? hipEventRecord(ss.recomputed[sgp & 1], ss.side) : err;
    };
    if ((e = recompute(L.nseg - 1)) != hipSuccess) return e;
    for (size_t sgp = L.nseg; sgp-- > 0;) {
        const int len = seg_len(sgp);
        if (sgp >= 1) {
            if (overlap && sgp + 1 < L.nseg &&
                (e = hipStreamWaitEvent(ss.side, ss.swept[(sgp + 1) & 1], 0)) != hipSuccess)
                return e;
            if (overlap && (e = recompute(sgp - 1)) != hipSuccess) return e;
        }
        if (overlap && (e = hipStreamWaitEvent(st, ss.recomputed[sgp & 1], 0)) != hipSuccess) return e;
        auto state_at = [&](int t) { return state_in(sgp, t); };
        for (int t = len - 1; t >= 0; --t) {
            St fw(state_at(t), cells);
            StA in(nu1, cells), out(nu0, cells);
            if (s.threads() <= 256)
                hipLaunchKernelGGL((reactive_adjoint_kernel<F, S, R, RA, CT, 256>), s.grid, s.block, 0, st, fw.f,
                                   fw.a, fw.b, in.f, in.a, in.b, out.f, out.a, out.b, s.cells, omega, wm, omega_c, k,
                                   part);
            else
                hipLaunchKernelGGL((reactive_adjoint_kernel<F, S, R, RA, CT, 1024>), s.grid, s.block, 0, st, fw.f,
                                   fw.a, fw.b, in.f, in.a, in.b, out.f, out.a, out.b, s.cells, omega, wm, omega_c, k,
                                   part);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            std::swap(nu0, nu1);
        }
        if (overlap && (e = hipEventRecord(ss.swept[sgp & 1], st)) != hipSuccess) return e;
        // in order on one stream, the previous segment is recomputed after this walk
        if (!overlap && sgp >= 1 && (e = recompute(sgp - 1)) != hipSuccess) return e;
    }
    // adjoints of the stored initial populations
    StA fin(nu1, cells);
    hipLaunchKernelGGL((reverse_pull_kernel<RA, R, F>), s.grid, s.block, 0, st, fin.f, gf, s.cells, RA(0));
    hipLaunchKernelGGL((reverse_pull_kernel<RA, R, S>), s.grid, s.block, 0, st, fin.a, ga, s.cells, inv_n);
    hipLaunchKernelGGL((reverse_pull_kernel<RA, R, S>), s.grid, s.block, 0, st, fin.b, gb, s.cells, RA(0));
    hipLaunchKernelGGL((sum_kernel<double>), dim3(1), dim3(256), 0, st, part, L.blocks, red + 2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    double h[3];
    if ((e = hipMemcpyAsync(h, red, 3 * sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    *J_out = (h[0] - h[1]) / static_cast<double>(cells);
    *djdwc_out = h[2];
    return hipSuccess;
}

}  // namespace lbc

namespace lbc {
// The adjoint's element counts (19 planes per state, its reductions) stay
// 32-bit: at most 2^29 cells per instance (23,170^2), beyond what its
// checkpoints would fit in HBM for any useful number of steps.
constexpr uint32_t kMaxAdjointCells = 1u << 29;
}  // namespace lbc

namespace lbc {

// The pairs the adjoint covers: D2Q9 + D2Q5, D3Q19 / D3Q27 + D3Q7, each with
// BGK, TRT, MRT or RLBM.
inline bool gradient_pair(int fluid, int collision) {
    if (fluid != kD2Q9 && fluid != kD3Q19 && fluid != kD3Q27) return false;
    return collision == kBGK || collision == kTRT || collision == kMRT || collision == kRLBM;
}

inline size_t state_planes(int fluid) {
    return fluid == kD2Q9 ? kStatePlanes<D2Q9, D2Q5> : fluid == kD3Q19 ? kStatePlanes<D3Q19, D3Q7>
                                                                       : kStatePlanes<D3Q27, D3Q7>;
}

hipError_t gradient_workspace(int fluid, int storage, uint64_t nx, uint64_t ny, uint64_t nz, int steps, int interval,
                              uint32_t block_x, uint32_t block_y, uint32_t block_z, uint64_t* bytes) {
    if (bytes == nullptr) return hipErrorInvalidValue;
    if (fluid != kD2Q9 && fluid != kD3Q19 && fluid != kD3Q27) return hipErrorInvalidValue;
    if (storage != kF32 && storage != kF64) return hipErrorInvalidValue;
    if (steps <= 0 || interval <= 0) return hipErrorInvalidValue;
    LaunchShape s;
    const hipError_t e = make_shape(fluid == kD2Q9, nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (s.cells.count > kMaxAdjointCells) return hipErrorInvalidValue;
    if (s.threads() % 64 != 0) return hipErrorInvalidConfiguration;
    *bytes = grad_layout(storage == kF64 ? 8 : 4, state_planes(fluid), s.cells.count, steps, interval, s.grid).total;
    return hipSuccess;
}

hipError_t gradient(int fluid, int collision, int storage, uint32_t flags, const void* f0, const void* a0,
                    const void* b0, uint64_t nx, uint64_t ny, uint64_t nz, int steps, int interval, double omega,
                    double omega_c, double k_react, double* loss, void* grad_f0, void* grad_a0, void* grad_b0,
                    double* dloss_domega_c, void* workspace, uint64_t workspace_bytes, uint32_t block_x,
                    uint32_t block_y, uint32_t block_z, void* stream) {
    if (f0 == nullptr || a0 == nullptr || b0 == nullptr || grad_f0 == nullptr || grad_a0 == nullptr ||
        grad_b0 == nullptr || workspace == nullptr)
        return hipErrorInvalidDevicePointer;
    if (loss == nullptr || dloss_domega_c == nullptr) return hipErrorInvalidValue;
    if (fluid != kD2Q9 && fluid != kD3Q19 && fluid != kD3Q27) return hipErrorInvalidValue;
    LaunchShape s;
    const hipError_t e = make_shape(fluid == kD2Q9, nx, ny, nz, block_x, block_y, block_z, s);
    if (e != hipSuccess) return e;
    if (s.cells.count > kMaxAdjointCells) return hipErrorInvalidValue;
    if (s.threads() % 64 != 0) return hipErrorInvalidConfiguration;  // whole waves: the block reduction
    if (storage != kF32 && storage != kF64) return hipErrorInvalidValue;
    if (!gradient_pair(fluid, collision)) return hipErrorInvalidValue;
    if (flags & ~static_cast<uint32_t>(LBC_FLAG_FAST | LBC_FLAG_F64_ADJOINT)) return hipErrorInvalidValue;
    if (steps <= 0 || interval <= 0) return hipErrorInvalidValue;
    const GradLayout L =
        grad_layout(storage == kF64 ? 8 : 4, state_planes(fluid), s.cells.count, steps, interval, s.grid);
    if (workspace_bytes < L.total) return hipErrorInvalidValue;
    char* ws = static_cast<char*>(workspace);
    const hipStream_t st = as_stream(stream);
    auto run = [&](auto fs, auto tag, auto adjoint_tag, auto ct) {
        using Pair = decltype(fs);
        using R = decltype(tag);
        return run_gradient<typename Pair::first_type, typename Pair::second_type, R, decltype(adjoint_tag),
                            decltype(ct)::value>(
            flags, static_cast<const R*>(f0), static_cast<const R*>(a0), static_cast<const R*>(b0), s, steps,
            interval, omega, omega_c, k_react, loss, static_cast<R*>(grad_f0), static_cast<R*>(grad_a0),
            static_cast<R*>(grad_b0), dloss_domega_c, ws, st);
    };
    using BGK = std::integral_constant<int, kBGK>;
    using TRT = std::integral_constant<int, kTRT>;
    using MRT = std::integral_constant<int, kMRT>;
    using RLBM = std::integral_constant<int, kRLBM>;
    // storage and adjoint precision: FP64; FP32; FP32 with the FP64 adjoint
    auto by_precision = [&](auto fs, auto ct) {
        if (storage == kF64) return run(fs, 0.0, 0.0, ct);
        if (flags & LBC_FLAG_F64_ADJOINT) return run(fs, 0.0f, 0.0, ct);
        return run(fs, 0.0f, 0.0f, ct);
    };
    auto by_collision = [&](auto fs) {
        if (collision == kBGK) return by_precision(fs, BGK{});
        if (collision == kTRT) return by_precision(fs, TRT{});
        if (collision == kMRT) return by_precision(fs, MRT{});
        return by_precision(fs, RLBM{});
    };
    if (fluid == kD2Q9) return by_collision(std::pair<D2Q9, D2Q5>{});
    if (fluid == kD3Q19) return by_collision(std::pair<D3Q19, D3Q7>{});
    return by_collision(std::pair<D3Q27, D3Q7>{});
}

}  // namespace lbc

extern "C" int lbc_reactive_gradient_workspace(int storage, uint64_t nx, uint64_t ny, int steps, int interval,
                                               uint32_t block_x, uint32_t block_y, uint64_t* bytes) {
    return lbc::gradient_workspace(lbc::kD2Q9, storage, nx, ny, 1, steps, interval, block_x, block_y, 1, bytes);
}

extern "C" int lbc_reactive_gradient(int collision, int storage, uint32_t flags, const void* f0, const void* a0,
                                     const void* b0, uint64_t nx, uint64_t ny, int steps, int interval, double omega,
                                     double omega_c, double k_react, double* loss, void* grad_f0, void* grad_a0,
                                     void* grad_b0, double* dloss_domega_c, void* workspace, uint64_t workspace_bytes,
                                     uint32_t block_x, uint32_t block_y, void* stream) {
    return lbc::gradient(lbc::kD2Q9, collision, storage, flags, f0, a0, b0, nx, ny, 1, steps, interval, omega,
                         omega_c, k_react, loss, grad_f0, grad_a0, grad_b0, dloss_domega_c, workspace,
                         workspace_bytes, block_x, block_y, 1, stream);
}

extern "C" int lbc_reactive_gradient_workspace_3d(int fluid_lattice, int storage, uint64_t nx, uint64_t ny,
                                                  uint64_t nz, int steps, int interval, uint32_t block_x,
                                                  uint32_t block_y, uint32_t block_z, uint64_t* bytes) {
    if (fluid_lattice != lbc::kD3Q19 && fluid_lattice != lbc::kD3Q27) return hipErrorInvalidValue;
    return lbc::gradient_workspace(fluid_lattice, storage, nx, ny, nz, steps, interval, block_x, block_y, block_z,
                                   bytes);
}

extern "C" int lbc_reactive_gradient_3d(int fluid_lattice, int collision, int storage, uint32_t flags,
                                        const void* f0, const void* a0, const void* b0, uint64_t nx, uint64_t ny,
                                        uint64_t nz, int steps, int interval, double omega, double omega_c,
                                        double k_react, double* loss, void* grad_f0, void* grad_a0, void* grad_b0,
                                        double* dloss_domega_c, void* workspace, uint64_t workspace_bytes,
                                        uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream) {
    if (fluid_lattice != lbc::kD3Q19 && fluid_lattice != lbc::kD3Q27) return hipErrorInvalidValue;
    return lbc::gradient(fluid_lattice, collision, storage, flags, f0, a0, b0, nx, ny, nz, steps, interval, omega,
                         omega_c, k_react, loss, grad_f0, grad_a0, grad_b0, dloss_domega_c, workspace,
                         workspace_bytes, block_x, block_y, block_z, stream);
}
