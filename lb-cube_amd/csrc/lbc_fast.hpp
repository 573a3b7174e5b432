// Deviation-form BGK and TRT for reduced-precision storage (the opt-in
// LBC_FLAG_FAST arithmetic of lbc_fluid_step_ex).
//
// Populations are carried as deviations d_i = f_i - w_i from the rest state,
// the quantity the f16/bf16 formats store.  With sum w_i = 1 and
// sum w_i c_i = 0:
//   rho = 1 + sum d_i,   j = sum c_i d_i,   u = j / rho,
//   feq_i - w_i = w_i (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u)),
// split for each pair (i, opp(i)) into an even part w_i A and an odd part
// w_i B with A = drho + rho (4.5 (c.u)^2 - 1.5 u.u), B = 3 rho c.u.  Written
// with explicit fused multiply-adds and a hardware reciprocal of rho, this is
// roughly half the VALU work of the reference-order form and keeps the
// rounding relative to the deviations rather than to f_i ~ w_i.  Results
// differ from the reference order by rounding; the tests bound the difference.
#pragma once

#include "lbc_physics.hpp"

namespace lbc {

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class L, class R> struct DevMacro {
    R drho, rho, jx, jy, jz, ux, uy, uz, m15uu;
};

template <class L, class R> __device__ __forceinline__ DevMacro<L, R> dev_macro(const R* d) {
    // pairwise sums over the opposite pairs keep the adds short and balanced
    R s = d[0], jx = R(0), jy = R(0), jz = R(0);
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        if (L::opp[i] < i) continue;
        const int j = L::opp[i];
        s += d[i] + d[j];
        const R diff = d[i] - d[j];
        if (L::cx[i] > 0) jx += diff;
        if (L::cx[i] < 0) jx -= diff;
        if (L::cy[i] > 0) jy += diff;
        if (L::cy[i] < 0) jy -= diff;
        if (L::cz[i] > 0) jz += diff;
        if (L::cz[i] < 0) jz -= diff;
    }
    DevMacro<L, R> m;
    m.drho = s;
    m.rho = R(1) + s;
    m.jx = jx;
    m.jy = jy;
    m.jz = jz;
    const R inv = fast_rcp(m.rho);
    m.ux = jx * inv;
    m.uy = jy * inv;
    m.uz = L::D == 3 ? jz * inv : R(0);
    const R uu = L::D == 3 ? fmadd(m.uz, m.uz, fmadd(m.uy, m.uy, m.ux * m.ux)) : fmadd(m.uy, m.uy, m.ux * m.ux);
    m.m15uu = R(-1.5) * uu;
    return m;
}

// c_i . u with the signs folded (same helper as the reference-order form).
template <class L, class R> __device__ __forceinline__ R dev_cu(int i, const DevMacro<L, R>& m) {
    return dir_dot<L, R>(i, m.ux, m.uy, m.uz);
}

// BGK on deviations: d_i <- d_i + omega (deq_i - d_i).
template <class L, class R, class Sink = NoSink>
__device__ __forceinline__ void collide_bgk_dev(R* d, R omega, Sink sink = {}) {
    const DevMacro<L, R> m = dev_macro<L, R>(d);
    const R r3 = R(3) * m.rho;
    {
        const R eq0 = R(L::w[0]) * fmadd(m.rho, m.m15uu, m.drho);
        d[0] = fmadd(omega, eq0 - d[0], d[0]);
        sink(0, d[0]);
    }
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        if (L::opp[i] < i) continue;
        const int j = L::opp[i];
        const R w = R(L::w[i]);
        const R cu = dev_cu<L, R>(i, m);
        const R t = fmadd(R(4.5) * cu, cu, m.m15uu);
        const R a = fmadd(m.rho, t, m.drho);
        const R wb = (w * r3) * cu;
        const R eqi = fmadd(w, a, wb);
        const R eqj = fmadd(w, a, -wb);
        d[i] = fmadd(omega, eqi - d[i], d[i]);
        d[j] = fmadd(omega, eqj - d[j], d[j]);
        sink(i, d[i]);
        sink(j, d[j]);
    }
}

// TRT on deviations, magic parameter through omega_minus (computed by the
// launcher exactly as compute_omega_minus, lattice_physics.hpp:336-344).
template <class L, class R, class Sink = NoSink>
__device__ __forceinline__ void collide_trt_dev(R* d, R wp, R wm, Sink sink = {}) {
    const DevMacro<L, R> m = dev_macro<L, R>(d);
    const R r3 = R(3) * m.rho;
    {
        // the rest population is its own opposite: only the even part relaxes
        const R eq0 = R(L::w[0]) * fmadd(m.rho, m.m15uu, m.drho);
        d[0] = fmadd(wp, eq0 - d[0], d[0]);
        sink(0, d[0]);
    }
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        if (L::opp[i] < i) continue;
        const int j = L::opp[i];
        const R w = R(L::w[i]);
        const R cu = dev_cu<L, R>(i, m);
        const R t = fmadd(R(4.5) * cu, cu, m.m15uu);
        const R wa = w * fmadd(m.rho, t, m.drho);
        const R wb = (w * r3) * cu;
        const R pe = fmadd(R(0.5), d[i] + d[j], -wa);
        const R po = fmadd(R(0.5), d[i] - d[j], -wb);
        const R ce = wp * pe;
        const R co = wm * po;
        d[i] = d[i] - ce - co;
        d[j] = d[j] - ce + co;
        sink(i, d[i]);
        sink(j, d[j]);
    }
}

// D3Q27 MRT on deviations.  The reference's D3Q27 equilibrium moments are the
// moments of feq (compute_equilibrium_moments_d3q27, lattice_mrt.hpp:1242-1268),
// so its collision f' = M^-1 [m - S (m - M feq)] equals f - M^-1 S M (f - feq):
// one separable forward transform of f_neq = f - feq (three passes of the 1D
// basis {1, c, 3c^2 - 2}, integer coefficients), the rates and the inverse
// norms 1/(n_a n_b n_c) folded into one factor per moment, and one separable
// inverse transform (integer coefficients again).  Rates as the reference's
// (:161-204, :1284-1305): bulk s_b 1.19 on m4 = X + Y + Z, shear s_nu on
// m5 = 2X - Y - Z, m6 = Y - Z and the three off-diagonal second moments,
// s_e 1.1, s_eps 1.2, s_q 1.2 on the rest; conserved moments do not relax.
// About a quarter of the reference order's VALU work (its inverse divides).
__host__ __device__ constexpr int g27(int a, int b, int c) { return (a * 3 + b) * 3 + c; }

template <class R, class Sink = NoSink>
__device__ __forceinline__ void collide_mrt_d3q27_dev(R* d, R s_nu, Sink sink = {}) {
    using L = D3Q27;
    const DevMacro<L, R> m = dev_macro<L, R>(d);
    const R r3 = R(3) * m.rho;
    R p[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) {
        const R cu = dev_cu<L, R>(i, m);
        const R w = R(L::w[i]);
        const R deq = fmadd(w, fmadd(m.rho, fmadd(R(4.5) * cu, cu, m.m15uu), m.drho), (w * r3) * cu);
        p[g27(L::cx[i] + 1, L::cy[i] + 1, L::cz[i] + 1)] = d[i] - deq;
    }
    // forward: along x, then y, then z; per line (q-1, q0, q+1) -> (sum, c, 3c^2 - 2)
    auto fwd = [](R& lo, R& mid, R& hi) {
        const R s = lo + hi;
        const R m0 = s + mid, m1 = hi - lo, m2 = fmadd(R(-2), mid, s);
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
    // p[g27(a, b, c)] now holds the raw moment (a, b, c) of f_neq.  Relaxed and
    // normalized amounts (n_0 = 3, n_1 = 2, n_2 = 6):
    constexpr double n[3] = {3.0, 2.0, 6.0};
    const R s_b = R(1.19), s_e = R(1.1), s_eps = R(1.2), s_q = R(1.2);
    {
        const R X = p[g27(2, 0, 0)], Y = p[g27(0, 2, 0)], Z = p[g27(0, 0, 2)];
        const R dm4 = s_b * (X + Y + Z), dm5 = s_nu * (R(2) * X - Y - Z), dm6 = s_nu * (Y - Z);
        constexpr double k = 1.0 / 54.0;  // 1 / (n_2 n_0 n_0)
        p[g27(2, 0, 0)] = (dm4 + dm5) * R(k / 3.0);
        p[g27(0, 2, 0)] = fmadd(R(3), dm6, R(2) * dm4 - dm5) * R(k / 6.0);
        p[g27(0, 0, 2)] = fmadd(R(-3), dm6, R(2) * dm4 - dm5) * R(k / 6.0);
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
                if (order <= 1) continue;                     // conserved
                if (order == 2 && nonzero == 1) continue;     // the diagonal second moments, above
                // the off-diagonal second moments are shear; (1,1,1) and the
                // other third-order moments are s_e or s_q as in the reference's
                // row order: rows 10-15 (a 2 and a 1) s_e, row 16 (1,1,1) s_eps,
                // rows 17-26 (orders >= 4, and (2,1,1)-type) s_q.
                R rate;
                if (order == 2) rate = s_nu;
                else if (order == 3 && nonzero == 2) rate = s_e;
                else if (order == 3) rate = s_eps;
                else rate = s_q;
                p[g27(a, b, c)] *= rate * R(1.0 / (n[a] * n[b] * n[c]));
            }
    // inverse: along z, then y, then x; per line (m0, m1, m2) -> (m0 - m1 + m2, m0 - 2 m2, m0 + m1 + m2)
    auto inv = [](R& lo, R& mid, R& hi) {
        const R m0 = lo, m1 = mid, m2 = hi;
        const R t = m0 + m2;
        lo = t - m1;
        mid = fmadd(R(-2), m2, m0);
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
    for (int i = 0; i < 27; ++i) {
        d[i] -= p[g27(L::cx[i] + 1, L::cy[i] + 1, L::cz[i] + 1)];
        sink(i, d[i]);
    }
}

// D3Q19 MRT on deviations, with the reference's rates and equilibrium
// moments (collide_mrt_d3q19, lattice_mrt.hpp:908-1009: rho and j conserved,
// the five shear moments 9-13 relaxed with omega toward their equilibria --
// including the reference's factor 3 on moment 9 -- every other moment set to
// its equilibrium, ghosts to zero).  Since M w = m_eq(1, 0) and the rest state
// has no shear, f' - w = M^-1 Delta with Delta_r = m'_r - m_r(w): the conserved
// deviations, eq_r(rho, u) - eq_r(1, 0) for the rate-1 moments, and
// (1 - omega) m_r(d) + omega eq_r for the shear moments.  Only the five shear
// moments of d are transformed; the inverse uses the 14 non-ghost rows, with
// coefficients M[r][i] / |M_r|^2 folded at compile time for the non-conserved
// rows.  About half the VALU work of the reference order.
template <class R, class Sink = NoSink>
__device__ __forceinline__ void collide_mrt_d3q19_dev(R* d, R omega, Sink sink = {}) {
    using L = D3Q19;
    constexpr const Basis<19>& B = kD3Q19Basis;
    const DevMacro<L, R> m = dev_macro<L, R>(d);
    const R rho = m.rho, ux = m.ux, uy = m.uy, uz = m.uz;
    const R ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const R usq = ux2 + uy2 + uz2;
    R delta[14];
    // the conserved moments pass through unchanged (not rho * u, whose
    // rounding would bias the momentum step after step)
    delta[0] = m.drho;
    delta[3] = m.jx;
    delta[5] = m.jy;
    delta[7] = m.jz;
    delta[1] = fmadd(R(19) * rho, usq, R(-11) * m.drho);
    delta[2] = fmadd(R(-5.5) * rho, usq, R(3) * m.drho);
    delta[4] = R(-2.0 / 3.0) * delta[3];
    delta[6] = R(-2.0 / 3.0) * delta[5];
    delta[8] = R(-2.0 / 3.0) * delta[7];
    const R eq[5] = {R(3) * rho * (R(2) * ux2 - uy2 - uz2), rho * (uy2 - uz2), rho * ux * uy, rho * uy * uz,
                     rho * ux * uz};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        R sm = R(0);
#pragma unroll
        for (int i = 0; i < 19; ++i)
            if (B.m[9 + k][i] != 0) sm = fmadd(R(B.m[9 + k][i]), d[i], sm);
        delta[9 + k] = fmadd(omega, eq[k] - sm, sm);
    }
    // Conserved moments: 1/|M_r|^2 (1/19, 1/10) rounded to R is off by up to
    // 1.5e-8 relative in FP32, which would scale mass and momentum by that
    // factor every step; split it into hi + lo so the scaled moment is rounded
    // once and the drift stays at rounding noise.  The other moments are reset
    // or relaxed each step, so their folded coefficients cannot accumulate.
    constexpr int kConserved[4] = {0, 3, 5, 7};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int r = kConserved[k];
        const double inv = 1.0 / B.norm[r];
        const R hi = R(inv);
        const R lo = R(inv - static_cast<double>(hi));
        delta[r] = fmadd(delta[r], hi, delta[r] * lo);
    }
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        R v = R(0);
#pragma unroll
        for (int r = 0; r < 14; ++r) {
            if (B.m[r][i] == 0) continue;
            const bool conserved = r == 0 || r == 3 || r == 5 || r == 7;
            const double coef = conserved ? B.m[r][i] : static_cast<double>(B.m[r][i]) / B.norm[r];
            v = fmadd(R(coef), delta[r], v);
        }
        d[i] = v;
        sink(i, v);
    }
}

// D2Q9 MRT on deviations, with the reference's rates (s_e 1.63, s_eps 1.14,
// s_q 1.92, s_nu on the two stress moments; collide_mrt_d2q9,
// lattice_mrt.hpp:548-603).  As for D3Q19, m(w) = m_eq(1, 0), so
// f' - w = M^-1 Delta with Delta_r the conserved deviations and
// m_r(d) + s_r (Deq_r - m_r(d)) for the six relaxed moments, Deq_r =
// eq_r(rho, u) - eq_r(1, 0).  The relaxed moments use the raw momentum j in
// their q equilibria (the reference's -rho u is j up to rounding), and the
// conserved rows are scaled by hi + lo splits of 1/9 and 1/6 so their
// rounding cannot accumulate.
template <class R, class Sink = NoSink>
__device__ __forceinline__ void collide_mrt_d2q9_dev(R* d, R s_nu, Sink sink = {}) {
    using L = D2Q9;
    constexpr const Basis<9>& B = kD2Q9Basis;
    const DevMacro<L, R> m = dev_macro<L, R>(d);
    const R rho = m.rho, ux = m.ux, uy = m.uy;
    const R rusq = rho * fmadd(uy, uy, ux * ux);
    R delta[9];
    delta[0] = m.drho;
    delta[3] = m.jx;
    delta[5] = m.jy;
    const R deq[9] = {R(0),   fmadd(R(3), rusq, R(-2) * m.drho), fmadd(R(-3), rusq, m.drho),
                      R(0),   -m.jx,
                      R(0),   -m.jy,
                      rho * (ux * ux - uy * uy), rho * ux * uy};
    const R rate[9] = {R(0), R(1.63), R(1.14), R(0), R(1.92), R(0), R(1.92), s_nu, s_nu};
    constexpr int kRelaxed[6] = {1, 2, 4, 6, 7, 8};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int r = kRelaxed[k];
        R sm = R(0);
#pragma unroll
        for (int i = 0; i < 9; ++i)
            if (B.m[r][i] != 0) sm = fmadd(R(B.m[r][i]), d[i], sm);
        delta[r] = fmadd(rate[r], deq[r] - sm, sm);
    }
    constexpr int kConserved[3] = {0, 3, 5};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int r = kConserved[k];
        const double inv = 1.0 / B.norm[r];
        const R hi = R(inv);
        const R lo = R(inv - static_cast<double>(hi));
        delta[r] = fmadd(delta[r], hi, delta[r] * lo);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        R v = R(0);
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            if (B.m[r][i] == 0) continue;
            const bool conserved = r == 0 || r == 3 || r == 5;
            const double coef = conserved ? B.m[r][i] : static_cast<double>(B.m[r][i]) / B.norm[r];
            v = fmadd(R(coef), delta[r], v);
        }
        d[i] = v;
        sink(i, v);
    }
}

// The deviation-form collision of LBC_FLAG_FAST for a lattice and operator
// (BGK, TRT, MRT).
template <class L, int CT, class R, class Sink = NoSink>
__device__ __forceinline__ void collide_dev(R* d, R omega, R omega_minus, Sink sink = {}) {
    static_assert(CT == kBGK || CT == kTRT || CT == kMRT, "the deviation form covers BGK, TRT and MRT");
    if constexpr (CT == kBGK) collide_bgk_dev<L, R>(d, omega, sink);
    else if constexpr (CT == kTRT) collide_trt_dev<L, R>(d, omega, omega_minus, sink);
    else if constexpr (L::Q == 9) collide_mrt_d2q9_dev<R>(d, omega, sink);
    else if constexpr (L::Q == 19) collide_mrt_d3q19_dev<R>(d, omega, sink);
    else collide_mrt_d3q27_dev<R>(d, omega, sink);
}

// ---- scalar ADR and A + B -> C for the fast fused reactive step

// g_i += omega_c (w_i C (1 + c.u / cs2) - g_i) + w_i S with 1/cs2 (3 for D2Q5,
// 4 for D3Q7, exact) multiplied instead of divided, and FMAs.  The moving
// populations' equilibria and source shares are formed first and the rest
// population takes the remainder (C - sum, S - sum): the D2Q5 weights 1/3 and
// 1/6 rounded to binary sum to 1 + 3e-8 in FP32, which in the plain form adds
// a systematic 3e-8 omega_c C of mass per step wherever g is off equilibrium.
template <class S, class R>
__device__ __forceinline__ void collide_scalar_fast(R* g, R conc, R ux, R uy, R uz, R omega_c, R source) {
    constexpr double inv_cs2 = S::cs2 == 0.25 ? 4.0 : 3.0;  // D3Q7 : D2Q5
    R eq_moving = R(0), src_moving = R(0);
#pragma unroll
    for (int i = 1; i < S::Q; ++i) {
        const R wc = R(S::w[i]) * conc;
        const R cu = dir_dot<S, R>(i, ux, uy, uz);
        const R eq = fmadd(wc, R(inv_cs2) * cu, wc);
        const R src = R(S::w[i]) * source;
        eq_moving += eq;
        src_moving += src;
        g[i] = fmadd(omega_c, eq - g[i], g[i] + src);
    }
    g[0] = fmadd(omega_c, (conc - eq_moving) - g[0], g[0] + (source - src_moving));
}

template <class R> struct SeriesLimit;
template <> struct SeriesLimit<float> { static constexpr float v = 0.0625f; };
template <> struct SeriesLimit<double> { static constexpr double v = 1e-3; };

// The exact one-step increment of dA/dt = dB/dt = -k A B, the quantity of
// reaction_source, in a form without cancellation.  With z = k (B - A),
// phi(z) = expm1(z) / z and N = k A phi(z) + expm1(z):
//   A' = A B e^{-z} / (B - A e^{-z}) = A / (1 + N),   S = A' - A = -A N / (1 + N).
// This is one expression for both of the reference's branches (z = 0 gives
// A / (1 + k A)), and it loses no digits at A ~ B, where the reference's
// B - A e^{-z} cancels (about eps A / |B - A| relative error).  phi is a
// degree-5 Taylor polynomial below SeriesLimit (truncation < z^6 / 5040:
// 1.2e-11 for float, 2e-22 for double), else expm1(z) / z.
template <class R> __device__ __forceinline__ R reaction_source_fast(R a, R b, R k) {
    const R z = k * (b - a);
    R phi, em1;
    if (fabs(z) < SeriesLimit<R>::v) {
        phi = fmadd(z, fmadd(z, fmadd(z, fmadd(z, fmadd(z, R(1.0 / 720.0), R(1.0 / 120.0)), R(1.0 / 24.0)),
                                      R(1.0 / 6.0)),
                             R(0.5)),
                    R(1.0));
        em1 = z * phi;
    } else {
        em1 = expm1(z);
        phi = em1 / z;
    }
    const R num = fmadd(k * a, phi, em1);
    return -a * num * fast_rcp(R(1.0) + num);
}

}  // namespace lbc
