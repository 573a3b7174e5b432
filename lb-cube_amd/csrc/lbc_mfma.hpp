// Deviation-form D3Q19 / D3Q27 MRT with both moment transforms on the matrix
// cores (LBC_FLAG_MFMA with LBC_FLAG_FAST; FP16 and BF16 storage).
//
// The collision of the deviation form (lbc_fast.hpp) written as an update of
// the stored deviations h = f - w:
//   h' = h - M^T p',   p' = R (M h - dmeq(rho, u)),
// M the lattice's integer moment basis, R the relaxation rates divided by the
// row norms |M_k|^2 (diagonal, but for D3Q27's three diagonal second moments,
// which relax in the combinations of the reference), dmeq = meq(rho, u) -
// meq(1, 0) (M w = meq(1, 0)).  rho and j are rows of M h, and R is zero on
// them, so mass and momentum change only by the FP32 rounding of the second
// product.  The two products are v_mfma_f32_32x32x16_{f16,bf16} over tiles of
// 32 cells (rows and columns padded to 32, K = 32 in two steps):
//   m   = M h   A = M   (moments x populations), B = h  (populations x cells)
//   out = M^T p' A = M^T (populations x moments), B = p' (moments x cells)
// The stored h are exact inputs of the first product and M's entries are
// small integers, exact in both formats, so m is M h accumulated in FP32;
// only p' is rounded to the 16-bit format (scaled by 2^12 in FP16, so that
// small moments stay normal), an error relative to the update, not to h.  The
// VALU keeps rho, u, dmeq and R: about a third of the separable form's work.
//
// Lane layouts of v_mfma_f32_32x32x16 (checked on gfx950): lane l holds
// A[l % 32][8 (l / 32) + j] and B[8 (l / 32) + j][l % 32] (j = 0..7), and its
// accumulator r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  Cells are
// one per lane (cell = wave base + lane): v_permlane32_swap of the register
// pair (x, y), which exchanges x's upper 32 lanes with y's lower 32, turns the
// lane-per-cell layout into the two tiles' B operands, and the tiles'
// accumulators back into one cell per lane, with one instruction per pair.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbc_fast.hpp"
#include "lbc_physics.hpp"
#include "lbc_storage.hpp"

namespace lbc {

typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 mf_b8 __attribute__((ext_vector_type(8)));
typedef float mf_f16 __attribute__((ext_vector_type(16)));
typedef uint32_t mf_u4 __attribute__((ext_vector_type(4)));

// 16-bit pattern of a small integer in FP16 (kF16) or BF16 (kBF16).
template <int ST> constexpr uint32_t small_int_bits(int n) {
    if (n == 0) return 0u;
    const uint32_t sign = n < 0 ? 0x8000u : 0u;
    const uint32_t a = static_cast<uint32_t>(n < 0 ? -n : n);
    int e = 0;
    while ((a >> (e + 1)) != 0u) ++e;
    if (ST == kF16) return sign | (static_cast<uint32_t>(e + 15) << 10) | ((a << (10 - e)) & 0x3FFu);
    return sign | (static_cast<uint32_t>(e + 127) << 7) | ((a << (7 - e)) & 0x7Fu);
}

// |n| below 2^11 (FP16) or 2^8 (BF16): every such integer is exact.
template <int ST> constexpr bool small_int_exact(int n) { return (n < 0 ? -n : n) < (ST == kF16 ? 2048 : 256); }

// Rows of the moment bases.  D3Q19: the reference's Gram-Schmidt basis
// (compute_moments_d3q19, lattice_mrt.hpp; rows 0 rho, 3/5/7 j, 9-13 shear,
// 14-18 ghosts).  D3Q27: row g27(a, b, c) = P_a(cx) P_b(cy) P_c(cz) of the 1D
// basis {1, c, 3c^2 - 2}, the raw moments the separable form transforms to.
template <class L> struct MfmaBasis;
template <> struct MfmaBasis<D3Q19> {
    static constexpr int row(int k, int i) { return kD3Q19Basis.m[k][i]; }
};
constexpr int tensor_poly(int a, int c) { return a == 0 ? 1 : a == 1 ? c : 3 * c * c - 2; }
template <> struct MfmaBasis<D3Q27> {
    static constexpr int row(int k, int i) {
        return tensor_poly(k / 9, D3Q27::cx[i]) * tensor_poly(k / 3 % 3, D3Q27::cy[i]) *
               tensor_poly(k % 3, D3Q27::cz[i]);
    }
};

// A operands of the two products in lane order: a[0] = M (row = moment,
// column = population), a[1] = M^T; per lane the 8 columns of k-step 0, then
// of k-step 1, as 16-bit pairs (the even column in the low half).
struct MfmaOperands {
    uint32_t a[2][64][8];
};

template <class L, int ST> constexpr MfmaOperands make_mfma_operands() {
    MfmaOperands t{};
    for (int which = 0; which < 2; ++which)
        for (int l = 0; l < 64; ++l)
            for (int s = 0; s < 2; ++s)
                for (int v = 0; v < 4; ++v) {
                    uint32_t word = 0u;
                    for (int half = 0; half < 2; ++half) {
                        const int row = l % 32, col = 16 * s + 8 * (l / 32) + 2 * v + half;
                        const int k = which == 0 ? row : col, i = which == 0 ? col : row;
                        const int n = (k < L::Q && i < L::Q) ? MfmaBasis<L>::row(k, i) : 0;
                        word |= small_int_bits<ST>(n) << (16 * half);
                    }
                    t.a[which][l][4 * s + v] = word;
                }
    return t;
}

template <class L, int ST> constexpr bool basis_exact() {
    for (int k = 0; k < L::Q; ++k)
        for (int i = 0; i < L::Q; ++i)
            if (!small_int_exact<ST>(MfmaBasis<L>::row(k, i))) return false;
    return true;
}
static_assert(basis_exact<D3Q19, kBF16>() && basis_exact<D3Q27, kBF16>(), "basis entries exact in BF16 (and FP16)");

static __constant__ const MfmaOperands kMfmaD3Q19F16 = make_mfma_operands<D3Q19, kF16>();
static __constant__ const MfmaOperands kMfmaD3Q19BF16 = make_mfma_operands<D3Q19, kBF16>();
static __constant__ const MfmaOperands kMfmaD3Q27F16 = make_mfma_operands<D3Q27, kF16>();
static __constant__ const MfmaOperands kMfmaD3Q27BF16 = make_mfma_operands<D3Q27, kBF16>();

template <class L, int ST> __device__ __forceinline__ const MfmaOperands& mfma_operands() {
    if constexpr (L::Q == 19) return ST == kF16 ? kMfmaD3Q19F16 : kMfmaD3Q19BF16;
    else return ST == kF16 ? kMfmaD3Q27F16 : kMfmaD3Q27BF16;
}

// Scale of p' before its rounding to 16 bits: FP16's normal range starts at
// 2^-14, and |p'| stays far below 16 (2^16 / 2^12) for any state the step
// keeps stable; BF16 has FP32's exponent range.
template <int ST> constexpr float kMfmaScale = ST == kF16 ? 4096.0f : 1.0f;

template <int ST> __device__ __forceinline__ mf_f16 mfma_32x32x16(mf_u4 a, mf_u4 b, mf_f16 c) {
    if constexpr (ST == kF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(mf_h8, a), __builtin_bit_cast(mf_h8, b), c,
                                                      0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mf_b8, a), __builtin_bit_cast(mf_b8, b), c,
                                                       0, 0, 0);
}

// Exchange of x's upper 32 lanes with y's lower 32.
__device__ __forceinline__ void swap_halves(uint32_t& x, uint32_t& y) {
    const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}

// Two 16-bit encodings packed (a in the low half), rounded to nearest even.
template <int ST> __device__ __forceinline__ uint32_t pack2(float a, float b) {
    if constexpr (ST == kF16) {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 v = {static_cast<_Float16>(a), static_cast<_Float16>(b)};
        return __builtin_bit_cast(uint32_t, v);
    } else {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        const b2 v = {static_cast<__bf16>(a), static_cast<__bf16>(b)};
        return __builtin_bit_cast(uint32_t, v);
    }
}

// One product over the wave's two tiles: `cols` are the 16 packed pairs of
// the B matrix of this lane's cell (rows 0..31), `a` the lane's A operand.
// Returns, per lane, the 32 result rows of its own cell.
template <int ST> __device__ __forceinline__ void mfma_product(const uint32_t* a_words, const uint32_t* cols,
                                                               float* rows) {
    mf_u4 b0[2], b1[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            uint32_t x = cols[8 * s + v], y = cols[8 * s + 4 + v];  // rows 16s + 2v.., 16s + 8 + 2v..
            swap_halves(x, y);
            b0[s][v] = x;  // tile 0 (cells 0..31 of the wave)
            b1[s][v] = y;  // tile 1 (cells 32..63)
        }
    const mf_u4 a0 = {a_words[0], a_words[1], a_words[2], a_words[3]};
    const mf_u4 a1 = {a_words[4], a_words[5], a_words[6], a_words[7]};
    mf_f16 d0 = {}, d1 = {};
    d0 = mfma_32x32x16<ST>(a0, b0[0], d0);
    d1 = mfma_32x32x16<ST>(a0, b1[0], d1);
    d0 = mfma_32x32x16<ST>(a1, b0[1], d0);
    d1 = mfma_32x32x16<ST>(a1, b1[1], d1);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        // (the element is copied out first: a bit cast of a vector element
        // expression reads element 0)
        const float e0 = d0[r], e1 = d1[r];
        uint32_t x = __builtin_bit_cast(uint32_t, e0), y = __builtin_bit_cast(uint32_t, e1);
        swap_halves(x, y);
        // now, for the lane's own cell: x = row 8 (r / 4) + r % 4, y = that row + 4
        rows[8 * (r / 4) + r % 4] = __builtin_bit_cast(float, x);
        rows[8 * (r / 4) + 4 + r % 4] = __builtin_bit_cast(float, y);
    }
}

// 432 x the coefficient of a monomial in meq_k / rho for D3Q27's second-order
// equilibrium (weights W / 216 with W = 64, 16, 4, 1 by |c|^2; twice the
// polynomial's coefficients, so that every term is an integer).  Monomials:
// 0: 1, 1-3: u_x, u_y, u_z, 4-6: u_x^2, u_y^2, u_z^2, 7-9: u_x u_y, u_x u_z,
// u_y u_z.
constexpr int meq27_coef432(int k, int mono) {
    int s = 0;
    for (int i = 0; i < 27; ++i) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const int c2 = cx * cx + cy * cy + cz * cz;
        const int W = c2 == 0 ? 64 : c2 == 1 ? 16 : c2 == 2 ? 4 : 1;
        const int p2 = mono == 0   ? 2
                       : mono == 1 ? 6 * cx
                       : mono == 2 ? 6 * cy
                       : mono == 3 ? 6 * cz
                       : mono == 4 ? 9 * cx * cx - 3
                       : mono == 5 ? 9 * cy * cy - 3
                       : mono == 6 ? 9 * cz * cz - 3
                       : mono == 7 ? 18 * cx * cy
                       : mono == 8 ? 18 * cx * cz
                                   : 18 * cy * cz;
        s += W * MfmaBasis<D3Q27>::row(k, i) * p2;
    }
    return s;
}
constexpr bool meq27_integral() {
    for (int k = 0; k < 27; ++k)
        for (int mono = 0; mono < 10; ++mono)
            if (meq27_coef432(k, mono) % 432 != 0) return false;
    return true;
}
static_assert(meq27_integral(), "D3Q27 equilibrium moments have integer coefficients in the tensor basis");
struct Meq27Table {
    int c[27][10];
};
constexpr Meq27Table make_meq27_table() {
    Meq27Table t{};
    for (int k = 0; k < 27; ++k)
        for (int mono = 0; mono < 10; ++mono) t.c[k][mono] = meq27_coef432(k, mono) / 432;
    return t;
}
// Evaluated by the compiler (the kernels index it with unrolled constants).
constexpr Meq27Table kMeq27 = make_meq27_table();

// p' * scale of D3Q27 from the raw moments m = M h: conserved rows 0; the
// rates of collide_mrt_d3q27_dev (the reference's, lattice_mrt.hpp:161-204,
// :1284-1305) and its combination of the diagonal second moments.
template <int ST> __device__ __forceinline__ void mfma_relax_d3q27(const float* m, float* p, float s_nu) {
    constexpr float S = kMfmaScale<ST>;
    const float drho = m[g27(0, 0, 0)], jx = m[g27(1, 0, 0)], jy = m[g27(0, 1, 0)], jz = m[g27(0, 0, 1)];
    const float inv = fast_rcp(1.0f + drho);
    const float ux = jx * inv, uy = jy * inv, uz = jz * inv;
    // rho u_a u_b as j_a u_b; rho u_a as j_a
    const float mono[10] = {drho, jx, jy, jz, jx * ux, jy * uy, jz * uz, jx * uy, jx * uz, jy * uz};
    float q[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        float dm = 0.0f;
#pragma unroll
        for (int t = 0; t < 10; ++t)
            if (kMeq27.c[k][t] != 0) dm = fmadd(static_cast<float>(kMeq27.c[k][t]), mono[t], dm);
        q[k] = m[k] - dm;
    }
    constexpr double n[3] = {3.0, 2.0, 6.0};
    const float s_b = 1.19f, s_e = 1.1f, s_eps = 1.2f, s_q = 1.2f;
#pragma unroll
    for (int k = 0; k < 32; ++k) p[k] = 0.0f;
    {
        const float X = q[g27(2, 0, 0)], Y = q[g27(0, 2, 0)], Z = q[g27(0, 0, 2)];
        const float dm4 = s_b * (X + Y + Z), dm5 = s_nu * (2.0f * X - Y - Z), dm6 = s_nu * (Y - Z);
        constexpr double kk = 1.0 / 54.0;
        p[g27(2, 0, 0)] = (dm4 + dm5) * static_cast<float>(S * kk / 3.0);
        p[g27(0, 2, 0)] = fmadd(3.0f, dm6, 2.0f * dm4 - dm5) * static_cast<float>(S * kk / 6.0);
        p[g27(0, 0, 2)] = fmadd(-3.0f, dm6, 2.0f * dm4 - dm5) * static_cast<float>(S * kk / 6.0);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int order = a + b + c;
                const int nonzero = (a > 0) + (b > 0) + (c > 0);
                if (order <= 1) continue;
                if (order == 2 && nonzero == 1) continue;
                const float scaled = static_cast<float>(S / (n[a] * n[b] * n[c]));
                float rate;
                if (order == 2) rate = s_nu;
                else if (order == 3 && nonzero == 2) rate = s_e;
                else if (order == 3) rate = s_eps;
                else rate = s_q;
                p[g27(a, b, c)] = q[g27(a, b, c)] * (rate * scaled);
            }
}

// p' * scale of D3Q19: the equilibria and rates of collide_mrt_d3q19_dev (the
// reference's collide_mrt_d3q19, lattice_mrt.hpp:908-1009): rows 0/3/5/7
// conserved, 9-13 relaxed with omega (the reference's factor 3 on 9), every
// other row with rate 1 toward its equilibrium (ghosts toward 0).
template <int ST> __device__ __forceinline__ void mfma_relax_d3q19(const float* m, float* p, float omega) {
    constexpr float S = kMfmaScale<ST>;
    constexpr const Basis<19>& B = kD3Q19Basis;
    const float drho = m[0], jx = m[3], jy = m[5], jz = m[7];
    const float rho = 1.0f + drho;
    const float inv = fast_rcp(rho);
    const float ux = jx * inv, uy = jy * inv, uz = jz * inv;
    const float ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz;
    const float usq = ux2 + uy2 + uz2;
    float deq[19] = {};
    deq[1] = fmadd(19.0f * rho, usq, -11.0f * drho);
    deq[2] = fmadd(-5.5f * rho, usq, 3.0f * drho);
    deq[4] = (-2.0f / 3.0f) * jx;
    deq[6] = (-2.0f / 3.0f) * jy;
    deq[8] = (-2.0f / 3.0f) * jz;
    deq[9] = 3.0f * rho * (2.0f * ux2 - uy2 - uz2);
    deq[10] = rho * (uy2 - uz2);
    deq[11] = rho * ux * uy;
    deq[12] = rho * uy * uz;
    deq[13] = rho * ux * uz;
#pragma unroll
    for (int k = 0; k < 32; ++k) p[k] = 0.0f;
#pragma unroll
    for (int r = 0; r < 19; ++r) {
        if (r == 0 || r == 3 || r == 5 || r == 7) continue;
        const float scaled = static_cast<float>(static_cast<double>(S) / B.norm[r]);
        const float rate = (r >= 9 && r <= 13) ? omega * scaled : scaled;
        p[r] = (m[r] - deq[r]) * rate;
    }
}

}  // namespace lbc
