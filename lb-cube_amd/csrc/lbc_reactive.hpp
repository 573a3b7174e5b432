// One cell of the fused fluid + A + B -> C reactive step (shared by the
// single-domain and the ensemble kernels).
#pragma once

#include "lbc_fast.hpp"
#include "lbc_physics.hpp"
#include "lbc_storage.hpp"

namespace lbc {

// Fluid step + A + B -> C step of one cell: `pf` locates the cell and its
// periodic sources in the fluid planes and `ps` in the species planes (the
// same pull unless the buffers are stored in tiles, whose gaps differ with the
// lattices' Q), `plane` is the element stride between population planes;
// WIDE selects 64-bit per-lane addresses (planes of more than 4 GiB).
// The reaction step of the reference reads the fluid populations that the
// fluid step has just written
// at the same cell (src/simulate_reaction.cpp:527-541); here they are still in
// registers.  With reduced-precision storage they are passed through the
// storage rounding first, so the velocity is the one the unfused pair sees.
// Shared by the single-domain kernel below and the ensemble kernel.
template <class F, int CT, class S, int ST, bool WIDE = false>
__device__ __forceinline__ void reactive_cell(const typename Storage<ST>::S* __restrict__ cur_f,
                                              typename Storage<ST>::S* __restrict__ nxt_f,
                                              const typename Storage<ST>::S* __restrict__ cur_a,
                                              typename Storage<ST>::S* __restrict__ nxt_a,
                                              const typename Storage<ST>::S* __restrict__ cur_b,
                                              typename Storage<ST>::S* __restrict__ nxt_b, const Pull& pf,
                                              const Pull& ps, size_t plane, typename Storage<ST>::R omega,
                                              typename Storage<ST>::R omega_minus, typename Storage<ST>::R omega_c,
                                              typename Storage<ST>::R k_react, typename Storage<ST>::R ref_a,
                                              typename Storage<ST>::R ref_b) {
    using St = Storage<ST>;
    using R = typename St::R;
    const uint32_t cell = pf.dst, scell = ps.dst;
    R f[F::Q];
#pragma unroll
    for (int i = 0; i < F::Q; ++i) {
        const uint32_t src = pf.src(F::cx[i], F::cy[i], F::cz[i]);
        f[i] = St::load(at<WIDE>(plane_ptr(cur_f, i, plane), src), kOffsetStorage<ST> ? R(F::w[i]) : R(0));
    }
    collide_cell<F, CT, R>(f, omega, omega_minus, [&](int i, R v) {
        St::store(at<WIDE>(plane_ptr(nxt_f, i, plane), cell), v, kOffsetStorage<ST> ? R(F::w[i]) : R(0));
    });
    if constexpr (kOffsetStorage<ST>) {
#pragma unroll
        for (int i = 0; i < F::Q; ++i) f[i] = St::roundtrip(f[i], R(F::w[i]));
    }
    const Macro<R> m = macro_state<F, R>(f);
    R ga[S::Q], gb[S::Q];
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        const uint32_t src = ps.src(S::cx[i], S::cy[i], S::cz[i]);
        ga[i] = St::load(at<WIDE>(plane_ptr(cur_a, i, plane), src), R(S::w[i]) * ref_a);
        gb[i] = St::load(at<WIDE>(plane_ptr(cur_b, i, plane), src), R(S::w[i]) * ref_b);
    }
    const R ca = concentration<S, R>(ga);
    const R cb = concentration<S, R>(gb);
    const R source = reaction_source<R>(ca, cb, k_react);
    collide_scalar<S, R>(ga, m.ux, m.uy, m.uz, omega_c, source);
    collide_scalar<S, R>(gb, m.ux, m.uy, m.uz, omega_c, source);
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        St::store(at<WIDE>(plane_ptr(nxt_a, i, plane), scell), ga[i], R(S::w[i]) * ref_a);
        St::store(at<WIDE>(plane_ptr(nxt_b, i, plane), scell), gb[i], R(S::w[i]) * ref_b);
    }
}

// The same step in the deviation form (LBC_FLAG_FAST; BGK, TRT, MRT): fluid
// populations carried as f_i - w_i with FMAs and a hardware reciprocal
// (lbc_fast.hpp), the velocity taken from the deviations as stored, the scalar
// collision with 1/cs2 multiplied and the rest equilibrium by remainder, the
// reaction increment without cancellation (lbc_fast.hpp).
template <class F, int CT, class S, int ST, bool WIDE = false>
__device__ __forceinline__ void reactive_cell_fast(const typename Storage<ST>::S* __restrict__ cur_f,
                                                   typename Storage<ST>::S* __restrict__ nxt_f,
                                                   const typename Storage<ST>::S* __restrict__ cur_a,
                                                   typename Storage<ST>::S* __restrict__ nxt_a,
                                                   const typename Storage<ST>::S* __restrict__ cur_b,
                                                   typename Storage<ST>::S* __restrict__ nxt_b, const Pull& pf,
                                                   const Pull& ps, size_t plane, typename Storage<ST>::R omega,
                                                   typename Storage<ST>::R omega_minus,
                                                   typename Storage<ST>::R omega_c, typename Storage<ST>::R k_react,
                                                   typename Storage<ST>::R ref_a, typename Storage<ST>::R ref_b) {
    using St = Storage<ST>;
    using R = typename St::R;
    const uint32_t cell = pf.dst, scell = ps.dst;
    R d[F::Q];
#pragma unroll
    for (int i = 0; i < F::Q; ++i) {
        const uint32_t src = pf.src(F::cx[i], F::cy[i], F::cz[i]);
        d[i] = St::load_dev(at<WIDE>(plane_ptr(cur_f, i, plane), src), R(F::w[i]));
    }
    collide_dev<F, CT, R>(d, omega, omega_minus,
                          [&](int i, R v) { St::store_dev(at<WIDE>(plane_ptr(nxt_f, i, plane), cell), v, R(F::w[i])); });
#pragma unroll
    for (int i = 0; i < F::Q; ++i) d[i] = St::roundtrip_dev(d[i], R(F::w[i]));
    const DevMacro<F, R> m = dev_macro<F, R>(d);
    R ga[S::Q], gb[S::Q];
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        const uint32_t src = ps.src(S::cx[i], S::cy[i], S::cz[i]);
        ga[i] = St::load(at<WIDE>(plane_ptr(cur_a, i, plane), src), R(S::w[i]) * ref_a);
        gb[i] = St::load(at<WIDE>(plane_ptr(cur_b, i, plane), src), R(S::w[i]) * ref_b);
    }
    const R ca = concentration<S, R>(ga);
    const R cb = concentration<S, R>(gb);
    const R source = reaction_source_fast<R>(ca, cb, k_react);
    collide_scalar_fast<S, R>(ga, ca, m.ux, m.uy, m.uz, omega_c, source);
    collide_scalar_fast<S, R>(gb, cb, m.ux, m.uy, m.uz, omega_c, source);
#pragma unroll
    for (int i = 0; i < S::Q; ++i) {
        St::store(at<WIDE>(plane_ptr(nxt_a, i, plane), scell), ga[i], R(S::w[i]) * ref_a);
        St::store(at<WIDE>(plane_ptr(nxt_b, i, plane), scell), gb[i], R(S::w[i]) * ref_b);
    }
}

}  // namespace lbc
