// Correctly rounded division by the constants 3 and 6 in a few operations.
//
// The reference's D3Q27 MRT inverse divides by the one-dimensional norms 3, 2
// and 6 inside every term (lattice_mrt.hpp:1130-1225, norms at :417-426); an
// IEEE division on gfx950 is a ten-instruction sequence (scale, reciprocal,
// refinement, fix-up), which made that kernel VALU-bound.  Here x / d is
//     q0 = RN(x * r)                 r = RN(1/d)
//     e  = fma(-q0, d, x)            the remainder x - d q0, exact
//     q  = fma(e, r, q0)             RN(q0 + e r)
// (Markstein's correction step; Muller et al., Handbook of Floating-Point
// Arithmetic, 2nd ed., section 4.7).  Why it is the correctly rounded x/d for
// d = 3 * 2^j: q0 lies within 1.5 ulp of x/d, so e is a small multiple of
// ulp(q0) and exact, and q0 + e r = x/d - e (1/d - r) differs from x/d by less
// than 2^-p ulp.  For a quotient in the normal range, x/d is never closer than
// ulp/6 to a rounding midpoint (x - d m is a nonzero multiple of the midpoint
// spacing), so both round alike.  For d = 3 the same holds for subnormal
// quotients (a/3 quanta is never a half-integer), but for d = 6 a subnormal
// quotient can be an exact tie (x = 9 * 2^-149 gives 1.5 * 2^-149), which the
// correction step can round away from even.  Hence:
//   float  / 3: the sequence in float;
//   float  / 6: the sequence in double on the widened x, rounded to float (x/6
//               is a normal double for every float x, an exact float tie is
//               exact in double, and no other float quotient lies within
//               2^-53 relative of a float midpoint);
//   double / 3: the sequence in double;
//   double / 6: the sequence in double, and the IEEE division for the
//               subnormal quotients (|x| < 6 * 2^-1022, never taken by the
//               moments of populations near the lattice weights).
// The kernels' quotients therefore equal the reference's division bit for
// bit; tests/constdiv_check.cpp checks every finite float and sampled doubles
// against the hardware division on the host.  Zero inputs may return +0 for
// -0 (values that are only ever added to a sum that starts at +0).
#pragma once

#include <hip/hip_runtime.h>

namespace lbc {

template <class R, int D> __host__ __device__ __forceinline__ R correction_step(R x) {
    constexpr R r = R(1) / R(D);
    const R q0 = x * r;
    if constexpr (sizeof(R) == 4) {
        const R e = __builtin_fmaf(-q0, R(D), x);
        return __builtin_fmaf(e, r, q0);
    } else {
        const R e = __builtin_fma(-q0, R(D), x);
        return __builtin_fma(e, r, q0);
    }
}

template <class R, int D> __host__ __device__ __forceinline__ R div_const(R x) {
    static_assert(D == 3 || D == 6, "only 3 and 6 are verified");
    if constexpr (D == 3) {
        return correction_step<R, 3>(x);
    } else if constexpr (sizeof(R) == 4) {
        return static_cast<float>(correction_step<double, 6>(static_cast<double>(x)));
    } else {
        if (__builtin_fabs(x) < 6.0 * 0x1p-1022) [[unlikely]]
            return x / 6.0;
        return correction_step<double, 6>(x);
    }
}

}  // namespace lbc
