// Population storage formats and the shared launch plumbing of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "lbc_lattice.hpp"

namespace lbc {

// kF32 / kF64: the population itself (LatticeMemory<L, Real> of
// include/lattice_memory.hpp).  kF16 / kBF16: (f_i - offset) rounded to
// nearest-even, offset = w_i for fluid populations and 0 for scalar ones,
// arithmetic in float.
template <int ST> struct Storage;

// Every population store is written once and read back only by the next
// launch, so it is issued as a streaming (non-temporal) store: it does not
// displace the lines the same launch still reads.
template <class T> __device__ __forceinline__ void put(T* p, T v) { __builtin_nontemporal_store(v, p); }
// Population loads stay ordinary cached loads: the x-shifted pulls of one wave
// share cache lines with its neighbours' loads.  (Non-temporal loads speed up
// a plain copy by ~5% on MI355X but measured ~1-3% slower here.)
template <class T> __device__ __forceinline__ T get(const T* p) { return *p; }

template <> struct Storage<kF32> {
    using S = float;
    using R = float;
    __device__ static R load(const S* p, R) { return get(p); }
    __device__ static void store(S* p, R v, R) { put(p, v); }
    // deviation form: d = f - w on load, f = w + d on store
    __device__ static R load_dev(const S* p, R w) { return get(p) - w; }
    __device__ static void store_dev(S* p, R d, R w) { put(p, static_cast<S>(w + d)); }
    __device__ static R roundtrip(R v, R) { return v; }
    // the deviation a later load_dev returns after store_dev(d, w)
    __device__ static R roundtrip_dev(R d, R w) { return (w + d) - w; }
};

template <> struct Storage<kF64> {
    using S = double;
    using R = double;
    __device__ static R load(const S* p, R) { return get(p); }
    __device__ static void store(S* p, R v, R) { put(p, v); }
    __device__ static R load_dev(const S* p, R w) { return get(p) - w; }
    __device__ static void store_dev(S* p, R d, R w) { put(p, static_cast<S>(w + d)); }
    __device__ static R roundtrip(R v, R) { return v; }
    __device__ static R roundtrip_dev(R d, R w) { return (w + d) - w; }
};

template <> struct Storage<kF16> {
    using S = uint16_t;
    using R = float;
    __device__ static R load(const S* p, R off) { return off + static_cast<float>(__builtin_bit_cast(_Float16, get(p))); }
    __device__ static void store(S* p, R v, R off) {
        put(p, __builtin_bit_cast(uint16_t, static_cast<_Float16>(v - off)));
    }
    // The value a later load of a stored v returns.
    __device__ static R roundtrip(R v, R off) { return off + static_cast<float>(static_cast<_Float16>(v - off)); }
    // fluid deviations are stored as they are
    __device__ static R load_dev(const S* p, R) { return static_cast<float>(__builtin_bit_cast(_Float16, get(p))); }
    __device__ static void store_dev(S* p, R d, R) { put(p, __builtin_bit_cast(uint16_t, static_cast<_Float16>(d))); }
    __device__ static R roundtrip_dev(R d, R) { return static_cast<float>(static_cast<_Float16>(d)); }
};

template <> struct Storage<kBF16> {
    using S = uint16_t;
    using R = float;
    __device__ static R load(const S* p, R off) {
        return off + __builtin_bit_cast(float, static_cast<uint32_t>(get(p)) << 16);
    }
    __device__ static uint16_t encode(float d) {
        uint32_t x = __builtin_bit_cast(uint32_t, d);
        if ((x & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((x >> 16) | 0x40u);
        x += 0x7fffu + ((x >> 16) & 1u);
        return static_cast<uint16_t>(x >> 16);
    }
    __device__ static void store(S* p, R v, R off) { put(p, encode(v - off)); }
    __device__ static R roundtrip(R v, R off) {
        return off + __builtin_bit_cast(float, static_cast<uint32_t>(encode(v - off)) << 16);
    }
    __device__ static R load_dev(const S* p, R) { return __builtin_bit_cast(float, static_cast<uint32_t>(get(p)) << 16); }
    __device__ static void store_dev(S* p, R d, R) { put(p, encode(d)); }
    __device__ static R roundtrip_dev(R d, R) {
        return __builtin_bit_cast(float, static_cast<uint32_t>(encode(d)) << 16);
    }
};

// Base of population plane i: a wave-uniform pointer, so the per-lane part of
// every address stays a 32-bit cell offset.
template <class T> __device__ __forceinline__ T* plane_ptr(T* base, int i, size_t plane) {
    return base + static_cast<size_t>(i) * plane;
}

// Base of population plane i for stores: the same address with the plane's
// byte offset passed through v_readfirstlane (a no-op on a value already in
// scalar registers).  The compiler otherwise sees the Q store addresses
// base + i * plane + cell as one chain, strength-reduces it to a per-lane
// 64-bit pointer stepped by the plane size, and spends two or three VALU
// instructions per store on 64-bit address arithmetic instead of the scalar
// base + 32-bit lane offset form.
template <class T> __device__ __forceinline__ T* plane_ptr_opaque(T* base, int i, size_t plane) {
    const uint64_t off = static_cast<uint64_t>(i) * plane * sizeof(T);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(off));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(off >> 32));
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T*>(reinterpret_cast<B*>(base) + ((static_cast<uint64_t>(hi) << 32) | lo));
}

// The same per-lane value, recomputed where it is used: an empty asm
// statement (no instruction) that the optimizer must assume changes it, so
// the stores' byte offsets are formed in the stores' own block as 32-bit
// values (the scalar base + 32-bit offset addressing mode) rather than taken
// over as a 64-bit value from an earlier block.
__device__ __forceinline__ uint32_t fresh_lane_value(uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Element `cell` of a plane.  Planes of at most 4 GiB (WIDE = false): uniform
// base + zero-extended 32-bit byte offset, the form the global-load/store
// "saddr" addressing mode takes directly (no per-lane 64-bit address
// arithmetic).  Larger planes (WIDE = true): a 64-bit per-lane address.  The
// launchers pick the instantiation from the plane size.
template <bool WIDE = false, class T> __device__ __forceinline__ T* at(T* plane_base, uint32_t cell) {
    if constexpr (WIDE) {
        return plane_base + static_cast<uint64_t>(cell);
    } else {
        using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
        return reinterpret_cast<T*>(reinterpret_cast<B*>(plane_base) + cell * static_cast<uint32_t>(sizeof(T)));
    }
}

template <int ST> constexpr bool kOffsetStorage = (ST == kF16 || ST == kBF16);

// Pull streaming with periodic wrap, as offsets from the destination cell: a
// population with velocity c is read from x - c (include/lattice_core.hpp:70-78),
// so c_x = +1 reads x - 1 (offset -1, or nx - 1 at x = 0) and c_x = -1 reads
// x + 1 (offset +1, or 1 - nx at the last column); y and z likewise in units
// of a row and of an xy-plane.  Six selects per cell, then each population's
// source is the destination plus up to three of them (no multiplies).  The
// sums are taken modulo 2^32, where every source lands back in [0, extent)
// (the plane's lane extent, below 2^32).
struct Pull {
    uint32_t dst;
    uint32_t xm, xp, ym, yp, zm, zp;
    // `cell`: the destination's element offset in its plane; `zgap`: elements
    // between consecutive xy-slabs of a plane beyond the slab itself (0 for
    // contiguous planes), so a z step is nx ny + zgap elements.
    __device__ __forceinline__ Pull(uint32_t cell, int x, int y, int z, int nx, int ny, int nz, uint32_t zgap = 0u)
        : dst(cell) {
        const uint32_t row = static_cast<uint32_t>(nx);
        const uint32_t slab = row * static_cast<uint32_t>(ny);
        const uint32_t zstep = slab + zgap;
        xm = x == 0 ? row - 1u : ~0u;
        xp = x + 1 == nx ? 1u - row : 1u;
        ym = y == 0 ? slab - row : 0u - row;
        yp = y + 1 == ny ? row - slab : row;
        zm = z == 0 ? zstep * static_cast<uint32_t>(nz - 1) : 0u - zstep;
        zp = z + 1 == nz ? 0u - zstep * static_cast<uint32_t>(nz - 1) : zstep;
    }
    // source cell of a population with velocity (cx, cy, cz)
    __device__ __forceinline__ uint32_t src(int cx, int cy, int cz) const {
        uint32_t s = dst;
        if (cx > 0) s += xm;
        if (cx < 0) s += xp;
        if (cy > 0) s += ym;
        if (cy < 0) s += yp;
        if (cz > 0) s += zm;
        if (cz < 0) s += zp;
        return s;
    }
};

}  // namespace lbc
