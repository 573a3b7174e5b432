// Launch validation shared by every C entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbc_lattice.hpp"

namespace lbc {

inline bool is_2d(int lattice) { return lattice == kD2Q9 || lattice == kD2Q5; }

struct LaunchShape {
    dim3 grid, block;
    int nx, ny, nz;
    unsigned threads() const { return block.x * block.y * block.z; }
};

// Kernels are compiled twice, for blocks of up to 256 work-items (the
// register budget the library tunes for) and up to 1024 (the largest block
// HIP accepts, which the reference's launcher also accepts).

// Validation order of src/lattice_cuda.cu:488-503: pointers, block, 2D
// normalisation, extents.  Then this library's own limits.
inline hipError_t make_shape(bool two_d, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t bx, uint32_t by,
                             uint32_t bz, LaunchShape& s) {
    if (bx == 0 || by == 0 || bz == 0) return hipErrorInvalidConfiguration;
    if (two_d) {
        nz = 1;
        bz = 1;
    }
    if (nx == 0 || ny == 0 || nz == 0) return hipErrorInvalidValue;
    if (static_cast<uint64_t>(bx) * by * bz > 1024) return hipErrorInvalidConfiguration;
    // Byte offsets inside one population plane are 32-bit (the global-load
    // saddr form), so a plane holds at most 2^29 cells (4 GiB of doubles).
    // (each factor and partial product checked first, so nothing overflows)
    constexpr uint64_t kMaxCells = 1ull << 29;
    if (nx > kMaxCells || ny > kMaxCells || nz > kMaxCells) return hipErrorInvalidValue;
    if (nx * ny > kMaxCells || nx * ny * nz > kMaxCells) return hipErrorInvalidValue;
    const uint64_t gx = (nx + bx - 1) / bx, gy = (ny + by - 1) / by, gz = (nz + bz - 1) / bz;
    if (gx > 0x7fffffffull || gy > 65535 || gz > 65535) return hipErrorInvalidConfiguration;
    s.grid = dim3(static_cast<unsigned>(gx), static_cast<unsigned>(gy), static_cast<unsigned>(gz));
    s.block = dim3(bx, by, bz);
    s.nx = static_cast<int>(nx);
    s.ny = static_cast<int>(ny);
    s.nz = static_cast<int>(nz);
    return hipSuccess;
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

}  // namespace lbc
