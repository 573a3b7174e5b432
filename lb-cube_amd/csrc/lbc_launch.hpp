// Launch validation shared by every C entry point.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "lbc_lattice.hpp"
#include "lbc_storage.hpp"

namespace lbc {

inline bool is_2d(int lattice) { return lattice == kD2Q9 || lattice == kD2Q5; }

// Populations per cell of a lattice code (0: unknown).
inline int q_of(int lattice) {
    switch (lattice) {
        case kD2Q9: return 9;
        case kD3Q19: return 19;
        case kD3Q27: return 27;
        case kD2Q5: return 5;
        case kD3Q7: return 7;
        default: return 0;
    }
}

// Division by a launch constant d (1 <= d < 2^32) of any dividend n < 2^32,
// with the add-back form of Granlund & Montgomery's round-up reciprocal
// (PLDI 1994, Fig. 4.1): l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1
// (< 2^32), t = hi32(m n), q = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0).
// The 33-bit multiplier 2^32 + m exceeds 2^(32+l) / d by at most one unit, so
// n (2^32 + m) / 2^(32+l) exceeds n / d by less than 2^-l <= 1/d and its floor
// is the quotient; the add-back evaluates floor((n + t) / 2^l) without a
// 33-bit intermediate.
struct FastDiv {
    uint32_t m, s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    const uint64_t m = ((((1ull << l) - d) << 32) / d) + 1;
    return FastDiv{static_cast<uint32_t>(m), l < 1 ? l : 1u, l > 1 ? l - 1 : 0u};
}

__host__ __device__ __forceinline__ uint32_t fast_div(uint32_t n, FastDiv f) {
    const uint32_t t = static_cast<uint32_t>((static_cast<uint64_t>(n) * f.m) >> 32);  // v_mul_hi_u32
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Cells of a domain numbered linearly, x fastest ((z * ny + y) * nx + x): the
// kernels map work-item p of the launch to cell p, so every wave covers 64
// consecutive cells whatever nx and the block shape are (no idle lanes at row
// ends beyond the last wave of the domain).
// Tiles.  A population buffer may be stored in tiles of `rows` rows (rows
// divides ny; rows = ny makes each xy-slab one tile), `tiles` = ny / rows of
// them per xy-slab.  zgap: elements between the end of one tile of a population
// plane and the start of the next (0 in the reference's [Q][Z][Y][X] planes;
// (Q - 1) x the tile's cells when each tile holds its Q planes back to back,
// e.g. the slab order [Z][Q][Y][X]).  Cell (x, y, z) of tile
// t = (z ny + y) / rows lies at element (z ny + y) nx + x + t zgap of its
// population plane.  Without tiles rows = ny and tiles = 1.
struct Cells {
    uint32_t nx, ny, nz, count;
    FastDiv by_nx, by_ny;
    uint32_t zgap;
    uint32_t rows, tiles;
    FastDiv by_rows;
};

struct CellCoord {
    int x, y, z;
};

__device__ __forceinline__ CellCoord cell_coord(uint32_t p, const Cells& c) {
    const uint32_t row = fast_div(p, c.by_nx);
    const uint32_t z = fast_div(row, c.by_ny);
    return CellCoord{static_cast<int>(p - row * c.nx), static_cast<int>(row - z * c.ny), static_cast<int>(z)};
}

// The pull sources of destination cell p at coordinates q, for planes whose
// tiles are spaced `zgap` elements apart (Cells); pull.dst is the cell's
// element offset within its plane.  A z step moves nx ny + tiles zgap; a y
// step that crosses a tile boundary moves zgap further (the wrap from the
// last row of a slab to its first crosses tiles - 1 of them).  zgap = 0 and
// one tile per slab give the reference's offsets; the y-tile arithmetic runs
// only when a slab holds several tiles (a wave-uniform branch).
template <int D>
__device__ __forceinline__ Pull locate(uint32_t p, const CellCoord& q, const Cells& c, uint32_t zgap) {
    const uint32_t slab_gap = c.tiles * zgap;  // the gaps of one xy-slab's tiles
    Pull pull(p + static_cast<uint32_t>(q.z) * slab_gap, q.x, q.y, q.z, static_cast<int>(c.nx),
              static_cast<int>(c.ny), static_cast<int>(c.nz), slab_gap);
    if (c.tiles > 1) {
        const uint32_t y = static_cast<uint32_t>(q.y);
        const uint32_t t = fast_div(y, c.by_rows);
        const uint32_t r = y - t * c.rows;
        const uint32_t row = c.nx;
        const uint32_t last = (c.tiles - 1u) * zgap + row * (c.ny - 1u);  // row 0 -> row ny - 1
        pull.dst += t * zgap;
        pull.ym = y == 0 ? last : (r == 0 ? 0u - row - zgap : 0u - row);
        pull.yp = y + 1 == c.ny ? 0u - last : (r + 1 == c.rows ? row + zgap : row);
    }
    return pull;
}

struct LaunchShape {
    dim3 grid, block;  // 1D: ceil(cells / threads) blocks of `threads` work-items
    int nx, ny, nz;
    Cells cells;
    unsigned threads() const { return block.x; }
};

// Kernels are compiled twice, for blocks of up to 256 work-items (the
// register budget the library tunes for) and up to 1024 (the largest block
// HIP accepts, which the reference's launcher also accepts).  The block shape
// the caller passes sets the work-items per block (bx * by * bz); the cells
// are covered linearly.

// Largest domain: 2^32 - 2^20 cells (4,293,918,720), e.g. 1,600^3; 1,200^3
// D3Q19 populations in FP32 ping-pong take 245 GiB of HBM.
constexpr uint64_t kMaxCells = (1ull << 32) - (1ull << 20);
// Coordinates are int.
constexpr uint64_t kMaxExtent = (1ull << 31) - 1;
// HIP's limit on blocks per launch along x.
constexpr uint64_t kMaxBlocks = (1ull << 31) - 1;

// Elements a lane's offset spans within a population plane: the cell count,
// or more when the slabs are spaced apart (zgap).
inline uint64_t lane_extent(const Cells& c) {
    return static_cast<uint64_t>(c.count) + (static_cast<uint64_t>(c.nz) * c.tiles - 1) * c.zgap;
}

// Population planes of more than 4 GiB take 64-bit per-lane addresses (the
// kernels' WIDE instantiation); smaller planes keep the 32-bit byte offsets.
template <class S> bool wide_planes(const LaunchShape& s) {
    return lane_extent(s.cells) * sizeof(S) > (1ull << 32);
}

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
    // Cell numbers are 32-bit, with room for a grid-stride step of the
    // statistics kernels (kMaxCells + 2^18 < 2^32); each factor and partial
    // product is checked first, so nothing overflows.
    if (nx > kMaxExtent || ny > kMaxExtent || nz > kMaxExtent) return hipErrorInvalidValue;
    if (nx * ny > kMaxCells || nx * ny * nz > kMaxCells) return hipErrorInvalidValue;
    const uint32_t threads = bx * by * bz;
    const uint64_t count = nx * ny * nz;
    if ((count + threads - 1) / threads > kMaxBlocks) return hipErrorInvalidConfiguration;
    s.block = dim3(threads, 1, 1);
    s.grid = dim3(static_cast<unsigned>((count + threads - 1) / threads), 1, 1);
    s.nx = static_cast<int>(nx);
    s.ny = static_cast<int>(ny);
    s.nz = static_cast<int>(nz);
    s.cells = Cells{static_cast<uint32_t>(nx), static_cast<uint32_t>(ny), static_cast<uint32_t>(nz),
                    static_cast<uint32_t>(count), make_fastdiv(static_cast<uint32_t>(nx)),
                    make_fastdiv(static_cast<uint32_t>(ny)), 0u, static_cast<uint32_t>(ny), 1u,
                    make_fastdiv(static_cast<uint32_t>(ny))};
    return hipSuccess;
}

// Dense tiles of `rows` rows for population buffers: 0 = the reference's
// planes (plane = cells, zgap = 0); otherwise each tile holds the Q planes of
// its cells back to back, so the plane stride is the tile's cell count and
// zgap = (Q - 1) x that.  rows divides ny (rows = ny in 3D: the slab order
// [Z][Q][Y][X]).  Sets the shape's tiles and the largest Q's gap (a buffer of
// q planes has gap (q - 1) plane) and returns the plane stride.  The lanes'
// offsets must stay below 2^32 elements for the largest Q.
inline hipError_t tile_layout(bool two_d, uint64_t rows, int q_max, LaunchShape& s, uint64_t& plane) {
    (void)two_d;
    Cells& c = s.cells;
    if (rows == 0) {
        plane = c.count;
        return hipSuccess;
    }
    if (c.ny % rows != 0) return hipErrorInvalidValue;
    plane = static_cast<uint64_t>(c.nx) * rows;
    c.rows = static_cast<uint32_t>(rows);
    c.tiles = static_cast<uint32_t>(c.ny / rows);
    c.by_rows = make_fastdiv(static_cast<uint32_t>(rows));
    if (static_cast<uint64_t>(q_max) * c.count > 0xFFFFFFFFull) return hipErrorInvalidValue;
    c.zgap = static_cast<uint32_t>((q_max - 1) * plane);
    return hipSuccess;
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Residency cap for the memory-bound 3D fluid kernels.  They need few
// registers (45-50 VGPRs), so the hardware keeps 8 waves per SIMD (32 per CU)
// resident; with 19-27 population streams per wave that many outstanding
// streams run the HBM below its best rate.  Reserving LDS the kernels do not
// use caps the blocks per CU: at 16 waves per CU the D3Q19 TRT FP32 step at
// 256^3 ran 4.7% faster (39.1k against 37.3k MLUPS), 512^3 ~10%, D3Q27 TRT
// FP32 and FP16 (deviation form) ~4%; 20 waves per CU gave the same, 24
// less, 12 lost ground.  Kernels whose collision keeps the VALU busy
// (reference-order FP16/BF16) lose from it, as did the 2D fused reactive
// step, so the launchers apply it per kernel.
constexpr uint32_t kCappedWavesPerCu = 16;

// LDS per CU and per block of the current device (queried once per device).
struct LdsLimits {
    uint32_t per_cu = 0, per_block = 0;
};

inline LdsLimits lds_limits() {
    static LdsLimits cache[64];
    static std::once_flag once[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return LdsLimits{};
    std::call_once(once[dev], [dev] {
        int cu = 0, block = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
            hipDeviceGetAttribute(&block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
            cu > 0 && block > 0)
            cache[dev] = LdsLimits{static_cast<uint32_t>(cu), static_cast<uint32_t>(block)};
    });
    return cache[dev];
}

// Dynamic LDS bytes per block that hold a launch of `threads`-wide blocks to
// `waves` resident waves per CU (0: no cap, e.g. blocks of more than half
// that many waves, or a device whose limits could not be read).
inline uint32_t capped_lds(uint32_t threads, uint32_t waves = kCappedWavesPerCu) {
    const uint32_t blocks = waves / ((threads + 63) / 64);
    if (blocks < 2) return 0;
    const LdsLimits l = lds_limits();
    const uint32_t lds = l.per_cu / blocks;
    return (lds == 0 || lds > l.per_block) ? 0 : lds;
}

// The dispatcher hands workgroup b to XCD b % 8 (8 XCDs, each with its own
// L2).  Blocks numbered linearly over the cells share the cache lines at their
// x edges with their neighbours, so with that dispatch every edge line is
// fetched into two L2s: +6.6% of the reads of a 2D FP32 reactive step, +13% in
// FP16 (rocprofv3 FETCH_SIZE).  This renumbering gives each XCD one contiguous
// run of blocks instead (a bijection for any grid size: XCDs x < nb % 8 take
// one block more).  3D grids at 256^3 and above have whole rows per block and
// no such lines, so only the 2D single-domain kernels use it (measured on the
// 2048^2 reactive step: FETCH_SIZE back to the algorithmic bytes, +3.5% FP32,
// +4% FP16).  The ensemble kernel keeps the dispatch order (renumbered, its
// 512 instances ran 12% slower), and so do 3D grids whose rows split blocks
// (renumbered, D3Q19 and D3Q27 at 384^3 ran 7-8% slower: each XCD's separate
// window of the 38 streams costs more than the edge lines it saves).
constexpr uint32_t kXcds = 8;

// Where the renumbering pays (measured, 2048^2 reactive step): 2D grids with
// elements of at most 4 bytes, whose 256-cell blocks span 8 lines per plane
// (FP32) or 4 (FP16), so an edge line is 1/8 or 1/4 of a block's reads.  With
// FP64 (16 lines per block) the saving is half as large and the renumbered
// FP64 step ran 5% slower.
template <int D, class S> __host__ __device__ constexpr bool renumber_blocks() { return D == 2 && sizeof(S) <= 4; }

__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb / kXcds, r = nb % kXcds;
    const uint32_t x = b % kXcds, slot = b / kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
}

}  // namespace lbc
