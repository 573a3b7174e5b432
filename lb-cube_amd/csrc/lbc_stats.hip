// Domain statistics on the device: the fluid integrals of log_diagnostics
// (include/lattice_io.hpp:182-234), the reactive statistics of
// compute_reactive_stats (include/lattice_io.hpp:256-338) and the stability
// guard's maximum speed (compute_max_velocity, src/turbulent_tgv_3d.cpp:149-173);
// and the per-cell macroscopic fields the VTK writers print
// (include/lattice_io.hpp:386-451, src/simulate_reaction.cpp:267-333).
//
// The reference sums serially on the host in Real after copying the whole
// population buffers back.  Here every work-item reduces a grid-stride slice
// of cells in double, waves combine with cross-lane shuffles, blocks through
// LDS, and one finishing block sums the per-block partials in a fixed order,
// so the result is deterministic for a given domain size.  Per-cell terms
// (density, |u|^2, concentrations) are formed in the storage's arithmetic type
// exactly as the reference forms them; only the accumulation is double.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/lbcube.h"
#include "lbc_launch.hpp"
#include "lbc_physics.hpp"
#include "lbc_storage.hpp"

namespace lbc {

constexpr int kStatBlocks = 1024;
constexpr int kStatThreads = 256;
constexpr int kStatSlots = 8;

// op: 0 = sum, 1 = max
template <int N>
__device__ __forceinline__ void block_reduce(double (&v)[N], const int (&op)[N], double* out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double o = __shfl_xor(v[k], off, 64);
            v[k] = op[k] ? fmax(v[k], o) : v[k] + o;
        }
    }
    __shared__ double part[kStatThreads / 64][N];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) part[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double acc = part[0][k];
            for (int w = 1; w < kStatThreads / 64; ++w) acc = op[k] ? fmax(acc, part[w][k]) : acc + part[w][k];
            out[k] = acc;
        }
    }
}

template <class L, int ST>
__global__ __launch_bounds__(kStatThreads) void fluid_diag_kernel(const typename Storage<ST>::S* __restrict__ f,
                                                                  uint32_t cells, double* partial) {
    using St = Storage<ST>;
    using R = typename St::R;
    double v[3] = {0.0, 0.0, 0.0};
    constexpr int op[3] = {0, 0, 1};
    const size_t plane = cells;
    for (uint32_t c = blockIdx.x * kStatThreads + threadIdx.x; c < cells; c += kStatBlocks * kStatThreads) {
        R loc[L::Q];
#pragma unroll
        for (int i = 0; i < L::Q; ++i) loc[i] = St::load(plane_ptr(f, i, plane) + c, kOffsetStorage<ST> ? R(L::w[i]) : R(0));
        const Macro<R> m = macro_state<L, R>(loc);
        const R sp2 = vel_sq<L, R>(m);
        v[0] += static_cast<double>(m.rho);
        v[1] += static_cast<double>(R(0.5) * m.rho * sp2);
        v[2] = fmax(v[2], static_cast<double>(sqrt(sp2)));
    }
    block_reduce<3>(v, op, partial + blockIdx.x * kStatSlots);
}

// The species may be stored in tiles (lbc_reactive_stats_tiled): cell c lies
// at element c + (c / tile) zgap of its plane, `plane` elements from one plane
// to the next (the reference's planes: plane = cells, zgap = 0).  The cells
// are visited in the same order either way, so the sums are the same bit for
// bit.
template <class S, int ST>
__global__ __launch_bounds__(kStatThreads) void reactive_stats_kernel(const typename Storage<ST>::S* __restrict__ a,
                                                                      const typename Storage<ST>::S* __restrict__ b,
                                                                      uint32_t cells, size_t plane, uint32_t zgap,
                                                                      FastDiv by_tile, typename Storage<ST>::R ref_a,
                                                                      typename Storage<ST>::R ref_b, double* partial) {
    using St = Storage<ST>;
    using R = typename St::R;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    constexpr int op[5] = {0, 0, 0, 0, 0};
    for (uint32_t c = blockIdx.x * kStatThreads + threadIdx.x; c < cells; c += kStatBlocks * kStatThreads) {
        const size_t e = c + static_cast<size_t>(fast_div(c, by_tile)) * zgap;
        R ga[S::Q], gb[S::Q];
#pragma unroll
        for (int i = 0; i < S::Q; ++i) {
            ga[i] = St::load(plane_ptr(a, i, plane) + e, R(S::w[i]) * ref_a);
            gb[i] = St::load(plane_ptr(b, i, plane) + e, R(S::w[i]) * ref_b);
        }
        const double ca = static_cast<double>(concentration<S, R>(ga));
        const double cb = static_cast<double>(concentration<S, R>(gb));
        v[0] += ca;
        v[1] += cb;
        v[2] += ca * ca;
        v[3] += cb * cb;
        v[4] += ca * cb;
    }
    block_reduce<5>(v, op, partial + blockIdx.x * kStatSlots);
}

// compute_max_velocity: the largest |u| over the cells, +inf as soon as one
// cell's speed is not finite (the reference returns infinity there so the
// guard stops the run).  Max is exact, so any reduction order gives the same
// value.
template <class L, int ST>
__global__ __launch_bounds__(kStatThreads) void max_speed_kernel(const typename Storage<ST>::S* __restrict__ f,
                                                                 uint32_t cells, double* partial) {
    using St = Storage<ST>;
    using R = typename St::R;
    double v[1] = {0.0};
    constexpr int op[1] = {1};
    const size_t plane = cells;
    for (uint32_t c = blockIdx.x * kStatThreads + threadIdx.x; c < cells; c += kStatBlocks * kStatThreads) {
        R loc[L::Q];
#pragma unroll
        for (int i = 0; i < L::Q; ++i) loc[i] = St::load(plane_ptr(f, i, plane) + c, kOffsetStorage<ST> ? R(L::w[i]) : R(0));
        const Macro<R> m = macro_state<L, R>(loc);
        const double speed = static_cast<double>(sqrt(vel_sq<L, R>(m)));
        v[0] = isfinite(speed) ? fmax(v[0], speed) : INFINITY;
    }
    block_reduce<1>(v, op, partial + blockIdx.x * kStatSlots);
}

// Sums the per-block partials in a fixed order and derives the reported
// quantities: N = 3 fluid diagnostics, N = 5 reactive statistics
// (formulas of include/lattice_io.hpp:309-337 evaluated in double; eps is the
// epsilon of the storage's arithmetic type, as the reference's
// numeric_limits<Real>::epsilon()).
template <int N>
__global__ __launch_bounds__(kStatThreads) void finish_kernel(const double* partial, double* result, uint32_t cells,
                                                              double k_react, double eps) {
    // N = 1: the maximum speed; N = 3: fluid diagnostics; N = 5: reactive
    constexpr int mode = N == 5 ? 1 : 0;
    double v[N];
    int op[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        v[k] = 0.0;
        op[k] = (N == 1 || (N == 3 && k == 2)) ? 1 : 0;
    }
    for (int blk = threadIdx.x; blk < kStatBlocks; blk += kStatThreads) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const double p = partial[blk * kStatSlots + k];
            v[k] = op[k] ? fmax(v[k], p) : v[k] + p;
        }
    }
    __shared__ double tot[N];
    int opc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) opc[k] = op[k];
    block_reduce<N>(v, opc, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if constexpr (mode == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) result[k] = tot[k];
        return;
    } else {
    const double inv = 1.0 / static_cast<double>(cells);
    const double ma = tot[0] * inv, mb = tot[1] * inv, maa = tot[2] * inv, mbb = tot[3] * inv, mab = tot[4] * inv;
    const double cov = mab - ma * mb;
    const double mp = ma * mb;
    result[0] = ma;
    result[1] = mb;
    result[2] = maa - ma * ma;
    result[3] = mbb - mb * mb;
    result[4] = cov;
    result[5] = fabs(mp) > eps ? cov / mp : 0.0;
    result[6] = k_react * mab;
    result[7] = k_react * ma * mb;
    }
}

// Per-device workspace: kStatBlocks partial slots plus the result slots,
// allocated on the first call on a device.  Every statistics call holds
// `stats_mutex` from launch to read-back (each call synchronizes its stream
// before returning), so calls from several host threads or streams never
// share the workspace concurrently.
std::mutex stats_mutex;

double* workspace(hipError_t& err) {
    static double* ws[64] = {nullptr};
    int dev = 0;
    err = hipGetDevice(&dev);
    if (err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= 64) {
        err = hipErrorInvalidDevice;
        return nullptr;
    }
    if (ws[dev] == nullptr) {
        err = hipMalloc(reinterpret_cast<void**>(&ws[dev]), (kStatBlocks + 1) * kStatSlots * sizeof(double));
        if (err != hipSuccess) {
            ws[dev] = nullptr;
            return nullptr;
        }
    }
    return ws[dev];
}

// Scratch of lbc_reactive_stats_noalloc: the same kStatBlocks partial slots
// plus result slots as the workspace above, as a static device array
// (65,600 bytes on every device the library runs on, whatever the domain
// size).  Every partial slot a finishing kernel reads is written by the
// reduction kernel of the same call, so nothing needs clearing between calls.
// Host code never names the array in a launch or copy: its device address is
// taken once per device (hipGetSymbolAddress, with that device current) and
// kept in scratch_ptr.
static __device__ __align__(256) double scratch[(kStatBlocks + 1) * kStatSlots];

double* scratch_address(hipError_t& err) {  // called with stats_mutex held
    static double* scratch_ptr[64] = {nullptr};
    int dev = 0;
    err = hipGetDevice(&dev);
    if (err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= 64) {
        err = hipErrorInvalidDevice;
        return nullptr;
    }
    if (scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return nullptr;
        scratch_ptr[dev] = static_cast<double*>(p);
    }
    return scratch_ptr[dev];
}

hipError_t read_back(const double* dev_result, double* host, int n, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(host, dev_result, n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

template <class L, int ST>
hipError_t run_fluid_diag(const void* f, uint32_t cells, double* out, hipStream_t st) {
    std::lock_guard<std::mutex> lock(stats_mutex);
    hipError_t e;
    double* ws = workspace(e);
    if (!ws) return e;
    hipLaunchKernelGGL((fluid_diag_kernel<L, ST>), dim3(kStatBlocks), dim3(kStatThreads), 0, st,
                       static_cast<const typename Storage<ST>::S*>(f), cells, ws);
    hipLaunchKernelGGL((finish_kernel<3>), dim3(1), dim3(kStatThreads), 0, st, ws, ws + kStatBlocks * kStatSlots,
                       cells, 0.0, 0.0);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return read_back(ws + kStatBlocks * kStatSlots, out, 3, st);
}

template <class L, int ST>
hipError_t run_max_speed(const void* f, uint32_t cells, double* out, hipStream_t st) {
    std::lock_guard<std::mutex> lock(stats_mutex);
    hipError_t e;
    double* ws = workspace(e);
    if (!ws) return e;
    hipLaunchKernelGGL((max_speed_kernel<L, ST>), dim3(kStatBlocks), dim3(kStatThreads), 0, st,
                       static_cast<const typename Storage<ST>::S*>(f), cells, ws);
    hipLaunchKernelGGL((finish_kernel<1>), dim3(1), dim3(kStatThreads), 0, st, ws, ws + kStatBlocks * kStatSlots,
                       cells, 0.0, 0.0);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return read_back(ws + kStatBlocks * kStatSlots, out, 1, st);
}

// Per-cell fields, one work-item per cell: density and velocity
// (compute_macro_state at the storage's arithmetic precision, as the
// reference's macro_at evaluates it in Real), or a scalar's concentration
// (compute_concentration), written as double in the cell order of the
// populations.  They replace the copy of every population to the host that
// the reference's writers start from: (1 + D) or 1 doubles per cell instead
// of Q values.
template <class L, int ST>
__global__ __launch_bounds__(kStatThreads) void macro_fields_kernel(const typename Storage<ST>::S* __restrict__ f,
                                                                    uint32_t cells, double* __restrict__ rho,
                                                                    double* __restrict__ u) {
    using St = Storage<ST>;
    using R = typename St::R;
    const uint32_t c = blockIdx.x * kStatThreads + threadIdx.x;
    if (c >= cells) return;
    const size_t plane = cells;
    R loc[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) loc[i] = St::load(plane_ptr(f, i, plane) + c, kOffsetStorage<ST> ? R(L::w[i]) : R(0));
    const Macro<R> m = macro_state<L, R>(loc);
    rho[c] = static_cast<double>(m.rho);
    u[c] = static_cast<double>(m.ux);
    u[plane + c] = static_cast<double>(m.uy);
    if constexpr (L::D == 3) u[2 * plane + c] = static_cast<double>(m.uz);
}

template <class S, int ST>
__global__ __launch_bounds__(kStatThreads) void scalar_fields_kernel(const typename Storage<ST>::S* __restrict__ g,
                                                                     uint32_t cells, typename Storage<ST>::R ref,
                                                                     double* __restrict__ conc) {
    using St = Storage<ST>;
    using R = typename St::R;
    const uint32_t c = blockIdx.x * kStatThreads + threadIdx.x;
    if (c >= cells) return;
    const size_t plane = cells;
    R loc[S::Q];
#pragma unroll
    for (int i = 0; i < S::Q; ++i) loc[i] = St::load(plane_ptr(g, i, plane) + c, R(S::w[i]) * ref);
    conc[c] = static_cast<double>(concentration<S, R>(loc));
}

template <class L, int ST>
hipError_t run_macro_fields(const void* f, uint32_t cells, double* rho, double* u, hipStream_t st) {
    hipLaunchKernelGGL((macro_fields_kernel<L, ST>), dim3((cells + kStatThreads - 1) / kStatThreads),
                       dim3(kStatThreads), 0, st, static_cast<const typename Storage<ST>::S*>(f), cells, rho, u);
    return hipGetLastError();
}

template <class S, int ST>
hipError_t run_scalar_fields(const void* g, uint32_t cells, double ref, double* conc, hipStream_t st) {
    hipLaunchKernelGGL((scalar_fields_kernel<S, ST>), dim3((cells + kStatThreads - 1) / kStatThreads),
                       dim3(kStatThreads), 0, st, static_cast<const typename Storage<ST>::S*>(g), cells,
                       static_cast<typename Storage<ST>::R>(ref), conc);
    return hipGetLastError();
}

// no_alloc: the partial sums go to the static scratch array instead of the
// workspace allocated on first use (lbc_reactive_stats_noalloc).
template <class S, int ST>
hipError_t run_reactive_stats(const void* a, const void* b, uint32_t cells, size_t plane, uint32_t zgap, double k,
                              double ra, double rb, double* out, hipStream_t st, bool no_alloc) {
    std::lock_guard<std::mutex> lock(stats_mutex);
    hipError_t e;
    double* ws = no_alloc ? scratch_address(e) : workspace(e);
    if (!ws) return e;
    using R = typename Storage<ST>::R;
    hipLaunchKernelGGL((reactive_stats_kernel<S, ST>), dim3(kStatBlocks), dim3(kStatThreads), 0, st,
                       static_cast<const typename Storage<ST>::S*>(a), static_cast<const typename Storage<ST>::S*>(b),
                       cells, plane, zgap, make_fastdiv(static_cast<uint32_t>(plane < cells ? plane : cells)),
                       static_cast<R>(ra), static_cast<R>(rb), ws);
    hipLaunchKernelGGL((finish_kernel<5>), dim3(1), dim3(kStatThreads), 0, st, ws, ws + kStatBlocks * kStatSlots,
                       cells, k, static_cast<double>(Eps<R>::v));
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return read_back(ws + kStatBlocks * kStatSlots, out, 8, st);
}

template <class L>
hipError_t fluid_diag_storage(int storage, const void* f, uint32_t cells, double* out, hipStream_t st) {
    switch (storage) {
        case kF32: return run_fluid_diag<L, kF32>(f, cells, out, st);
        case kF64: return run_fluid_diag<L, kF64>(f, cells, out, st);
        case kF16: return run_fluid_diag<L, kF16>(f, cells, out, st);
        case kBF16: return run_fluid_diag<L, kBF16>(f, cells, out, st);
        default: return hipErrorInvalidValue;
    }
}

// Storage dispatch: RUNNER<LAT, storage code>(args...).
#define LBC_STORAGE_DISPATCH(RUNNER, LAT, storage, ...)                  \
    switch (storage) {                                                   \
        case kF32: return RUNNER<LAT, kF32>(__VA_ARGS__);                \
        case kF64: return RUNNER<LAT, kF64>(__VA_ARGS__);                \
        case kF16: return RUNNER<LAT, kF16>(__VA_ARGS__);                \
        case kBF16: return RUNNER<LAT, kBF16>(__VA_ARGS__);              \
        default: return hipErrorInvalidValue;                            \
    }

template <class L> hipError_t max_speed_storage(int storage, const void* f, uint32_t cells, double* out, hipStream_t st) {
    LBC_STORAGE_DISPATCH(run_max_speed, L, storage, f, cells, out, st)
}

template <class L>
hipError_t macro_fields_storage(int storage, const void* f, uint32_t cells, double* rho, double* u, hipStream_t st) {
    LBC_STORAGE_DISPATCH(run_macro_fields, L, storage, f, cells, rho, u, st)
}

template <class S>
hipError_t scalar_fields_storage(int storage, const void* g, uint32_t cells, double ref, double* conc,
                                 hipStream_t st) {
    LBC_STORAGE_DISPATCH(run_scalar_fields, S, storage, g, cells, ref, conc, st)
}

template <class S>
hipError_t stats_storage(int storage, const void* a, const void* b, uint32_t cells, size_t plane, uint32_t zgap,
                         double k, double ra, double rb, double* out, hipStream_t st, bool no_alloc) {
    switch (storage) {
        case kF32: return run_reactive_stats<S, kF32>(a, b, cells, plane, zgap, k, ra, rb, out, st, no_alloc);
        case kF64: return run_reactive_stats<S, kF64>(a, b, cells, plane, zgap, k, ra, rb, out, st, no_alloc);
        case kF16: return run_reactive_stats<S, kF16>(a, b, cells, plane, zgap, k, ra, rb, out, st, no_alloc);
        case kBF16: return run_reactive_stats<S, kBF16>(a, b, cells, plane, zgap, k, ra, rb, out, st, no_alloc);
        default: return hipErrorInvalidValue;
    }
}

// tile_rows: 0 for the reference's planes, else the species' dense tiles
// (tile_layout).
hipError_t reactive_stats_entry(int scalar_lattice, int storage, const void* species_a, const void* species_b,
                                uint64_t nx, uint64_t ny, uint64_t nz, double k_react, double reference_a,
                                double reference_b, double* out8, void* stream, bool no_alloc,
                                uint64_t tile_rows = 0) {
    if (species_a == nullptr || species_b == nullptr || out8 == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    hipError_t e = make_shape(is_2d(scalar_lattice), nx, ny, nz, kStatThreads, 1, 1, s);
    if (e != hipSuccess) return e;
    const uint32_t cells = static_cast<uint32_t>(s.nx) * s.ny * s.nz;
    if (scalar_lattice != kD2Q5 && scalar_lattice != kD3Q7) return hipErrorInvalidValue;
    uint64_t plane = cells;
    e = tile_layout(is_2d(scalar_lattice), tile_rows, q_of(scalar_lattice), s, plane);
    if (e != hipSuccess) return e;
    const uint32_t zgap = s.cells.zgap;
    const hipStream_t st = as_stream(stream);
    if (scalar_lattice == kD2Q5)
        return stats_storage<D2Q5>(storage, species_a, species_b, cells, plane, zgap, k_react, reference_a,
                                   reference_b, out8, st, no_alloc);
    return stats_storage<D3Q7>(storage, species_a, species_b, cells, plane, zgap, k_react, reference_a, reference_b,
                               out8, st, no_alloc);
}

}  // namespace lbc

extern "C" int lbc_fluid_diagnostics(int lattice, int storage, const void* populations, uint64_t nx, uint64_t ny,
                                     uint64_t nz, double* out3, void* stream) {
    using namespace lbc;
    if (populations == nullptr || out3 == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    const hipError_t e = make_shape(is_2d(lattice), nx, ny, nz, kStatThreads, 1, 1, s);
    if (e != hipSuccess) return e;
    const uint32_t cells = static_cast<uint32_t>(s.nx) * s.ny * s.nz;
    const hipStream_t st = as_stream(stream);
    switch (lattice) {
        case kD2Q9: return fluid_diag_storage<D2Q9>(storage, populations, cells, out3, st);
        case kD3Q19: return fluid_diag_storage<D3Q19>(storage, populations, cells, out3, st);
        case kD3Q27: return fluid_diag_storage<D3Q27>(storage, populations, cells, out3, st);
        default: return hipErrorInvalidValue;
    }
}

extern "C" int lbc_reactive_stats(int scalar_lattice, int storage, const void* species_a, const void* species_b,
                                  uint64_t nx, uint64_t ny, uint64_t nz, double k_react, double reference_a,
                                  double reference_b, double* out8, void* stream) {
    return lbc::reactive_stats_entry(scalar_lattice, storage, species_a, species_b, nx, ny, nz, k_react, reference_a,
                                     reference_b, out8, stream, false);
}

extern "C" int lbc_reactive_stats_tiled(int scalar_lattice, int storage, const void* species_a,
                                        const void* species_b, uint64_t nx, uint64_t ny, uint64_t nz,
                                        uint64_t tile_rows, double k_react, double reference_a, double reference_b,
                                        double* out8, void* stream) {
    return lbc::reactive_stats_entry(scalar_lattice, storage, species_a, species_b, nx, ny, nz, k_react, reference_a,
                                     reference_b, out8, stream, false, tile_rows);
}

extern "C" int lbc_reactive_stats_noalloc(int scalar_lattice, int storage, const void* species_a,
                                          const void* species_b, uint64_t nx, uint64_t ny, uint64_t nz,
                                          double k_react, double reference_a, double reference_b, double* out8,
                                          void* stream) {
    return lbc::reactive_stats_entry(scalar_lattice, storage, species_a, species_b, nx, ny, nz, k_react, reference_a,
                                     reference_b, out8, stream, true);
}

extern "C" int lbc_max_speed(int lattice, int storage, const void* populations, uint64_t nx, uint64_t ny,
                             uint64_t nz, double* out1, void* stream) {
    using namespace lbc;
    if (populations == nullptr || out1 == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    const hipError_t e = make_shape(is_2d(lattice), nx, ny, nz, kStatThreads, 1, 1, s);
    if (e != hipSuccess) return e;
    const uint32_t cells = static_cast<uint32_t>(s.nx) * s.ny * s.nz;
    const hipStream_t st = as_stream(stream);
    switch (lattice) {
        case kD2Q9: return max_speed_storage<D2Q9>(storage, populations, cells, out1, st);
        case kD3Q19: return max_speed_storage<D3Q19>(storage, populations, cells, out1, st);
        case kD3Q27: return max_speed_storage<D3Q27>(storage, populations, cells, out1, st);
        default: return hipErrorInvalidValue;
    }
}

extern "C" int lbc_macro_fields(int lattice, int storage, const void* populations, uint64_t nx, uint64_t ny,
                                uint64_t nz, double* rho, double* u, void* stream) {
    using namespace lbc;
    if (populations == nullptr || rho == nullptr || u == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    const hipError_t e = make_shape(is_2d(lattice), nx, ny, nz, kStatThreads, 1, 1, s);
    if (e != hipSuccess) return e;
    const uint32_t cells = static_cast<uint32_t>(s.nx) * s.ny * s.nz;
    const hipStream_t st = as_stream(stream);
    switch (lattice) {
        case kD2Q9: return macro_fields_storage<D2Q9>(storage, populations, cells, rho, u, st);
        case kD3Q19: return macro_fields_storage<D3Q19>(storage, populations, cells, rho, u, st);
        case kD3Q27: return macro_fields_storage<D3Q27>(storage, populations, cells, rho, u, st);
        default: return hipErrorInvalidValue;
    }
}

extern "C" int lbc_scalar_fields(int scalar_lattice, int storage, const void* populations, uint64_t nx, uint64_t ny,
                                 uint64_t nz, double reference, double* concentration, void* stream) {
    using namespace lbc;
    if (populations == nullptr || concentration == nullptr) return hipErrorInvalidDevicePointer;
    LaunchShape s;
    const hipError_t e = make_shape(is_2d(scalar_lattice), nx, ny, nz, kStatThreads, 1, 1, s);
    if (e != hipSuccess) return e;
    const uint32_t cells = static_cast<uint32_t>(s.nx) * s.ny * s.nz;
    const hipStream_t st = as_stream(stream);
    switch (scalar_lattice) {
        case kD2Q5: return scalar_fields_storage<D2Q5>(storage, populations, cells, reference, concentration, st);
        case kD3Q7: return scalar_fields_storage<D3Q7>(storage, populations, cells, reference, concentration, st);
        default: return hipErrorInvalidValue;
    }
}

extern "C" void lbc_default_block(int lattice, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t* block3) {
    (void)lattice;
    (void)nx;
    (void)ny;
    (void)nz;
    // 256 work-items (4 waves) per block; cells are covered linearly, so the
    // shape of the block does not matter, only its size.
    block3[0] = 256;
    block3[1] = 1;
    block3[2] = 1;
}

extern "C" const char* lbc_status_string(int status) { return hipGetErrorString(static_cast<hipError_t>(status)); }

extern "C" int lbc_abi_version(void) { return 2; }
