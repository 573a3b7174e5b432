// Batched ensembles of independent 2D reactive-mixing instances (D2Q9 BGK or
// TRT + two D2Q5 species, A + B -> C), each with its own extents and
// parameters, advanced by ONE launch per step over a packed index space.
// No counterpart in the reference (its drivers run one domain per process).
//
// Layout: three packed buffers per time level (fluid, species A, species B);
// instance i owns Q planes of `plane_i` elements (nx_i * ny_i cells rounded up
// to 64) of each, starting at element offset off_i.
// Work decomposition: the cells of instance i are numbered linearly (x
// fastest) and cut into ceil(cells_i / threads) blocks of `threads`
// work-items; a per-block table gives each block its instance (one uniform
// load), so instances of every size share one grid and only the last block of
// each instance has idle lanes.  Every cell runs the same code as
// lbc_reactive_fused_step(_ex), so an instance advances bit-identically to a
// standalone run.  The block size is chosen by timing (lbc_ensemble_autotune),
// and long runs replay a captured hipGraph of consecutive steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <new>
#include <vector>

#include "../../include/lbcube.h"
#include "lbc_launch.hpp"
#include "lbc_reactive.hpp"

namespace lbc {

struct EnsembleInstance {
    uint64_t f_off, s_off;  // element offsets of this instance in the fluid / species buffers
    uint32_t nx, ny, cells, block_start;
    uint32_t plane;  // elements between population planes: cells rounded up to kPlaneAlign
    FastDiv by_nx;
    float omega, omega_minus, omega_c, k;
    double omega_d, omega_minus_d, omega_c_d, k_d;
    // reference concentrations of the species storage (F16 / BF16 hold g - w c_ref)
    float ref_a, ref_b;
    double ref_a_d, ref_b_d;
};

constexpr uint32_t kEnsembleThreads[4] = {128, 256, 512, 1024};

// Population planes start on multiples of 64 cells (128 B at FP16, 512 B at
// FP64), so every wave's 64 unshifted cells fill whole cache lines: with
// planes packed back to back an instance of arbitrary size leaves every store
// of a wave straddling two lines, and the ensemble ran at two thirds of the
// bandwidth of equal, aligned instances.
constexpr uint32_t kPlaneAlign = 64;

template <int CT, int ST, bool FAST, int THREADS>
__global__ __launch_bounds__(THREADS) void ensemble_kernel(const typename Storage<ST>::S* __restrict__ cf,
                                                           typename Storage<ST>::S* __restrict__ nf,
                                                           const typename Storage<ST>::S* __restrict__ ca,
                                                           typename Storage<ST>::S* __restrict__ na,
                                                           const typename Storage<ST>::S* __restrict__ cb,
                                                           typename Storage<ST>::S* __restrict__ nb,
                                                           const EnsembleInstance* __restrict__ table,
                                                           const uint32_t* __restrict__ block_instance) {
    using R = typename Storage<ST>::R;
    const EnsembleInstance& in = table[block_instance[blockIdx.x]];
    const uint32_t p = (blockIdx.x - in.block_start) * THREADS + threadIdx.x;
    if (p >= in.cells) return;
    const uint32_t y = fast_div(p, in.by_nx);
    const int x = static_cast<int>(p - y * in.nx);
    const R omega = std::is_same<R, double>::value ? R(in.omega_d) : R(in.omega);
    const R wm = std::is_same<R, double>::value ? R(in.omega_minus_d) : R(in.omega_minus);
    const R wc = std::is_same<R, double>::value ? R(in.omega_c_d) : R(in.omega_c);
    const R k = std::is_same<R, double>::value ? R(in.k_d) : R(in.k);
    const R ra = std::is_same<R, double>::value ? R(in.ref_a_d) : R(in.ref_a);
    const R rb = std::is_same<R, double>::value ? R(in.ref_b_d) : R(in.ref_b);
    const Pull pull(p, x, static_cast<int>(y), 0, static_cast<int>(in.nx), static_cast<int>(in.ny), 1);
    if constexpr (FAST)
        reactive_cell_fast<D2Q9, CT, D2Q5, ST>(cf + in.f_off, nf + in.f_off, ca + in.s_off, na + in.s_off,
                                               cb + in.s_off, nb + in.s_off, pull, pull, in.plane, omega, wm, wc, k,
                                               ra, rb);
    else
        reactive_cell<D2Q9, CT, D2Q5, ST>(cf + in.f_off, nf + in.f_off, ca + in.s_off, na + in.s_off, cb + in.s_off,
                                          nb + in.s_off, pull, pull, in.plane, omega, wm, wc, k, ra, rb);
}

}  // namespace lbc

struct lbc_ensemble {
    int collision, storage, count;
    std::vector<lbc::EnsembleInstance> host;
    lbc::EnsembleInstance* table = nullptr;
    uint32_t* block_instance = nullptr;  // sized for the smallest block (the most blocks)
    uint64_t f_elems = 0, s_elems = 0, blocks = 0;
    uint32_t threads = 256;
    uint32_t flags = 0;
    hipStream_t capture = nullptr;
    hipGraphExec_t graph = nullptr;
    const void* graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int graph_steps = 0;
};

namespace lbc {

static uint64_t blocks_for(const std::vector<EnsembleInstance>& host, uint32_t threads) {
    uint64_t blocks = 0;
    for (const auto& in : host) blocks += (in.cells + threads - 1) / threads;
    return blocks;
}

// Lays out the blocks for `threads` work-items per block and uploads the
// instance and block tables.
static hipError_t build_table(lbc_ensemble* e, uint32_t threads) {
    // The tables are rewritten in place: no launch that reads them may be in flight.
    hipError_t err = hipDeviceSynchronize();
    if (err != hipSuccess) return err;
    std::vector<uint32_t> owner;
    owner.reserve(blocks_for(e->host, threads));
    for (uint32_t i = 0; i < e->host.size(); ++i) {
        EnsembleInstance& in = e->host[i];
        in.block_start = static_cast<uint32_t>(owner.size());
        owner.insert(owner.end(), (in.cells + threads - 1) / threads, i);
    }
    e->blocks = owner.size();
    e->threads = threads;
    err = hipMemcpy(e->table, e->host.data(), e->host.size() * sizeof(EnsembleInstance), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(e->block_instance, owner.data(), owner.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    return err;
}

template <int CT, int ST, bool FAST>
static void launch_ensemble(const lbc_ensemble* e, const void* cf, void* nf, const void* ca, void* na, const void* cb,
                            void* nb, hipStream_t st) {
    using T = typename Storage<ST>::S;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(e->blocks)), dim3(e->threads), 0, st,
                           static_cast<const T*>(cf), static_cast<T*>(nf), static_cast<const T*>(ca),
                           static_cast<T*>(na), static_cast<const T*>(cb), static_cast<T*>(nb), e->table,
                           e->block_instance);
    };
    switch (e->threads) {
        case 128: go(ensemble_kernel<CT, ST, FAST, 128>); break;
        case 256: go(ensemble_kernel<CT, ST, FAST, 256>); break;
        case 512: go(ensemble_kernel<CT, ST, FAST, 512>); break;
        default: go(ensemble_kernel<CT, ST, FAST, 1024>); break;
    }
}

template <int CT, bool FAST>
static hipError_t step_ct(const lbc_ensemble* e, const void* cf, void* nf, const void* ca, void* na, const void* cb,
                          void* nb, hipStream_t st) {
    switch (e->storage) {
        case kF32: launch_ensemble<CT, kF32, FAST>(e, cf, nf, ca, na, cb, nb, st); break;
        case kF64: launch_ensemble<CT, kF64, FAST>(e, cf, nf, ca, na, cb, nb, st); break;
        case kF16: launch_ensemble<CT, kF16, FAST>(e, cf, nf, ca, na, cb, nb, st); break;
        case kBF16: launch_ensemble<CT, kBF16, FAST>(e, cf, nf, ca, na, cb, nb, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t ensemble_step(const lbc_ensemble* e, const void* cf, void* nf, const void* ca, void* na,
                                const void* cb, void* nb, hipStream_t st) {
    const bool fast = (e->flags & LBC_FLAG_FAST) != 0;
    if (e->collision == kBGK)
        return fast ? step_ct<kBGK, true>(e, cf, nf, ca, na, cb, nb, st)
                    : step_ct<kBGK, false>(e, cf, nf, ca, na, cb, nb, st);
    return fast ? step_ct<kTRT, true>(e, cf, nf, ca, na, cb, nb, st)
                : step_ct<kTRT, false>(e, cf, nf, ca, na, cb, nb, st);
}

}  // namespace lbc

extern "C" int lbc_ensemble_create(int collision, int storage, int count, const uint32_t* nx, const uint32_t* ny,
                                   const double* omega, const double* omega_c, const double* k_react,
                                   lbc_ensemble** out) {
    using namespace lbc;
    if (out == nullptr || nx == nullptr || ny == nullptr || omega == nullptr || omega_c == nullptr ||
        k_react == nullptr)
        return hipErrorInvalidValue;
    *out = nullptr;
    if (count <= 0 || (collision != kBGK && collision != kTRT)) return hipErrorInvalidValue;
    if (storage < kF32 || storage > kBF16) return hipErrorInvalidValue;
    auto* e = new (std::nothrow) lbc_ensemble();
    if (e == nullptr) return hipErrorOutOfMemory;
    e->collision = collision;
    e->storage = storage;
    e->count = count;
    e->host.resize(count);
    uint64_t fo = 0, so = 0;
    for (int i = 0; i < count; ++i) {
        // an instance is one plane of at most 2^29 cells (32-bit offsets, exact division)
        if (nx[i] == 0 || ny[i] == 0 || static_cast<uint64_t>(nx[i]) * ny[i] > (1ull << 29)) {
            delete e;
            return hipErrorInvalidValue;
        }
        EnsembleInstance& in = e->host[i];
        in.nx = nx[i];
        in.ny = ny[i];
        in.cells = nx[i] * ny[i];
        in.plane = (in.cells + kPlaneAlign - 1) / kPlaneAlign * kPlaneAlign;
        in.by_nx = make_fastdiv(nx[i]);
        in.f_off = fo;
        in.s_off = so;
        fo += 9ull * in.plane;
        so += 5ull * in.plane;
        in.omega_d = omega[i];
        in.omega_c_d = omega_c[i];
        in.k_d = k_react[i];
        in.omega_minus_d = collision == kTRT ? trt_omega_minus<double>(omega[i]) : 0.0;
        in.omega = static_cast<float>(omega[i]);
        in.omega_c = static_cast<float>(omega_c[i]);
        in.k = static_cast<float>(k_react[i]);
        in.omega_minus = collision == kTRT ? trt_omega_minus<float>(static_cast<float>(omega[i])) : 0.0f;
    }
    e->f_elems = fo;
    e->s_elems = so;
    const uint64_t max_blocks = blocks_for(e->host, kEnsembleThreads[0]);
    hipError_t err = max_blocks > 0x7fffffffull ? hipErrorInvalidValue : hipSuccess;
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&e->table), count * sizeof(EnsembleInstance));
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&e->block_instance), max_blocks * sizeof(uint32_t));
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&e->capture, hipStreamNonBlocking);
    if (err == hipSuccess) err = build_table(e, 256);
    if (err != hipSuccess) {
        if (e->table) (void)hipFree(e->table);
        if (e->block_instance) (void)hipFree(e->block_instance);
        if (e->capture) (void)hipStreamDestroy(e->capture);
        delete e;
        return err;
    }
    *out = e;
    return hipSuccess;
}

extern "C" int lbc_ensemble_layout(const lbc_ensemble* e, uint64_t* fluid_elems, uint64_t* species_elems,
                                   uint64_t* fluid_offsets, uint64_t* species_offsets, uint64_t* plane_strides) {
    if (e == nullptr) return hipErrorInvalidValue;
    if (fluid_elems) *fluid_elems = e->f_elems;
    if (species_elems) *species_elems = e->s_elems;
    for (int i = 0; i < e->count; ++i) {
        if (fluid_offsets) fluid_offsets[i] = e->host[i].f_off;
        if (species_offsets) species_offsets[i] = e->host[i].s_off;
        if (plane_strides) plane_strides[i] = e->host[i].plane;
    }
    return hipSuccess;
}

extern "C" int lbc_ensemble_set_block(lbc_ensemble* e, uint32_t threads) {
    if (e == nullptr) return hipErrorInvalidValue;
    if (std::find(std::begin(lbc::kEnsembleThreads), std::end(lbc::kEnsembleThreads), threads) ==
        std::end(lbc::kEnsembleThreads))
        return hipErrorInvalidConfiguration;
    return lbc::build_table(e, threads);
}

extern "C" int lbc_ensemble_set_flags(lbc_ensemble* e, uint32_t flags) {
    if (e == nullptr) return hipErrorInvalidValue;
    if (flags & ~static_cast<uint32_t>(LBC_FLAG_FAST)) return hipErrorInvalidValue;
    const hipError_t err = hipDeviceSynchronize();  // a captured graph may still be replaying
    if (err != hipSuccess) return err;
    if (e->graph) {
        (void)hipGraphExecDestroy(e->graph);
        e->graph = nullptr;
    }
    e->flags = flags;
    return hipSuccess;
}

extern "C" int lbc_ensemble_set_references(lbc_ensemble* e, const double* reference_a, const double* reference_b) {
    using namespace lbc;
    if (e == nullptr) return hipErrorInvalidValue;
    for (int i = 0; i < e->count; ++i) {
        EnsembleInstance& in = e->host[i];
        in.ref_a_d = reference_a ? reference_a[i] : 0.0;
        in.ref_b_d = reference_b ? reference_b[i] : 0.0;
        in.ref_a = static_cast<float>(in.ref_a_d);
        in.ref_b = static_cast<float>(in.ref_b_d);
    }
    return build_table(e, e->threads);  // waits for the device, rewrites the tables, drops a captured graph
}

extern "C" int lbc_ensemble_step(lbc_ensemble* e, const void* cur_f, void* next_f, const void* cur_a, void* next_a,
                                 const void* cur_b, void* next_b, void* stream) {
    if (e == nullptr) return hipErrorInvalidValue;
    if (!cur_f || !next_f || !cur_a || !next_a || !cur_b || !next_b) return hipErrorInvalidDevicePointer;
    if (cur_f == next_f || cur_a == next_a || cur_b == next_b) return hipErrorInvalidValue;
    return lbc::ensemble_step(e, cur_f, next_f, cur_a, next_a, cur_b, next_b, lbc::as_stream(stream));
}

// Advances `steps` steps ping-ponging between buffer sets 0 and 1; the result
// is in set (steps % 2).  With use_graph, chunks of up to 64 steps (an even
// number) are captured once into a hipGraph on a private stream and replayed;
// the caller's stream is ordered before and after through events.
extern "C" int lbc_ensemble_run(lbc_ensemble* e, void* f0, void* f1, void* a0, void* a1, void* b0, void* b1,
                                int steps, int use_graph, void* stream) {
    using namespace lbc;
    if (e == nullptr || steps < 0) return hipErrorInvalidValue;
    if (!f0 || !f1 || !a0 || !a1 || !b0 || !b1) return hipErrorInvalidDevicePointer;
    if (f0 == f1 || a0 == a1 || b0 == b1) return hipErrorInvalidValue;
    const hipStream_t st = as_stream(stream);
    void* F[2] = {f0, f1};
    void* A[2] = {a0, a1};
    void* B[2] = {b0, b1};
    int done = 0;
    hipError_t err = hipSuccess;
    if (use_graph && steps >= 2) {
        const int chunk = std::min(64, steps - steps % 2);
        const void* key[6] = {f0, f1, a0, a1, b0, b1};
        bool same = e->graph != nullptr && e->graph_steps == chunk;
        for (int i = 0; i < 6 && same; ++i) same = e->graph_key[i] == key[i];
        if (!same) {
            if (e->graph) {
                (void)hipGraphExecDestroy(e->graph);
                e->graph = nullptr;
            }
            hipGraph_t g;
            err = hipStreamBeginCapture(e->capture, hipStreamCaptureModeThreadLocal);
            if (err != hipSuccess) return err;
            for (int s = 0; s < chunk && err == hipSuccess; ++s)
                err = ensemble_step(e, F[s & 1], F[(s + 1) & 1], A[s & 1], A[(s + 1) & 1], B[s & 1], B[(s + 1) & 1],
                                    e->capture);
            const hipError_t end = hipStreamEndCapture(e->capture, &g);
            if (end == hipSuccess && err != hipSuccess) (void)hipGraphDestroy(g);
            if (err != hipSuccess) return err;
            if (end != hipSuccess) return end;
            err = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (err != hipSuccess) return err;
            e->graph_steps = chunk;
            for (int i = 0; i < 6; ++i) e->graph_key[i] = key[i];
        }
        hipEvent_t before, after;
        if ((err = hipEventCreateWithFlags(&before, hipEventDisableTiming)) != hipSuccess) return err;
        if ((err = hipEventCreateWithFlags(&after, hipEventDisableTiming)) != hipSuccess) {
            (void)hipEventDestroy(before);
            return err;
        }
        err = hipEventRecord(before, st);
        if (err == hipSuccess) err = hipStreamWaitEvent(e->capture, before, 0);
        while (err == hipSuccess && steps - done >= chunk) {
            err = hipGraphLaunch(e->graph, e->capture);
            done += chunk;
        }
        if (err == hipSuccess) err = hipEventRecord(after, e->capture);
        if (err == hipSuccess) err = hipStreamWaitEvent(st, after, 0);
        (void)hipEventDestroy(before);
        (void)hipEventDestroy(after);
        if (err != hipSuccess) return err;
    }
    for (; done < steps && err == hipSuccess; ++done)
        err = ensemble_step(e, F[done & 1], F[(done + 1) & 1], A[done & 1], A[(done + 1) & 1], B[done & 1],
                            B[(done + 1) & 1], st);
    return err;
}

// Times every block size in {128, 256, 512, 1024} for 2 * trial_steps steps
// (after 2 warm-up steps) on the given buffers (their contents advance by an
// even number of steps) and keeps the fastest.  Writes the chosen block size
// and, if ms is non-null, the 4 measured times in ms per step.
extern "C" int lbc_ensemble_autotune(lbc_ensemble* e, void* f0, void* f1, void* a0, void* a1, void* b0, void* b1,
                                     int trial_steps, uint32_t* chosen, float* ms, void* stream) {
    using namespace lbc;
    if (e == nullptr || trial_steps <= 0) return hipErrorInvalidValue;
    const hipStream_t st = as_stream(stream);
    hipEvent_t t0, t1;
    hipError_t err;
    if ((err = hipEventCreate(&t0)) != hipSuccess) return err;
    if ((err = hipEventCreate(&t1)) != hipSuccess) {
        (void)hipEventDestroy(t0);
        return err;
    }
    float best = 1e30f;
    uint32_t bt = 256;
    for (int idx = 0; idx < 4 && err == hipSuccess; ++idx) {
        const uint32_t t = kEnsembleThreads[idx];
        float m = -1.0f;
        err = build_table(e, t);
        if (err == hipSuccess) err = static_cast<hipError_t>(lbc_ensemble_run(e, f0, f1, a0, a1, b0, b1, 2, 0, stream));
        if (err == hipSuccess) err = hipEventRecord(t0, st);
        if (err == hipSuccess)
            err = static_cast<hipError_t>(lbc_ensemble_run(e, f0, f1, a0, a1, b0, b1, 2 * trial_steps, 0, stream));
        if (err == hipSuccess) err = hipEventRecord(t1, st);
        if (err == hipSuccess) err = hipEventSynchronize(t1);
        if (err == hipSuccess) err = hipEventElapsedTime(&m, t0, t1);
        if (err != hipSuccess) break;
        m /= 2.0f * trial_steps;
        if (m < best) {
            best = m;
            bt = t;
        }
        if (ms) ms[idx] = m;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (err != hipSuccess) return err;
    if (chosen) *chosen = bt;
    return build_table(e, bt);
}

extern "C" int lbc_ensemble_destroy(lbc_ensemble* e) {
    if (e == nullptr) return hipSuccess;
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    if (e->capture) (void)hipStreamDestroy(e->capture);
    if (e->table) (void)hipFree(e->table);
    if (e->block_instance) (void)hipFree(e->block_instance);
    delete e;
    return hipSuccess;
}
