/*
 * lbcube.h — C ABI of the MI355X (gfx950) lattice Boltzmann kernels.
 *
 * This is the drop-in boundary for LB-Cube's GPU launchers
 * (include/lattice_cuda.cuh of LB-Cube).  Every entry point:
 *   - takes device pointers to SoA population buffers laid out [Q][Z][Y][X]
 *     with x fastest (include/lattice_memory.hpp:67-272, src/lattice_cuda.cu:69-78);
 *     the *_pitched, *_strided and *_tiled extensions take the layouts they
 *     document,
 *   - validates in the reference's order (src/lattice_cuda.cu:488-503):
 *       null pointer            -> hipErrorInvalidDevicePointer (17)
 *       any zero block extent   -> hipErrorInvalidConfiguration (9)
 *       2D lattice              -> z extent and block z forced to 1
 *       any zero domain extent  -> hipErrorInvalidValue (1)
 *     then enqueues on `stream` (NULL = the default stream) and returns
 *     hipGetLastError() without synchronizing,
 *   - never allocates, copies or frees caller memory, and never swaps
 *     buffers: the caller swaps current/next after each call
 *     (src/main.cpp:587, src/simulate_reaction.cpp:527,542-543).
 * Checks the reference does not make, all reported as hipErrorInvalidValue:
 * an unsupported lattice/collision/storage code, current == next (the
 * ping-pong buffers must not alias), an extent of 2^31 or more, and a domain
 * of more than 2^32 - 2^20 cells (4,293,918,720: cell numbers are 32-bit; the
 * adjoint entry points keep at most 2^29 cells); a block size that would need
 * more than 2^31 - 1 blocks is hipErrorInvalidConfiguration.  Population planes
 * of more than 4 GiB are addressed with 64-bit per-lane addresses, smaller
 * ones with 32-bit offsets; results are identical either way.  Return values
 * are hipError_t codes as int (0 = hipSuccess).
 *
 * Streams are passed as void* (a hipStream_t) so this header needs no HIP
 * headers.  Thread safety: every entry point is stateless and re-entrant,
 * except the statistics entry points, which share a small per-device
 * workspace and must not run concurrently on the same device.
 */
#ifndef LBCUBE_H
#define LBCUBE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Lattices (include/lattice_traits.hpp:62-305 of LB-Cube). */
enum lbc_lattice { LBC_D2Q9 = 0, LBC_D3Q19 = 1, LBC_D3Q27 = 2, LBC_D2Q5 = 3, LBC_D3Q7 = 4 };
/* Collision operators (CollisionType, include/lattice_core.hpp:33-42). */
enum lbc_collision { LBC_BGK = 0, LBC_TRT = 1, LBC_MRT = 2, LBC_RLBM = 3 };
/* Population storage.  F32/F64 store populations as float/double with
 * arithmetic in the same type (the reference's Real).  F16/BF16 store
 * (f_i - w_i) for fluid populations and g_i for scalar populations as IEEE
 * half / bfloat16 bit patterns (uint16_t), rounded to nearest-even, with
 * float arithmetic. */
enum lbc_storage { LBC_F32 = 0, LBC_F64 = 1, LBC_F16 = 2, LBC_BF16 = 3 };

/* Fused pull-stream + collide step of one fluid lattice.
 * Replaces launch_step_gpu<Lattice, Real, CT> (include/lattice_cuda.cuh:179-190,
 * src/lattice_cuda.cu:476-520) and its LatticeMemory overload (:310-319,
 * src/lattice_cuda.cu:680-698).  omega is 1/tau (TRT: omega_plus, with
 * omega_minus from the magic parameter 1/4; MRT: the shear rate s_nu). */
int lbc_fluid_step(int lattice, int collision, int storage, const void* current, void* next,
                   uint64_t nx, uint64_t ny, uint64_t nz, double omega,
                   uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream);

/* Flags of lbc_fluid_step_ex. */
enum lbc_flags {
    /* BGK, TRT and MRT on every lattice (ignored by RLBM): deviation-form
     * arithmetic (f_i - w_i carried through the collision) with fused
     * multiply-adds and a hardware reciprocal of the density; D2Q9 and D3Q19
     * MRT transform only the relaxed moments of the deviations (the rest
     * state's moments are the rest equilibrium's); D3Q27 MRT as
     * f - M^-1 S M (f - feq) with separable transforms.  Up to half (D3Q27
     * MRT: a quarter) of the VALU work of the reference-order form; results
     * differ from it by rounding only (bounded by the tests). */
    LBC_FLAG_FAST = 1u,
    /* lbc_reactive_gradient with LBC_F32 storage only: the forward (and its
     * checkpoints) in FP32 as without the flag, the adjoint state and its
     * arithmetic in FP64 on that FP32 trajectory; the gradients are rounded
     * to FP32.  Ignored with LBC_F64 storage; rejected by the step entry
     * points. */
    LBC_FLAG_F64_ADJOINT = 2u,
    /* lbc_fluid_step_ex / _pitched with LBC_FLAG_FAST, MRT, D3Q19 or D3Q27
     * and LBC_F16 or LBC_BF16 storage only: the deviation-form MRT with its
     * two moment transforms (M h and M^T p') on the matrix cores.  The
     * relaxed non-equilibrium moments p' are rounded to the storage format
     * before the second product (an error relative to the update); results
     * agree with LBC_FLAG_FAST alone to that rounding (bounded by the tests).
     * hipErrorInvalidValue for any other lattice, collision, storage or
     * without LBC_FLAG_FAST; hipErrorInvalidConfiguration unless the block
     * holds a whole number of 64-wide wavefronts. */
    LBC_FLAG_MFMA = 4u
};

/* lbc_fluid_step with flags (0 = lbc_fluid_step exactly).  Unknown flag bits
 * -> hipErrorInvalidValue.  No counterpart in the reference. */
int lbc_fluid_step_ex(int lattice, int collision, int storage, uint32_t flags, const void* current, void* next,
                      uint64_t nx, uint64_t ny, uint64_t nz, double omega, uint32_t block_x, uint32_t block_y,
                      uint32_t block_z, void* stream);

/* lbc_fluid_step_ex on pitched population buffers: plane i of `current` and
 * `next` starts plane_pitch elements after plane i - 1 (plane_pitch = 0: the
 * contiguous [Q][Z][Y][X] layout of the reference, pitch = nx * ny * nz;
 * otherwise at least the cell count and at most 2^51).  Cells inside a plane
 * keep the reference's order, and elements past the cells of a plane are
 * neither read nor written.  Padding the planes moves the Q streams of a
 * launch off a common power-of-two stride.  Results equal those of the
 * contiguous layout bit for bit.  No counterpart in the reference. */
int lbc_fluid_step_pitched(int lattice, int collision, int storage, uint32_t flags, const void* current, void* next,
                           uint64_t nx, uint64_t ny, uint64_t nz, uint64_t plane_pitch, double omega,
                           uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream);

/* The pull-streaming step alone: next[i][x] = current[i][x - c_i] (periodic)
 * for every population of a fluid lattice, the step of BGK at omega = 0 with
 * the collision left out.  Same addressing, store form, launch shape and
 * residency as the collide-stream kernels, so its rate is the ceiling of
 * their memory access pattern (reported by the benchmark beside the spec
 * peak).  plane_pitch as for lbc_fluid_step_pitched.  Validation as
 * lbc_fluid_step.  No counterpart in the reference. */
int lbc_pull_stream(int lattice, int storage, const void* current, void* next, uint64_t nx, uint64_t ny, uint64_t nz,
                    uint64_t plane_pitch, uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream);

/* The same 2 Q plane streams without the pull shifts: next[i][x] =
 * current[i][x] for every population plane, with lbc_pull_stream's
 * addressing, store form, launch shape and residency.  Its rate against
 * lbc_pull_stream's separates what the shifted reads cost from what the
 * number of concurrent plane streams costs (lbc_stream_copy: two streams).
 * Arguments and validation as lbc_pull_stream.  No counterpart in the
 * reference. */
int lbc_plane_copy(int lattice, int storage, const void* current, void* next, uint64_t nx, uint64_t ny, uint64_t nz,
                   uint64_t plane_pitch, uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream);

/* Strided population layouts.  Population i of cell (x, y, z) lies at element
 *   i * plane_stride + z * slab_stride + y * nx + x
 * of a buffer.  Two families are accepted:
 *  - slab_stride 0 or nx * ny: the reference's [Q][Z][Y][X] planes, with
 *    plane_stride 0 (the cell count) or at least the cell count (pitched, as
 *    lbc_fluid_step_pitched);
 *  - a larger slab_stride: the planes stored slab by slab, with plane_stride
 *    0 (= nx * ny) or at least nx * ny and slab_stride at least Q times
 *    plane_stride; plane_stride = nx * ny and slab_stride = Q * nx * ny is the
 *    dense [Z][Q][Y][X] order, in which the Q planes of one xy-slab lie
 *    side by side.
 * Anything else, or a layout whose per-plane offsets reach 2^32 elements
 * ((nz - 1) * slab_stride + nx * ny > 2^32 - 1), -> hipErrorInvalidValue.
 * Elements outside the cells are neither read nor written. */

/* lbc_fluid_step_ex on a strided layout.  Every cell's result equals the
 * contiguous layout's bit for bit; only the placement differs.  The dense
 * [Z][Q][Y][X] order keeps a launch's 2 Q streams within Q * nx * ny elements
 * of each other (the reference's planes put them a whole plane apart), which
 * the HBM serves at a higher rate.  No counterpart in the reference: a
 * caller that holds the reference's planes converts once with lbc_relayout
 * and steps in the slab order. */
int lbc_fluid_step_strided(int lattice, int collision, int storage, uint32_t flags, const void* current, void* next,
                           uint64_t nx, uint64_t ny, uint64_t nz, uint64_t plane_stride, uint64_t slab_stride,
                           double omega, uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream);

/* lbc_pull_stream (shift != 0) or lbc_plane_copy (shift == 0) on a strided
 * layout: the access-pattern ceilings of lbc_fluid_step_strided.  No
 * counterpart in the reference. */
int lbc_pull_stream_strided(int lattice, int storage, int shift, const void* current, void* next, uint64_t nx,
                            uint64_t ny, uint64_t nz, uint64_t plane_stride, uint64_t slab_stride, uint32_t block_x,
                            uint32_t block_y, uint32_t block_z, void* stream);

/* Copies a population set from one strided layout to another (both as
 * above; element size from `storage`), e.g. the reference's planes to the
 * [Z][Q][Y][X] order and back.  src == dst -> hipErrorInvalidValue; the two
 * buffers must not overlap.  No counterpart in the reference. */
int lbc_relayout(int lattice, int storage, const void* src, uint64_t src_plane_stride, uint64_t src_slab_stride,
                 void* dst, uint64_t dst_plane_stride, uint64_t dst_slab_stride, uint64_t nx, uint64_t ny,
                 uint64_t nz, void* stream);

/* Tiles.  A population buffer in tiles of tile_rows rows (tile_rows divides
 * ny) holds, tile after tile, the Q planes of the tile's cells back to back:
 * population i of cell (x, y, z) lies at
 *   t * Q * T + i * T + ((z * ny + y) * nx + x - t * T)
 * with T = tile_rows * nx the cells of a tile and t = (z * ny + y) / tile_rows
 * the cell's tile: [nz][ny / tile_rows][Q][tile_rows][X].  In 3D tile_rows =
 * ny is the slab order [Z][Q][Y][X] (lbc_fluid_step_strided with
 * plane_stride nx * ny and slab_stride Q * nx * ny); in 2D nz = 1.
 * tile_rows = 0 selects the reference's planes.  Other values, or buffers of
 * 2^32 elements or more (Q times the cells), -> hipErrorInvalidValue.  The
 * kernels give every cell the result of the reference's planes bit for bit;
 * tiles keep the planes a step reads within a few tiles of each other, which
 * the HBM serves faster than planes a whole domain apart.  No counterpart in
 * the reference. */

/* lbc_relayout between the reference's planes (rows 0) and tiles, or between
 * two tilings, of any lattice's populations (fluid or species). */
int lbc_relayout_tiled(int lattice, int storage, const void* src, uint64_t src_tile_rows, void* dst,
                       uint64_t dst_tile_rows, uint64_t nx, uint64_t ny, uint64_t nz, void* stream);

/* lbc_fluid_step_ex on populations in tiles (fluid lattices). */
int lbc_fluid_step_tiled(int lattice, int collision, int storage, uint32_t flags, const void* current, void* next,
                         uint64_t nx, uint64_t ny, uint64_t nz, uint64_t tile_rows, double omega, uint32_t block_x,
                         uint32_t block_y, uint32_t block_z, void* stream);

/* lbc_reactive_fused_step_ex with the fluid and both species in tiles of the
 * same rows (each buffer with its own lattice's Q per tile). */
int lbc_reactive_fused_step_tiled(int fluid_lattice, int collision, int scalar_lattice, int storage, uint32_t flags,
                                  const void* current_fluid, void* next_fluid, const void* current_a, void* next_a,
                                  const void* current_b, void* next_b, uint64_t nx, uint64_t ny, uint64_t nz,
                                  uint64_t tile_rows, double omega, double omega_c, double k_react,
                                  double reference_a, double reference_b, uint32_t block_x, uint32_t block_y,
                                  uint32_t block_z, void* stream);

/* lbc_scalar_step and lbc_reaction_ab_step (the reference's separate scalar
 * and reaction launches) with the fluid and the species in tiles of the same
 * rows (each buffer with its own lattice's Q per tile); tile_rows 0 is the
 * reference's planes.  Validation as those calls, then as
 * lbc_reactive_fused_step_tiled. */
int lbc_scalar_step_tiled(int fluid_lattice, int scalar_lattice, int storage, const void* current_scalar,
                          void* next_scalar, const void* current_fluid, uint64_t nx, uint64_t ny, uint64_t nz,
                          uint64_t tile_rows, double omega_c, double source_term, double reference, uint32_t block_x,
                          uint32_t block_y, uint32_t block_z, void* stream);
int lbc_reaction_ab_step_tiled(int fluid_lattice, int scalar_lattice, int storage, const void* current_a,
                               void* next_a, const void* current_b, void* next_b, const void* current_fluid,
                               uint64_t nx, uint64_t ny, uint64_t nz, uint64_t tile_rows, double omega_c,
                               double k_react, double reference_a, double reference_b, uint32_t block_x,
                               uint32_t block_y, uint32_t block_z, void* stream);

/* The fused step's memory pattern alone (the ceiling of
 * lbc_reactive_fused_step(_tiled) on the same buffers): every fluid and
 * species population pulled from x - c_i (shift != 0) or read in place
 * (shift == 0) and stored, with the fused kernel's addressing, block order,
 * launch shape and residency, and no arithmetic.  tile_rows as for
 * lbc_reactive_fused_step_tiled (0: the reference's planes); validation as
 * that call.  No counterpart in the reference. */
int lbc_reactive_pull_stream(int fluid_lattice, int scalar_lattice, int storage, int shift,
                             const void* current_fluid, void* next_fluid, const void* current_a, void* next_a,
                             const void* current_b, void* next_b, uint64_t nx, uint64_t ny, uint64_t nz,
                             uint64_t tile_rows, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                             void* stream);

/* lbc_reactive_stats of species in tiles: the cells are summed in the same
 * order, so the results equal lbc_reactive_stats' on the reference's planes
 * bit for bit. */
int lbc_reactive_stats_tiled(int scalar_lattice, int storage, const void* species_a, const void* species_b,
                             uint64_t nx, uint64_t ny, uint64_t nz, uint64_t tile_rows, double k_react,
                             double reference_a, double reference_b, double* out8, void* stream);

/* Reference concentrations of scalar fields.  With LBC_F16 / LBC_BF16
 * storage a scalar field's populations are stored as g_i - w_i c_ref, with
 * c_ref the field's reference concentration (w_i alone is the fluid's rest
 * state, not a scalar's): concentrations near c_ref keep the full 11 (8)
 * significand bits on their deviation.  Every call that reads or writes a
 * field must pass the same c_ref for it (0 stores g_i as it is); FP32/FP64
 * storage ignores it.  The reference has no reduced-precision storage. */

/* Passive-scalar advection-diffusion step with a uniform source.
 * Replaces launch_scalar_step_gpu<F, S, Real> (include/lattice_cuda.cuh:209-222,
 * src/lattice_cuda.cu:539-589).  Pairs: D2Q9/D2Q5, D3Q19/D3Q7, D3Q27/D3Q7.
 * `reference`: the scalar field's reference concentration (above). */
int lbc_scalar_step(int fluid_lattice, int scalar_lattice, int storage, const void* current_scalar,
                    void* next_scalar, const void* current_fluid, uint64_t nx, uint64_t ny, uint64_t nz,
                    double omega_c, double source_term, double reference, uint32_t block_x, uint32_t block_y,
                    uint32_t block_z, void* stream);

/* Fused A + B -> C advection-diffusion-reaction step of two species.
 * Replaces launch_reaction_AB_gpu<F, S, Real> (include/lattice_cuda.cuh:276-291,
 * src/lattice_cuda.cu:610-666).  reference_a/b: the species' reference
 * concentrations (above). */
int lbc_reaction_ab_step(int fluid_lattice, int scalar_lattice, int storage, const void* current_a, void* next_a,
                         const void* current_b, void* next_b, const void* current_fluid, uint64_t nx, uint64_t ny,
                         uint64_t nz, double omega_c, double k_react, double reference_a, double reference_b,
                         uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream);

/* One pass that performs lbc_fluid_step followed by lbc_reaction_ab_step on
 * the freshly written fluid state (the order of src/simulate_reaction.cpp:516-546),
 * reading and writing every population once.  Results are identical to the
 * two separate calls.  No counterpart in the reference: it fuses the two
 * launches the reference's reactive drivers make per step. */
int lbc_reactive_fused_step(int fluid_lattice, int collision, int scalar_lattice, int storage,
                            const void* current_fluid, void* next_fluid, const void* current_a, void* next_a,
                            const void* current_b, void* next_b, uint64_t nx, uint64_t ny, uint64_t nz,
                            double omega, double omega_c, double k_react, double reference_a, double reference_b,
                            uint32_t block_x, uint32_t block_y, uint32_t block_z, void* stream);

/* lbc_reactive_fused_step with flags (0 = lbc_reactive_fused_step exactly).
 * LBC_FLAG_FAST (BGK, TRT and MRT; ignored by RLBM): the fluid part in the
 * deviation form of lbc_fluid_step_ex, the velocity taken from the stored
 * deviations, the scalar collision with 1/cs2 multiplied and FMAs, and the
 * reaction increment of the same exact one-step map written without the
 * cancellation of the reference's formula at A ~ B
 * (S = -A N / (1 + N), N = k A expm1(z) / z + expm1(z), z = k (B - A)).
 * Results differ from the reference order by rounding only, and near A = B
 * the fast form is the more accurate one (bounded by the tests).
 * Unknown flag bits -> hipErrorInvalidValue.  No counterpart in the reference. */
int lbc_reactive_fused_step_ex(int fluid_lattice, int collision, int scalar_lattice, int storage, uint32_t flags,
                               const void* current_fluid, void* next_fluid, const void* current_a, void* next_a,
                               const void* current_b, void* next_b, uint64_t nx, uint64_t ny, uint64_t nz,
                               double omega, double omega_c, double k_react, double reference_a,
                               double reference_b, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                               void* stream);

/* Gradient of the 2D reactive loop (D2Q9 BGK, TRT, MRT or RLBM + two D2Q5 species,
 * A + B -> C; `steps` calls of lbc_reactive_fused_step from f0, a0, b0) by the
 * discrete adjoint.  loss = <C>(T) = <A>(0) - <A>(T) (domain means).  Outputs:
 * *loss, the adjoints dloss/d(f0, a0, b0) of the initial populations (same
 * layout and type as the inputs) and dloss/d(omega_c).  States are
 * checkpointed every `interval` steps and each segment is recomputed before
 * it is walked backward; the recompute of a segment runs on a private stream
 * of the library (default priority, created and destroyed by the call) while
 * `stream` walks the following segment, ordered by events.  storage: LBC_F32
 * or LBC_F64.  flags: 0 (forward =
 * lbc_reactive_fused_step) or LBC_FLAG_FAST (forward =
 * lbc_reactive_fused_step_ex(..., LBC_FLAG_FAST, ...)); the adjoint sweep is
 * the same for both; with LBC_F32, LBC_FLAG_F64_ADJOINT carries the adjoint
 * in FP64 on the FP32 trajectory.  The caller passes a
 * device workspace of at least lbc_reactive_gradient_workspace() bytes
 * (checkpoints plus two segment buffers); no device memory is allocated.  block_x * block_y must be a multiple of 64 (whole
 * wavefronts), else hipErrorInvalidConfiguration.  Synchronizes `stream`.
 * No counterpart in the reference. */
int lbc_reactive_gradient_workspace(int storage, uint64_t nx, uint64_t ny, int steps, int interval,
                                    uint32_t block_x, uint32_t block_y, uint64_t* bytes);
int lbc_reactive_gradient(int collision, int storage, uint32_t flags, const void* f0, const void* a0, const void* b0,
                          uint64_t nx,
                          uint64_t ny, int steps, int interval, double omega, double omega_c, double k_react,
                          double* loss, void* grad_f0, void* grad_a0, void* grad_b0, double* dloss_domega_c,
                          void* workspace, uint64_t workspace_bytes, uint32_t block_x, uint32_t block_y,
                          void* stream);

/* The same for the 3D reactive loop: a D3Q19 or D3Q27 fluid (fluid_lattice
 * LBC_D3Q19 / LBC_D3Q27; BGK, TRT, MRT or RLBM, else hipErrorInvalidValue) plus two D3Q7
 * species, `steps` calls of lbc_reactive_fused_step(fluid_lattice,
 * collision, LBC_D3Q7, ...) from f0, a0, b0; loss <A>(0) - <A>(T).  The
 * workspace holds the checkpoints and two segment buffers of states of
 * Q + 14 planes.  Flags, storage, checkpoints, streams, block rule and
 * synchronisation as lbc_reactive_gradient.  No counterpart in the reference. */
int lbc_reactive_gradient_workspace_3d(int fluid_lattice, int storage, uint64_t nx, uint64_t ny, uint64_t nz,
                                       int steps, int interval, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                                       uint64_t* bytes);
int lbc_reactive_gradient_3d(int fluid_lattice, int collision, int storage, uint32_t flags, const void* f0,
                             const void* a0, const void* b0, uint64_t nx, uint64_t ny, uint64_t nz, int steps,
                             int interval, double omega, double omega_c, double k_react, double* loss,
                             void* grad_f0, void* grad_a0, void* grad_b0, double* dloss_domega_c, void* workspace,
                             uint64_t workspace_bytes, uint32_t block_x, uint32_t block_y, uint32_t block_z,
                             void* stream);

/* Ensembles of independent 2D reactive instances (D2Q9 BGK/TRT + two D2Q5
 * species, A + B -> C), each with its own extents, omega, omega_c and k, in
 * three packed buffers per time level (fluid, A, B) and advanced by one launch
 * per step.  Instance i starts at the element offsets lbc_ensemble_layout
 * reports; its Q population planes follow each other plane_strides[i]
 * elements apart (nx_i * ny_i rounded up to a multiple of 64, so that planes
 * start on whole cache lines), each holding the [ny_i][nx_i] cells; buffers
 * must be 256-byte aligned.  Each instance advances bit-identically to
 * lbc_reactive_fused_step(_ex) on it alone.  create allocates the instance
 * and block tables and a capture stream; step/run/autotune allocate nothing.
 * No counterpart in the reference. */
typedef struct lbc_ensemble lbc_ensemble;
int lbc_ensemble_create(int collision, int storage, int count, const uint32_t* nx, const uint32_t* ny,
                        const double* omega, const double* omega_c, const double* k_react, lbc_ensemble** out);
int lbc_ensemble_layout(const lbc_ensemble* ensemble, uint64_t* fluid_elems, uint64_t* species_elems,
                        uint64_t* fluid_offsets, uint64_t* species_offsets, uint64_t* plane_strides);
/* Work-items per block, threads in {128, 256, 512, 1024} (each instance's
 * cells are covered linearly by whole blocks of its own).  Waits for the device
 * first (the tables are rewritten in place). */
int lbc_ensemble_set_block(lbc_ensemble* ensemble, uint32_t threads);
/* 0 or LBC_FLAG_FAST: every instance then advances as
 * lbc_reactive_fused_step_ex(..., LBC_FLAG_FAST, ...) on it alone.  Waits for
 * the device first. */
int lbc_ensemble_set_flags(lbc_ensemble* ensemble, uint32_t flags);
/* Per-instance reference concentrations of the species A and B (arrays of
 * `count`; NULL = 0 for every instance).  Defaults to 0.  Waits for the
 * device first (the tables are rewritten in place). */
int lbc_ensemble_set_references(lbc_ensemble* ensemble, const double* reference_a, const double* reference_b);
int lbc_ensemble_step(lbc_ensemble* ensemble, const void* current_fluid, void* next_fluid, const void* current_a,
                      void* next_a, const void* current_b, void* next_b, void* stream);
/* `steps` steps between buffer sets 0 and 1 (result in set steps % 2); with
 * use_graph, even chunks of up to 64 steps are captured once into a hipGraph
 * and replayed. */
int lbc_ensemble_run(lbc_ensemble* ensemble, void* fluid0, void* fluid1, void* a0, void* a1, void* b0, void* b1,
                     int steps, int use_graph, void* stream);
/* Times the 4 block sizes over 2*trial_steps steps each, after 2 warm-up
 * steps (advancing the buffers by an even count), keeps the fastest;
 * *chosen = its work-items per block; ms[4] (optional) = ms per step of each
 * of 128, 256, 512, 1024. */
int lbc_ensemble_autotune(lbc_ensemble* ensemble, void* fluid0, void* fluid1, void* a0, void* a1, void* b0, void* b1,
                          int trial_steps, uint32_t* chosen, float* ms, void* stream);
int lbc_ensemble_destroy(lbc_ensemble* ensemble);

/* Fluid integrals of log_diagnostics (include/lattice_io.hpp:182-234):
 * out[0] = sum rho, out[1] = sum rho |u|^2 / 2, out[2] = max |u|.
 * Accumulated in double on the device in a fixed order (deterministic);
 * synchronizes `stream`.  Allocates a small per-device workspace on its first
 * call on a device; statistics calls are serialized across host threads. */
int lbc_fluid_diagnostics(int lattice, int storage, const void* populations, uint64_t nx, uint64_t ny, uint64_t nz,
                          double* out3, void* stream);

/* Reactive statistics of compute_reactive_stats (include/lattice_io.hpp:256-338):
 * out = {mean_A, mean_B, var_A, var_B, covariance, segregation_intensity,
 *        true_reaction_rate, mixed_reaction_rate}.
 * Accumulated in double on the device in a fixed order; synchronizes `stream`;
 * shares the workspace and serialization of lbc_fluid_diagnostics. */
int lbc_reactive_stats(int scalar_lattice, int storage, const void* species_a, const void* species_b, uint64_t nx,
                       uint64_t ny, uint64_t nz, double k_react, double reference_a, double reference_b,
                       double* out8, void* stream);

/* lbc_reactive_stats without any device allocation: same arguments, kernels,
 * validation and results (bit for bit), but the per-block partial sums live
 * in a static device array of the library (65,600 bytes per device, the same
 * for every domain size; the domain limits are lbc_reactive_stats') instead
 * of a workspace allocated on the first call.  It calls no hipMalloc or
 * hipFree and allocates nothing else on the device.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code in a tight
 * serving or sampling loop, harnesses that forbid allocation after set-up).
 * All calls on one device share the array; calls from several host threads
 * of one process are serialized like lbc_reactive_stats, while calls from
 * several processes on one device that overlap race on it. */
int lbc_reactive_stats_noalloc(int scalar_lattice, int storage, const void* species_a, const void* species_b,
                               uint64_t nx, uint64_t ny, uint64_t nz, double k_react, double reference_a,
                               double reference_b, double* out8, void* stream);

/* Stability guard of compute_max_velocity (src/turbulent_tgv_3d.cpp:149-173):
 * out[0] = max |u| over the cells, +infinity if any cell's speed is not
 * finite.  Synchronizes `stream`; shares the workspace and serialization of
 * lbc_fluid_diagnostics. */
int lbc_max_speed(int lattice, int storage, const void* populations, uint64_t nx, uint64_t ny, uint64_t nz,
                  double* out1, void* stream);

/* Macroscopic fields per cell, the values the VTK writers print
 * (write_vtk, include/lattice_io.hpp:386-451; write_reaction_vtk,
 * src/simulate_reaction.cpp:267-333): rho[cell] and u[d * cells + cell]
 * (d < D) from compute_macro_state (include/lattice_physics.hpp:227-249) at
 * the storage's arithmetic precision, stored as double in device memory, in
 * the cells' [Z][Y][X] order.  Copying these back instead of the populations
 * moves (1 + D) doubles per cell instead of Q values.  Asynchronous on
 * `stream`; allocates nothing. */
int lbc_macro_fields(int lattice, int storage, const void* populations, uint64_t nx, uint64_t ny, uint64_t nz,
                     double* rho, double* u, void* stream);

/* Concentration per cell of a scalar population buffer (compute_concentration,
 * include/lattice_physics.hpp:618-627), as double in device memory; `reference`
 * is the field's reference concentration.  Asynchronous on `stream`; allocates
 * nothing. */
int lbc_scalar_fields(int scalar_lattice, int storage, const void* populations, uint64_t nx, uint64_t ny, uint64_t nz,
                      double reference, double* concentration, void* stream);

/* Streaming device copy of `bytes` (a multiple of 16) from src to dst: the
 * measured HBM ceiling the benchmark reports next to the spec peak.  No
 * counterpart in the reference. */
int lbc_stream_copy(const void* src, void* dst, uint64_t bytes, void* stream);

/* Block shape the library picks for a domain: {256, 1, 1}.  Every launcher
 * covers the cells linearly (x fastest), so only bx * by * bz (the work-items
 * per block) affects the launch, never the results.  The reference's default
 * is {16, 8, 1} (include/lattice_cuda.cuh:52-54). */
void lbc_default_block(int lattice, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t* block3);

/* Human-readable name of a returned status code. */
const char* lbc_status_string(int status);

/* ABI version: bumped when an entry point changes. */
int lbc_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LBCUBE_H */
