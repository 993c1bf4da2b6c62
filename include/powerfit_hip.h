/*
 * powerfit_hip.h — C-ABI of libpowerfit_hip.so, the MI355X (gfx950) engine for
 * PowerFit's exhaustive 6-D local-cross-correlation (LCC) search.
 *
 * Conventions
 *   - Every function returns int: 0 = ok, < 0 = error (PF_ERR_*); the message
 *     of the last error on the calling thread is pf_last_error().
 *   - Arrays are C-order (z, y, x), x fastest, matching numpy.
 *   - Host buffers are caller-owned and only read/written during the call.
 *   - *_dev functions take device pointers and a hipStream_t (as void*,
 *     NULL = default stream); they enqueue and return without syncing.
 *   - A pf_ctx owns its device memory and stream; one host thread at a time
 *     may use a ctx; several ctxs may share a device.
 */
#ifndef POWERFIT_HIP_H
#define POWERFIT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_ABI_VERSION 2

#define PF_OK 0
#define PF_ERR_ARG (-1)
#define PF_ERR_HIP (-2)
#define PF_ERR_STATE (-3)
#define PF_ERR_MEMORY (-4)
#define PF_ERR_UNSUPPORTED (-5)

/* Scan modes for pf_scan. */
#define PF_MODE_PLAIN 0   /* one rotation per launch chain, unfused kernels        */
#define PF_MODE_BATCHED 1 /* `batch` rotations per launch, same unfused kernels     */
#define PF_MODE_FUSED 2   /* `batch` rotations per launch through the three fused
                             kernels (cubic grids with a compiled size only)        */
#define PF_MODE_AUTO 3    /* PF_MODE_FUSED where available, else PF_MODE_BATCHED   */
/* Stage-mixing modes for testing: one fused stage, the rest unfused. */
#define PF_MODE_MIX_FWD 4  /* fused rotate + forward xy                             */
#define PF_MODE_MIX_Z 5    /* fused z correlate                                     */
#define PF_MODE_MIX_INV 6  /* fused inverse xy + LCC                                */
/* The three fused stages with the grid shape taken at run time (FFT passes
 * compiled per length): grids without a compiled fused size
 * (pf_ctx_fused_available returns 2) whose ny x (nx/2+1) complex plane (for odd
 * nx, its rows packed in pairs) fits LDS twice, ny nx <= 16384.  PF_MODE_AUTO
 * takes it where the compiled pipeline does not exist. */
#define PF_MODE_FUSED_RT 7

/* pf_ctx_create flags */
#define PF_FLAG_NO_BINARY_MASK 1u /* never use the W^2 = W shortcut for 0/1 masks  */
/* Split-precision FFT: the first radix-16 stage runs as a DFT-as-GEMM on the
 * matrix cores (FP16 hi + lo operands, FP32 accumulation) -- in the y and z
 * passes of the unfused pipeline (needs ny, nz divisible by 16 and <= 320),
 * and in the forward z transforms of the fused pipeline at the compiled sizes
 * whose z plan starts with radix 16 (128^3).  With this flag AUTO runs the
 * fused pipeline there and the batched kernels elsewhere; the run-time planned
 * pipeline is not used.  No counterpart in the reference: its precision
 * option. */
#define PF_FLAG_SPLIT_DFT 2u

int pf_abi_version(void);
const char *pf_last_error(void);
int pf_device_count(int *count);

/* ---- Kernel-level drop-ins ------------------------------------------------ */

/* Replaces _extensions.rotate_grid3d(grid, rotmat, radius, out, nearest)
 * (src/powerfit_em/_extensions.c:7-196, bound by powerfitter.py:339-347).
 * Host FP64 buffers; rotates about voxel (0,0,0) inside |v| <= radius,
 * trilinear (nearest = 0) or nearest (nearest != 0, C round), writing only
 * in-sphere voxels of out.  Requires 0 <= radius <= min(all dims) - 2. */
int pf_rotate_grid3d(const double *grid, int gz, int gy, int gx, const double *rotmat,
                     int radius, double *out, int oz, int oy, int ox, int nearest);

/* Replaces _powerfit.conj_multiply(in1, in2, out) (src/powerfit_em/_powerfit.pyx:45-53;
 * OpenCL conj_multiply powerfitter.py:568-571): out = conj(in1) * in2, complex64. */
int pf_conj_multiply_dev(const void *in1, const void *in2, void *out, int64_t n,
                         void *stream);

/* Replaces _powerfit.calc_lcc(gcc, ave, ave2, mask, lcc) (src/powerfit_em/_powerfit.pyx:56-72):
 * lcc[i] = gcc[i] / sqrt(ave2[i] - ave[i]^2) where mask[i] != 0 (float32, the
 * variance formed in FP64); other entries untouched. */
int pf_calc_lcc_dev(const float *gcc, const float *ave, const float *ave2,
                    const uint8_t *mask, float *lcc, int64_t n, void *stream);

/* Replaces numpy.fft.rfftn / irfftn(s=shape) as used by CPUCorrelator
 * (powerfitter.py:306-315, 375-393) and clFFT's grfftn_builder (:605-638):
 * `batch` volumes (nz, ny, nx) float32 <-> (nz, ny, nx/2+1) complex64.
 * pf_irfftn_dev is normalised like numpy (1/V) and overwrites its input. */
int pf_rfftn_dev(const float *in, void *out, int nz, int ny, int nx, int batch, void *stream);
int pf_irfftn_dev(void *in, float *out, int nz, int ny, int nx, int batch, void *stream);

/* ---- Correlator context (BaseCorrelator / CPUCorrelator / GPUCorrelator
 *      protocol, powerfitter.py:166-556) ------------------------------------ */
typedef struct pf_ctx pf_ctx;

/* Allocates the per-target state for a (nz, ny, nx) grid on `device`.
 * flags: PF_FLAG_* bits (0 for defaults). */
int pf_ctx_create(int device, int nz, int ny, int nx, uint32_t flags, pf_ctx **out);
int pf_ctx_destroy(pf_ctx *ctx);
/* Use an external hipStream_t (e.g. torch.cuda.Stream.cuda_stream); NULL
 * restores the ctx's own stream. */
int pf_ctx_set_stream(pf_ctx *ctx, void *stream);

/* Target: the already normalised and (optionally) Laplace-filtered map Tf
 * (BaseCorrelator.__init__ :169-176, CPUCorrelator.__init__ :268-277) and the
 * lcc mask T > 0.05 max(T) (:178-180).  Computes F[Tf], F[Tf^2] on device and
 * keeps them resident. */
int pf_set_target(pf_ctx *ctx, const double *target_filtered, const uint8_t *lcc_mask);

/* Template and mask as the mask setter leaves them (:190-220): template
 * filtered, masked, normalised and masked again; mask possibly core-weighted;
 * norm_factor N = count(mask != 0); rmax = min(shape) // 2 (:176). */
int pf_set_template_mask(pf_ctx *ctx, const double *template_prepared, const double *mask,
                         double norm_factor, int rmax);

/* Rotation set (rotations setter :226-230): K row-major 3x3 FP64 matrices. */
int pf_set_rotations(pf_ctx *ctx, const double *rotmats, int64_t count);

/* Best LCC := 0, rotation index := 0 (CPUCorrelator.scan :320-321). */
int pf_reset_result(pf_ctx *ctx);

/* Scan rotations [first, first + count) into the running best/argmax
 * (CPUCorrelator.scan :323-333; GPUCorrelator.scan :513-538): strict '>',
 * ties keep the lowest index, global rotation indices.  `batch` rotations per
 * launch (at most count; 256 measured best on MI355X at every fused size).
 * Enqueues only. */
int pf_scan(pf_ctx *ctx, int64_t first, int64_t count, int batch, int mode);

/* pf_scan recorded once into a HIP graph (per ctx, for the last
 * (first, count, batch, mode) and tuning used) and replayed: same launches,
 * one graph launch.  Needs the work buffers sized beforehand (done here,
 * outside the capture) and kernel timing off.  pf_set_rotations,
 * pf_set_template_mask and a work-buffer reallocation discard the captured
 * graph (its launches name those buffers and values); the next call captures
 * again.  Enqueues only. */
int pf_scan_graph(pf_ctx *ctx, int64_t first, int64_t count, int batch, int mode);

/* pf_scan that allocates nothing: the per-batch work buffers live in a static
 * device array of PF_SCRATCH_BYTES per GPU (each GPU the library runs on
 * holds it until the process ends) instead of ctx allocations, so the call
 * makes no hipMalloc/hipFree and records no timing events.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, modes and results as pf_scan; rejects with PF_ERR_ARG a batch
 * whose work buffers exceed the scratch: 24 H bytes per rotation fused, plus
 * 12 V bytes with an unfused stage (at 128^3 fused: up to 5 rotations per
 * launch).  Every call on one GPU shares that GPU's scratch, whatever its ctx
 * or stream, and the scratch keeps the last call's contents (no call reads a
 * region it has not written): calls that overlap on one GPU race, so callers
 * serialise them (one stream, or wait for the previous call).  Enqueues only. */
#define PF_SCRATCH_BYTES ((size_t)128 << 20)
int pf_scan_noalloc(pf_ctx *ctx, int64_t first, int64_t count, int batch, int mode);

/* Copy out best LCC (float32) and rotation index (int32), shape (nz, ny, nx)
 * — GPUCorrelator's lcc/rot (:536-537).  Synchronises the ctx stream. */
int pf_get_result(pf_ctx *ctx, float *best_lcc, int32_t *best_rot);
int pf_sync(pf_ctx *ctx);
/* Non-blocking: 1 when all work enqueued on the ctx stream has finished,
 * 0 when some is still running (hipStreamQuery). */
int pf_ctx_query(pf_ctx *ctx);
/* Peak device bytes of a ctx for an (nz, ny, nx) grid created with `flags`,
 * holding `nrot` rotations and scanning with `batch` rotations per launch in
 * `mode` (resident inputs and keys, plus the larger of the scan's work
 * buffers and pf_set_target's transient buffers), for admission control.
 * No counterpart in the reference (its GPU path allocates per correlator,
 * powerfitter.py:432-443, without a budget). */
int64_t pf_ctx_bytes(int nz, int ny, int nx, int batch, int mode, uint32_t flags, int64_t nrot);

/* Progress marks on the ctx stream, for callers that keep several batches in
 * flight (the serving loop): pf_ctx_mark records a point after everything
 * enqueued so far and returns its ticket (increasing from 1); pf_ctx_reached
 * returns 1 once the work before that mark has finished, else 0 (it does not
 * block).  No counterpart in the reference, whose GPU loop blocks on
 * queue.finish() per rotation (powerfitter.py:513-538). */
int pf_ctx_mark(pf_ctx *ctx, uint64_t *ticket);
int pf_ctx_reached(pf_ctx *ctx, uint64_t ticket);

/* Kernel timing of subsequent pf_scan calls with HIP events on the ctx
 * stream (0 = off).  pf_ctx_kernel_time returns, for kernel class `which`
 * (PF_K_*), the summed milliseconds and launch count since enabling;
 * it synchronises the stream. */
#define PF_K_ROTATE 0
#define PF_K_R2C_X 1
#define PF_K_C2C_FWD 2
#define PF_K_CONJ_MUL 3
#define PF_K_C2C_INV 4
#define PF_K_C2R_X 5
#define PF_K_LCC 6
#define PF_K_FWD_PLANE 7
#define PF_K_ZCORR 8
#define PF_K_INV_PLANE 9
#define PF_K_COUNT 10
/* 1 when PF_MODE_FUSED (compiled kernels) is available for this ctx's grid,
 * 2 when only PF_MODE_FUSED_RT (run-time plans) is, 0 when neither is. */
int pf_ctx_fused_available(pf_ctx *ctx);
/* Rotations per block in the fused z-correlation and LCC kernels (0 = choose
 * from the batch size so that each launch has at least 4 blocks per CU).
 * The results do not depend on these values. */
int pf_ctx_set_tuning(pf_ctx *ctx, int chunk_z, int chunk_lcc);
/* Schedule of the fused z stage, 0 .. PF_Z_SCHEDULES - 1.  0-3: bit 0
 * selects the radix plan of its FFTs (the default, or an alternative with a
 * different element count per thread and radix sequence), bit 1 the width of
 * the (ky, kx) column tile per block (the default, or an alternative width);
 * all four are compiled for every fused size.  4 and 5: the z stage of the
 * run-time planned pipeline (columns in
 * registers with per-stage butterfly counts), with its default (4) or
 * alternative (5) register plan; PF_MODE_FUSED_RT always runs that stage and
 * reads bit 0 only (the alternative plan, 8 slots per thread instead of 16,
 * where the z length has one).  A new context starts at its size's measured
 * best (4 at 64^3, 72^3, 96^3, 160^3 and 192^3, else 0); the autotuner
 * picks per inputs.  The
 * plans change results by FP32 rounding of the FFTs, the tile width not at
 * all.  No counterpart in the reference. */
#define PF_Z_SCHEDULES 6
int pf_ctx_set_schedule(pf_ctx *ctx, int z_schedule);
int pf_ctx_profile(pf_ctx *ctx, int enable);
int pf_ctx_kernel_time(pf_ctx *ctx, int which, double *ms, int64_t *launches);

/* ---- template and mask generation (the step before the search) ----------
 * Replace _powerfit.blur_points (_powerfit.pyx:75-140) and
 * _powerfit.dilate_points (_powerfit.pyx:143-206), called by
 * structure_to_shape_like (volume.py:192-224) for the CLI's template and mask
 * (powerfit.py:245-255).  Same arguments: points is (3, n) row-major in grid
 * units (x, y, z rows), out a (nz, ny, nx) grid updated in place (host
 * memory); blur adds weights[k] exp(-d^2 / (2 sigma^2)) within 4 sigma of
 * each atom, dilate sets 1 within radii[k]; an atom's box is clipped below at
 * 0 (or at -dim + 1 with wraparound, negative indices wrapping once) and above
 * at dim - 1.  Runs on `device` and returns when out holds the result. */
int pf_blur_points(int device, const double *points, const double *weights, int64_t n,
                   double sigma, double *out, int nz, int ny, int nx, int wraparound);
int pf_dilate_points(int device, const double *points, const double *radii, int64_t n,
                     double *out, int nz, int ny, int nx, int wraparound);

/* ---- peak positions (the step after the search) ---------------------------
 * Replaces the labelling in Analyzer._watershed (analyzer.py:80-96):
 * scipy.ndimage.label(corr >= cutoff) with the 6-connected cross, no wrap,
 * then scipy.ndimage.maximum_position over every label, for each of the
 * ncut cutoffs.  corr and cutoffs are host arrays of dtype 0 = float32 or
 * 1 = float64 (corr C-order (nz, ny, nx)).  positions receives the flat
 * index of every component's maximum (the lowest index among equal maxima),
 * cutoff after cutoff, up to capacity entries; *count is the number found,
 * which may exceed capacity (call again with a larger buffer). */
int pf_peak_positions(int device, const void *corr, int dtype, int nz, int ny, int nx,
                      const void *cutoffs, int ncut, int64_t *positions, int64_t capacity,
                      int64_t *count);

#ifdef __cplusplus
}
#endif
#endif
