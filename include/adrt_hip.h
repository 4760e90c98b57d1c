/*
 * adrt_hip.h -- C ABI of the MI355X (gfx950) ADRT transform core.
 *
 * Every entry point takes plain device pointers and sizes, launches
 * asynchronously on the given HIP stream (a hipStream_t passed as void*;
 * NULL means the default stream) and returns an int status (ADRT_HIP_OK or
 * one of the ADRT_HIP_ERR_* codes below).  Inputs are never modified,
 * outputs are written in full (adrt_hip_bdrt_rows: the rows asked for), and
 * nothing is allocated inside the library: temporary space comes from the
 * caller through the `workspace` argument, sized by
 * adrt_hip_workspace_bytes().  No torch or Python types appear.
 *
 * Each transform entry point replaces one function of the reference's
 * CPython extension adrt._adrt_cdefs (src/adrt/adrt_cdefs_py.cpp, method
 * table at :852-863) and performs the same shape/step validation
 * (src/adrt/adrt_cdefs_common.cpp:176-330) before launching.
 *
 * Array layouts are C-contiguous with one optional leading batch dimension,
 * exactly as the reference's NumPy arrays:
 *   image        (batch, n, n)
 *   ADRT output  (batch, 4, 2n-1, n)
 */
#ifndef ADRT_HIP_H
#define ADRT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of the floating-point transforms. */
#define ADRT_HIP_F32 0
#define ADRT_HIP_F64 1

/* Status codes. */
#define ADRT_HIP_OK 0
#define ADRT_HIP_ERR_SHAPE 1        /* shape invalid for the operation          */
#define ADRT_HIP_ERR_STEP 2         /* step outside [0, num_iters(n))            */
#define ADRT_HIP_ERR_DTYPE 3        /* unsupported element type                  */
#define ADRT_HIP_ERR_TOO_BIG 4      /* element count overflows size_t            */
#define ADRT_HIP_ERR_NULL 5         /* required pointer is NULL                  */
#define ADRT_HIP_ERR_WORKSPACE 6    /* workspace smaller than required           */
#define ADRT_HIP_ERR_DEVICE 7       /* HIP runtime / launch failure              */
#define ADRT_HIP_ERR_ARG 8          /* other invalid argument                    */
#define ADRT_HIP_ERR_SCRATCH 9      /* needs more than the no-allocation scratch */

/* Operation ids for adrt_hip_workspace_bytes(). */
#define ADRT_HIP_OP_ADRT 0
#define ADRT_HIP_OP_BDRT 1
#define ADRT_HIP_OP_IADRT 2

/* ---- library information ------------------------------------------------ */

/* Static string describing a status code. */
const char *adrt_hip_status_string(int status);

/* Text of the last HIP runtime error seen on the calling thread ("" if none). */
const char *adrt_hip_last_error(void);

/* Library ABI version (major * 100 + minor). */
int adrt_hip_abi_version(void);

/* Number of visible HIP devices (0 when none or the runtime is unusable). */
int adrt_hip_device_count(void);

/* ---- shape rules (no device work) --------------------------------------- */

/* num_iters(n) = bit_width(n) - has_single_bit(n); adrt_cdefs_common.cpp:140-142 */
int adrt_hip_num_iters(size_t n);

/* (batch, n, n) is a valid adrt input: all > 0, square, n a power of two,
 * n <= 2^63.  adrt_cdefs_common.cpp:176-182 */
int adrt_hip_valid_adrt_shape(const size_t shape[3]);

/* (batch, 4, 2n-1, n) is a valid ADRT-output shape (bdrt, iadrt, adrt_step,
 * bdrt_step input).  adrt_cdefs_common.cpp:185-192 */
int adrt_hip_valid_output_shape(const size_t shape[4]);

/* 0 <= step < num_iters(shape[3]).  adrt_cdefs_common.cpp:195-197 */
int adrt_hip_valid_step(const size_t shape[4], int step);

/* Restriction input: (batch, 4, 2n-1, n) with n even, n >= 2.
 * adrt_cdefs_common.cpp:289-296 */
int adrt_hip_valid_restriction_shape(const size_t shape[4]);

/* Prolongation input: all > 0 and both image dims <= SIZE_MAX/2.
 * adrt_cdefs_common.cpp:307-312 */
int adrt_hip_valid_prolongation_shape(const size_t shape[3]);

/* High-pass input: all > 0 and both image dims >= 2.
 * adrt_cdefs_common.cpp:322-326 */
int adrt_hip_valid_highpass_shape(const size_t shape[3]);

/* Bytes of workspace `op` needs for `batch` images of side n (0 = none). */
size_t adrt_hip_workspace_bytes(int op, int dtype, size_t batch, size_t n);

/* ---- transforms ----------------------------------------------------------- */

/* Forward ADRT: (batch, n, n) -> (batch, 4, 2n-1, n).
 * Replaces adrt._adrt_cdefs.adrt (adrt_cdefs_py.cpp:275-341 ->
 * adrt_basic, adrt_cdefs_adrt.hpp:100-212).  Bit-exact to the reference. */
int adrt_hip_adrt(const void *in, void *out, void *workspace, size_t workspace_bytes,
                  int dtype, size_t batch, size_t n, void *stream);

/* Forward ADRT that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, validation and bit-identical results as adrt_hip_adrt, without
 * the workspace: the fused kernels' intermediate lives in a library-owned
 * static device array of ADRT_HIP_NOALLOC_SCRATCH_BYTES (68 MiB) on each
 * device the library runs on, held until the process ends; launch groups are
 * sized to it.  Sides whose intermediate for one image exceeds it (float64 at
 * n = 2048; the streaming path when its output-sized buffer exceeds it) are
 * refused with ADRT_HIP_ERR_SCRATCH before any device work.  No hipMalloc,
 * hipFree or allocating library call is made.  All calls on one device share
 * its array (nothing in it needs clearing between calls): calls that overlap
 * on one device, from other streams or threads, race and must be serialised
 * by the caller. */
#define ADRT_HIP_NOALLOC_SCRATCH_BYTES (68u << 20)
int adrt_hip_adrt_noalloc(const void *in, void *out, int dtype, size_t batch, size_t n,
                          void *stream);

/* One forward stage on the ADRT-output layout, 0 <= step < num_iters(n).
 * Replaces adrt._adrt_cdefs.adrt_step (adrt_cdefs_py.cpp:343-412 ->
 * adrt_step, adrt_cdefs_adrt.hpp:214-258). */
int adrt_hip_adrt_step(const void *in, void *out, int dtype, size_t batch, size_t n,
                       int step, void *stream);

/* Back-projection: (batch, 4, 2n-1, n) -> same shape.
 * Replaces adrt._adrt_cdefs.bdrt (adrt_cdefs_py.cpp:482-548 ->
 * bdrt_basic, adrt_cdefs_bdrt.hpp:120-187). */
int adrt_hip_bdrt(const void *in, void *out, void *workspace, size_t workspace_bytes,
                  int dtype, size_t batch, size_t n, void *stream);

/* adrt_hip_bdrt for the rows [0, rows) of every quadrant only (1 <= rows <=
 * 2n-1, else ADRT_HIP_ERR_ARG): those rows equal adrt_hip_bdrt's, the rows
 * from `rows` on are left unspecified (written or not).  Same workspace.  No
 * reference counterpart: the reference's transpose of adrt,
 * truncate(bdrt(y)).mean (docs/examples.cginverse.md:51,63; truncate at
 * utils.py:191-245), reads rows [0, n) of each quadrant only, and the fused
 * kernels then do only the work those rows need (about half of the second
 * pass, a part of the first). */
int adrt_hip_bdrt_rows(const void *in, void *out, void *workspace, size_t workspace_bytes,
                       int dtype, size_t batch, size_t n, size_t rows, void *stream);

/* Forward ADRT and back-projection of several batches of small sides in one
 * persistent launch (no reference counterpart: the reference runs one call
 * per array; this is the uneven-size batch of many small images).  For each
 * entry, images (batch, n, n) give adrt = adrt_hip_adrt(images) and
 * bdrt = adrt_hip_bdrt(adrt), both (batch, 4, 2n-1, n), bit-identical to
 * those calls; either output may be NULL (not stored), not both.  n a power
 * of two <= 64; 1 <= count <= 8; outputs must not overlap any input or each
 * other; no workspace, no allocation, capturable in a graph.  Entries of batch 0 are skipped.  max_groups > 0 caps the persistent
 * grid (rounded down to a multiple of 32, at least 32) to leave CUs to other
 * streams; 0 = every work-group that fits at once. */
typedef struct adrt_hip_batch {
    const void *images;
    void *adrt;
    void *bdrt;
    size_t batch;
    size_t n;
} adrt_hip_batch;

int adrt_hip_adrt_bdrt_small(const adrt_hip_batch *batches, int count, int dtype,
                             int max_groups, void *stream);

/* Forward ADRT and back-projection of several batches of any power-of-two
 * sides n <= 1024 in ONE persistent launch with a stage-ordered work queue
 * (no reference counterpart: the reference runs one call per array,
 * adrt_cdefs_py.cpp:275-341 and 482-548 -> adrt_cdefs_adrt.hpp:111-211,
 * adrt_cdefs_bdrt.hpp:120-187; this replaces that loop over arrays for an
 * uneven-size batch).  Per entry the outputs are bit-identical to
 * adrt_hip_adrt(images) and adrt_hip_bdrt(adrt).  Sides n >= 64 run the
 * fused passes (forward pass 1, pass 2, back-projection pass A, pass B, each
 * over every image before the next, an entry waiting only for its own
 * image's previous pass); sides n <= 32 run whole in LDS.  For n >= 64 the
 * adrt output is required (the back-projection reads it) and bdrt may be
 * NULL; for n <= 32 either may be NULL, not both.  1 <= count <= 16, batch
 * <= 2^20 per entry, entries of batch 0 skipped; outputs must not overlap
 * inputs or each other.
 *
 *   adrt_hip_queue_workspace_bytes / _state_bytes: sizes of the caller's
 *     device workspace (intermediates of every image of side >= 64) and
 *     device state (queue header, segment table, completion counters);
 *     0 when the batches are invalid.
 *   adrt_hip_queue_plan: validates, builds the queue for these pointers and
 *     copies it into `state` (synchronous hipMemcpy; no allocation); fills
 *     *plan.  bundle = fused items per queue entry: 1, 2, 4 or 8 (0 = 8);
 *     waves = register budget in waves per SIMD: 8 (float32 only) or 4
 *     (0 = 8 for float32, 4 for float64).
 *   adrt_hip_queue_run: the one launch, on `stream`; no allocation, no
 *     memset (the kernel resets its counters on exit), capturable in a graph
 *     and replayable.  Runs on one state must not overlap.  max_groups > 0
 *     caps the grid; 0 = every work-group that fits at once.
 *   adrt_hip_queue_timeouts: number of dependency waits abandoned after
 *     their bound since the last call (then reset; synchronous).  Nonzero
 *     means the results of those runs are invalid; it stays 0 unless the
 *     device stalls for seconds. */
typedef struct adrt_hip_queue {
    void *state;
    int dtype;
    int entries;
    int segments;
    int counters;
    int lds_bytes;
    int waves;
} adrt_hip_queue;

size_t adrt_hip_queue_workspace_bytes(const adrt_hip_batch *batches, int count, int dtype);
size_t adrt_hip_queue_state_bytes(const adrt_hip_batch *batches, int count);
int adrt_hip_queue_plan(const adrt_hip_batch *batches, int count, int dtype, int bundle,
                        int waves, void *workspace, size_t workspace_bytes, void *state,
                        size_t state_bytes, adrt_hip_queue *plan);
int adrt_hip_queue_run(const adrt_hip_queue *plan, int max_groups, void *stream);
int adrt_hip_queue_timeouts(const adrt_hip_queue *plan, int *count);

/* One back-projection stage.  Replaces adrt._adrt_cdefs.bdrt_step
 * (adrt_cdefs_py.cpp:550-619 -> bdrt_step, adrt_cdefs_bdrt.hpp:189-244). */
int adrt_hip_bdrt_step(const void *in, void *out, int dtype, size_t batch, size_t n,
                       int step, void *stream);

/* Exact inverse: (batch, 4, 2n-1, n) -> same shape.
 * Replaces adrt._adrt_cdefs.iadrt (adrt_cdefs_py.cpp:414-480 ->
 * iadrt_basic, adrt_cdefs_iadrt.hpp:109-176). */
int adrt_hip_iadrt(const void *in, void *out, void *workspace, size_t workspace_bytes,
                   int dtype, size_t batch, size_t n, void *stream);

/* Multigrid restriction (batch, 4, 2n-1, n) -> (batch, 4, n-1, n/2).
 * Replaces adrt._adrt_cdefs.press_fmg_restriction (adrt_cdefs_py.cpp:684-738
 * -> fmg_restriction, adrt_cdefs_fmg.hpp:53-73). */
int adrt_hip_fmg_restriction(const void *in, void *out, int dtype, size_t batch, size_t n,
                             void *stream);

/* Multigrid prolongation (batch, h, w) -> (batch, 2h, 2w).
 * Replaces adrt._adrt_cdefs.press_fmg_prolongation (adrt_cdefs_py.cpp:740-794
 * -> fmg_prolongation, adrt_cdefs_fmg.hpp:75-95). */
int adrt_hip_fmg_prolongation(const void *in, void *out, int dtype, size_t batch, size_t h,
                              size_t w, void *stream);

/* Multigrid high-pass filter (batch, h, w) -> same shape, h, w >= 2.
 * Replaces adrt._adrt_cdefs.press_fmg_highpass (adrt_cdefs_py.cpp:796-850
 * -> fmg_highpass, adrt_cdefs_fmg.hpp:97-171). */
int adrt_hip_fmg_highpass(const void *in, void *out, int dtype, size_t batch, size_t h,
                          size_t w, void *stream);

/* Forward stages first, first+1, ..., last-1 applied in turn to an
 * ADRT-layout array (batch, 4, 2n-1, n): the same operations as that many
 * adrt_hip_adrt_step calls (adrt_cdefs_adrt.hpp:214-258), without the host
 * round trips.  0 <= first < last <= num_iters(n).  The whole chain
 * (first = 0, last = num_iters(n)) is the transpose of adrt_hip_bdrt on that
 * layout, since bdrt_step(k) is the transpose of adrt_step(L-1-k)
 * (tests/test_bdrt_step.py:76-92): the back-projection's adjoint.  Needs
 * adrt_hip_stage_chain_workspace() bytes when last - first > 1. */
size_t adrt_hip_stage_chain_workspace(int dtype, size_t batch, size_t n);
int adrt_hip_stage_chain(const void *in, void *out, void *workspace, size_t workspace_bytes,
                         int dtype, size_t batch, size_t n, int first, int last, void *stream);

/* adrt.utils.interp_to_cart: nearest-neighbour interpolation of an ADRT
 * output (batch, 4, 2n-1, n) onto the regular grid (batch, n, 4n).
 * Replaces adrt._adrt_cdefs.interp_to_cart (adrt_cdefs_py.cpp:621-682 ->
 * interp_adrtcart<T, float>, adrt_cdefs_interp_adrtcart.hpp:60-114).  The
 * angle- and offset-dependent quantities are computed on the host by
 * adrt_hip_interp_table into a caller buffer of adrt_hip_interp_table_bytes
 * (a plain CPU function, same float math as the reference); the caller
 * copies that table to the device once per (dtype, n) and passes it to
 * adrt_hip_interp_to_cart, which launches the per-element kernel.  n must
 * be a power of two > 1 and 4n exact in float (else ADRT_HIP_ERR_TOO_BIG). */
size_t adrt_hip_interp_table_bytes(int dtype, size_t n);
int adrt_hip_interp_table(int dtype, size_t n, void *host_buf, size_t bytes);
int adrt_hip_interp_to_cart(const void *in, void *out, const void *table, size_t table_bytes,
                            int dtype, size_t batch, size_t n, void *stream);

/* ---- layout helpers (any element size: 1, 2, 4 or 8 bytes) --------------- */

/* adrt.core.adrt_init (core.py:123-176): (batch, n, n) -> (batch, 4, 2n-1, n),
 * quadrant flips into the top n offsets, zero elsewhere. */
int adrt_hip_adrt_init(const void *in, void *out, size_t elem_bytes, size_t batch, size_t n,
                       void *stream);

/* adrt.utils.truncate (utils.py:191-242): (outer, 4, 2n-1, n) -> (outer, 4, n, n). */
int adrt_hip_truncate(const void *in, void *out, size_t elem_bytes, size_t outer, size_t n,
                      void *stream);

/* adrt.utils.stitch_adrt (utils.py:65-130): (outer, 4, 2n-1, n) ->
 * (outer, 3n-2, 4 * vc), vc = n (remove_repeated 0) or n - 1 (1, n >= 2);
 * any n >= 1.  Quadrants side by side, odd ones flipped in both axes, zero
 * fill elsewhere. */
int adrt_hip_stitch(const void *in, void *out, size_t elem_bytes, size_t outer, size_t n,
                    int remove_repeated, void *stream);

/* adrt.utils.unstitch_adrt (utils.py:133-188): the inverse,
 * (outer, 3n-2, 4 * vc) -> (outer, 4, 2n-1, n); with removed_repeated 1
 * (vc = n - 1) the dropped columns are restored from the neighbouring
 * quadrant as the reference does. */
int adrt_hip_unstitch(const void *in, void *out, size_t elem_bytes, size_t outer, size_t n,
                      int removed_repeated, void *stream);

/* ---- fused helpers for the solvers (no reference counterpart) ------------ */

/* out[b] = combine_q(truncate(in)[b, q] / divisor), in (batch, 4, 2n-1, n),
 * out (batch, n, n).  mode 0: ((q0+q1)+q2)+q3 then / 4 (NumPy mean over the
 * quadrant axis, as in core.py:329); mode 1: plain sum ((q0+q1)+q2)+q3.
 * divisor == 1 skips the per-element divide. */
int adrt_hip_truncate_combine(const void *in, void *out, int dtype, size_t batch, size_t n,
                              double divisor, int mode, void *stream);

/* ---- batched CG vector operations (no reference counterpart) ------------- *
 * For B independent systems of m unknowns (C-contiguous, B <= 65535), with
 * float64 scalars per system and deterministic float64 sums (fixed
 * 16384-element chunks: a system's sums do not depend on B or on the other
 * systems):
 *   adrt_hip_batched_dot:  out[b] = sum_i u[b][i] * v[b][i]
 *   adrt_hip_cg_update:    x += a_b p, r -= a_b q (a_b = alpha[b] rounded to
 *                          the element type), rr[b] = sum_i r[b][i]^2 after
 *   adrt_hip_cg_direction: p = z + c_b p (c_b = beta[b] rounded)
 * alpha, beta, out and rr are device arrays of B doubles; the workspace is
 * adrt_hip_cg_workspace(B, m) bytes.  Used by the solvers; capturable in a
 * graph (no allocation, no memset). */
size_t adrt_hip_cg_workspace(size_t batch, size_t m);
int adrt_hip_batched_dot(const void *u, const void *v, double *out, void *workspace,
                         size_t workspace_bytes, int dtype, size_t batch, size_t m, void *stream);
int adrt_hip_cg_update(void *x, void *r, const void *p, const void *q, const double *alpha,
                       double *rr, void *workspace, size_t workspace_bytes, int dtype,
                       size_t batch, size_t m, void *stream);
int adrt_hip_cg_direction(void *p, const void *z, const double *beta, int dtype, size_t batch,
                          size_t m, void *stream);

/* One (preconditioned) CG iteration's scalar logic on the device, over a
 * per-image state of adrt_hip_pcg_state_bytes(B) bytes laid out as
 *   double rz[B], pq[B], rr[B], tol[B], alpha[B], beta[B];
 *   int32  active[B], next[B], use_p[B], iters[B], count
 * (the caller fills rz, tol, active, use_p and iters before the first step).
 *   adrt_hip_pcg_update:    pq = p.q; alpha = active && pq > 0 ? rz / pq : 0;
 *                           x += alpha p; r -= alpha q; rr = r.r;
 *                           iters += active; next = active && sqrt(rr) > tol
 *   adrt_hip_pcg_direction: z == NULL (plain CG) or use_p == 0: rz' = rr,
 *                           beta = rr / rz; otherwise rz' = r.z and, if
 *                           next && (rz' <= 0 || pq <= 0), use_p = 0, rz' = rr,
 *                           beta = 0, else beta = rz' / rz (rz == 0 reads as
 *                           1); p = (use_p ? z : r) + beta p; rz = rz';
 *                           active = next; count = number of active images.
 * An inactive image keeps x and r.  Workspace as adrt_hip_cg_update; no
 * allocation, no memset, capturable in a graph. */
size_t adrt_hip_pcg_state_bytes(size_t batch);
int adrt_hip_pcg_update(void *x, void *r, const void *p, const void *q, void *state,
                        void *workspace, size_t workspace_bytes, int dtype, size_t batch,
                        size_t m, void *stream);
int adrt_hip_pcg_direction(void *p, const void *r, const void *z, void *state, void *workspace,
                           size_t workspace_bytes, int dtype, size_t batch, size_t m,
                           void *stream);

/* ---- kernel configuration (no reference counterpart) --------------------- */

/* Variants of the forward and back-projection paths for one size:
 *   DEFAULT    n <= 64: SINGLE; log2 n in [7, 12]: the fused two-pass kernels
 *              with the default stage split of the L = log2 n stages (they
 *              differ for odd L): forward pass 1 takes floor(L/2) stages for
 *              float64 or L <= 9, else ceil(L/2); back-projection pass A
 *              takes ceil(L/2) for L <= 9, else floor(L/2);
 *              n > 4096: streaming;
 *   ALT_SPLIT  log2 n in [6, 12]: the fused kernels with the other split
 *              (at n = 64 the fused kernels' only split);
 *   STREAMING  one kernel per stage;
 *   SINGLE     n <= 64: one launch with every stage in LDS (no workspace).
 * and of the inverse (ADRT_HIP_OP_IADRT; the group budget does not apply):
 *   DEFAULT    the last min(L, 5) stages in one fused pass, the others one
 *              kernel each;
 *   ALT_SPLIT  a fused tail of min(L, 2) stages;
 *   STREAMING  one kernel per stage.
 * Every variant gives bit-identical results. */
#define ADRT_HIP_VARIANT_DEFAULT 0
#define ADRT_HIP_VARIANT_ALT_SPLIT 1
#define ADRT_HIP_VARIANT_STREAMING 2
#define ADRT_HIP_VARIANT_SINGLE 3

/* Selects the variant and the per-launch-group workspace budget (bytes of
 * intermediate per group of images; 0 = the default 136 MiB; a group holds
 * at least one image) used from now on for `op` (ADRT_HIP_OP_ADRT,
 * ADRT_HIP_OP_BDRT or ADRT_HIP_OP_IADRT) at side n.  The setting is process-wide (all devices
 * and threads) and changes adrt_hip_workspace_bytes for that (op, n):
 * query the workspace after configuring.  A launch that meets a workspace
 * sized for another configuration returns ADRT_HIP_ERR_WORKSPACE.  Used by
 * the autotuner; ALT_SPLIT is rejected (ADRT_HIP_ERR_ARG) where n has no
 * fused path, SINGLE for n > 64 and for ADRT_HIP_OP_IADRT. */
int adrt_hip_set_config(int op, size_t n, int variant, size_t group_bytes);

/* Reads the configuration of (op, n) (group_bytes: the budget in effect). */
int adrt_hip_get_config(int op, size_t n, int *variant, size_t *group_bytes);

/* ---- learned CG initial guess / preconditioner (no reference counterpart) --
 *
 * Kernels of the 4-level U-Net that supplies CG's initial guess, its
 * training, the multiscale linear preconditioner and the on-device phantom
 * generator.  Activations are NHWC bf16 (2-byte elements) with a pixel
 * stride in elements (a multiple of 8, >= the channel count); weights are
 * bf16 [cout][9][cin] (cin a multiple of 8); biases and gradients fp32.
 * Pointers to bf16 data must be 16-byte aligned. */

/* 3x3 convolution with zero padding, on the matrix cores (bf16 in, fp32
 * accumulation).  mode 0: y = relu(conv + bias); 1: y32[p] = resid[p] +
 * conv[p][0] + bias[0] (fp32 output, channel 0); 2: y = conv * (mask > 0);
 * 3: y = conv (y 16-byte aligned).  Modes 2 and 3 with the weights of
 * adrt_hip_nn_transpose_weights give the input gradient.  cout % 16 == 0,
 * batch*h*w_ % 64 == 0. */
int adrt_hip_nn_conv3x3(const void *x, int x_stride, const void *w, const float *bias, void *y,
                        int y_stride, float *y32, const float *resid, const void *mask,
                        int mask_stride, int batch, int h, int w_, int cin, int cout, int mode,
                        void *stream);

/* The same convolution (modes 0 and 3) over the channel concatenation of x
 * (channels [0, c_skip), h x w_) and the nearest 2x upsampling of up
 * (channels [c_skip, cin), h/2 x w_/2): the U-Net decoder's input, read in
 * place instead of materialised.  Needs h % 4 == 0, w_ % 16 == 0 (and
 * w_ <= 64 or w_ % 64 == 0), cin >= 16 with cin % 32 != 24, c_skip % 8 == 0;
 * the result equals
 * adrt_hip_nn_conv3x3 on the concatenated buffer bit for bit. */
int adrt_hip_nn_conv3x3_upcat(const void *x, int x_stride, int c_skip, const void *up,
                              int up_stride, const void *w, const float *bias, void *y,
                              int y_stride, int batch, int h, int w_, int cin, int cout, int mode,
                              void *stream);

/* Workspace bytes of adrt_hip_nn_wgrad. */
size_t adrt_hip_nn_wgrad_workspace(int batch, int h, int w, int cin, int cout);

/* Weight and bias gradients of a 3x3 convolution: dw[co][t][ci] =
 * sum_p dz[p][co] * x[p + tap t][ci], dbias[co] = sum_p dz[p][co]
 * (overwritten, fp32). */
int adrt_hip_nn_wgrad(const void *x, int x_stride, const void *dz, int dz_stride, float *dw,
                      float *dbias, void *workspace, size_t workspace_bytes, int batch, int h,
                      int w, int cin, int cout, void *stream);

/* wt[ci][8 - t][co] = w[co][t][ci]: weights of the input-gradient convolution. */
int adrt_hip_nn_transpose_weights(const void *w, void *wt, int cin, int cout, void *stream);

/* 2x2 average pooling (h, w: input size, even). */
int adrt_hip_nn_pool(const void *x, int x_stride, void *y, int y_stride, int batch, int h, int w,
                     int c, void *stream);

/* ds = (dp[parent] / 4 + dskip) * (s > 0), at the fine size h x w. */
int adrt_hip_nn_pool_backward(const void *dp, int dp_stride, const void *dskip, int dskip_stride,
                              const void *s, int s_stride, void *ds, int ds_stride, int batch,
                              int h, int w, int c, void *stream);

/* Nearest 2x upsampling of an h x w input. */
int adrt_hip_nn_upsample(const void *x, int x_stride, void *y, int y_stride, int batch, int h,
                         int w, int c, void *stream);

/* dx = sum over the 2x2 children of dy (h x w: the coarse size). */
int adrt_hip_nn_upsample_backward(const void *dy, int dy_stride, void *dx, int dx_stride,
                                  int batch, int h, int w, int c, void *stream);

/* fp32 images -> 8-channel bf16 NHWC (channel 0 = scale * x, others 0). */
int adrt_hip_nn_pack_input(const float *x, void *y, int batch, int h, int w, float scale,
                           void *stream);

/* The network's first and last 3x3 convolutions (zero padding) on the vector
 * ALU, fp32 fused multiply-adds in tap (then channel) order:
 *   conv_in:  y[p][co] = relu(bias[co] + sum_t w[co][t][0] * bf16(scale * z[p + t])),
 *             co < 16, z fp32 (batch, h, w_), w bf16 [16][9][cinw], y bf16
 *             NHWC with pixel stride y_stride (16-byte aligned);
 *   conv_out: y32[p] = resid[p] + (sum_{t, ci < 16} w[0][t][ci] * x[p + t][ci]
 *             + bias[0]), x bf16 NHWC (16 channels, pixel stride x_stride),
 *             w bf16 [cout][9][16] (row 0 used). */
int adrt_hip_nn_conv_in(const float *z, float scale, const void *w, int cinw, const float *bias,
                        void *y, int y_stride, int batch, int h, int w_, void *stream);
int adrt_hip_nn_conv_out(const void *x, int x_stride, const void *w, const float *bias, float *y32,
                         const float *resid, int batch, int h, int w_, void *stream);
/* The encoder's full-resolution level for inference in one launch: e0a
 * (adrt_hip_nn_conv_in's arithmetic, never stored), e0b (16 -> 16, ReLU,
 * stored as the skip connection) and its 2x2 average pooling (stored as the
 * next level's input); bit-identical to conv_in + conv3x3 + pool.  h a
 * multiple of 4, w_ a multiple of 64. */
int adrt_hip_nn_encoder0(const float *z, float scale, const void *w1, int cinw1, const float *b1,
                         const void *w2, const float *b2, void *skip, int skip_stride,
                         void *pooled, int pooled_stride, int batch, int h, int w_,
                         void *stream);

/* 16-channel bf16 gradient: channel 0 = scale * (pred - target), others 0. */
int adrt_hip_nn_loss_grad(const float *pred, const float *target, void *dz, int batch, int h,
                          int w, float scale, void *stream);

/* One Adam step on count fp32 parameters (c1, c2: the bias corrections
 * 1 - beta^t); w16, if not NULL, receives the bf16 copy. */
int adrt_hip_nn_adam(float *w, const float *g, float *m, float *v, void *w16, size_t count,
                     float lr, float beta1, float beta2, float eps, float c1, float c2,
                     void *stream);

/* Multiscale preconditioner pieces (fp32 images, batch x h x w):
 * 2x2 averaging; y = coef25 (5x5, zero padding) applied to x, plus
 * coarse[parent] / 4 when coarse is not NULL; and partial[block][o] =
 * sum_p e[p] * x[p + o] over o in the 5x5 window (parts blocks).  The
 * pooling input must be 8-byte aligned. */
int adrt_hip_pyr_pool(const float *x, float *y, int batch, int h, int w, void *stream);
int adrt_hip_pyr_stencil_up(const float *x, const float *coarse, float *y, const float *coef25,
                            int batch, int h, int w, void *stream);
int adrt_hip_pyr_correlate(const float *e, const float *x, float *partial, int parts, int batch,
                           int h, int w, void *stream);

/* Radial DCT-domain preconditioner of the learned solver: for float32
 * images (batch, n, n), n a power of two in [2, 1024], with C the
 * orthonormal 2-D DCT-II (coefficient [u][v]: u along rows, v along
 * columns, as scipy.fft.dctn(..., norm="ortho")):
 *   mode 0: y = C x
 *   mode 1: y = C^T (H .* C x), H an (n, n) symmetric float32 table
 * (symmetric positive weights make the filter symmetric positive definite).
 * table: adrt_hip_dct_table_len(n) float2 twiddles from adrt_hip_dct_table
 * (a host function), copied to the device; tmp: batch * n * n floats of
 * workspace; x, y, tmp distinct.  Three row passes (two for mode 0) of
 * in-LDS FFTs (Makhoul's DCT). */
size_t adrt_hip_dct_table_len(size_t n);
int adrt_hip_dct_table(size_t n, float *host_table, size_t floats);
int adrt_hip_dct_filter(const float *x, float *y, const float *table, const float *H, float *tmp,
                        size_t batch, size_t n, int mode, void *stream);

/* Ellipse phantoms: out[b][i][j] = sum of val over the ellipses of image b
 * (params[b][k] = cx, cy, cos, sin, ax, ay, val as float64; counts[b] <=
 * max_ellipses) whose closed interior holds the pixel centre
 * (-1 + (2j+1)/n, 1 - (2i+1)/n). */
int adrt_hip_phantom_raster(const double *params, const int *counts, int max_ellipses, float *out,
                            int batch, int n, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ADRT_HIP_H */
