/*
 * ambient_amd.h - C ABI of the MI355X (gfx950) TGV-DCT deshade hot path.
 *
 * Drop-in replacement for the CUDA launchers of josefkoller/ambient_image_processor
 * (reference paths relative to its repository root).  Conventions that differ
 * from the reference, and why:
 *   - every entry point returns an amb_status instead of calling exit() from
 *     cudaCheckError (reference source/modules/cuda/cuda_helper.in:11-22);
 *     amb_last_error() gives the message of the calling thread's last failure;
 *   - the caller owns every buffer: outputs are written into caller-provided
 *     arrays instead of being returned as `new Pixel[N]`
 *     (reference tgv_deshade/TGVDeshadeProcessor.cpp:146-157);
 *   - buffers may live on the host or in device memory (amb_memory); the
 *     reference's launchers take host arrays only;
 *   - tau, sigma and theta are parameters (the reference hard-codes
 *     tau = sigma = 1/sqrt(20), theta = 1, tgv_deshade/tgv2_l1_deshade.cu:148-150;
 *     amb_tgv_params_default() sets exactly those values).
 * Device state (the iteration fields, DCT scratch) lives in an amb_context,
 * which owns one HIP stream; a context must not be used by two threads at once.
 * Layout of every volume: float32, x fastest, idx = z*W*H + y*W + x
 * (reference source/modules/itk/ITKImage.cpp:426-429).
 * include/ambient_amd_reference_compat.hpp re-exposes the reference's own
 * C++ launcher templates on top of this ABI.
 */
#ifndef AMBIENT_AMD_H
#define AMBIENT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMB_API __attribute__((visibility("default")))

typedef enum {
    AMB_OK = 0,
    AMB_ERR_INVALID_ARGUMENT = 1,
    AMB_ERR_HIP = 2,
    AMB_ERR_OUT_OF_MEMORY = 3,
    AMB_ERR_NOT_SUPPORTED = 4
} amb_status;

/* Where the caller's buffers live. */
typedef enum { AMB_MEM_HOST = 0, AMB_MEM_DEVICE = 1 } amb_memory;

/* Data term of the u-update: L1 = tgv_kernel_part4_tgv2_l1 (reference
 * tgv_cuda/tgv2_l1.cu:27-60), L2 = tgv_kernel_part4_tgv2_l2 (tgv_cuda/tgv2_l2.cu:27-56). */
typedef enum { AMB_DATA_L1 = 0, AMB_DATA_L2 = 1 } amb_data_term;

/* Storage type of the iteration fields u, u_bar, v, v_bar, p, q.  Arithmetic
 * is FP32 in both cases; f is always stored FP32.  BF16 stores round to
 * nearest, ties to even. */
typedef enum { AMB_STORAGE_F32 = 0, AMB_STORAGE_BF16 = 1 } amb_storage;

/* Arithmetic of the separable DCT-II/III passes.
 *   REFERENCE: the reference's numerics - FP64 cosines, FP64 products, FP32
 *              accumulator rounded after every term, reference term order
 *              (cuda_image_operations/cosine_transform.cu:142-223);
 *   F32:       per-axis dense products on FP32 MFMA, FP32 accumulate;
 *   BF16X3:    per-axis dense products on BF16 MFMA with hi/lo splits of both
 *              operands (hi*hi + hi*lo + lo*hi), FP32 accumulate. */
typedef enum { AMB_DCT_REFERENCE = 0, AMB_DCT_F32 = 1, AMB_DCT_BF16X3 = 2 } amb_dct_precision;

/* amb_tgv_params.flags */
#define AMB_FLAG_SET_NEGATIVE_ZERO    0x1u  /* r < 0 -> 0 (clamp_negative_values.cu:27) */
#define AMB_FLAG_MASK_SHADING         0x2u  /* mask also multiplies l (TGVDeshadeProcessor.cpp:83-87) */
#define AMB_FLAG_ADD_BACKGROUND_BACK  0x4u  /* r += u*(1-mask) (TGVDeshadeProcessor.cpp:275-279) */
#define AMB_FLAG_POISSON_INDEX_QUIRK  0x8u  /* recover voxel coordinates through float arithmetic as
                                               solve_poisson_in_cosine_domain.cu:36-40 does (differs
                                               from exact coordinates only for linear index >= 2^24) */
#define AMB_FLAG_MASKED_ITERATION     0x10u /* amb_tgv2_deshade_pipeline with a mask: the masked
                                               deshade composition of
                                               TGVDeshadeMaskedProcessor.cpp:76-210 - f unmasked,
                                               updates only where mask >= 1, l and r times the
                                               mask, clamp, then add-back or r times the mask;
                                               at depth 1 the 2-D loop (:93-98), as with
                                               AMB_FLAG_PLANAR */
#define AMB_FLAG_PLANAR               0x20u /* depth 1 only: the reference's 2-D loop
                                               (tgv_deshade/tgv2_l1_deshade_2d.cu:33-215,
                                               tgv_deshade_masked/tgv2_l1_deshade_masked_2d.cu)
                                               - eps_B(v_bar) without z backward differences, so
                                               v_z, p_z and the q z-components stay 0; the
                                               widget (TGVDeshadeWidget.cpp:100-127) and the
                                               masked processor (TGVDeshadeMaskedProcessor.cpp:93)
                                               take this loop for every depth-1 image */

typedef struct {
    float lambda;                /* data-term weight */
    float alpha0;                /* second-order weight (q projection, tgv2_common.cu:262) */
    float alpha1;                /* first-order weight (p projection, tgv2_common.cu:124) */
    float tau, sigma, theta;     /* primal/dual step sizes, over-relaxation */
    uint32_t iterations;
    uint32_t data_term;          /* amb_data_term */
    uint32_t storage;            /* amb_storage */
    uint32_t dct_precision;      /* amb_dct_precision */
    uint32_t callback_interval;  /* paint_iteration_interval; 0 = never */
    uint32_t flags;              /* AMB_FLAG_* */
} amb_tgv_params;

/* Called on the calling thread when iteration_index > 0 and
 * iteration_index % callback_interval == 0, with host copies of u and v that
 * are valid only during the call (reference tgv_cuda/tgv2_common.cu:374-408).
 * vx/vy/vz are NULL for the denoise entry points.  Return nonzero to stop. */
typedef int (*amb_iteration_callback)(uint32_t iteration_index, uint32_t iteration_count,
                                      const float* u, const float* vx, const float* vy,
                                      const float* vz, void* user);

typedef struct amb_context amb_context;

/* ---- context, memory ---------------------------------------------------- */
AMB_API amb_status amb_context_create(int device, amb_context** ctx);
AMB_API amb_status amb_context_destroy(amb_context* ctx);
AMB_API const char* amb_last_error(void);
AMB_API const char* amb_version(void);
/* Reference defaults: tau = sigma = 1/sqrt(20), theta = 1, lambda = 1,
 * alpha0 = 2, alpha1 = 1 (TGVDeshadeWidget.ui:44,94,119), L1, FP32, REFERENCE DCT,
 * 10000 iterations, no callback, AMB_FLAG_SET_NEGATIVE_ZERO. */
AMB_API void amb_tgv_params_default(amb_tgv_params* params);
AMB_API amb_status amb_device_alloc(amb_context* ctx, size_t bytes, void** ptr);
AMB_API amb_status amb_device_free(amb_context* ctx, void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device (stream-ordered, then synchronised) */
AMB_API amb_status amb_memcpy(amb_context* ctx, void* dst, const void* src, size_t bytes, int kind);
AMB_API amb_status amb_synchronize(amb_context* ctx);
/* The context's HIP stream (a hipStream_t), for callers that order their own work after ours. */
AMB_API void* amb_context_stream(amb_context* ctx);
/* Launch-plan overrides for tests and tuning (no counterpart in the reference,
 * whose launchers take a cuda_block_dimension that changes only the block
 * size).  Results never depend on them: every plan computes the same values.
 * They belong to the context and apply from the next call that begins an
 * iteration (deshade, denoise, begin, TGV-k); a value < 0 unsets one.  Names:
 *   tile (0-7), zchunk, sweep_pf (0/1)        the two-sweep kernels
 *   fused (0 = off), fused_vx (1 on odd widths, 2, 4), fused_ty, fused_zchunk,
 *   fused_tx, fused_pf
 *                                             the fused TGV2 sweep
 *   fused_image (0 = off), fused_image_tx, fused_image_zchunk   the 2-D row march
 *   index64 (1 = the 64-bit-offset kernel forms on any volume)
 *   tgvk_zchunk, tgvk_fused (0 = off), tgvk_fused_tx, tgvk_fused_ty, tgvk_fused_zchunk
 *   adjoint_vx (1 = one voxel per thread in the adjoint's reverse phases)
 * An unknown name returns AMB_ERR_INVALID_ARGUMENT. */
AMB_API amb_status amb_context_set_plan_override(amb_context* ctx, const char* name,
                                                 int32_t value);
AMB_API amb_status amb_context_clear_plan_overrides(amb_context* ctx);

/* ---- A13/A19: TGV2 deshade iteration ------------------------------------
 * Replaces tgv2_l1_deshade_launch<float>(f_host, width, height, depth, lambda,
 * iteration_count, paint_iteration_interval, cuda_block_dimension, callback,
 * alpha0, alpha1, &v_x, &v_y, &v_z)  (reference tgv_deshade/tgv2_l1_deshade.cu:33-303,
 * declared TGVDeshadeProcessor.cpp:29-39).  data_term = AMB_DATA_L2 gives the
 * same loop with the L2 u-update (the reference has no L2 deshade launcher).
 * Writes u and v (x, y, z) into the caller's buffers. */
AMB_API amb_status amb_tgv2_deshade(amb_context* ctx, amb_memory mem, const float* f,
                                    uint32_t width, uint32_t height, uint32_t depth,
                                    const amb_tgv_params* params, float* u, float* v_x,
                                    float* v_y, float* v_z, amb_iteration_callback callback,
                                    void* user);

/* ---- A20: TGV2 denoise ----------------------------------------------------
 * Replaces tgv2_l1_launch<float> / tgv2_l2_launch<float>(f_host, width, height,
 * depth, lambda, iteration_count, paint_iteration_interval, callback, alpha0,
 * alpha1)  (reference tgv_cuda/tgv2_l1.cu:64-321, tgv_cuda/tgv2_l2.cu:60-316,
 * declared tgv/TGVProcessor.cpp:52-69).  The callback receives u only. */
AMB_API amb_status amb_tgv2_denoise(amb_context* ctx, amb_memory mem, const float* f,
                                    uint32_t width, uint32_t height, uint32_t depth,
                                    const amb_tgv_params* params, float* u,
                                    amb_iteration_callback callback, void* user);

/* ---- A14-A18: DCT Poisson bias solve ------------------------------------- */
/* Replaces divergence_kernel_launch<float>(dx, dy, dz, w, h, d, is_host_data)
 * (reference cuda_image_operations/divergence.cu:59-141): out = B_x dx + B_y dy + B_z dz. */
AMB_API amb_status amb_divergence(amb_context* ctx, amb_memory mem, const float* dx,
                                  const float* dy, const float* dz, uint32_t width,
                                  uint32_t height, uint32_t depth, float* out);
/* Replaces cosine_transform_kernel_launch<float>(image, w, h, d)
 * (reference cuda_image_operations/cosine_transform.cu:226-293). */
AMB_API amb_status amb_cosine_transform(amb_context* ctx, amb_memory mem, const float* image,
                                        uint32_t width, uint32_t height, uint32_t depth,
                                        uint32_t dct_precision, float* out);
/* Replaces inverse_cosine_transform_kernel_launch<float>(image, w, h, d)
 * (reference cuda_image_operations/inverse_cosine_transform.cu:167-227). */
AMB_API amb_status amb_inverse_cosine_transform(amb_context* ctx, amb_memory mem,
                                                const float* image, uint32_t width,
                                                uint32_t height, uint32_t depth,
                                                uint32_t dct_precision, float* out);
/* Replaces solve_poisson_in_cosine_domain_kernel_launch<float>(image, w, h, d)
 * (reference cuda_image_operations/solve_poisson_in_cosine_domain.cu:86-130).
 * flags: AMB_FLAG_POISSON_INDEX_QUIRK. */
AMB_API amb_status amb_solve_poisson_in_cosine_domain(amb_context* ctx, amb_memory mem,
                                                      const float* image, uint32_t width,
                                                      uint32_t height, uint32_t depth,
                                                      uint32_t flags, float* out);
/* Replaces TGVDeshadeProcessor::integrate_image_gradients_poisson_cosine_transform
 * (reference tgv_deshade/TGVDeshadeProcessor.cpp:386-408): l = iDCT(P(DCT(B-div v))). */
AMB_API amb_status amb_integrate_gradients_poisson(amb_context* ctx, amb_memory mem,
                                                   const float* v_x, const float* v_y,
                                                   const float* v_z, uint32_t width,
                                                   uint32_t height, uint32_t depth,
                                                   uint32_t dct_precision, uint32_t flags,
                                                   float* shading);

/* ---- A19 epilogue alone ----------------------------------------------------
 * Replaces the image-operation chain after the Poisson integration
 * (reference TGVDeshadeProcessor.cpp:83-93 subtract / multiply /
 * clamp_negative_values, :251-255 add-back; masked composition
 * TGVDeshadeMaskedProcessor.cpp:119-135): r = u - l; with a mask, r times the
 * mask (and l times the mask with MASK_SHADING or MASKED_ITERATION); clamp
 * with SET_NEGATIVE_ZERO; then add-back of u*(1 - mask) with
 * ADD_BACKGROUND_BACK, else with MASKED_ITERATION r times the mask again.
 * shading_out (optional) receives l after masking; mask may be NULL. */
AMB_API amb_status amb_deshade_epilogue(amb_context* ctx, amb_memory mem, const float* u,
                                        const float* shading, const float* mask, uint32_t width,
                                        uint32_t height, uint32_t depth, uint32_t flags,
                                        float* shading_out, float* deshaded);

/* ---- masked TGV2-L1 deshade ----------------------------------------------
 * Replaces tgv2_l1_deshade_masked_launch (reference
 * tgv_deshade_masked/tgv2_l1_deshade_masked.cu:27-305): the amb_tgv2_deshade
 * iteration, but the dual and primal updates run only at voxels with
 * active[i] != 0 (the reference's masked_pixel_indices,
 * TGVDeshadeMaskedProcessor.cpp:337-385); the stencils are the unmasked
 * ones, so voxels outside the set keep their initial values (u = u_bar = f,
 * v = p = q = 0) and are read as neighbours.  active: width*height*depth bytes
 * in `mem` space.  Same outputs and callback as amb_tgv2_deshade. */
AMB_API amb_status amb_tgv2_deshade_masked(amb_context* ctx, amb_memory mem, const float* f,
                                           uint32_t width, uint32_t height, uint32_t depth,
                                           const amb_tgv_params* params, const uint8_t* active,
                                           float* u, float* v_x, float* v_y, float* v_z,
                                           amb_iteration_callback callback, void* user);

/* ---- A19: full deshade ----------------------------------------------------
 * Replaces TGVDeshadeProcessor::processTGV2L1GPUCuda (reference
 * TGVDeshadeProcessor.cpp:122-166 with flags = SET_NEGATIVE_ZERO and without
 * MASK_SHADING; the callback overload :285-314 -> :224-282 with MASK_SHADING and
 * optionally ADD_BACKGROUND_BACK): iteration, Poisson integration of v, then
 * r = u - l, mask, clamp, add-back.  mask may be NULL.  With a mask and
 * AMB_FLAG_MASKED_ITERATION: TGVDeshadeMaskedProcessor::processTGV2L1GPUCuda
 * (reference TGVDeshadeMaskedProcessor.cpp:76-210; MASK_SHADING implied).
 * The callback (if any) receives u and v as in amb_tgv2_deshade.  Any of u,
 * shading, deshaded may be NULL if not wanted. */
AMB_API amb_status amb_tgv2_deshade_pipeline(amb_context* ctx, amb_memory mem, const float* f,
                                             uint32_t width, uint32_t height, uint32_t depth,
                                             const amb_tgv_params* params, const float* mask,
                                             float* u, float* shading, float* deshaded,
                                             amb_iteration_callback callback, void* user);

/* ---- A21 (no counterpart in the reference): FP64 energy and gap ----------
 * Evaluates, for the iteration state left in ctx by the last amb_tgv2_* call
 * (or set by amb_tgv2_set_state), with FP64 accumulation on the GPU:
 *   out[0] E = G(u) + alpha1 sum|D+u - v|_2 + alpha0 sum|eps_B v|_w
 *   out[1] G(u)   (lambda|u-f|_1 or lambda/2 |u-f|^2)
 *   out[2] alpha1 term, out[3] alpha0 term
 *   out[4] gap = out[5] + out[6] + out[7]  (Fenchel-Young brackets, >= 0)
 *   out[5] data bracket with s = -(D+)^T p (true transpose)
 *   out[6] alpha1 sum|D+u - v| - <D+u - v, p> ; out[7] alpha0 sum|eps v|_w - <eps v, q>_w
 *   out[8] L1 dual infeasibility sum max(0, |s| - lambda) (0 for L2)
 *   out[9] ||eps_B^T q - p||_2  (v stationarity residual) */
AMB_API amb_status amb_tgv2_energy(amb_context* ctx, double out[10]);

/* ---- A21 in the iteration (no counterpart in the reference) ----------------
 * With interval > 0, every TGV2 iteration call of ctx (deshade, denoise,
 * masked, pipeline, iterate) evaluates amb_tgv2_energy's ten quantities after
 * every iteration count c (1-based, within the call) with c % interval == 0,
 * after its last iteration and where a callback stops it.  Where iteration c's
 * primal half runs in the fused sweep (3-D loop, volumes under 2^30 voxels)
 * the sweep reduces them itself, in FP64 from the values it holds (wave64
 * shuffles, one row of partial sums per block, a fixed-order final sum): no
 * extra pass over the state.  Elsewhere the standalone pass runs.  0 = off
 * (the default).  The setting stays with the context. */
AMB_API amb_status amb_tgv2_set_energy_interval(amb_context* ctx, uint32_t interval);
/* The rows of the last call: *rows = their number; out (may be NULL) receives
 * min(rows, max_rows) rows of 11 doubles - the iteration count c, then
 * out[0..9] of amb_tgv2_energy for the state after iteration c. */
AMB_API amb_status amb_tgv2_energy_trace(amb_context* ctx, double* out, uint32_t max_rows,
                                         uint32_t* rows);

/* ---- amb_tgv2_energy without allocation (no counterpart in the reference) --
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The ten
 * quantities of amb_tgv2_energy for a state the caller holds in device memory,
 * enqueued on `stream` (a hipStream_t of the calling thread's current device;
 * NULL = the null stream) with no context and no hipMalloc / hipFree:
 *   f        W*H*D FP32 input (device);
 *   fields   17 device pointers in the amb_tgv2_get_state order (FP32, or
 *            bfloat16 values when params->storage is AMB_STORAGE_BF16);
 *   params   lambda, alpha0, alpha1, data_term, storage and AMB_FLAG_PLANAR;
 *   out_dev  10 doubles in device memory, written by the stream's work.
 * Scratch: one fixed device array per device of 4096 rows of 10 FP64 partial
 * sums (327,808 bytes, held until the process ends); any volume the library
 * accepts (at most 2^31 - 1 voxels) fits it.  Calls that overlap on one device
 * (on different streams or threads) share that array and race: order them. */
AMB_API amb_status amb_tgv2_energy_noalloc(const float* f, const void* const* fields,
                                           uint32_t width, uint32_t height, uint32_t depth,
                                           const amb_tgv_params* params, void* stream,
                                           double* out_dev);

/* ---- adjoint of the unrolled TGV2-L2 denoise (no counterpart in the reference)
 * Runs the amb_tgv2_denoise iteration (data_term must be AMB_DATA_L2, storage
 * AMB_STORAGE_F32) with a checkpoint of the state every `checkpoint_every`
 * iterations (0 = 20), then the hand-written reverse sweeps, recomputing each
 * segment from its checkpoint.  Given dL/du for the final u, returns dL/df and
 * dL_dparams = (dL/dalpha0, dL/dalpha1, dL/dlambda) (FP64 reductions).  u_out
 * (may be NULL) receives the forward result.  checkpoint_every = 1 keeps every
 * state (where the fused sweep runs) and recomputes nothing.  Workspace:
 * (ceil(iterations / checkpoint_every) + checkpoint_every + 2) * 17 + 8 volumes
 * of FP32.  The context's iteration state after the call is not specified. */
AMB_API amb_status amb_tgv2_denoise_adjoint(amb_context* ctx, amb_memory mem, const float* f,
                                            uint32_t width, uint32_t height, uint32_t depth,
                                            const amb_tgv_params* params, const float* dL_du,
                                            float* u_out, float* dL_df, double dL_dparams[3],
                                            uint32_t checkpoint_every);

/* ---- stepwise access to the iteration (tests, checkpointing) -------------
 * fields order: u, u_bar, v_x, v_y, v_z, vbar_x, vbar_y, vbar_z, p_x, p_y, p_z,
 *               q_xx, q_yy, q_zz, q_xy, q_xz, q_yz  (17 arrays of W*H*D floats). */
AMB_API amb_status amb_tgv2_begin(amb_context* ctx, amb_memory mem, const float* f,
                                  uint32_t width, uint32_t height, uint32_t depth,
                                  const amb_tgv_params* params);
AMB_API amb_status amb_tgv2_set_state(amb_context* ctx, amb_memory mem, const float* const* fields);
AMB_API amb_status amb_tgv2_get_state(amb_context* ctx, amb_memory mem, float* const* fields);
AMB_API amb_status amb_tgv2_iterate(amb_context* ctx, uint32_t iterations);
/* One half-iteration only: which = 0 dual sweep (p, q), 1 primal sweep (u, v). */
AMB_API amb_status amb_tgv2_sweep(amb_context* ctx, int which);

/* ---- TGV-k deshade (k >= 3) ----------------------------------------------
 * Replaces tgvk_l1_deshade_launch<float>(f_host, width, height, depth, lambda,
 * iteration_count, paint_iteration_interval, callback, order, alpha, &v_x, &v_y,
 * &v_z) (reference tgvk_deshade/tgvk_l1_deshade.cu:35-45, declared
 * TGVKDeshadeProcessor.cpp:27-36).  order = k in 3 .. 1000; alpha: k positive
 * host weights, alpha[0] bounding p, alpha[1] q and alpha[i + 2] the dual of
 * level i.  From params: lambda, tau, sigma, theta, iterations,
 * callback_interval and AMB_FLAG_PLANAR (depth 1: tgvk_l1_deshade_launch_2d,
 * tgvk_deshade/tgvk_l1_deshade_2d.cu); data_term must be L1 and storage F32;
 * params->alpha0/alpha1 are not read.  Writes u and v (x, y, z). */
AMB_API amb_status amb_tgvk_deshade(amb_context* ctx, amb_memory mem, const float* f,
                                    uint32_t width, uint32_t height, uint32_t depth,
                                    const amb_tgv_params* params, uint32_t order,
                                    const float* alpha, float* u, float* v_x, float* v_y,
                                    float* v_z, amb_iteration_callback callback, void* user);

/* Replaces tgvk_l1_deshade_masked_launch<float>(..., masked_pixel_indices, twelve edge
 * lists) (reference tgvk_deshade_masked/tgvk_l1_deshade_masked.cu:29-…, declared
 * TGVKDeshadeMaskedProcessor.cpp:27-50): the TGV-k loop with every update only where
 * active[i] != 0 (one byte per voxel); the stencils are the unmasked ones. */
AMB_API amb_status amb_tgvk_deshade_masked(amb_context* ctx, amb_memory mem, const float* f,
                                           uint32_t width, uint32_t height, uint32_t depth,
                                           const amb_tgv_params* params, uint32_t order,
                                           const float* alpha, const uint8_t* active, float* u,
                                           float* v_x, float* v_y, float* v_z,
                                           amb_iteration_callback callback, void* user);

/* TGVKDeshadeProcessor::processTGVKL1Cuda (TGVKDeshadeProcessor.cpp:76-174):
 * order 2 runs the TGV2 loop with alpha0 = alpha[1], alpha1 = alpha[0], order >= 3
 * the TGV-k loop, depth 1 the 2-D loops; then l from the Poisson integration of
 * v, r = u - l, the mask (may be NULL) multiplying l and r, the clamp
 * (AMB_FLAG_SET_NEGATIVE_ZERO) and the add-back of u*(1 - mask)
 * (AMB_FLAG_ADD_BACKGROUND_BACK).  params->dct_precision selects the DCT.
 * With a mask and AMB_FLAG_MASKED_ITERATION: TGVKDeshadeMaskedProcessor::processTGVKL1Cuda
 * (TGVKDeshadeMaskedProcessor.cpp:76-226) - order 2 the masked TGV2 composition, else the
 * masked TGV-k loop on the unmasked f (updates where mask >= 1), then l and r times the mask,
 * clamp, and the add-back or r times the mask again. */
AMB_API amb_status amb_tgvk_deshade_pipeline(amb_context* ctx, amb_memory mem, const float* f,
                                             uint32_t width, uint32_t height, uint32_t depth,
                                             const amb_tgv_params* params, uint32_t order,
                                             const float* alpha, const float* mask, float* u,
                                             float* shading, float* deshaded,
                                             amb_iteration_callback callback, void* user);

/* ---- resampling and the downsampled deshade ----------------------------------
 * ResizeProcessor::process(image, width, height, depth, BSpline3)
 * (reference resize/ResizeProcessor.cpp:54-152, through ITK 4.8.1's
 * ResampleImageFilter + BSplineInterpolateImageFunction of order 3): cubic
 * B-spline resampling of in (width x height x depth) to out (ow x oh x od), FP64
 * coefficients with mirror boundaries, output rounded to float.  As the
 * reference, an upsampling resize resamples with 2 more voxels per axis and
 * keeps the first ow x oh x od; samples outside the input are 0. */
AMB_API amb_status amb_resize_bspline3(amb_context* ctx, amb_memory mem, const float* in,
                                       uint32_t width, uint32_t height, uint32_t depth,
                                       float* out, uint32_t ow, uint32_t oh, uint32_t od);
/* ResizeProcessor::process(image, size_factor, ...) target extents
 * (ResizeProcessor.cpp:41-52): max(1, ceil(float(extent) * factor)) per axis. */
AMB_API void amb_resize_extents(uint32_t width, uint32_t height, uint32_t depth, float factor,
                                uint32_t out[3]);

/* TGVKDeshadeDownsampledProcessor::processTGVKL1Cuda
 * (tgvk_deshade_downsampled/TGVKDeshadeDownsampledProcessor.cpp:29-129, the
 * widget's and tgvk_deshade_downsampled_main.cpp's path) with
 * AMB_FLAG_MASKED_ITERATION: the input times the mask is resized by `factor`
 * (cubic B-splines; factor 1 skips the resizes), so is the mask; the masked TGV-k
 * processor (amb_tgvk_deshade_pipeline) runs there; u and l are resized back;
 * r = f*mask - l, l and r times the mask, plus f*(1 - mask) with
 * AMB_FLAG_ADD_BACKGROUND_BACK, then the clamp (AMB_FLAG_SET_NEGATIVE_ZERO).
 * Without the flag: TGVKDeshadeDownsampledNotMaskedProcessor.cpp:29-120 - f
 * itself goes down, the unmasked TGV-k processor, r = f - l (times the mask),
 * the add-back and the clamp.  mask may be NULL.  Writes u, shading and
 * deshaded at the input's extents; the callback sees the resized volume. */
AMB_API amb_status amb_tgvk_deshade_downsampled(amb_context* ctx, amb_memory mem, const float* f,
                                                uint32_t width, uint32_t height, uint32_t depth,
                                                const amb_tgv_params* params, float factor,
                                                uint32_t order, const float* alpha,
                                                const float* mask, float* u, float* shading,
                                                float* deshaded, amb_iteration_callback callback,
                                                void* user);

/* ---- timing ----------------------------------------------------------------
 * With profiling on, every kernel launch of the context is bracketed by HIP
 * events on the context's stream; amb_profile_read returns the summed device
 * time (ms) and launch count per kernel class:
 *   0 dual sweep, 1 primal sweep, 2 DCT passes, 3 Poisson divide,
 *   4 divergence, 5 epilogue, 6 init, 7 energy, 8 adjoint (whole call),
 *   9 fused sweep (primal half of one iteration + dual half of the next),
 *   10 TGV-k dual sweep, 11 TGV-k primal sweep, 12 B-spline resampling,
 *   13 fused TGV-k sweep (primal half of one iteration + dual half of the next),
 *   14 fused sweep with the in-sweep FP64 energy/gap (and its final sum). */
AMB_API amb_status amb_profile_enable(amb_context* ctx, int enable);
AMB_API amb_status amb_profile_read(amb_context* ctx, int kernel_class, double* total_ms,
                                    uint64_t* launches);
/* enable = k >= 2 times one launch in k of each class (the events' own cost
 * then stays out of k - 1 launches in k); amb_profile_read reports the timed
 * launches, amb_profile_count every launch of the class since the last reset. */
AMB_API amb_status amb_profile_count(amb_context* ctx, int kernel_class, uint64_t* launches);
/* Tiles of the 3-D fused sweep (class 9) launched while profiling since the
 * last reset: *swept those that ran, *total those in the launches' grids.  A
 * masked run skips, from its second fused launch on, the tiles that own no
 * active voxel; otherwise swept == total.  For pricing a masked launch's bytes
 * (the reference counts launches only; this has no counterpart there). */
AMB_API amb_status amb_profile_fused_tiles(amb_context* ctx, uint64_t* swept, uint64_t* total);
AMB_API amb_status amb_profile_reset(amb_context* ctx);

#ifdef __cplusplus
}
#endif
#endif /* AMBIENT_AMD_H */
