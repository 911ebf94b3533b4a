/*
 * pytme_hip.h - C ABI of libpytme_hip.so, the gfx950 HIP kernels behind the
 * `hip` matching backend of pytme_amd.
 *
 * Conventions
 *   - Every function returns TM_OK (0) or an error status; tm_last_error()
 *     gives the message (thread-local).  TM_ERR_VALUE = bad argument/shape,
 *     TM_ERR_OOM = device allocation failure, TM_ERR_HIP = other HIP error.
 *   - Array arguments are raw device pointers to C-ordered data: float32,
 *     int32, double, or interleaved complex64 (passed as void*).
 *   - `stream` is a hipStream_t (NULL = default stream); work is enqueued
 *     asynchronously and ordered on that stream.
 *   - No torch types appear in the ABI.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the pyTME repository).  The Python side binds these with ctypes in
 * pytme_amd/_lib.py.
 */
#ifndef PYTME_HIP_H
#define PYTME_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TM_OK 0
#define TM_ERR_VALUE 1
#define TM_ERR_OOM 2
#define TM_ERR_HIP 3

typedef struct tm_fft_plan tm_fft_plan;

/* ---- library ----------------------------------------------------------- */
int tm_abi_version(void);
const char* tm_last_error(void);
/* Replaces MatchingBackend.device_count (tme/backends/matching_backend.py:1262). */
int tm_device_count(int* n);
/* Wait for `stream`; surfaces asynchronous kernel faults as TM_ERR_HIP. */
int tm_stream_sync(void* stream);
/* Optional per-kernel timing: while enabled, every library launch is
 * bracketed by HIP events on its stream; query returns total device ms,
 * launch count and algorithmic HBM bytes per kernel name.  No counterpart
 * in the reference (it has no per-stage timing, see tme/matching_exhaustive.py). */
int tm_timing_enable(int on);
int tm_timing_names(char* buf, int len);
int tm_timing_query(const char* name, double* total_ms, int64_t* count, double* bytes);

/* ---- FFT: numpy.fft.rfftn / irfftn semantics ------------------------------
 * Replaces MatchingBackend.build_fft / rfftn / irfftn
 * (tme/backends/matching_backend.py:1089-1140; numpy backend pyFFTW plans,
 * tme/backends/npfftw_backend.py:243-306).  `shape` is the real shape of one
 * transform (rank 1..3, each axis <= 4096 with prime factors <= 13); `batch`
 * transforms are contiguous.  Forward is unnormalised; the inverse applies
 * 1/N and overwrites its complex input (pyFFTW overwrite_input=True). */
int tm_fft_plan_create(int ndim, const int64_t* shape, int64_t batch, tm_fft_plan** plan);
int tm_fft_plan_destroy(tm_fft_plan* plan);
int tm_fft_r2c(const tm_fft_plan* plan, const float* in, void* out, void* stream);
int tm_fft_c2r(const tm_fft_plan* plan, void* in, float* out, void* stream);
/* Radix plan of a length (host only, for inspection). */
int tm_fft_factor(int n, int* radices, int max_out, int* n_out);

/* ---- rotation ---------------------------------------------------------------
 * Replaces NumpyFFTWBackend.rigid_transform / _rigid_transform at spline
 * order 1, mode="constant" (tme/backends/npfftw_backend.py:326-449;
 * scipy.ndimage.affine_transform semantics).  Writes the shape[0..2] box of
 * each output; optionally also rotates a mask and returns the FP64 sum of
 * each rotated mask (the FLC n_obs, tme/matching_scores.py:743). */
int64_t tm_rotate_workspace(const int* shape, int nbatch);
int tm_rotate_trilinear(const float* src, const float* src_mask, const int* shape, const float* rots_dev,
                        const double* mats_dev, const double* center, const double* translation, int nbatch,
                        float* dst, float* dst_mask, const int64_t* dst_strides, int64_t dst_batch_stride,
                        double* n_obs_out, double* workspace, void* stream);

/* ---- pointwise / reductions ---------------------------------------------- */
int64_t tm_reduce_workspace(void);
/* Replaces MatchingBackend.fill (matching_backend.py:629-640). */
int tm_fill_f32(float* p, float value, int64_t n, void* stream);
int tm_fill_i32(int* p, int value, int64_t n, void* stream);
/* Spectral products of the scoring loop: be.multiply on complex spectra
 * (tme/matching_scores.py:619, 747-753) and the Fourier filter multiply
 * (tme/matching_scores.py:121-123).  Operand b / f repeats with period nb/nf. */
int tm_cmul_cc(const void* a, const void* b, void* out, int64_t n, int64_t nb, void* stream);
int tm_cmul_cr(const void* a, const float* f, void* out, int64_t n, int64_t nf, void* stream);
/* be.add/subtract/multiply/divide on float32 (op 0..3); b == NULL uses scalar s
 * (tme/matching_scores.py:622-623 numerator / inv_denominator). */
int tm_binary_f32(int op, const float* a, const float* b, float s, float* out, int64_t n, int64_t nb,
                  void* stream);
/* Replaces NumpyFFTWBackend.norm_scores (tme/backends/npfftw_backend.py:537-587);
 * arr == NULL means arr = 1 (flcSphericalMask_setup, tme/matching_scores.py:438). */
int tm_norm_scores(const float* arr, const float* exp_sq, const float* sq_exp, const double* n_obs_dev,
                   double n_obs_host, float eps, float* out, int64_t n, int64_t n_ex, void* stream);
/* Replaces NumpyFFTWBackend.max_score_over_rotations (npfftw_backend.py:503-535). */
int tm_max_update(const float* scores, float* max_scores, int* rotations, int rotation_index, int64_t n,
                  void* stream);
/* be.sum over a float32 array, accumulated deterministically in FP64
 * (tme/matching_scores.py:411-413, 557, 743). */
int tm_sum_f64(const float* x, int64_t n, double* out_dev, double* workspace, void* stream);
/* Replaces normalize_template (tme/matching_utils.py:61-99): FP64 moments. */
int tm_normalize_template(float* t, const float* m, int64_t n, const double* n_obs_dev, double n_obs_host,
                          double* workspace, void* stream);

/* ---- fused batched scan ----------------------------------------------------
 * Replaces the per-rotation loops of corr_scoring (tme/matching_scores.py:
 * 600-624, mode 1: CC / LCC / FLCSphericalMask) and flc_scoring (:727-757,
 * mode 0: FLC) together with MaxScoreOverRotations.__call__ ->
 * max_score_over_rotations (tme/analyzer/aggregation.py:142-189,
 * tme/backends/npfftw_backend.py:503-535) for a batch of rotations: rotate,
 * filter, normalise, correlate through the FFT and fold into the max/argmax
 * state without materialising per-rotation score volumes.  Inputs are the
 * setup products of flc_setup / flcSphericalMask_setup / cc_setup
 * (tme/matching_scores.py:153-189, 357-450): the reversed template and mask,
 * the template-size filter (or NULL), the target spectra and, for
 * FLCSphericalMask, the precomputed inverse denominator. */
typedef struct tm_scan_plan tm_scan_plan;
int64_t tm_scan_workspace_bytes(int mode, const int64_t* fast_shape, const int* template_shape, int max_batch,
                                int with_filter);
int tm_scan_plan_create(int mode, const int64_t* fast_shape, const int* template_shape, const float* templ,
                        const float* mask, const float* filter, const void* ft_target, const void* ft_target2,
                        const float* inv_denominator, int normalize, double mask_sum, int max_batch,
                        tm_scan_plan** plan);
int tm_scan_plan_destroy(tm_scan_plan* plan);
/* Refresh the plan's copy of the target spectra (same shapes as at creation;
 * stream-ordered).  The plan keeps the spectra at a padded row stride, so a
 * caller whose spectra change while a plan is kept (a new target split in
 * tme/matching_exhaustive.py:305-327) calls this before the next batch. */
int tm_scan_plan_set_target(tm_scan_plan* plan, const void* ft_target, const void* ft_target2, void* stream);
int tm_scan_batch(tm_scan_plan* plan, const float* rotations, const int* rotation_indices, int nbatch,
                  float* max_scores, int* max_rotations, void* stream);
/* No counterpart in the reference: replay the batch's kernel sequence as a
 * hipGraph on a plan-owned stream (event-ordered with the caller's stream);
 * the first batch of a (batch size, state arrays) configuration runs eagerly,
 * the second is captured, later ones replay.  Off by default; per-kernel
 * timing forces eager. */
int tm_scan_plan_use_graph(tm_scan_plan* plan, int enable);
/* No counterpart in the reference: how many of the plan's graph-mode
 * batches were captured, replayed and run eagerly.  The plan keeps the
 * captured sequences of two configurations (a scan's full batches and its
 * remainder batch), so alternating full and partial batches keep replaying. */
int tm_scan_plan_graph_stats(const tm_scan_plan* plan, int64_t* captures, int64_t* replays, int64_t* eager);
/* No counterpart in the reference: tile shapes of the batch passes, chosen
 * by the backend's autotuner.  pass_a_columns 16 or 8 (0: 16 where the rows
 * are padded to 16 bins), pass_c_rows 5 or 10 (0: 5; 10 only for an x length
 * of 200); unsupported values return an error.  Results do not depend on
 * them. */
int tm_scan_plan_set_tiles(tm_scan_plan* plan, int pass_a_columns, int pass_c_rows);
/* No counterpart in the reference (pyFFTW plans with FFTW_MEASURE,
 * tme/backends/npfftw_backend.py:243-306, choose their own algorithm): the
 * radix plan variant of the batch passes' compile-time FFTs, chosen by the
 * backend's autotuner.  Variant 0 takes radix-8 stages first (512 = 8*8*8,
 * 100 = 4*5*5), variant 1 the reverse order, or radix-4 stages for lengths
 * of radix-8 stages only (512 = 4*4*4*8).  Results agree to rounding.
 * tm_fft_radix_plan reports the forward radices of a length in a variant
 * (count 0: no compile-time plan for n). */
int tm_scan_plan_set_radix(tm_scan_plan* plan, int variant);
int tm_fft_radix_plan(int n, int variant, int* radices, int max_out, int* count);
/* A group of scan plans (several templates of uneven size against one
 * target, each with its own state arrays) stepped one batch per plan per
 * call: tm_scan_group_batch runs, in plan order, one batch of every plan
 * whose nbatch[i] > 0 (rotations[i], rotation_indices[i] into max_scores[i],
 * max_rotations[i]; the per-plan arguments of tm_scan_batch); the first call
 * of a configuration (batch sizes, state arrays, tile shapes) runs eagerly,
 * the second captures all plans' kernel sequences into one hipGraph, later
 * calls replay it on a group stream event-ordered with the caller's stream.
 * Replaces the per-template invocations of scripts/match_template.py (one
 * template each, tme/matching_exhaustive.py:298-465) for a multi-template
 * schedule.  The plans stay owned by the caller and must outlive the group;
 * a plan is not to be used through tm_scan_batch while its group runs. */
typedef struct tm_scan_group tm_scan_group;
int tm_scan_group_create(tm_scan_plan* const* plans, int n, tm_scan_group** group);
int tm_scan_group_destroy(tm_scan_group* group);
int tm_scan_group_batch(tm_scan_group* group, const float* const* rotations, const int* const* rotation_indices,
                        const int* nbatch, float* const* max_scores, int* const* max_rotations, void* stream);

/* ---- local FLC refinement --------------------------------------------------
 * tm_flc_local_fwd replaces the score evaluation of the derivative-free
 * refinement, FLC.__call__ after _MatchDensityToDensity.score
 * (tme/matching_optimization.py:488-548, 198-238), for a batch of poses: it
 * returns per pose the six moments (n, sum g m, sum g^2 m, sum g m f,
 * sum m f, sum m f^2) of the rotated standardised template g and mask m over
 * the template box with corner v in the target f; the host assembles
 * overlap / (n sd_f).  The rotated mask goes into the mask (the reference
 * writes it into the template buffer, :162-163).  tm_flc_local_bwd has no
 * counterpart in the reference (which is derivative-free): given
 * w = dScore/d(moment) it returns dScore/d(R^T) and dScore/ds.
 * poses: float32 [n][12] = R^T row-major (9) and the sub-voxel shift s (3);
 * corners: int32 [n][3]; moments float64 [n][6]; grad float64 [n][12]. */
typedef struct tm_flc_local_plan tm_flc_local_plan;
int tm_flc_local_plan_create(const float* template_std, const float* mask, const int* template_shape, int max_poses,
                             tm_flc_local_plan** plan);
int tm_flc_local_plan_destroy(tm_flc_local_plan* plan);
int tm_flc_local_fwd(tm_flc_local_plan* plan, const float* target, const int64_t* target_shape, const float* poses,
                     const int* corners, int n_poses, double* moments, void* stream);
int tm_flc_local_bwd(tm_flc_local_plan* plan, const float* target, const int64_t* target_shape, const float* poses,
                     const int* corners, int n_poses, const double* weights, double* grad, void* stream);
/* No counterpart in the reference: the moments and their Jacobian in one
 * pass, out float64 [n][78] = six moments, then per moment the 12
 * derivatives d(moment)/d(R^T) (9, row-major) and d(moment)/ds (3); the
 * gradient of any function of the moments is then a 6-term contraction. */
int tm_flc_local_fwd_jac(tm_flc_local_plan* plan, const float* target, const int64_t* target_shape,
                         const float* poses, const int* corners, int n_poses, double* out, void* stream);
/* No counterpart in the reference (its refinement is a derivative-free
 * scipy loop per peak, scripts/postprocess.py:588-613): the batched Adam
 * step of the gradient refinement on the device.  tm_refine_prepare turns
 * translations t and zyz angles (float64 [n][3], radians) into poses and
 * corners (v = trunc(t), shift -R^T (t - v) or t - v); tm_refine_adam takes
 * the moments-and-Jacobian output, tracks the best score and pose, and
 * applies one Adam step (torch's update, ascending the score) to t and the
 * angles, with adam_m / adam_v float64 [n][6] state and t kept within
 * bound of t0 when bound >= 0.  The two-pass form of the same step:
 * tm_refine_weights takes tm_flc_local_fwd's moments (float64 [n][6]),
 * tracks the best score and pose and, when update, writes the weights
 * dScore/d(moment) (float64 [n][6]) for tm_flc_local_bwd, whose gradient
 * (float64 [n][12]) tm_refine_adam_grad turns into the Adam step. */
int tm_refine_prepare(const double* t, const double* angles, int n, int rigid, float* poses, int* corners,
                      void* stream);
int tm_refine_adam(const double* moments_jac, double* t, double* angles, const double* t0, double* adam_m,
                   double* adam_v, double* best_score, double* best_t, double* best_angles, int n, int rigid,
                   int update, double lr_t, double lr_a, double beta1, double beta2, double eps, int adam_step,
                   double bound, double eps32, void* stream);
int tm_refine_weights(const double* moments, const double* t, const double* angles, double* best_score,
                      double* best_t, double* best_angles, double* weights, int n, int update, double eps32,
                      void* stream);
int tm_refine_adam_grad(const double* grad, double* t, double* angles, const double* t0, double* adam_m,
                        double* adam_v, int n, int rigid, double lr_t, double lr_a, double beta1, double beta2,
                        double eps, int adam_step, double bound, void* stream);

/* ---- peak extraction on the max-score volume -------------------------------
 * tm_peaks_max_filter + tm_peaks_write replace max_filter_coordinates
 * (tme/backends/npfftw_backend.py:200-206), the PeakCallerMaximumFilter
 * primitive (tme/analyzer/peaks.py:662-670): scipy maximum_filter of edge
 * size (1..64) with mode="constant" (zero borders), peaks where the score
 * equals the filtered score, as int64 (z, y, x) rows in C order.  Buffers of
 * a volume of F voxels, nblk = ceil(F / 4096): tmp F floats (used for edges
 * above 32 only, may be NULL otherwise), mask 64 nblk words, cnt nblk and
 * off nblk + 1 int64; off[nblk] is the peak count.
 * tm_topk_f32 replaces topk_indices (npfftw_backend.py:176-181, used by
 * PeakCaller._top_peaks and PeakCallerSort, peaks.py:501-529, 644-659):
 * indices of the k largest values, largest first, ties by lower index, NaN
 * above every number; k <= 65536.
 * tm_peak_filter replaces the greedy pass of filter_points_indices
 * (peaks.py:68-106 with find_candidate_indices, tme/external/bindings.cpp:
 * 84-124, int64 coordinates): rows taken in the given order are kept unless
 * trunc(sqrt(d2)) <= min_distance to a kept row; returns the first max_keep
 * kept rows (n <= 16384: any max_keep, workspace from
 * tm_peak_filter_workspace_bytes; larger n: max_keep <= 4096). */
int tm_peaks_max_filter(const float* scores, const int64_t* shape, int size, float* tmp, unsigned long long* mask,
                        int64_t* cnt, int64_t* off, void* stream);
int tm_peaks_write(const unsigned long long* mask, const int64_t* off, const int64_t* shape, int64_t* coords,
                   void* stream);
int64_t tm_topk_workspace_bytes(int64_t n, int64_t k);
int tm_topk_f32(const float* values, int64_t n, int64_t k, int64_t* out, void* work, int64_t work_bytes,
                void* stream);
int64_t tm_peak_filter_workspace_bytes(int64_t n);
int tm_peak_filter(const int64_t* coords, int64_t n, int64_t min_distance, int64_t max_keep, int64_t* keep,
                   int64_t* n_keep, void* work, int64_t work_bytes, void* stream);
/* tm_peak_caller_fold replaces the rest of PeakCaller.__call__ after
 * call_peaks (tme/analyzer/peaks.py:327-384 with _get_peak_mask :241-290 and
 * _update :531-583) for one score volume: boundary and score-window mask,
 * the top-2 num_peaks preselection when there are more than num_peaks
 * candidates, merge with the state (placeholder rows reduced to one),
 * sort by score, distance filter, the first num_peaks rows.  Counts stay on
 * the device (counts[0] valid candidates: 0 means the state is unchanged,
 * counts[1] new rows, counts[2] output rows), so one call needs no host
 * synchronisation.  State rows + new rows <= 8192. */
int64_t tm_peak_caller_fold_workspace_bytes(int64_t n_cand, int64_t n_state, int64_t num_peaks);
int tm_peak_caller_fold(const int64_t* cand, int64_t n_cand, const float* scores, const int64_t* shape,
                        const int64_t* st_pos, const float* st_rot, const float* st_score, const float* st_det,
                        int64_t n_state, const float* rotation, int64_t num_peaks, int64_t min_distance,
                        int64_t min_boundary_distance, int has_min, float min_score, int has_max, float max_score,
                        int64_t* out_pos, float* out_rot, float* out_score, float* out_det, int64_t* counts,
                        void* work, int64_t work_bytes, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tm_peak_caller_fold with the same arguments, results
 * and errors but no workspace argument; its workspace is a static scratch
 * array of the library, one copy per GPU (tm_peak_caller_fold_noalloc_max_bytes
 * = 8 MiB: with num_peaks 1000 and a full state about 1.8 million
 * candidates), held until the process ends.  Steps whose
 * tm_peak_caller_fold_workspace_bytes exceeds it return TM_ERR_VALUE before
 * any launch.  It calls no allocator.  Calls that overlap on one GPU race on
 * the shared scratch. */
int64_t tm_peak_caller_fold_noalloc_max_bytes(void);
int tm_peak_caller_fold_noalloc(const int64_t* cand, int64_t n_cand, const float* scores, const int64_t* shape,
                                const int64_t* st_pos, const float* st_rot, const float* st_score,
                                const float* st_det, int64_t n_state, const float* rotation, int64_t num_peaks,
                                int64_t min_distance, int64_t min_boundary_distance, int has_min, float min_score,
                                int has_max, float max_score, int64_t* out_pos, float* out_rot, float* out_score,
                                float* out_det, int64_t* counts, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYTME_HIP_H */
