/*
 * sad_hip.h - C ABI of the MI355X (gfx950) SAD fit/render library, libsad_hip.so.
 *
 * One entry point per kernel launcher of the reference's CUDA backend, with the
 * same argument order and meaning; each comment names the reference launcher it
 * replaces.  Conventions, as in the reference:
 *   - all buffers are device pointers allocated by the caller; launchers do not
 *     allocate (the fit/render drivers at the end of this file own their buffers);
 *   - work is enqueued on `stream` and is asynchronous; a NULL stream is the
 *     device's null stream;
 *   - layouts: Site AoS (40 B), AdamState AoS (80 B), ten SoA fp32 gradient
 *     arrays, candidate field as two planes of 4 x uint32 per cell, images as
 *     packed float3 [H][W][3], mask float [H][W], pairs as (key, value) uint32.
 * Differences from the reference ABI:
 *   - every entry point returns an int status: 0 on success, a positive
 *     hipError_t from the launch, or a negative SADH_ERR_* validation code
 *     (the reference returns void and reports errors at the next sync);
 *   - the stream type is sadh_stream_t, layout-compatible with hipStream_t.
 */
#ifndef SAD_HIP_H
#define SAD_HIP_H

#include <stdint.h>
#include "sad_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* sadh_stream_t;

#define SADH_OK 0
#define SADH_ERR_INVALID (-1)      /* bad size, null pointer or out-of-range scalar */
#define SADH_ERR_UNSUPPORTED (-2)  /* option this build does not implement */
#define SADH_ERR_CAPACITY (-3)     /* request exceeds a buffer capacity */
#define SADH_ERR_NO_DEVICE (-4)    /* no usable gfx950 device */

/* Message for the last non-zero status returned on the calling thread. */
const char* sadh_last_error(void);
/* Library version string. */
const char* sadh_version(void);

/* ----------------------------------------------------------- render ----- */

/* launchRenderVoronoi, backends/cuda/src/render.cu:159-179 */
int sadh_render_voronoi(const uint32_t* cand0, const uint32_t* cand1, float* output,
                        const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                        int width, int height, int cand_width, int cand_height,
                        sadh_stream_t stream);

/* renderVoronoiColoring, backends/metal/shaders/encoders/render_encoder.metal:73-134
 * (encodeColoring, backends/metal/sources/encoders/render_encoder.swift:52-71): the
 * render with each site's colour replaced by a hash of its id (the _cells.png
 * picture of --render-cells and --include-debug-mask). */
int sadh_render_cells(const uint32_t* cand0, const uint32_t* cand1, float* output,
                      const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                      int width, int height, int cand_width, int cand_height,
                      sadh_stream_t stream);

/* renderCentroidsTauHeatmap, backends/metal/shaders/encoders/render_encoder.metal:136-190
 * (encodeTauHeatmap, render_encoder.swift:73-93): dots of radius dot_radius at the
 * candidate sites, blue -> white -> red over min_tau -> mean_tau -> max_tau
 * (the _tau_heatmap.png picture of --include-debug-mask). */
int sadh_render_tau_heatmap(const uint32_t* cand0, const uint32_t* cand1, float* output,
                            const SadSite* sites, uint32_t site_count, float min_tau,
                            float mean_tau, float max_tau, float dot_radius, int width, int height,
                            int cand_width, int cand_height, sadh_stream_t stream);

/* render_sad_backward_kernel launch, backends/pytorch/torch_ext/cuda/sad_bindings.cu:100-207,733-746.
 * grad_output [H][W][3]; grad_sites [N][10] (Site field order), accumulated into (caller zeroes). */
int sadh_render_backward(const uint32_t* cand0, const uint32_t* cand1, const float* grad_output,
                         float* grad_sites, const SadSite* sites, float inv_scale_sq,
                         uint32_t site_count, int width, int height, int cand_width,
                         int cand_height, sadh_stream_t stream);

/* ----------------------------------------------------- candidate field ---- */

/* launchInitCandidates, render.cu:181-197 */
int sadh_init_candidates(uint32_t* cand0, uint32_t* cand1, uint32_t site_count, uint32_t seed,
                         int per_pixel_mode, int width, int height, sadh_stream_t stream);

/* launchClearCandidates, render.cu:199-212 */
int sadh_clear_candidates(uint32_t* cand0, uint32_t* cand1, int width, int height,
                          sadh_stream_t stream);

/* launchJFASeed, jfa.cu:391-406.  When several active sites share a home cell
 * the largest site id is written (the reference leaves this to thread timing). */
int sadh_jfa_seed(uint32_t* cand0, const SadSite* sites, uint32_t site_count, int width,
                  int height, int cand_downscale, sadh_stream_t stream);

/* launchJFAFlood, jfa.cu:408-429 */
int sadh_jfa_flood(const uint32_t* in_cand0, uint32_t* out_cand0, const SadSite* sites,
                   uint32_t site_count, uint32_t step_size, float inv_scale_sq, int width,
                   int height, int cand_downscale, int target_width, int target_height,
                   sadh_stream_t stream);

/* launchPackCandidateSites, candidate_pack.cu:245-253 */
int sadh_pack_candidate_sites(const SadSite* sites, SadPackedSite* packed, uint32_t site_count,
                              sadh_stream_t stream);

/* launchUpdateCandidates (one VPT pass), jfa.cu:431-464 */
int sadh_update_candidates(const uint32_t* in_cand0, const uint32_t* in_cand1,
                           uint32_t* out_cand0, uint32_t* out_cand1,
                           const SadPackedSite* sites, uint32_t site_count, uint32_t step,
                           float inv_scale_sq, uint32_t step_high, float radius_scale,
                           uint32_t radius_probes, uint32_t inject_count,
                           const uint32_t* hilbert_order, const uint32_t* hilbert_pos,
                           uint32_t hilbert_probes, uint32_t hilbert_window, int cand_downscale,
                           int target_width, int target_height, int width, int height,
                           sadh_stream_t stream);

/* computeHilbertPairsKernel launch, sad_cuda.cu:1547-1571 / sad_bindings.cu:233-257 */
int sadh_hilbert_pairs(const SadSite* sites, SadPair* pairs, uint32_t site_count,
                       uint32_t padded_count, int width, int height, int bits,
                       sadh_stream_t stream);

/* writeHilbertOrderKernel launch, sad_cuda.cu:1573-1584 / sad_bindings.cu:259-270 */
int sadh_hilbert_order(const SadPair* sorted_pairs, uint32_t* order, uint32_t* pos,
                       uint32_t site_count, sadh_stream_t stream);

/* --------------------------------------------------------- gradients ----- */

/* launchComputeGradientsTiled, gradients.cu:491-531.  `rendered` (the render of
 * the current sites) is read only by the SSIM term, ssim_weight > 0: each
 * masked pixel adds -ssim_weight * dSSIM/dx_k * (0.299, 0.587, 0.114) to
 * dL/dpred, the derivative of its own 3x3 window's SSIM through its own luma
 * (gradients.cu:236-266). */
int sadh_compute_gradients_tiled(const uint32_t* cand0, const uint32_t* cand1,
                                 const float* target, const float* rendered, const float* mask,
                                 float* grad_pos_x, float* grad_pos_y, float* grad_log_tau,
                                 float* grad_radius, float* grad_color_r, float* grad_color_g,
                                 float* grad_color_b, float* grad_dir_x, float* grad_dir_y,
                                 float* grad_log_aniso, const SadSite* sites, float inv_scale_sq,
                                 uint32_t site_count, float* removal_delta,
                                 uint32_t compute_removal, float ssim_weight, int width,
                                 int height, int cand_width, int cand_height,
                                 sadh_stream_t stream);

/* sadh_compute_gradients_tiled without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Same arguments, results bit-identical
 * to sadh_compute_gradients_tiled.  The fixed-point lines live in a device
 * array of the library, shared by every call on one device: at most 131,072
 * sites (larger site_count returns SADH_ERR_CAPACITY); calls that overlap on
 * one device race. */
int sadh_compute_gradients_tiled_noalloc(const uint32_t* cand0, const uint32_t* cand1,
                                         const float* target, const float* rendered, const float* mask,
                                         float* grad_pos_x, float* grad_pos_y, float* grad_log_tau,
                                         float* grad_radius, float* grad_color_r, float* grad_color_g,
                                         float* grad_color_b, float* grad_dir_x, float* grad_dir_y,
                                         float* grad_log_aniso, const SadSite* sites, float inv_scale_sq,
                                         uint32_t site_count, float* removal_delta,
                                         uint32_t compute_removal, float ssim_weight, int width,
                                         int height, int cand_width, int cand_height,
                                         sadh_stream_t stream);

/* The tiled gradient of sadh_compute_gradients_tiled accumulated into AoS
 * lines: grad_lines is [site_count][16] floats, values 0-9 of site i in Site
 * field order (pos.x, pos.y, log_tau, radius, r, g, b, dir.x, dir.y,
 * log_aniso), 10-15 unused.  No counterpart in the reference, which adds into
 * ten arrays: here one site's ten atomic adds from one tile are one 64-byte
 * memory request.  The fit uses this layout internally. */
int sadh_compute_gradients_tiled_aos(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                                     const float* mask, float* grad_lines, const SadSite* sites, float inv_scale_sq,
                                     uint32_t site_count, float* removal_delta, uint32_t compute_removal, int width,
                                     int height, int cand_width, int cand_height, sadh_stream_t stream);

/* launchTauDiffuse, tau_diffusion.cu:388-401 */
int sadh_tau_diffuse(const uint32_t* cand0, const uint32_t* cand1, const SadSite* sites,
                     const float* grad_raw, const float* grad_in, float* grad_out,
                     uint32_t site_count, float lambda, int width, int height,
                     int cand_downscale, sadh_stream_t stream);

/* ----------------------------------------------------------- optimizer --- */

/* The trainer's per-iteration tail (sad_cuda.cu:2333-2389): the four
 * tau-diffusion passes (raw = grad_log_tau before the first pass) and Adam
 * with the gradient clear, the fourth pass and Adam fused into one kernel.
 * tau_scratch holds 10 * site_count 32-bit words (two float arrays and the
 * 8-id neighbour list of every site, built by the first pass and read by the
 * others).  Bit-identical to sadh_tau_diffuse
 * x4 (grad_log_tau -> tmp -> grad_log_tau -> tmp -> grad_log_tau) followed by
 * sadh_adam_update_clear. */
int sadh_tau_diffuse_adam(const uint32_t* cand0, const uint32_t* cand1, SadSite* sites, SadAdamState* adam,
                          float* grad_pos_x, float* grad_pos_y, float* grad_log_tau, float* grad_radius,
                          float* grad_color_r, float* grad_color_g, float* grad_color_b, float* grad_dir_x,
                          float* grad_dir_y, float* grad_log_aniso, float* tau_scratch, float lambda, int cand_width,
                          int cand_height, int cand_downscale, float lr_pos, float lr_tau, float lr_radius,
                          float lr_color, float lr_dir, float lr_aniso, float beta1, float beta2, float eps, uint32_t t,
                          uint32_t site_count, int width, int height, sadh_stream_t stream);

/* launchAdamUpdate, adam.cu:224-251 */
int sadh_adam_update(SadSite* sites, SadAdamState* adam, const float* grad_pos_x,
                     const float* grad_pos_y, const float* grad_log_tau, const float* grad_radius,
                     const float* grad_color_r, const float* grad_color_g,
                     const float* grad_color_b, const float* grad_dir_x, const float* grad_dir_y,
                     const float* grad_log_aniso, float lr_pos, float lr_tau, float lr_radius,
                     float lr_color, float lr_dir, float lr_aniso, float beta1, float beta2,
                     float eps, uint32_t t, uint32_t site_count, int width, int height,
                     sadh_stream_t stream);

/* Adam followed by zeroing of the ten gradient arrays in the same pass, the
 * fused form of the Metal backend (backends/metal/shaders/encoders/adam_encoder.metal:117-127). */
int sadh_adam_update_clear(SadSite* sites, SadAdamState* adam, float* grad_pos_x,
                           float* grad_pos_y, float* grad_log_tau, float* grad_radius,
                           float* grad_color_r, float* grad_color_g, float* grad_color_b,
                           float* grad_dir_x, float* grad_dir_y, float* grad_log_aniso,
                           float lr_pos, float lr_tau, float lr_radius, float lr_color,
                           float lr_dir, float lr_aniso, float beta1, float beta2, float eps,
                           uint32_t t, uint32_t site_count, int width, int height,
                           sadh_stream_t stream);

/* launchClearGradients, adam.cu:253-270 */
int sadh_clear_gradients(float* grad_pos_x, float* grad_pos_y, float* grad_log_tau,
                         float* grad_radius, float* grad_color_r, float* grad_color_g,
                         float* grad_color_b, float* grad_dir_x, float* grad_dir_y,
                         float* grad_log_aniso, uint32_t count, sadh_stream_t stream);

/* launchClearBuffer, adam.cu:272-281 */
int sadh_clear_buffer(float* buffer, uint32_t count, sadh_stream_t stream);

/* ------------------------------------------------------ densify / prune --- */

/* launchComputeSiteStatsSimple, utils.cu:653-689 (clears the eight arrays first) */
int sadh_compute_site_stats_simple(const uint32_t* cand0, const uint32_t* cand1,
                                   const float* target, const float* mask, const SadSite* sites,
                                   float inv_scale_sq, uint32_t site_count, float* mass,
                                   float* energy, float* err_w, float* err_wx, float* err_wy,
                                   float* err_wxx, float* err_wxy, float* err_wyy, int width,
                                   int height, int cand_width, int cand_height,
                                   sadh_stream_t stream);

/* launchComputeDensifyScorePairs, densify_prune.cu:359-376 */
int sadh_compute_densify_score_pairs(const SadSite* sites, const float* mass, const float* energy,
                                     SadPair* pairs, uint32_t site_count, float min_mass,
                                     float score_alpha, uint32_t pair_count, sadh_stream_t stream);

/* launchComputePruneScorePairs, densify_prune.cu:378-393 */
int sadh_compute_prune_score_pairs(const SadSite* sites, const float* removal_delta,
                                   SadPair* pairs, uint32_t site_count, float delta_norm,
                                   uint32_t pair_count, sadh_stream_t stream);

/* launchRadixSortUInt2, radix_sort.cu:279-336: stable ascending sort of
 * `padded_count` pairs by key; 2 passes (low 16 key bits) when
 * max_key_exclusive <= 2^16, else 4.  scratch: padded_count pairs;
 * hist: 256 * ceil(padded_count / 1024) words; block_sums: ceil(padded_count / 1024) words. */
int sadh_radix_sort_pairs(SadPair* data, SadPair* scratch, uint32_t* hist, uint32_t* block_sums,
                          uint32_t padded_count, uint32_t max_key_exclusive,
                          sadh_stream_t stream);

/* launchWriteSplitIndicesFromSorted, densify_prune.cu:395-407 */
int sadh_write_split_indices(const SadPair* sorted_pairs, uint32_t* indices, uint32_t count,
                             sadh_stream_t stream);

/* launchSplitSites, densify_prune.cu:335-357.  Parents >= current_site_count
 * (sentinel pairs) are skipped. */
int sadh_split_sites(SadSite* sites, SadAdamState* adam, const uint32_t* split_indices,
                     uint32_t num_to_split, const float* mass, const float* err_w,
                     const float* err_wx, const float* err_wy, const float* err_wxx,
                     const float* err_wxy, const float* err_wyy, uint32_t current_site_count,
                     const float* target, int width, int height, sadh_stream_t stream);

/* launchPruneSitesByIndex, densify_prune.cu:409-421 */
int sadh_prune_sites_by_index(SadSite* sites, const uint32_t* indices, uint32_t count,
                              sadh_stream_t stream);

/* ------------------------------------------------------------- metrics --- */

/* launchComputePSNR, utils.cu:597-615: *mse_accum = sum over masked pixels of |r - t|^2 */
int sadh_compute_psnr(const float* rendered, const float* target, const float* mask,
                      float* mse_accum, int width, int height, sadh_stream_t stream);

/* launchComputeSSIM, utils.cu:279-297: *ssim_accum = sum over masked pixels of
 * the SSIM of the pixel's 3x3 luma window (clamped at the borders). */
int sadh_compute_ssim(const float* rendered, const float* target, const float* mask,
                      float* ssim_accum, int width, int height, sadh_stream_t stream);

/* makePackedInferenceSites, backends/metal/sources/utils/packed_inference.swift:86-199,
 * on the device: `quant` (device, one struct) receives the ranges of log_tau,
 * radius and colour over the active sites, `packed` one 16-byte site per input
 * site (inactive ones flagged).  The anisotropy angle uses this library's
 * atan2 (sad_math.h); Swift's is not reproduced bit for bit. */
int sadh_pack_inference_sites(const SadSite* sites, uint32_t site_count, int width, int height,
                              SadInferenceSite* packed, SadInferenceQuant* quant, sadh_stream_t stream);

/* decodePackedSite, backends/metal/shaders/sad_common.metal:71-120: back to
 * Site form (device arrays); an inactive site decodes with pos = (-1, -1). */
int sadh_unpack_inference_sites(const SadInferenceSite* packed, const SadInferenceQuant* quant,
                                uint32_t site_count, int width, int height, SadSite* sites,
                                sadh_stream_t stream);

/* launchCountActiveSites, utils.cu:637-651 */
int sadh_count_active_sites(const SadSite* sites, uint32_t* count, uint32_t site_count,
                            sadh_stream_t stream);

/* launchInitGradientWeighted, densify_prune.cu:5-83,312-334: num_sites sites at
 * the first accepted seeds of the R2 sequence (mask > 0 and mask-weighted Sobel
 * magnitude of the luma > grad_threshold; the reference CLI passes 0.01 * alpha
 * and 256 attempts).  *seed_counter (device, nullable = 0) is the first seed and
 * receives one past the last seed drawn.  The reference's sites race for seeds
 * on one atomic counter; here site i draws after site i-1, so sites come in
 * seed order and the result is deterministic.  A site that draws max_attempts
 * rejected seeds is inactive (pos -1, colour 0.5).  mask may be NULL (all ones). */
int sadh_init_gradient_weighted(SadSite* sites, uint32_t num_sites, uint32_t* seed_counter,
                                const float* target, const float* mask, float grad_threshold,
                                uint32_t max_attempts, float init_log_tau, float init_radius,
                                int width, int height, sadh_stream_t stream);

/* ------------------------------------------------------- fit / render ---- */

/* Training options: the TrainingOptions fields of sad_cuda.cu:528-576 that the
 * fit schedule reads.  init_mode: 0 = sites given by the caller
 * (--init-from-sites), 1 = one site per pixel (--init-per-pixel), 2 = n_sites
 * gradient-weighted sites made on the device (--init-gradient, the reference
 * CLI's default; sad_cuda.cu:2078-2095). */
typedef struct SadhFitOptions {
    int iterations;
    float target_bpp;                 /* <= 0: keep the percentiles below */
    int init_mode;
    int n_sites;                      /* initial sites in `sites_io` (ignored for per-pixel) */
    int max_sites;                    /* <= 0: the reference's capacity rule */
    float prune_percentile;
    int prune_start, prune_freq, prune_end;
    int prune_during_densify;
    int densify_enabled;
    int densify_start, densify_freq, densify_end;
    float densify_percentile, densify_score_alpha;
    int cand_update_freq, cand_update_passes, cand_downscale;
    float cand_radius_scale;
    uint32_t cand_radius_probes, cand_inject_count, cand_hilbert_window, cand_hilbert_probes;
    float lr_scale, lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso;
    float beta1, beta2, eps;
    int log_freq;
    int profile_gradients;            /* 1: time every gradient launch with HIP events */
    float init_gradient_alpha;        /* init_mode 2: threshold = 0.01 * alpha */
    float init_log_tau, init_radius;  /* init_mode 2: initial log_tau and radius */
    float ssim_weight;                /* > 0: loss L2 - w * SSIM, with a render before each gradient
                                         (--ssim-weight, sad_cuda.cu:2307-2330) */
    int ssim_metric;                  /* 1: log the mean SSIM beside the PSNR (--ssim, sad_cuda.cu:2432-2447) */
    int packed_psnr;                  /* 1: after the fit, render the quantised inference sites with the final
                                         candidates and report that PSNR (--packed-psnr, training.swift:676-703) */
    int graph_ticks;                  /* sadh_fit_batch and the server: 1 replays each steady iteration (no
                                         candidate update, densify, prune, log or profiling) as a captured
                                         hipGraph of its gradient and tau/Adam launches; the same results.
                                         No reference counterpart; sadh_fit ignores it */
    int autotune;                     /* sadh_fit_batch and the server: 1 times steady iterations (from the
                                         41st) under each launch schedule that leaves every result unchanged
                                         (128- or 64-slot tile hash, 256- or 64-thread tau blocks) and keeps
                                         the fastest for the rest of the run.  No reference counterpart;
                                         sadh_fit ignores it */
} SadhFitOptions;

typedef struct SadhFitResult {
    float final_psnr, best_psnr;
    int actual_sites, active_estimate, active_count;
    float densify_percentile, prune_percentile;
    int capacity;
    int n_logs, n_densify, n_prune, n_cand_updates;
    double train_seconds;             /* host wall time of the iteration loop */
    double grad_kernel_ms;            /* summed gradient-kernel time (profile_gradients) */
    int grad_launches;
    double grad_bytes;                /* summed algorithmic bytes of those launches */
    float* grad_ms_trace;             /* in: optional host array; out: per-launch gradient ms */
    int grad_ms_trace_cap;            /* in: entries of grad_ms_trace */
    float* final_render;              /* in: optional host [H][W][3]; out: render of the final sites with the
                                         final candidates (sad_cuda.cu:2482-2496) */
    float final_ssim, best_ssim;      /* ssim_metric: mean SSIM at the last log, and the best */
    float* log_ssim;                  /* in: optional host array of log_cap entries; out: the logged SSIMs */
    float packed_psnr;                /* packed_psnr: PSNR of the quantised inference sites */
    float* packed_render;             /* in: optional host [H][W][3]; out: that render (packed_psnr) */
    float* cells_render;              /* in: optional host [H][W][3]; out: sadh_render_cells of the final sites
                                         (--include-debug-mask, training.swift:636-671) */
    float* tau_heatmap;               /* in: optional host [H][W][3]; out: sadh_render_tau_heatmap of the final
                                         sites over their tau range, dot radius 2 (training.swift:641-652) */
} SadhFitResult;

/* Resolve densify/prune percentiles and buffer capacity as trainVoronoi does
 * (sad_cuda.cu:1871-1905, solveTargetBpp :962-1066, planSiteCapacity :1079-1098). */
/* One image of a batched fit.  All images share one SadhFitOptions (its
 * n_sites and init_mode are ignored: each job brings n_sites initial sites). */
typedef struct SadhBatchJob {
    const float* target;              /* host [height][width][3] */
    const float* mask;                /* host [height][width] or NULL */
    int width, height;
    SadSite* sites_io;                /* host: n_sites initial sites in; capacity sites out */
    int n_sites;
    int sites_cap;                    /* entries of sites_io (>= the job's plan capacity) */
    float* log_psnr;                  /* optional PSNR log */
    int log_cap;
    SadhFitResult* result;            /* optional per-job result */
} SadhBatchJob;

typedef struct SadhBatchResult {
    double train_seconds;             /* host wall time of the whole batch's iteration loop */
    double grad_kernel_ms;            /* summed batched-gradient launch time (profile_gradients) */
    int grad_launches;
    double grad_bytes;                /* summed algorithmic bytes of those launches */
    double total_pixels;              /* sum of W*H over the jobs */
    int graph_ticks, graph_captures;  /* graph_ticks: iterations replayed as a graph, graphs captured */
    int tuned_hash_slots, tuned_tau_block; /* autotune: the schedule kept (0, 0 when the run did not tune) */
} SadhBatchResult;

int sadh_fit_plan(const SadhFitOptions* opts, int width, int height, float* densify_percentile,
                  float* prune_percentile, int* capacity, int* pair_count);

/* The training loop of trainVoronoi, sad_cuda.cu:1860-2533, on the current
 * device.  target [H][W][3] and mask [H][W] (NULL = all ones) are host arrays;
 * sites_io is a host array of `sites_cap` sites (>= capacity from
 * sadh_fit_plan) holding the initial sites on entry and the final state on
 * return; log_psnr (nullable) receives up to log_cap logged PSNR values. */
int sadh_fit(const SadhFitOptions* opts, const float* target, const float* mask, int width,
             int height, SadSite* sites_io, int sites_cap, float* log_psnr, int log_cap,
             SadhFitResult* result);

/* Batched fit (no reference counterpart: the reference fits one image per run,
 * sad_cuda.cu:1860-2533).  Every job follows the single-image schedule of
 * sadh_fit; each iteration's VPT pass, gradient and tau/Adam kernels run as
 * one launch over all jobs (a device table of per-image descriptors). */
int sadh_fit_batch(const SadhFitOptions* opts, SadhBatchJob* jobs, int n_jobs, SadhBatchResult* result);

/* Continuous-batching server (no reference counterpart: the reference serves
 * one fit or render per process).  Fit requests (SadhBatchJob, options shared
 * through SadhServerOptions.fit, each with its own iteration counter) and
 * render requests (sites -> image: clear, JFA seed + flood, VPT passes,
 * render) wait in arrival order and are admitted while their projected device
 * memory fits the budget and a fit slot is free.  sadh_server_step admits what
 * fits, advances every running fit by one iteration in batched launches, lets
 * render requests run on their own stream (overlapping the fits), and retires
 * finished requests; sadh_server_poll returns their events.  The host buffers
 * of a request must stay valid until its event. */
typedef struct SadhServerOptions {
    SadhFitOptions fit;               /* shared by every fit request (n_sites, init_mode ignored) */
    double memory_budget;             /* bytes; <= 0: 0.8 x the device's memory */
    int max_fit_jobs;                 /* running fits at most; <= 0: 64 */
    int render_passes;                /* VPT passes of a render request after JFA; <= 0: 16 */
} SadhServerOptions;

typedef struct SadhServerEvent {
    long long id;
    int kind;                         /* 0 fit, 1 render */
    int status;                       /* SADH_OK */
    double submit_s, admit_s, done_s; /* seconds on the server's clock */
} SadhServerEvent;

typedef struct SadhServer* sadh_server_t;

int sadh_server_create(const SadhServerOptions* opts, sadh_server_t* server);
int sadh_server_submit_fit(sadh_server_t server, const SadhBatchJob* job, long long* id);
int sadh_server_submit_render(sadh_server_t server, const SadSite* sites, uint32_t site_count, int width, int height,
                              float* output, long long* id);
int sadh_server_step(sadh_server_t server, int* outstanding);
int sadh_server_poll(sadh_server_t server, SadhServerEvent* events, int cap, int* count);
double sadh_server_time(sadh_server_t server);   /* seconds since creation (the events' clock) */
int sadh_server_destroy(sadh_server_t server);

/* Render mode of renderVoronoiFromSites, sad_cuda.cu:1678-1858: JFA (jfa_rounds
 * rounds of clear + seed + log2 floods) and `cand_passes` VPT passes, then the
 * softmax render.  sites and output are host arrays; timings in ms (nullable);
 * cells_output (nullable host [H][W][3]) also receives sadh_render_cells of the
 * same candidates (--render-cells, backends/metal/sources/rendering/renderer.swift:213-218). */
int sadh_render_sites(const SadSite* sites, uint32_t site_count, int width, int height,
                      int cand_passes, int use_jfa, int jfa_rounds, int cand_downscale,
                      float radius_scale, uint32_t radius_probes, uint32_t inject_count,
                      uint32_t hilbert_window, uint32_t hilbert_probes, float* output,
                      float* cand_ms, float* render_ms, float* cells_output);

/* renderPackedInference, backends/metal/sources/rendering/packed_renderer.swift:15-137
 * (render mode --render-packed): the render job of sadh_render_sites on
 * quantised inference sites (host arrays), decoded on the device first. */
int sadh_render_packed_sites(const SadInferenceSite* packed, const SadInferenceQuant* quant,
                             uint32_t site_count, int width, int height, int cand_passes, int use_jfa,
                             int jfa_rounds, int cand_downscale, float radius_scale,
                             uint32_t radius_probes, uint32_t inject_count, uint32_t hilbert_window,
                             uint32_t hilbert_probes, float* output, float* cand_ms, float* render_ms,
                             float* cells_output);

#ifdef __cplusplus
}
#endif

#endif /* SAD_HIP_H */
