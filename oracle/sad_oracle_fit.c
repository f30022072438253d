/*
 * sad_oracle_fit.c - CPU restatement of the reference's training schedule
 * (TEST INFRASTRUCTURE: the checker and CPU baseline, never the product path).
 *
 * orc_fit follows trainVoronoi, backends/cuda/src/sad_cuda.cu:1860-2533, step by
 * step, calling the kernel restatements in sad_oracle.c.  The candidate update
 * that the reference runs on a side stream is applied with the same data
 * semantics: the gradient pass of an update iteration reads the field from
 * before the update (unless the iteration densifies, in which case the update is
 * synchronous), tau diffusion and everything after it read the updated field
 * (sad_cuda.cu:2171-2230, 2342); with several passes the gradient still reads
 * the field from before the update (the reference's side stream would be
 * racing its own passes there).  Gradients, removal deltas, densify
 * statistics and the logged squared error are summed as the HIP library sums
 * them (the tile-ordered fixed-point sums of sad_oracle.c), so a fit here and
 * the library's fit on the same inputs end with the same bits.
 * solveTargetBpp / simulateFinalSites
 * (sad_cuda.cu:904-1066) and planSiteCapacity (sad_cuda.cu:1079-1098) are
 * restated for the site budget.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#include "../sad_amd/csrc/sad_math.h"

#define ORC_API __attribute__((visibility("default")))

/* kernels from sad_oracle.c */
void orc_render(const uint32_t*, const uint32_t*, float*, const SadSite*, float, uint32_t, int, int, int, int);
void orc_init_candidates(uint32_t*, uint32_t*, uint32_t, uint32_t, int, int, int);
void orc_jfa_seed(uint32_t*, const SadSite*, uint32_t, int, int, int);
void orc_pack_sites(const SadSite*, SadPackedSite*, uint32_t);
void orc_update_candidates(const uint32_t*, const uint32_t*, uint32_t*, uint32_t*, const SadPackedSite*,
                           uint32_t, uint32_t, float, uint32_t, float, uint32_t, uint32_t, const uint32_t*,
                           const uint32_t*, uint32_t, uint32_t, int, int, int, int, int);
void orc_tau_diffuse(const uint32_t*, const uint32_t*, const SadSite*, const float*, const float*, float*,
                     uint32_t, float, int, int, int);
void orc_adam(SadSite*, SadAdamState*, float* const[10], float, float, float, float, float, float, float,
              float, float, uint32_t, uint32_t, int, int);
void orc_densify_pairs(const SadSite*, const float*, const float*, SadPair*, uint32_t, float, float, uint32_t);
uint32_t orc_init_gradient_weighted(SadSite*, uint32_t, uint32_t, const float*, const float*, float, uint32_t, float,
                                    float, int, int);
void orc_prune_pairs(const SadSite*, const float*, SadPair*, uint32_t, float, uint32_t);
void orc_radix_sort(SadPair*, uint32_t, uint32_t);
void orc_write_indices(const SadPair*, uint32_t*, uint32_t);
void orc_prune_sites(SadSite*, const uint32_t*, uint32_t);
void orc_split_sites(SadSite*, SadAdamState*, const uint32_t*, uint32_t, const float*, const float*,
                     const float*, const float*, const float*, const float*, const float*, uint32_t,
                     const float*, int, int);
void orc_hilbert_pairs(const SadSite*, SadPair*, uint32_t, uint32_t, int, int, int);
void orc_hilbert_order(const SadPair*, uint32_t*, uint32_t*, uint32_t);
void orc_gradients_fix(const uint32_t*, const uint32_t*, const float*, const float*, const SadSite*, float, uint32_t,
                       int, int, int, int, int, const float*, float, int64_t*, int64_t*, int32_t*);
void orc_site_stats_fix(const uint32_t*, const uint32_t*, const float*, const float*, const SadSite*, float, uint32_t,
                        int, int, int, int, int64_t*, int32_t*);
float orc_sse_fix(const float*, const float*, const float*, int, int);
float orc_fix_value(int64_t, int64_t);
double orc_ssim_sum(const float*, const float*, const float*, int, int, float*);

/* Option block: the fields of TrainingOptions (sad_cuda.cu:528-576) that the
 * fit schedule reads.  init_mode 0 = sites given by the caller, 1 = per-pixel,
 * 2 = gradient-weighted (n_sites made by orc_init_gradient_weighted). */
typedef struct OrcFitConfig {
    int iterations;
    float target_bpp;
    int init_mode;
    int n_sites;
    int max_sites;
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
    float init_gradient_alpha, init_log_tau, init_radius;
    float ssim_weight;             /* > 0: loss L2 - w SSIM, render before each gradient */
    int ssim_metric;               /* 1: log the mean SSIM */
} OrcFitConfig;

typedef struct OrcFitResult {
    float final_psnr, best_psnr;
    int actual_sites, active_estimate;
    float densify_percentile, prune_percentile;
    int capacity;
    int n_logs;
    int n_densify, n_prune, n_cand_updates;
    float final_ssim, best_ssim;
} OrcFitResult;

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* simulateFinalSites, sad_cuda.cu:904-960 */
ORC_API int orc_simulate_final(int init_sites, int max_sites, int iters, int densify_enabled,
                              int d_start, int d_end, int d_freq, float d_pct, int prune_during,
                              int p_start, int p_end, int p_freq, float p_pct, int max_split)
{
    int actual = init_sites, active = init_sites;
    int maxc = imax(max_sites, init_sites);
    int eff_ps = p_start;
    if (densify_enabled && !prune_during && p_start < d_end) eff_ps = d_end;
    if (iters <= 0) return active;
    for (int it = 0; it < iters; ++it) {
        if (densify_enabled && d_pct > 0.0f && it >= d_start && it <= d_end &&
            it % imax(1, d_freq) == 0 && actual < maxc) {
            int desired = (int)((float)active * d_pct);
            int k = imin(desired, imin(maxc - actual, max_split));
            if (k > 0) { actual += k; active += k; }
        }
        if (p_pct > 0.0f && it >= eff_ps && it < p_end && it % imax(1, p_freq) == 0) {
            int desired = (int)((float)active * p_pct);
            int k = imin(desired, max_split);
            if (k > 0) active = imax(0, active - k);
        }
    }
    return active;
}

/* solveTargetBpp, sad_cuda.cu:962-1066 */
ORC_API void orc_solve_target_bpp(float target_bpp, int width, int height, int init_sites, int max_sites,
                                  int iters, int densify_enabled, int d_start, int d_end, int d_freq,
                                  float base_d, int prune_during, int p_start, int p_end, int p_freq,
                                  float base_p, int max_split, float* out_d, float* out_p, int* out_sites)
{
    const float bits_per_site = 16.0f * 8.0f;
    int target = imax(1, (int)roundf(target_bpp * (float)(width * height) / bits_per_site));
    float max_base = fmaxf(base_d, base_p);
#define EVAL(dp, pp) orc_simulate_final(init_sites, max_sites, iters, densify_enabled, d_start, d_end, \
                                        d_freq, (dp), prune_during, p_start, p_end, p_freq, (pp), max_split)
    if (max_base <= 0.0f) {
        *out_d = 0.0f; *out_p = 0.0f; *out_sites = EVAL(0.0f, 0.0f);
        return;
    }
    const float max_pct = 0.95f;
    float smax = max_pct / max_base;
    if (smax > 50.0f) smax = 50.0f;
#define EVAL_SCALE(s) EVAL(densify_enabled ? fminf(max_pct, base_d * (s)) : 0.0f, fminf(max_pct, base_p * (s)))
    float best_scale = 0.0f;
    int best_sites = EVAL_SCALE(0.0f);
    int best_err = abs(best_sites - target);
    const int samples = 80;
    for (int i = 0; i <= samples; ++i) {
        float s = smax * (float)i / (float)samples;
        int st = EVAL_SCALE(s);
        int err = abs(st - target);
        if (err < best_err) { best_err = err; best_scale = s; best_sites = st; }
    }
    float step = smax / (float)samples;
    for (int i = 0; i < 20; ++i) {
        int improved = 0;
        float s0 = best_scale - step, s1 = best_scale + step;
        if (s0 >= 0.0f) {
            int st = EVAL_SCALE(s0);
            int err = abs(st - target);
            if (err < best_err) { best_err = err; best_scale = s0; best_sites = st; improved = 1; }
        }
        if (s1 <= smax) {
            int st = EVAL_SCALE(s1);
            int err = abs(st - target);
            if (err < best_err) { best_err = err; best_scale = s1; best_sites = st; improved = 1; }
        }
        if (!improved) step *= 0.5f;
    }
    *out_d = densify_enabled ? fminf(max_pct, base_d * best_scale) : 0.0f;
    *out_p = fminf(max_pct, base_p * best_scale);
    *out_sites = best_sites;
#undef EVAL_SCALE
#undef EVAL
}

/* Resolve the fit's densify/prune percentiles and buffer capacity exactly as
 * trainVoronoi does before allocating (sad_cuda.cu:1871-1905, 1079-1098). */
ORC_API void orc_fit_plan(const OrcFitConfig* cfg, int width, int height, float* d_pct, float* p_pct,
                          int* capacity, int* pairs_count)
{
    int np = width * height;
    int n_sites = cfg->init_mode == 1 ? np : cfg->n_sites;
    float dp = cfg->densify_percentile, pp = cfg->prune_percentile;
    if (cfg->target_bpp > 0.0f) {
        int p_end = cfg->prune_end > 0 ? cfg->prune_end : cfg->iterations - 1;
        int maxc;
        if (cfg->max_sites > 0) maxc = cfg->max_sites;
        else if (cfg->densify_enabled) maxc = imin(np * 2, imax(n_sites * 8, 8192));
        else maxc = np * 2;
        int fs;
        orc_solve_target_bpp(cfg->target_bpp, width, height, n_sites, maxc, cfg->iterations,
                             cfg->densify_enabled, cfg->densify_start, cfg->densify_end,
                             imax(1, cfg->densify_freq), dp, cfg->prune_during_densify,
                             cfg->prune_start, p_end, imax(1, cfg->prune_freq), pp, 65536, &dp, &pp, &fs);
    }
    int needs_pairs = cfg->densify_enabled || pp > 0.0f;
    int maxcap;
    if (cfg->max_sites > 0) maxcap = cfg->max_sites;
    else if (cfg->densify_enabled) maxcap = imin(np * 2, imax(n_sites * 8, 8192));
    else maxcap = np * 2;
    int requested = cfg->densify_enabled ? maxcap : n_sites;
    int cap = imax(n_sites, requested);
    *d_pct = dp;
    *p_pct = pp;
    *capacity = cap;
    *pairs_count = needs_pairs ? imax(1, cap) : 0;
}

/* trainVoronoi hot loop, sad_cuda.cu:2158-2479.
 * sites_io: capacity-sized array; on entry the first n_sites hold the initial
 * sites, on exit all capacity entries hold the final state.
 * log_psnr (nullable, log_cap entries) receives each logged PSNR. */
ORC_API int orc_fit(const OrcFitConfig* cfg, const float* target, const float* mask_in,
                    int width, int height, SadSite* sites_io, int sites_cap,
                    float* log_psnr, int log_cap, OrcFitResult* res, float* log_ssim)
{
    int np = width * height;
    float d_pct, p_pct;
    int cap, pairs_count;
    orc_fit_plan(cfg, width, height, &d_pct, &p_pct, &cap, &pairs_count);
    if (sites_cap < cap) return -1;
    int per_pixel = cfg->init_mode == 1;
    int n0 = per_pixel ? np : cfg->n_sites;
    int ds = imax(1, cfg->cand_downscale);
    int cw = imax(1, (width + ds - 1) / ds), ch = imax(1, (height + ds - 1) / ds);
    size_t cells = (size_t)cw * ch;

    float* mask = (float*)malloc(sizeof(float) * np);
    float mask_sum = 0.0f;
    if (mask_in) {
        memcpy(mask, mask_in, sizeof(float) * np);
        for (int i = 0; i < np; ++i) mask_sum += mask[i] > 0.0f ? 1.0f : 0.0f;   /* sad_cuda.cu:207-213 */
    } else {
        for (int i = 0; i < np; ++i) mask[i] = 1.0f;
        mask_sum = (float)np;
    }
    SadSite* sites = sites_io;
    for (int i = n0; i < cap; ++i) {
        memset(&sites[i], 0, sizeof(SadSite));
        sites[i].pos_x = -1.0f; sites[i].pos_y = -1.0f;
    }
    SadAdamState* adam = (SadAdamState*)calloc(cap, sizeof(SadAdamState));
    float* grads[10];
    for (int j = 0; j < 10; ++j) grads[j] = (float*)calloc(cap, sizeof(float));
    float* tau_raw = (float*)calloc(cap, sizeof(float));
    float* tau_tmp = (float*)calloc(cap, sizeof(float));
    float* removal = (float*)calloc(cap, sizeof(float));
    int64_t* gfix = (int64_t*)malloc(sizeof(int64_t) * 20 * (size_t)cap);
    int64_t* rfix = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)cap);
    int64_t* sfix = (int64_t*)malloc(sizeof(int64_t) * 16 * (size_t)cap);
    float* st[8];
    for (int j = 0; j < 8; ++j) st[j] = (float*)calloc(cap, sizeof(float));
    SadPair* pairs = pairs_count ? (SadPair*)malloc(sizeof(SadPair) * pairs_count) : NULL;
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * 65536);
    uint32_t* cA0 = (uint32_t*)malloc(cells * 16), *cA1 = (uint32_t*)malloc(cells * 16);
    uint32_t* cB0 = (uint32_t*)malloc(cells * 16), *cB1 = (uint32_t*)malloc(cells * 16);
    /* the field the gradient reads while a several-pass update runs */
    uint32_t* cP0 = (uint32_t*)malloc(cells * 16), *cP1 = (uint32_t*)malloc(cells * 16);
    SadPackedSite* packed = (SadPackedSite*)malloc(sizeof(SadPackedSite) * cap);
    int use_hil = cfg->cand_hilbert_probes > 0 && cfg->cand_hilbert_window > 0;
    uint32_t hil_padded = ((uint32_t)imax(1, cap) + 1023u) / 1024u * 1024u;
    SadPair* hpairs = (SadPair*)malloc(sizeof(SadPair) * hil_padded);
    uint32_t* horder = (uint32_t*)calloc(cap, sizeof(uint32_t));
    uint32_t* hpos = (uint32_t*)calloc(cap, sizeof(uint32_t));
    int hil_ready = 0;
    uint32_t hil_count = 0;
    float* render = (float*)malloc(sizeof(float) * 3 * np);

    int actual = n0;
    uint32_t n = (uint32_t)actual;
    int active_est = 0;
    if (cfg->init_mode == 2) {
        /* sad_cuda.cu:2078-2095: threshold 0.01 * alpha, 256 attempts; the
         * reference counts its n default-constructed host sites as active */
        orc_init_gradient_weighted(sites, n, 0u, target, mask, 0.01f * cfg->init_gradient_alpha, 256u,
                                   cfg->init_log_tau, cfg->init_radius, width, height);
        active_est = n0;
    } else {
        for (int i = 0; i < n0; ++i) if (sites[i].pos_x >= 0.0f) active_est++;
    }

    orc_init_candidates(cA0, cA1, n, 0u, per_pixel, cw, ch);
    orc_jfa_seed(cA0, sites, n, cw, ch, ds);

    int md = imax(width, height);
    float iss = 1.0f / ((float)md * (float)md);
    float delta_norm = 1.0f / fmaxf(1.0f, mask_sum);
    const int tau_passes = 4;
    const float tau_lambda0 = 0.05f;
    float lr_pos = cfg->lr_pos * cfg->lr_scale, lr_tau = cfg->lr_tau * cfg->lr_scale;
    float lr_rad = cfg->lr_radius * cfg->lr_scale, lr_col = cfg->lr_color * cfg->lr_scale;
    float lr_dir = cfg->lr_dir * cfg->lr_scale, lr_an = cfg->lr_aniso * cfg->lr_scale;
    int eff_ps = cfg->prune_start;
    if (cfg->densify_enabled && !cfg->prune_during_densify && cfg->prune_start < cfg->densify_end)
        eff_ps = cfg->densify_end;
    int eff_pe = cfg->prune_end > 0 ? cfg->prune_end : cfg->iterations - 1;
    int needs_pairs = pairs_count > 0;
    int needs_prune = p_pct > 0.0f;
    uint32_t jump_index = 0;
    float best = 0.0f, final_psnr = 0.0f, best_ssim = 0.0f, final_ssim = 0.0f;
    int n_logs = 0, n_dens = 0, n_prune = 0, n_upd = 0;

    for (int iter = 0; iter < cfg->iterations; ++iter) {
        int upd = 0, passes = 0;
        if (!(per_pixel && iter < eff_ps)) {
            upd = cfg->cand_update_freq > 0 ? (iter % cfg->cand_update_freq == 0) : 0;
            passes = upd ? imax(1, cfg->cand_update_passes) : 0;
        }
        int desired = imax(0, (int)((float)active_est * d_pct));
        int densify = cfg->densify_enabled && iter >= cfg->densify_start && iter <= cfg->densify_end &&
                      (iter % imax(1, cfg->densify_freq) == 0) && actual < cap && desired > 0;
        uint32_t* gA0 = cA0;             /* field the gradient pass reads */
        uint32_t* gA1 = cA1;
        if (upd && passes > 0) {
            n_upd++;
            if (use_hil && (!hil_ready || hil_count != n)) {
                int bits = sad_hilbert_bits(width, height);
                uint32_t pc = (n + 1023u) / 1024u * 1024u;
                if (pc > hil_padded) pc = hil_padded;
                orc_hilbert_pairs(sites, hpairs, n, pc, width, height, bits);
                uint32_t mk = (bits >= 16) ? 0xFFFFFFFFu : (1u << (2 * bits));
                orc_radix_sort(hpairs, pc, mk);
                orc_hilbert_order(hpairs, horder, hpos, n);
                hil_ready = 1;
                hil_count = n;
            }
            orc_pack_sites(sites, packed, n);
            if (passes > 1 && !densify) {
                memcpy(cP0, cA0, cells * 16);
                memcpy(cP1, cA1, cells * 16);
                gA0 = cP0;
                gA1 = cP1;
            }
            for (int p = 0; p < passes; ++p) {
                uint32_t step = sad_pack_jump_step(jump_index, cw, ch);
                orc_update_candidates(cA0, cA1, cB0, cB1, packed, n, step, iss, jump_index >> 16,
                                      cfg->cand_radius_scale, cfg->cand_radius_probes, cfg->cand_inject_count,
                                      horder, hpos, use_hil ? cfg->cand_hilbert_probes : 0u,
                                      use_hil ? cfg->cand_hilbert_window : 0u, ds, width, height, cw, ch);
                jump_index++;
                uint32_t* t0 = cA0; cA0 = cB0; cB0 = t0;
                uint32_t* t1 = cA1; cA1 = cB1; cB1 = t1;
            }
            if (densify) { gA0 = cA0; gA1 = cA1; }
        }
        if (densify && cfg->densify_enabled && needs_pairs) {
            n_dens++;
            memset(sfix, 0, sizeof(int64_t) * 16 * n);
            orc_site_stats_fix(cA0, cA1, target, mask, sites, iss, n, width, height, cw, ch, sfix, NULL);
            for (uint32_t i = 0; i < n; ++i)
                for (int j = 0; j < 8; ++j) st[j][i] = orc_fix_value(sfix[16 * (size_t)i + 2 * j], sfix[16 * (size_t)i + 2 * j + 1]);
            orc_densify_pairs(sites, st[0], st[1], pairs, n, 1.0f, cfg->densify_score_alpha, (uint32_t)pairs_count);
            orc_radix_sort(pairs, (uint32_t)pairs_count, 0xFFFFFFFFu);
            if (desired > 0 && actual < cap) {
                int k = imin(desired, imin(cap - actual, 65536));
                if (k > 0) {
                    orc_write_indices(pairs, idx, (uint32_t)k);
                    orc_split_sites(sites, adam, idx, (uint32_t)k, st[0], st[2], st[3], st[4], st[5],
                                    st[6], st[7], (uint32_t)actual, target, width, height);
                    actual += k;
                    n = (uint32_t)actual;
                    active_est += k;
                }
            }
        }
        int prune = p_pct > 0.0f && iter >= eff_ps && iter < eff_pe &&
                    (iter % imax(1, cfg->prune_freq) == 0) && needs_pairs && needs_prune;
        if (prune) memset(removal, 0, sizeof(float) * n);

        memset(gfix, 0, sizeof(int64_t) * 20 * n);
        if (prune) memset(rfix, 0, sizeof(int64_t) * 2 * n);
        if (cfg->ssim_weight > 0.0f) {
            /* sad_cuda.cu:2307-2330: render the current state, then the gradient reads it */
            orc_render(gA0, gA1, render, sites, iss, n, width, height, cw, ch);
        }
        orc_gradients_fix(gA0, gA1, target, mask, sites, iss, n, prune, width, height, cw, ch,
                          cfg->ssim_weight > 0.0f ? render : NULL, cfg->ssim_weight, gfix, prune ? rfix : NULL, NULL);
        for (uint32_t i = 0; i < n; ++i) {
            for (int j = 0; j < 10; ++j)
                grads[j][i] += orc_fix_value(gfix[20 * (size_t)i + 2 * j], gfix[20 * (size_t)i + 2 * j + 1]);
            if (prune) removal[i] += orc_fix_value(rfix[2 * (size_t)i], rfix[2 * (size_t)i + 1]);
        }

        memcpy(tau_raw, grads[2], sizeof(float) * n);
        float blend = (float)iter / (float)imax(1, cfg->iterations);
        float lambda = tau_lambda0 * (0.1f + 0.9f * blend);
        float* tin = grads[2];
        float* tout = tau_tmp;
        for (int p = 0; p < tau_passes; ++p) {
            orc_tau_diffuse(cA0, cA1, sites, tau_raw, tin, tout, n, lambda, cw, ch, ds);
            float* t = tin; tin = tout; tout = t;
        }

        orc_adam(sites, adam, grads, lr_pos, lr_tau, lr_rad, lr_col, lr_dir, lr_an,
                 cfg->beta1, cfg->beta2, cfg->eps, (uint32_t)(iter + 1), n, width, height);
        for (int j = 0; j < 10; ++j) memset(grads[j], 0, sizeof(float) * n);

        if (prune) {
            n_prune++;
            orc_prune_pairs(sites, removal, pairs, n, delta_norm, (uint32_t)pairs_count);
            orc_radix_sort(pairs, (uint32_t)pairs_count, 0xFFFFFFFFu);
            int dp = imax(0, (int)((float)active_est * p_pct));
            int k = imin(dp, 65536);
            if (k > 0) {
                orc_write_indices(pairs, idx, (uint32_t)k);
                orc_prune_sites(sites, idx, (uint32_t)k);
                active_est = imax(0, active_est - k);
            }
        }

        if (iter % imax(1, cfg->log_freq) == 0 || iter == cfg->iterations - 1) {
            orc_render(cA0, cA1, render, sites, iss, n, width, height, cw, ch);
            float sse = orc_sse_fix(render, target, mask, width, height);
            float mse = sse / (fmaxf(1.0f, mask_sum) * 3.0f);
            float psnr = mse > 0.0f ? 20.0f * log10f(1.0f / sqrtf(mse)) : 100.0f;
            if (log_psnr && n_logs < log_cap) log_psnr[n_logs] = psnr;
            if (cfg->ssim_metric) {
                float ssim = (float)orc_ssim_sum(render, target, mask, width, height, NULL) / fmaxf(1.0f, mask_sum);
                if (log_ssim && n_logs < log_cap) log_ssim[n_logs] = ssim;
                final_ssim = ssim;
                if (ssim > best_ssim) best_ssim = ssim;
            }
            n_logs++;
            final_psnr = psnr;
            if (psnr > best) best = psnr;
        }
    }

    if (res) {
        res->final_psnr = final_psnr;
        res->best_psnr = best;
        res->final_ssim = final_ssim;
        res->best_ssim = best_ssim;
        res->actual_sites = actual;
        res->active_estimate = active_est;
        res->densify_percentile = d_pct;
        res->prune_percentile = p_pct;
        res->capacity = cap;
        res->n_logs = n_logs;
        res->n_densify = n_dens;
        res->n_prune = n_prune;
        res->n_cand_updates = n_upd;
    }
    free(mask); free(adam);
    for (int j = 0; j < 10; ++j) free(grads[j]);
    for (int j = 0; j < 8; ++j) free(st[j]);
    free(tau_raw); free(tau_tmp); free(removal); free(gfix); free(rfix); free(sfix); free(pairs); free(idx);
    free(cA0); free(cA1); free(cB0); free(cB1); free(cP0); free(cP1); free(packed); free(hpairs); free(horder); free(hpos); free(render);
    return 0;
}
