/*
 * fit.cpp - native fit and render drivers of libsad_hip.so.
 *
 * sadh_fit runs the per-iteration schedule of trainVoronoi
 * (backends/cuda/src/sad_cuda.cu:1860-2533) on the HIP kernels of this library:
 *   candidate update every CAND_UPDATE_FREQ iterations (pack + VPT passes),
 *   densify (site stats, score pairs, radix sort, split), tiled gradients,
 *   tau-gradient diffusion (4 Jacobi passes), Adam with fused gradient clear,
 *   prune (score pairs, radix sort, mark), and render + PSNR on log iterations.
 * Schedule semantics kept from the reference:
 *   - a non-densify candidate update runs on a second stream, concurrently with
 *     the gradient pass, which therefore reads the field from before the update;
 *     tau diffusion and later steps read the updated field (sad_cuda.cu:2171-2230, 2342);
 *   - densify iterations update the field synchronously first (sad_cuda.cu:2176-2183);
 *   - split/prune counts come from the host-side active-site estimate, so no
 *     device->host copy is needed except on log iterations.
 * Differences of implementation (same results): the VPT pass reads 32-byte
 * search sites with the per-site key terms precomputed from the fp16-rounded
 * values, bit-identical to the 16-byte pack; Adam clears the gradients in the
 * same pass; the copy of the raw tau gradient is folded into the first
 * diffusion pass.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fit_common.h"

namespace sadh {

/* simulateFinalSites, sad_cuda.cu:904-960 */
static int simulate_final_sites(int init_sites, int max_sites, int iters, bool densify_enabled, int d_start, int d_end,
                                int d_freq, float d_pct, bool prune_during, int p_start, int p_end, int p_freq,
                                float p_pct, int max_split)
{
    int actual = init_sites, active = init_sites;
    int maxc = std::max(max_sites, init_sites);
    int eff_ps = p_start;
    if (densify_enabled && !prune_during && p_start < d_end) eff_ps = d_end;
    for (int it = 0; it < iters; ++it) {
        if (densify_enabled && d_pct > 0.0f && it >= d_start && it <= d_end && it % std::max(1, d_freq) == 0 &&
            actual < maxc) {
            int k = std::min((int)((float)active * d_pct), std::min(maxc - actual, max_split));
            if (k > 0) { actual += k; active += k; }
        }
        if (p_pct > 0.0f && it >= eff_ps && it < p_end && it % std::max(1, p_freq) == 0) {
            int k = std::min((int)((float)active * p_pct), max_split);
            if (k > 0) active = std::max(0, active - k);
        }
    }
    return active;
}

/* solveTargetBpp, sad_cuda.cu:962-1066: pick a common scale of the densify and
 * prune percentiles so the simulated final site count meets target_bpp at 128
 * bits per site (81-point grid, then 20 refinement steps). */
static void solve_target_bpp(float target_bpp, int width, int height, int init_sites, int max_sites, int iters,
                             bool densify_enabled, int d_start, int d_end, int d_freq, float base_d,
                             bool prune_during, int p_start, int p_end, int p_freq, float base_p, int max_split,
                             float* out_d, float* out_p)
{
    const float bits_per_site = 128.0f;
    int target = std::max(1, (int)std::round(target_bpp * (float)(width * height) / bits_per_site));
    float max_base = std::max(base_d, base_p);
    if (max_base <= 0.0f) { *out_d = 0.0f; *out_p = 0.0f; return; }
    const float max_pct = 0.95f;
    float smax = std::min(max_pct / max_base, 50.0f);
    auto eval = [&](float s) {
        float d = densify_enabled ? std::min(max_pct, base_d * s) : 0.0f;
        float p = std::min(max_pct, base_p * s);
        return simulate_final_sites(init_sites, max_sites, iters, densify_enabled, d_start, d_end, d_freq, d, prune_during,
                                    p_start, p_end, p_freq, p, max_split);
    };
    float best_scale = 0.0f;
    int best_err = std::abs(eval(0.0f) - target);
    for (int i = 0; i <= 80; ++i) {
        float s = smax * (float)i / 80.0f;
        int err = std::abs(eval(s) - target);
        if (err < best_err) { best_err = err; best_scale = s; }
    }
    float step = smax / 80.0f;
    for (int i = 0; i < 20; ++i) {
        bool improved = false;
        float s0 = best_scale - step, s1 = best_scale + step;
        if (s0 >= 0.0f) {
            int err = std::abs(eval(s0) - target);
            if (err < best_err) { best_err = err; best_scale = s0; improved = true; }
        }
        if (s1 <= smax) {
            int err = std::abs(eval(s1) - target);
            if (err < best_err) { best_err = err; best_scale = s1; improved = true; }
        }
        if (!improved) step *= 0.5f;
    }
    *out_d = densify_enabled ? std::min(max_pct, base_d * best_scale) : 0.0f;
    *out_p = std::min(max_pct, base_p * best_scale);
}

static int max_sites_capacity(const SadhFitOptions& o, int np, int n_sites)
{
    if (o.max_sites > 0) return o.max_sites;
    if (o.densify_enabled) return std::min(np * 2, std::max(n_sites * 8, 8192));
    return np * 2;
}

int plan(const SadhFitOptions& o, int width, int height, float* d_pct, float* p_pct, int* cap, int* pairs)
{
    if (width <= 0 || height <= 0 || o.iterations < 0) return fail(SADH_ERR_INVALID, "fit: bad image size or iterations");
    int np = width * height;
    int n_sites = o.init_mode == 1 ? np : o.n_sites;
    if (n_sites <= 0) return fail(SADH_ERR_INVALID, "fit: n_sites must be positive");
    float dp = o.densify_percentile, pp = o.prune_percentile;
    if (o.target_bpp > 0.0f) {
        int p_end = o.prune_end > 0 ? o.prune_end : o.iterations - 1;
        solve_target_bpp(o.target_bpp, width, height, n_sites, max_sites_capacity(o, np, n_sites), o.iterations,
                         o.densify_enabled != 0, o.densify_start, o.densify_end, std::max(1, o.densify_freq), dp,
                         o.prune_during_densify != 0, o.prune_start, p_end, std::max(1, o.prune_freq), pp, 65536, &dp,
                         &pp);
    }
    bool needs_pairs = o.densify_enabled || pp > 0.0f;
    int requested = o.densify_enabled ? max_sites_capacity(o, np, n_sites) : n_sites;
    *cap = std::max(n_sites, requested);
    *pairs = needs_pairs ? std::max(1, *cap) : 0;
    *d_pct = dp;
    *p_pct = pp;
    return SADH_OK;
}

static int fit_impl(const SadhFitOptions& o, const float* target_h, const float* mask_h, int width, int height,
                    SadSite* sites_io, int sites_cap, float* log_psnr, int log_cap, SadhFitResult* res)
{
    float d_pct, p_pct;
    int cap, pairs_count;
    SAD_TRY(plan(o, width, height, &d_pct, &p_pct, &cap, &pairs_count));
    if (!target_h || !sites_io) return fail(SADH_ERR_INVALID, "fit: null target or sites");
    if (sites_cap < cap) return fail(SADH_ERR_CAPACITY, "fit: sites_io holds " + std::to_string(sites_cap) +
                                                           " sites, the plan needs " + std::to_string(cap));
    const int np = width * height;
    if (o.init_mode < 0 || o.init_mode > 2) return fail(SADH_ERR_INVALID, "fit: init_mode must be 0, 1 or 2");
    const bool per_pixel = o.init_mode == 1;
    const bool grad_init = o.init_mode == 2;
    const int n0 = per_pixel ? np : o.n_sites;
    const int ds = std::max(1, o.cand_downscale);
    const int cw = std::max(1, (width + ds - 1) / ds), ch = std::max(1, (height + ds - 1) / ds);
    const size_t cells = (size_t)cw * ch;

    std::vector<float> mask(np, 1.0f);
    float mask_sum = (float)np;
    if (mask_h) {
        mask_sum = 0.0f;
        /* loadMask, sad_cuda.cu:207-213: the normaliser counts pixels with a positive mask */
        for (int i = 0; i < np; ++i) { mask[i] = mask_h[i]; mask_sum += mask_h[i] > 0.0f ? 1.0f : 0.0f; }
    }
    std::vector<SadSite> host_sites(cap);
    for (int i = 0; i < cap; ++i) {
        if (i < n0 && !grad_init) host_sites[i] = sites_io[i];
        else { std::memset(&host_sites[i], 0, sizeof(SadSite)); host_sites[i].pos_x = host_sites[i].pos_y = -1.0f; }
    }

    DevArena mem;
    float *d_target, *d_mask, *d_render, *tau_raw, *tau_tmp, *tau_tmp2;
    long long* d_mse;
    SadSite* d_sites;
    SadAdamState* d_adam;
    /* Cross-tile sums in exact fixed point: gradients as lines of FIX_GRAD_STRIDE int64 words per site,
     * removal deltas as one (lo, hi) pair per site, site statistics as lines of FIX_STAT_STRIDE words.  Each
     * tile adds one float partial per site and value as two integer atomics; rounded once to float where
     * they are read, the values do not depend on the order in which tiles finish, so a fit repeats exactly. */
    long long* gacc;
    long long* removal;
    long long* st = nullptr;
    SadPair *pairs = nullptr, *sort_scratch = nullptr;
    uint32_t *sort_hist = nullptr, *split_idx = nullptr, *prune_idx = nullptr, *d_count;
    uint32_t *cA0, *cA1, *cB0, *cB1;
    SadSearchSite* search;
    SAD_TRY(mem.alloc(&d_target, (size_t)np * 3));
    SAD_TRY(mem.alloc(&d_mask, (size_t)np));
    SAD_TRY(mem.alloc(&d_render, (size_t)np * 3));
    SAD_TRY(mem.alloc(&d_mse, 2));
    long long* d_ssim;
    SAD_TRY(mem.alloc(&d_ssim, 2));
    SAD_TRY(mem.alloc(&d_count, 1));
    SAD_TRY(mem.alloc(&d_sites, cap));
    SAD_TRY(mem.alloc(&d_adam, cap));
    SAD_TRY(mem.alloc(&gacc, (size_t)cap * FIX_GRAD_STRIDE));
    SAD_TRY(mem.alloc(&tau_raw, cap));
    SAD_TRY(mem.alloc(&tau_tmp, cap));
    SAD_TRY(mem.alloc(&tau_tmp2, cap));
    uint32_t* tau_nbr;   /* per-site neighbour lists shared by the four tau passes */
    SAD_TRY(mem.alloc(&tau_nbr, (size_t)cap * 8));
    SAD_TRY(mem.alloc(&removal, (size_t)cap * 2));
    if (o.densify_enabled) SAD_TRY(mem.alloc(&st, (size_t)cap * FIX_STAT_STRIDE));
    if (pairs_count > 0) {
        SAD_TRY(mem.alloc(&pairs, pairs_count));
        SAD_TRY(mem.alloc(&sort_scratch, pairs_count));
        SAD_TRY(mem.alloc(&sort_hist, 256u * div_up((uint32_t)pairs_count, 1024u)));
        SAD_TRY(mem.alloc(&split_idx, 65536));
        SAD_TRY(mem.alloc(&prune_idx, 65536));
    }
    SAD_TRY(mem.alloc(&cA0, cells * 4));
    SAD_TRY(mem.alloc(&cA1, cells * 4));
    SAD_TRY(mem.alloc(&cB0, cells * 4));
    SAD_TRY(mem.alloc(&cB1, cells * 4));
    /* An update of several passes on the second stream ping-pongs through a third field, so no pass writes
     * the field the concurrent gradient reads (with two buffers an even pass count would end there). */
    uint32_t *cC0 = nullptr, *cC1 = nullptr;
    if (o.cand_update_passes > 1) {
        SAD_TRY(mem.alloc(&cC0, cells * 4));
        SAD_TRY(mem.alloc(&cC1, cells * 4));
    }
    SAD_TRY(mem.alloc(&search, cap));
    const bool use_hil = o.cand_hilbert_probes > 0 && o.cand_hilbert_window > 0;
    const uint32_t hil_padded = div_up((uint32_t)std::max(1, cap), 1024u) * 1024u;
    uint32_t *h_order = nullptr, *h_pos = nullptr, *h_hist = nullptr;
    SadPair *h_pairs = nullptr, *h_scratch = nullptr;
    if (use_hil) {
        SAD_TRY(mem.alloc(&h_order, cap));
        SAD_TRY(mem.alloc(&h_pos, cap));
        SAD_TRY(mem.alloc(&h_pairs, hil_padded));
        SAD_TRY(mem.alloc(&h_scratch, hil_padded));
        SAD_TRY(mem.alloc(&h_hist, 256u * (hil_padded / 1024u)));
    }

    StreamPair sp;
    HIP_TRY(hipStreamCreateWithFlags(&sp.main, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&sp.cand, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&sp.cand_done, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&sp.sites_ready, hipEventDisableTiming));
    hipStream_t S = sp.main;

    HIP_TRY(hipMemcpyAsync(d_target, target_h, sizeof(float) * 3 * np, hipMemcpyHostToDevice, S));
    HIP_TRY(hipMemcpyAsync(d_mask, mask.data(), sizeof(float) * np, hipMemcpyHostToDevice, S));
    HIP_TRY(hipMemcpyAsync(d_sites, host_sites.data(), sizeof(SadSite) * cap, hipMemcpyHostToDevice, S));
    HIP_TRY(hipMemsetAsync(d_adam, 0, sizeof(SadAdamState) * cap, S));
    HIP_TRY(hipMemsetAsync(gacc, 0, sizeof(long long) * FIX_GRAD_STRIDE * cap, S));
    HIP_TRY(hipMemsetAsync(removal, 0, sizeof(long long) * 2 * cap, S));

    int actual = n0;
    uint32_t n = (uint32_t)actual;
    int active_est = 0;
    if (grad_init) {
        /* sad_cuda.cu:2078-2095; the host copy the reference counts holds n
         * default sites at (0, 0), so its active estimate is n */
        HIP_TRY(hipMemsetAsync(d_count, 0, sizeof(uint32_t), S));
        SAD_TRY(launch_init_gradient(d_sites, n, d_count, d_target, d_mask, 0.01f * o.init_gradient_alpha, 256u,
                                     o.init_log_tau, o.init_radius, width, height, S));
        active_est = n0;
    } else {
        for (int i = 0; i < n0; ++i) active_est += host_sites[i].pos_x >= 0.0f ? 1 : 0;
    }

    SAD_TRY(sadh_init_candidates(cA0, cA1, n, 0u, per_pixel ? 1 : 0, cw, ch, (sadh_stream_t)S));
    SAD_TRY(sadh_jfa_seed(cA0, d_sites, n, cw, ch, ds, (sadh_stream_t)S));
    HIP_TRY(hipEventRecord(sp.sites_ready, S));

    const int md = std::max(width, height);
    const float iss = 1.0f / ((float)md * (float)md);
    const float delta_norm = 1.0f / std::max(1.0f, mask_sum);
    const float lr_pos = o.lr_pos * o.lr_scale, lr_tau = o.lr_tau * o.lr_scale, lr_rad = o.lr_radius * o.lr_scale;
    const float lr_col = o.lr_color * o.lr_scale, lr_dir = o.lr_dir * o.lr_scale, lr_an = o.lr_aniso * o.lr_scale;
    int eff_ps = o.prune_start;
    if (o.densify_enabled && !o.prune_during_densify && o.prune_start < o.densify_end) eff_ps = o.densify_end;
    const int eff_pe = o.prune_end > 0 ? o.prune_end : o.iterations - 1;
    const bool needs_pairs = pairs_count > 0, needs_prune = p_pct > 0.0f;
    uint32_t jump_index = 0;
    bool hil_ready = false;
    uint32_t hil_count = 0;
    float best = 0.0f, final_psnr = 0.0f, best_ssim = 0.0f, final_ssim = 0.0f;
    int n_logs = 0, n_dens = 0, n_prune = 0, n_upd = 0;
    uint32_t active_count = 0;

    std::vector<hipEvent_t> gev;
    std::vector<double> gbytes;
    struct EventGuard {
        std::vector<hipEvent_t>& v;
        ~EventGuard() { for (auto e : v) (void)hipEventDestroy(e); }
    } gev_guard{gev};

    auto t0 = std::chrono::steady_clock::now();
    for (int iter = 0; iter < o.iterations; ++iter) {
        bool upd = false;
        int passes = 0;
        if (!(per_pixel && iter < eff_ps)) {
            upd = o.cand_update_freq > 0 && iter % o.cand_update_freq == 0;
            passes = upd ? std::max(1, o.cand_update_passes) : 0;
        }
        const int desired = std::max(0, (int)((float)active_est * d_pct));
        const bool densify = o.densify_enabled && iter >= o.densify_start && iter <= o.densify_end &&
                             iter % std::max(1, o.densify_freq) == 0 && actual < cap && desired > 0;
        uint32_t *nA0 = cA0, *nA1 = cA1, *nB0 = cB0, *nB1 = cB1;   /* field after this iteration's update */
        bool pending = false;
        if (upd && passes > 0) {
            n_upd++;
            hipStream_t U = densify ? S : sp.cand;
            if (!densify) HIP_TRY(hipStreamWaitEvent(sp.cand, sp.sites_ready, 0));
            if (use_hil && (!hil_ready || hil_count != n)) {
                int bits = sad_hilbert_bits(width, height);
                uint32_t pc = std::min(div_up(n, 1024u) * 1024u, hil_padded);
                uint32_t mk = bits >= 16 ? 0xFFFFFFFFu : (1u << (2 * bits));
                SAD_TRY(launch_hilbert_pairs(d_sites, h_pairs, n, pc, width, height, bits, U));
                SAD_TRY(launch_radix_sort(h_pairs, h_scratch, h_hist, pc, mk, U));
                SAD_TRY(launch_hilbert_order(h_pairs, h_order, h_pos, n, U));
                hil_ready = true;
                hil_count = n;
            }
            SAD_TRY(launch_search_sites(d_sites, search, n, iss, true, U));
            for (int p = 0; p < passes; ++p) {
                uint32_t step = sad_pack_jump_step(jump_index, cw, ch);
                SAD_TRY(launch_vpt_search(nA0, nA1, nB0, nB1, search, n, step, iss, jump_index >> 16,
                                          o.cand_radius_scale, o.cand_radius_probes, o.cand_inject_count,
                                          use_hil ? h_order : nullptr, use_hil ? h_pos : nullptr,
                                          use_hil ? o.cand_hilbert_probes : 0u, use_hil ? o.cand_hilbert_window : 0u,
                                          ds, width, height, cw, ch, U));
                jump_index++;
                if (!densify && p == 0 && passes > 1) {
                    /* A (read by the gradient) -> B, then B <-> C; A becomes the spare */
                    uint32_t* a0 = nA0; uint32_t* a1 = nA1;
                    nA0 = nB0; nA1 = nB1;
                    nB0 = cC0; nB1 = cC1;
                    cC0 = a0; cC1 = a1;
                } else {
                    std::swap(nA0, nB0);
                    std::swap(nA1, nB1);
                }
            }
            if (densify) {
                cA0 = nA0; cA1 = nA1; cB0 = nB0; cB1 = nB1;
            } else {
                HIP_TRY(hipEventRecord(sp.cand_done, sp.cand));
                pending = true;
            }
        }
        if (densify && o.densify_enabled && needs_pairs) {
            n_dens++;
            SAD_TRY(launch_clear_fix(st, (size_t)n * FIX_STAT_STRIDE, S));
            SAD_TRY(launch_site_stats_fix(cA0, cA1, d_target, d_mask, d_sites, iss, n, st, width, height, cw, ch, S));
            SAD_TRY(launch_densify_pairs_fix(d_sites, st, pairs, n, 1.0f, o.densify_score_alpha,
                                             (uint32_t)pairs_count, S));
            /* pairs at index >= n hold the (0xFFFFFFFF, 0xFFFFFFFF) sentinel: the stable sort of the
             * whole buffer is the sorted first n pairs followed by them, so only those n are sorted */
            SAD_TRY(launch_radix_sort(pairs, sort_scratch, sort_hist, std::min(n, (uint32_t)pairs_count), 0xFFFFFFFFu, S));
            int k = std::min(desired, std::min(cap - actual, 65536));
            if (k > 0) {
                SAD_TRY(launch_write_indices(pairs, split_idx, (uint32_t)k, S));
                SAD_TRY(launch_split_fix(d_sites, d_adam, split_idx, (uint32_t)k, st, (uint32_t)actual, d_target,
                                         width, height, S));
                actual += k;
                n = (uint32_t)actual;
                active_est += k;
            }
        }
        const float blend = (float)iter / (float)std::max(1, o.iterations);
        const float lambda = 0.05f * (0.1f + 0.9f * blend);
        const bool prune = p_pct > 0.0f && iter >= eff_ps && iter < eff_pe && iter % std::max(1, o.prune_freq) == 0 &&
                           needs_pairs && needs_prune;
        if (prune) SAD_TRY(launch_clear_fix(removal, 2u * (size_t)n, S));

        if (o.profile_gradients) {
            hipEvent_t a, b;
            HIP_TRY(hipEventCreate(&a));
            gev.push_back(a);
            HIP_TRY(hipEventCreate(&b));
            gev.push_back(b);
            HIP_TRY(hipEventRecord(a, S));
        }
        /* the SSIM loss term reads a render of the current state (sad_cuda.cu:2307-2318) */
        if (o.ssim_weight > 0.0f) SAD_TRY(launch_render(cA0, cA1, d_render, d_sites, iss, n, width, height, cw, ch, S));
        /* the tail's neighbour lists are written by extra blocks of the gradient launch when the tail uses the
         * gradient's field; after a candidate update on the second stream the first tau pass writes them */
        SAD_TRY(launch_gradients_fix(cA0, cA1, d_target, d_mask, gacc, d_sites, iss, n, removal, prune, width,
                                     height, cw, ch, S, d_render, o.ssim_weight > 0.0f ? o.ssim_weight : 0.0f,
                                     pending ? nullptr : tau_nbr, lambda, ds));
        if (o.profile_gradients) {
            HIP_TRY(hipEventRecord(gev.back(), S));
            gbytes.push_back(48.0 * np + 120.0 * n + (prune ? 8.0 * n : 0.0));
        }

        if (pending) {
            HIP_TRY(hipStreamWaitEvent(S, sp.cand_done, 0));
            cA0 = nA0; cA1 = nA1; cB0 = nB0; cB1 = nB1;
        }
        /* 4 tau passes (raw saved by the first), the last fused with Adam + clear of the gradient lines */
        SAD_TRY(launch_tau_adam_fix(cA0, cA1, gacc, tau_raw, tau_tmp, tau_tmp2, lambda, cw, ch, ds, d_sites, d_adam,
                                    lr_pos, lr_tau, lr_rad, lr_col, lr_dir, lr_an, o.beta1, o.beta2, o.eps,
                                    (uint32_t)(iter + 1), n, width, height, S, tau_nbr, !pending));

        if (prune) {
            n_prune++;
            SAD_TRY(launch_prune_pairs_fix(d_sites, removal, pairs, n, delta_norm, (uint32_t)pairs_count, S));
            SAD_TRY(launch_radix_sort(pairs, sort_scratch, sort_hist, std::min(n, (uint32_t)pairs_count), 0xFFFFFFFFu, S));
            int k = std::min(std::max(0, (int)((float)active_est * p_pct)), 65536);
            if (k > 0) {
                SAD_TRY(launch_write_indices(pairs, prune_idx, (uint32_t)k, S));
                SAD_TRY(launch_prune_sites(d_sites, prune_idx, (uint32_t)k, S));
                active_est = std::max(0, active_est - k);
            }
        }
        /* the next iteration's candidate update (second stream) reads the sites Adam just wrote; an event
         * record drains the queue for a few microseconds, so record it only before such an iteration */
        if (o.cand_update_freq > 0 && (iter + 1) % o.cand_update_freq == 0)
            HIP_TRY(hipEventRecord(sp.sites_ready, S));

        if (iter % std::max(1, o.log_freq) == 0 || iter == o.iterations - 1) {
            SAD_TRY(launch_render(cA0, cA1, d_render, d_sites, iss, n, width, height, cw, ch, S));
            SAD_TRY(launch_sse_fix(d_render, d_target, d_mask, d_mse, width, height, S));
            if (o.ssim_metric) SAD_TRY(launch_ssim_fix(d_render, d_target, d_mask, d_ssim, width, height, S));
            SAD_TRY(launch_count_active(d_sites, d_count, n, S));
            long long sse_fx[2] = {0, 0}, ssim_fx[2] = {0, 0};
            HIP_TRY(hipMemcpyAsync(sse_fx, d_mse, sizeof(sse_fx), hipMemcpyDeviceToHost, S));
            if (o.ssim_metric) HIP_TRY(hipMemcpyAsync(ssim_fx, d_ssim, sizeof(ssim_fx), hipMemcpyDeviceToHost, S));
            HIP_TRY(hipMemcpyAsync(&active_count, d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, S));
            HIP_TRY(hipStreamSynchronize(S));
            const float sse = fix_value_host(sse_fx), ssim_sum = fix_value_host(ssim_fx);
            float mse = sse / (std::max(1.0f, mask_sum) * 3.0f);
            float psnr = mse > 0.0f ? 20.0f * log10f(1.0f / std::sqrt(mse)) : 100.0f;
            if (log_psnr && n_logs < log_cap) log_psnr[n_logs] = psnr;
            if (o.ssim_metric) {
                const float ssim = ssim_sum / std::max(1.0f, mask_sum);   /* sad_cuda.cu:2443-2448 */
                if (res && res->log_ssim && n_logs < log_cap) res->log_ssim[n_logs] = ssim;
                final_ssim = ssim;
                best_ssim = std::max(best_ssim, ssim);
            }
            n_logs++;
            final_psnr = psnr;
            best = std::max(best, psnr);
        }
    }
    HIP_TRY(hipStreamSynchronize(S));
    HIP_TRY(hipStreamSynchronize(sp.cand));
    auto t1 = std::chrono::steady_clock::now();

    double gms = 0.0, gb = 0.0;
    for (size_t i = 0; i + 1 < gev.size(); i += 2) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, gev[i], gev[i + 1]));
        gms += ms;
        if (res && res->grad_ms_trace && (int)(i / 2) < res->grad_ms_trace_cap) res->grad_ms_trace[i / 2] = ms;
    }
    for (double b : gbytes) gb += b;

    HIP_TRY(hipMemcpy(host_sites.data(), d_sites, sizeof(SadSite) * cap, hipMemcpyDeviceToHost));
    std::memcpy(sites_io, host_sites.data(), sizeof(SadSite) * cap);
    float packed_psnr = 0.0f;
    if (o.packed_psnr) {
        /* training.swift:676-703: quantise the final sites and render them with the final candidates */
        SadInferenceSite* d_packed;
        SadInferenceQuant* d_quant;
        SadSite* d_dec;
        SAD_TRY(mem.alloc(&d_packed, std::max(n, 1u)));
        SAD_TRY(mem.alloc(&d_quant, 1));
        SAD_TRY(mem.alloc(&d_dec, std::max(n, 1u)));
        SAD_TRY(launch_pack_inference(d_sites, n, width, height, d_packed, d_quant, S));
        SAD_TRY(launch_unpack_inference(d_packed, d_quant, n, width, height, d_dec, S));
        SAD_TRY(launch_render(cA0, cA1, d_render, d_dec, iss, n, width, height, cw, ch, S));
        SAD_TRY(launch_sse_fix(d_render, d_target, d_mask, d_mse, width, height, S));
        long long sse_fx[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(sse_fx, d_mse, sizeof(sse_fx), hipMemcpyDeviceToHost, S));
        if (res && res->packed_render)
            HIP_TRY(hipMemcpyAsync(res->packed_render, d_render, sizeof(float) * 3 * np, hipMemcpyDeviceToHost, S));
        HIP_TRY(hipStreamSynchronize(S));
        const float mse = fix_value_host(sse_fx) / (std::max(1.0f, mask_sum) * 3.0f);
        packed_psnr = mse > 0.0f ? 20.0f * log10f(1.0f / std::sqrt(mse)) : 100.0f;
    }
    if (res && res->final_render) {
        SAD_TRY(launch_render(cA0, cA1, d_render, d_sites, iss, n, width, height, cw, ch, S));
        HIP_TRY(hipMemcpyAsync(res->final_render, d_render, sizeof(float) * 3 * np, hipMemcpyDeviceToHost, S));
        HIP_TRY(hipStreamSynchronize(S));
    }
    SAD_TRY(debug_renders(cA0, cA1, d_sites, host_sites.data(), n, iss, width, height, cw, ch, d_render, res, S));
    if (res) {
        res->final_psnr = final_psnr;
        res->best_psnr = best;
        res->final_ssim = final_ssim;
        res->best_ssim = best_ssim;
        res->packed_psnr = packed_psnr;
        res->actual_sites = actual;
        res->active_estimate = active_est;
        res->active_count = (int)active_count;
        res->densify_percentile = d_pct;
        res->prune_percentile = p_pct;
        res->capacity = cap;
        res->n_logs = n_logs;
        res->n_densify = n_dens;
        res->n_prune = n_prune;
        res->n_cand_updates = n_upd;
        res->train_seconds = std::chrono::duration<double>(t1 - t0).count();
        res->grad_kernel_ms = gms;
        res->grad_launches = (int)(gev.size() / 2);
        res->grad_bytes = gb;
    }
    return SADH_OK;
}

void tau_range(const SadSite* sites, uint32_t n, float* tmin, float* tmean, float* tmax)
{
    float mn = INFINITY, mx = -INFINITY, sum = 0.0f;
    uint32_t count = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!(sites[i].pos_x >= 0.0f)) continue;
        const float tau = sad_expf(sites[i].log_tau);
        mn = std::min(mn, tau);
        mx = std::max(mx, tau);
        sum += tau;
        count++;
    }
    if (count == 0) mn = mx = 0.0f;
    *tmin = mn;
    *tmax = mx;
    *tmean = count > 0 ? sum / (float)count : 0.0f;
}

int debug_renders(const uint32_t* c0, const uint32_t* c1, const SadSite* d_sites, const SadSite* host_sites,
                  uint32_t n, float iss, int width, int height, int cw, int ch, float* d_scratch, SadhFitResult* r,
                  hipStream_t stream)
{
    if (!r || !(r->cells_render || r->tau_heatmap)) return SADH_OK;
    const size_t img = sizeof(float) * 3 * (size_t)width * height;
    if (r->cells_render) {
        SAD_TRY(launch_render_cells(c0, c1, d_scratch, d_sites, iss, n, width, height, cw, ch, stream));
        HIP_TRY(hipMemcpyAsync(r->cells_render, d_scratch, img, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    }
    if (r->tau_heatmap) {
        float tmin, tmean, tmax;
        tau_range(host_sites, n, &tmin, &tmean, &tmax);
        SAD_TRY(launch_tau_heatmap(c0, c1, d_scratch, d_sites, n, tmin, tmean, tmax, 2.0f, width, height, cw, ch,
                                   stream));
        HIP_TRY(hipMemcpyAsync(r->tau_heatmap, d_scratch, img, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    }
    return SADH_OK;
}

/* sites_h, or packed_h + quant_h (quantised inference sites, decoded on the device first) */
static int render_impl(const SadSite* sites_h, const SadInferenceSite* packed_h, const SadInferenceQuant* quant_h,
                       uint32_t n, int width, int height, int cand_passes, int use_jfa, int jfa_rounds,
                       int cand_downscale, float radius_scale, uint32_t radius_probes, uint32_t inject_count,
                       uint32_t hil_window, uint32_t hil_probes, float* out_h, float* cand_ms, float* render_ms,
                       float* cells_h)
{
    if (!(sites_h || (packed_h && quant_h)) || !out_h || n == 0 || width <= 0 || height <= 0)
        return fail(SADH_ERR_INVALID, "render_sites: bad arguments");
    const int ds = std::max(1, cand_downscale);
    const int cw = std::max(1, (width + ds - 1) / ds), ch = std::max(1, (height + ds - 1) / ds);
    const size_t cells = (size_t)cw * ch;
    const size_t np = (size_t)width * height;
    DevArena mem;
    SadSite* d_sites;
    SadSearchSite* search;
    uint32_t *a0, *a1, *b0, *b1;
    float* d_out;
    SAD_TRY(mem.alloc(&d_sites, n));
    SAD_TRY(mem.alloc(&search, n));
    SAD_TRY(mem.alloc(&a0, cells * 4));
    SAD_TRY(mem.alloc(&a1, cells * 4));
    SAD_TRY(mem.alloc(&b0, cells * 4));
    SAD_TRY(mem.alloc(&b1, cells * 4));
    SAD_TRY(mem.alloc(&d_out, np * 3));
    const bool use_hil = hil_probes > 0 && hil_window > 0;
    const uint32_t padded = div_up(n, 1024u) * 1024u;
    uint32_t *h_order = nullptr, *h_pos = nullptr, *h_hist = nullptr;
    SadPair *h_pairs = nullptr, *h_scratch = nullptr;
    if (use_hil) {
        SAD_TRY(mem.alloc(&h_order, n));
        SAD_TRY(mem.alloc(&h_pos, n));
        SAD_TRY(mem.alloc(&h_pairs, padded));
        SAD_TRY(mem.alloc(&h_scratch, padded));
        SAD_TRY(mem.alloc(&h_hist, 256u * (padded / 1024u)));
    }
    StreamPair sp;
    HIP_TRY(hipStreamCreateWithFlags(&sp.main, hipStreamNonBlocking));
    hipStream_t S = sp.main;
    hipEvent_t e0, e1, e2;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventCreate(&e2));
    struct Ev { hipEvent_t a, b, c; ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); } } evg{e0, e1, e2};
    if (sites_h) {
        HIP_TRY(hipMemcpyAsync(d_sites, sites_h, sizeof(SadSite) * n, hipMemcpyHostToDevice, S));
    } else {
        /* renderPackedInference, packed_renderer.swift:15-137: the kernels there decode the
         * 16-byte sites at every use; the decode is a pure function of the bits, so one
         * decode pass into Site form feeds the same kernels as the full-precision job */
        SadInferenceSite* d_packed;
        SadInferenceQuant* d_quant;
        SAD_TRY(mem.alloc(&d_packed, n));
        SAD_TRY(mem.alloc(&d_quant, 1));
        HIP_TRY(hipMemcpyAsync(d_packed, packed_h, sizeof(SadInferenceSite) * n, hipMemcpyHostToDevice, S));
        HIP_TRY(hipMemcpyAsync(d_quant, quant_h, sizeof(SadInferenceQuant), hipMemcpyHostToDevice, S));
        SAD_TRY(launch_unpack_inference(d_packed, d_quant, n, width, height, d_sites, S));
    }

    const int md = std::max(width, height);
    const float iss = 1.0f / ((float)md * (float)md);
    uint32_t jump_index = 0;
    if (!use_jfa) {
        SAD_TRY(sadh_init_candidates(a0, a1, n, 12345u, 0, cw, ch, (sadh_stream_t)S));
        SAD_TRY(sadh_jfa_seed(a0, d_sites, n, cw, ch, ds, (sadh_stream_t)S));
    }
    const int passes = std::max(0, cand_passes);
    const int rounds = use_jfa ? std::max(1, jfa_rounds) : 1;
    const int base_passes = passes / rounds, remainder = passes % rounds;
    HIP_TRY(hipEventRecord(e0, S));
    if (use_hil) {
        int bits = sad_hilbert_bits(width, height);
        uint32_t mk = bits >= 16 ? 0xFFFFFFFFu : (1u << (2 * bits));
        SAD_TRY(launch_hilbert_pairs(d_sites, h_pairs, n, padded, width, height, bits, S));
        SAD_TRY(launch_radix_sort(h_pairs, h_scratch, h_hist, padded, mk, S));
        SAD_TRY(launch_hilbert_order(h_pairs, h_order, h_pos, n, S));
    }
    for (int round = 0; round < rounds; ++round) {
        if (use_jfa) {
            /* runJFA, sad_cuda.cu:1616-1655: plane 1 is the flood ping-pong buffer */
            SAD_TRY(sadh_clear_candidates(a0, a1, cw, ch, (sadh_stream_t)S));
            SAD_TRY(sadh_jfa_seed(a0, d_sites, n, cw, ch, ds, (sadh_stream_t)S));
            SAD_TRY(launch_search_sites(d_sites, search, n, iss, false, S));
            int step = 1;
            while (step < std::max(cw, ch)) step <<= 1;
            uint32_t *in = a0, *outp = a1;
            for (int s = step / 2; s >= 1; s /= 2) {
                SAD_TRY(launch_jfa_flood_search(in, outp, search, n, (uint32_t)s, iss, cw, ch, ds, width, height, S));
                std::swap(in, outp);
            }
            if (in != a0) HIP_TRY(hipMemcpyAsync(a0, in, cells * 16, hipMemcpyDeviceToDevice, S));
        }
        int pt = base_passes + (round < remainder ? 1 : 0);
        if (pt > 0) {
            SAD_TRY(launch_search_sites(d_sites, search, n, iss, true, S));
            for (int p = 0; p < pt; ++p) {
                uint32_t stp = sad_pack_jump_step(jump_index, cw, ch);
                SAD_TRY(launch_vpt_search(a0, a1, b0, b1, search, n, stp, iss, jump_index >> 16, radius_scale,
                                          radius_probes, inject_count, h_order, h_pos, use_hil ? hil_probes : 0u,
                                          use_hil ? hil_window : 0u, ds, width, height, cw, ch, S));
                jump_index++;
                std::swap(a0, b0);
                std::swap(a1, b1);
            }
        }
    }
    HIP_TRY(hipEventRecord(e1, S));
    SAD_TRY(launch_render(a0, a1, d_out, d_sites, iss, n, width, height, cw, ch, S));
    HIP_TRY(hipEventRecord(e2, S));
    HIP_TRY(hipMemcpyAsync(out_h, d_out, sizeof(float) * 3 * np, hipMemcpyDeviceToHost, S));
    HIP_TRY(hipStreamSynchronize(S));
    if (cells_h) {
        /* --render-cells (renderer.swift:213-218): the same candidates, hashed colours */
        SAD_TRY(launch_render_cells(a0, a1, d_out, d_sites, iss, n, width, height, cw, ch, S));
        HIP_TRY(hipMemcpyAsync(cells_h, d_out, sizeof(float) * 3 * np, hipMemcpyDeviceToHost, S));
        HIP_TRY(hipStreamSynchronize(S));
    }
    float ms = 0.0f;
    if (cand_ms) { HIP_TRY(hipEventElapsedTime(&ms, e0, e1)); *cand_ms = ms; }
    if (render_ms) { HIP_TRY(hipEventElapsedTime(&ms, e1, e2)); *render_ms = ms; }
    return SADH_OK;
}

}  // namespace sadh

using namespace sadh;

extern "C" {

int sadh_fit_plan(const SadhFitOptions* opts, int width, int height, float* densify_percentile,
                  float* prune_percentile, int* capacity, int* pair_count)
{
    SADH_REQUIRE(opts && densify_percentile && prune_percentile && capacity && pair_count, "fit_plan: null pointer");
    return plan(*opts, width, height, densify_percentile, prune_percentile, capacity, pair_count);
}

int sadh_fit(const SadhFitOptions* opts, const float* target, const float* mask, int width, int height,
             SadSite* sites_io, int sites_cap, float* log_psnr, int log_cap, SadhFitResult* result)
{
    SADH_REQUIRE(opts, "fit: null options");
    return fit_impl(*opts, target, mask, width, height, sites_io, sites_cap, log_psnr, log_cap, result);
}

int sadh_render_sites(const SadSite* sites, uint32_t site_count, int width, int height, int cand_passes, int use_jfa,
                      int jfa_rounds, int cand_downscale, float radius_scale, uint32_t radius_probes,
                      uint32_t inject_count, uint32_t hilbert_window, uint32_t hilbert_probes, float* output,
                      float* cand_ms, float* render_ms, float* cells_output)
{
    return render_impl(sites, nullptr, nullptr, site_count, width, height, cand_passes, use_jfa, jfa_rounds,
                       cand_downscale, radius_scale, radius_probes, inject_count, hilbert_window, hilbert_probes,
                       output, cand_ms, render_ms, cells_output);
}

int sadh_render_packed_sites(const SadInferenceSite* packed, const SadInferenceQuant* quant, uint32_t site_count,
                             int width, int height, int cand_passes, int use_jfa, int jfa_rounds, int cand_downscale,
                             float radius_scale, uint32_t radius_probes, uint32_t inject_count,
                             uint32_t hilbert_window, uint32_t hilbert_probes, float* output, float* cand_ms,
                             float* render_ms, float* cells_output)
{
    if (!packed || !quant) return fail(SADH_ERR_INVALID, "render_packed_sites: null packed sites or quant");
    return render_impl(nullptr, packed, quant, site_count, width, height, cand_passes, use_jfa, jfa_rounds,
                       cand_downscale, radius_scale, radius_probes, inject_count, hilbert_window, hilbert_probes,
                       output, cand_ms, render_ms, cells_output);
}

}  // extern "C"
