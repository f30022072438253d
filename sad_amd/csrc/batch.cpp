/*
 * batch.cpp - batched fit: many images, one shared option set, advanced in
 * lock step (the reference fits one image per process, sad_cuda.cu:1860-2533).
 *
 * Every iteration issues one launch per kernel for the whole batch: the VPT
 * pass (with its search-site preparation), the tiled gradient and the four tau
 * passes (the last fused with Adam + clear) each run over the concatenation of
 * all jobs' tiles or sites, finding their job through the device descriptor
 * table (batch.h).  The rare iterations keep per-job launches: densify (site
 * statistics, score pairs, sort, split), prune, Hilbert tables and the log
 * render + PSNR.
 *
 * Each job follows exactly the schedule of the single-image fit (fit.cpp),
 * including the overlap semantics: on a candidate-update iteration a job that
 * does not densify computes its gradient from the field before the update
 * (the VPT runs on a second stream), a job that densifies updates its field
 * first.  Which field is current is a per-job flag in the descriptor, flipped
 * in stream order, so a mixed batch keeps both behaviours.
 */
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "fit_common.h"

namespace sadh {

namespace {

struct JobRun {
    SadhBatchJob* io = nullptr;
    float d_pct = 0.0f, p_pct = 0.0f;
    int cap = 0, pairs_count = 0;
    int width = 0, height = 0, np = 0, ds = 1, cw = 0, ch = 0;
    size_t cells = 0;
    float mask_sum = 0.0f, iss = 0.0f, delta_norm = 0.0f;
    DevArena mem;
    float *d_target = nullptr, *d_mask = nullptr, *d_render = nullptr, *d_mse = nullptr;
    float *tau_raw = nullptr, *tau_tmp = nullptr, *tau_tmp2 = nullptr, *removal = nullptr;
    SadSite* d_sites = nullptr;
    SadAdamState* d_adam = nullptr;
    float* gacc = nullptr;
    float* st[8] = {};
    SadPair *pairs = nullptr, *sort_scratch = nullptr;
    uint32_t *sort_hist = nullptr, *split_idx = nullptr, *prune_idx = nullptr, *d_count = nullptr;
    uint32_t *fa0 = nullptr, *fa1 = nullptr, *fb0 = nullptr, *fb1 = nullptr;
    SadSearchSite* search = nullptr;
    uint32_t *h_order = nullptr, *h_pos = nullptr, *h_hist = nullptr;
    SadPair *h_pairs = nullptr, *h_scratch = nullptr;
    uint32_t hil_padded = 0;
    /* host state */
    int actual = 0, active_est = 0;
    uint32_t n = 0, cur = 0, hil_count = 0;
    bool hil_ready = false;
    float best = 0.0f, final_psnr = 0.0f;
    int n_logs = 0, n_dens = 0, n_prune = 0, n_upd = 0;
    uint32_t active_count = 0;

    uint32_t* c0() const { return cur ? fb0 : fa0; }
    uint32_t* c1() const { return cur ? fb1 : fa1; }
};

int setup_job(const SadhFitOptions& base, JobRun& J, hipStream_t S, bool use_hil)
{
    SadhBatchJob& io = *J.io;
    SadhFitOptions o = base;
    o.n_sites = io.n_sites;
    o.init_mode = 0;
    J.width = io.width;
    J.height = io.height;
    SAD_TRY(plan(o, J.width, J.height, &J.d_pct, &J.p_pct, &J.cap, &J.pairs_count));
    if (!io.target || !io.sites_io) return fail(SADH_ERR_INVALID, "fit_batch: null target or sites");
    if (io.sites_cap < J.cap)
        return fail(SADH_ERR_CAPACITY, "fit_batch: a job's sites_io holds " + std::to_string(io.sites_cap) +
                                           " sites, its plan needs " + std::to_string(J.cap));
    J.np = J.width * J.height;
    J.ds = std::max(1, o.cand_downscale);
    J.cw = std::max(1, (J.width + J.ds - 1) / J.ds);
    J.ch = std::max(1, (J.height + J.ds - 1) / J.ds);
    J.cells = (size_t)J.cw * J.ch;
    const int md = std::max(J.width, J.height);
    J.iss = 1.0f / ((float)md * (float)md);

    std::vector<float> mask(J.np, 1.0f);
    J.mask_sum = (float)J.np;
    if (io.mask) {
        J.mask_sum = 0.0f;
        for (int i = 0; i < J.np; ++i) { mask[i] = io.mask[i]; J.mask_sum += io.mask[i]; }
    }
    J.delta_norm = 1.0f / std::max(1.0f, J.mask_sum);
    std::vector<SadSite> hs(J.cap);
    for (int i = 0; i < J.cap; ++i) {
        if (i < io.n_sites) hs[i] = io.sites_io[i];
        else { std::memset(&hs[i], 0, sizeof(SadSite)); hs[i].pos_x = hs[i].pos_y = -1.0f; }
    }
    const int cap = J.cap;
    SAD_TRY(J.mem.alloc(&J.d_target, (size_t)J.np * 3));
    SAD_TRY(J.mem.alloc(&J.d_mask, (size_t)J.np));
    SAD_TRY(J.mem.alloc(&J.d_render, (size_t)J.np * 3));
    SAD_TRY(J.mem.alloc(&J.d_mse, 1));
    SAD_TRY(J.mem.alloc(&J.d_count, 1));
    SAD_TRY(J.mem.alloc(&J.d_sites, cap));
    SAD_TRY(J.mem.alloc(&J.d_adam, cap));
    SAD_TRY(J.mem.alloc(&J.gacc, (size_t)cap * 16));
    SAD_TRY(J.mem.alloc(&J.tau_raw, cap));
    SAD_TRY(J.mem.alloc(&J.tau_tmp, cap));
    SAD_TRY(J.mem.alloc(&J.tau_tmp2, cap));
    SAD_TRY(J.mem.alloc(&J.removal, cap));
    if (o.densify_enabled) for (int j = 0; j < 8; ++j) SAD_TRY(J.mem.alloc(&J.st[j], cap));
    if (J.pairs_count > 0) {
        SAD_TRY(J.mem.alloc(&J.pairs, J.pairs_count));
        SAD_TRY(J.mem.alloc(&J.sort_scratch, J.pairs_count));
        SAD_TRY(J.mem.alloc(&J.sort_hist, 256u * div_up((uint32_t)J.pairs_count, 1024u)));
        SAD_TRY(J.mem.alloc(&J.split_idx, 65536));
        SAD_TRY(J.mem.alloc(&J.prune_idx, 65536));
    }
    SAD_TRY(J.mem.alloc(&J.fa0, J.cells * 4));
    SAD_TRY(J.mem.alloc(&J.fa1, J.cells * 4));
    SAD_TRY(J.mem.alloc(&J.fb0, J.cells * 4));
    SAD_TRY(J.mem.alloc(&J.fb1, J.cells * 4));
    SAD_TRY(J.mem.alloc(&J.search, cap));
    if (use_hil) {
        J.hil_padded = div_up((uint32_t)std::max(1, cap), 1024u) * 1024u;
        SAD_TRY(J.mem.alloc(&J.h_order, cap));
        SAD_TRY(J.mem.alloc(&J.h_pos, cap));
        SAD_TRY(J.mem.alloc(&J.h_pairs, J.hil_padded));
        SAD_TRY(J.mem.alloc(&J.h_scratch, J.hil_padded));
        SAD_TRY(J.mem.alloc(&J.h_hist, 256u * (J.hil_padded / 1024u)));
    }
    HIP_TRY(hipMemcpyAsync(J.d_target, io.target, sizeof(float) * 3 * J.np, hipMemcpyHostToDevice, S));
    HIP_TRY(hipMemcpyAsync(J.d_mask, mask.data(), sizeof(float) * J.np, hipMemcpyHostToDevice, S));
    HIP_TRY(hipMemcpyAsync(J.d_sites, hs.data(), sizeof(SadSite) * cap, hipMemcpyHostToDevice, S));
    HIP_TRY(hipMemsetAsync(J.d_adam, 0, sizeof(SadAdamState) * cap, S));
    HIP_TRY(hipMemsetAsync(J.gacc, 0, sizeof(float) * 16 * cap, S));
    HIP_TRY(hipMemsetAsync(J.removal, 0, sizeof(float) * cap, S));
    J.actual = io.n_sites;
    J.n = (uint32_t)J.actual;
    J.active_est = 0;
    for (int i = 0; i < io.n_sites; ++i) J.active_est += hs[i].pos_x >= 0.0f ? 1 : 0;
    J.cur = 0;
    SAD_TRY(sadh_init_candidates(J.fa0, J.fa1, J.n, 0u, 0, J.cw, J.ch, (sadh_stream_t)S));
    SAD_TRY(sadh_jfa_seed(J.fa0, J.d_sites, J.n, J.cw, J.ch, J.ds, (sadh_stream_t)S));
    /* the host vectors above are read by the asynchronous copies */
    HIP_TRY(hipStreamSynchronize(S));
    return SADH_OK;
}

void describe(const JobRun& J, BatchJob& d, uint32_t tile_begin, uint32_t ctile_begin, uint32_t site_begin)
{
    std::memset(&d, 0, sizeof(d));
    d.target = J.d_target;
    d.mask = J.d_mask;
    d.width = J.width;
    d.height = J.height;
    d.cw = J.cw;
    d.ch = J.ch;
    d.ds = J.ds;
    d.iss = J.iss;
    d.inv_scale = sqrtf(J.iss);
    d.norm = 1.0f / (float)(J.width * J.height);
    d.maxx = (float)(J.width - 1);
    d.maxy = (float)(J.height - 1);
    d.fa0 = J.fa0; d.fa1 = J.fa1; d.fb0 = J.fb0; d.fb1 = J.fb1;
    d.sites = J.d_sites;
    d.adam = J.d_adam;
    d.search = J.search;
    d.gacc = J.gacc;
    d.tau_raw = J.tau_raw;
    d.tau_tmp = J.tau_tmp;
    d.tau_tmp2 = J.tau_tmp2;
    d.removal = J.removal;
    d.h_order = J.h_order;
    d.h_pos = J.h_pos;
    d.n = J.n;
    d.cur = J.cur;
    d.cap = (uint32_t)J.cap;
    d.tiles_x = div_up((uint32_t)J.width, 16u);
    d.ntiles = d.tiles_x * div_up((uint32_t)J.height, 16u);
    d.tile_begin = tile_begin;
    d.ctiles_x = div_up((uint32_t)J.cw, 16u);
    d.cntiles = d.ctiles_x * div_up((uint32_t)J.ch, 16u);
    d.ctile_begin = ctile_begin;
    d.site_begin = site_begin;
}

int flip_jobs(BatchJob* d_jobs, const std::vector<uint32_t>& which, hipStream_t S)
{
    if (which.empty()) return SADH_OK;
    std::vector<uint32_t> keep(which.size(), 0xFFFFFFFFu), one(which.size(), 1u);
    return launch_update_jobs(d_jobs, which.data(), keep.data(), one.data(), (uint32_t)which.size(), S);
}

}  // namespace

static int fit_batch_impl(const SadhFitOptions& o, SadhBatchJob* ios, int n_jobs, SadhBatchResult* bres)
{
    if (n_jobs <= 0 || !ios) return fail(SADH_ERR_INVALID, "fit_batch: no jobs");
    if (o.init_mode != 0) return fail(SADH_ERR_UNSUPPORTED, "fit_batch: only caller-provided initial sites");
    const bool use_hil = o.cand_hilbert_probes > 0 && o.cand_hilbert_window > 0;
    StreamPair sp;
    HIP_TRY(hipStreamCreateWithFlags(&sp.main, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&sp.cand, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&sp.cand_done, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&sp.sites_ready, hipEventDisableTiming));
    hipStream_t S = sp.main;

    std::vector<std::unique_ptr<JobRun>> jobs;
    for (int j = 0; j < n_jobs; ++j) {
        jobs.emplace_back(new JobRun());
        jobs.back()->io = &ios[j];
        int rc = setup_job(o, *jobs.back(), S, use_hil);
        if (rc) return rc;
    }
    /* descriptor table */
    std::vector<BatchJob> desc(n_jobs);
    uint32_t total_tiles = 0, total_ctiles = 0, total_sites = 0;
    for (int j = 0; j < n_jobs; ++j) {
        describe(*jobs[j], desc[j], total_tiles, total_ctiles, total_sites);
        total_tiles += desc[j].ntiles;
        total_ctiles += desc[j].cntiles;
        total_sites += desc[j].cap;
    }
    DevArena tmem;
    BatchJob* d_jobs;
    SAD_TRY(tmem.alloc(&d_jobs, (size_t)n_jobs));
    HIP_TRY(hipMemcpy(d_jobs, desc.data(), sizeof(BatchJob) * n_jobs, hipMemcpyHostToDevice));
    HIP_TRY(hipEventRecord(sp.sites_ready, S));

    const float lr_pos = o.lr_pos * o.lr_scale, lr_tau = o.lr_tau * o.lr_scale, lr_rad = o.lr_radius * o.lr_scale;
    const float lr_col = o.lr_color * o.lr_scale, lr_dir = o.lr_dir * o.lr_scale, lr_an = o.lr_aniso * o.lr_scale;
    int eff_ps = o.prune_start;
    if (o.densify_enabled && !o.prune_during_densify && o.prune_start < o.densify_end) eff_ps = o.densify_end;
    const int eff_pe = o.prune_end > 0 ? o.prune_end : o.iterations - 1;
    uint32_t jump_index = 0;

    std::vector<hipEvent_t> gev;
    std::vector<double> gbytes;
    struct EventGuard {
        std::vector<hipEvent_t>& v;
        ~EventGuard() { for (auto e : v) (void)hipEventDestroy(e); }
    } gev_guard{gev};
    std::vector<uint32_t> upd_idx, upd_n, upd_flip, late_flip;
    std::vector<char> dens(n_jobs), prune(n_jobs);

    auto t0 = std::chrono::steady_clock::now();
    for (int iter = 0; iter < o.iterations; ++iter) {
        const bool upd = o.cand_update_freq > 0 && iter % o.cand_update_freq == 0;
        const int passes = upd ? std::max(1, o.cand_update_passes) : 0;
        const bool dens_iter = o.densify_enabled && iter >= o.densify_start && iter <= o.densify_end &&
                               iter % std::max(1, o.densify_freq) == 0;
        const bool prune_iter = iter >= eff_ps && iter < eff_pe && iter % std::max(1, o.prune_freq) == 0;
        bool any_dens = false, any_prune = false;
        for (int j = 0; j < n_jobs; ++j) {
            JobRun& J = *jobs[j];
            const int desired = std::max(0, (int)((float)J.active_est * J.d_pct));
            dens[j] = dens_iter && J.actual < J.cap && desired > 0;
            prune[j] = prune_iter && J.p_pct > 0.0f && J.pairs_count > 0;
            any_dens = any_dens || dens[j];
            any_prune = any_prune || prune[j];
        }
        bool pending = false;
        late_flip.clear();
        if (passes > 0) {
            hipStream_t U = any_dens ? S : sp.cand;
            if (!any_dens) HIP_TRY(hipStreamWaitEvent(sp.cand, sp.sites_ready, 0));
            for (int j = 0; j < n_jobs; ++j) {
                JobRun& J = *jobs[j];
                J.n_upd++;
                if (use_hil && (!J.hil_ready || J.hil_count != J.n)) {
                    int bits = sad_hilbert_bits(J.width, J.height);
                    uint32_t pc = std::min(div_up(J.n, 1024u) * 1024u, J.hil_padded);
                    uint32_t mk = bits >= 16 ? 0xFFFFFFFFu : (1u << (2 * bits));
                    SAD_TRY(launch_hilbert_pairs(J.d_sites, J.h_pairs, J.n, pc, J.width, J.height, bits, U));
                    SAD_TRY(launch_radix_sort(J.h_pairs, J.h_scratch, J.h_hist, pc, mk, U));
                    SAD_TRY(launch_hilbert_order(J.h_pairs, J.h_order, J.h_pos, J.n, U));
                    J.hil_ready = true;
                    J.hil_count = J.n;
                }
            }
            for (int p = 0; p < passes; ++p) {
                BatchStep stp = {};
                stp.jump_index = jump_index;
                stp.vpt_parity = (uint32_t)(p & 1);
                stp.radius_scale = o.cand_radius_scale;
                stp.radius_probes = o.cand_radius_probes;
                stp.inject_count = o.cand_inject_count;
                stp.hil_probes = use_hil ? o.cand_hilbert_probes : 0u;
                stp.hil_window = use_hil ? o.cand_hilbert_window : 0u;
                SAD_TRY(launch_vpt_batch(d_jobs, (uint32_t)n_jobs, total_ctiles, total_sites, stp, U));
                jump_index++;
            }
            /* an odd number of passes leaves the result in each job's other field: densifying jobs
             * switch to it now, the others after their gradient (which reads the field before the update) */
            if (passes & 1) {
                std::vector<uint32_t> now;
                for (int j = 0; j < n_jobs; ++j) ((any_dens && dens[j]) ? now : late_flip).push_back((uint32_t)j);
                SAD_TRY(flip_jobs(d_jobs, now, S));
                for (uint32_t j : now) jobs[j]->cur ^= 1u;
            }
            if (!any_dens) {
                HIP_TRY(hipEventRecord(sp.cand_done, sp.cand));
                pending = true;
            }
        }
        /* densify (per job, rare) */
        upd_idx.clear(); upd_n.clear(); upd_flip.clear();
        for (int j = 0; j < n_jobs; ++j) {
            JobRun& J = *jobs[j];
            if (!dens[j] || J.pairs_count <= 0) continue;
            J.n_dens++;
            const int desired = std::max(0, (int)((float)J.active_est * J.d_pct));
            SAD_TRY(launch_clear(J.st, 8, J.n, S));
            SAD_TRY(launch_site_stats(J.c0(), J.c1(), J.d_target, J.d_mask, J.d_sites, J.iss, J.n, J.st, J.width,
                                      J.height, J.cw, J.ch, S));
            SAD_TRY(launch_densify_pairs(J.d_sites, J.st[0], J.st[1], J.pairs, J.n, 1.0f, o.densify_score_alpha,
                                         (uint32_t)J.pairs_count, S));
            SAD_TRY(launch_radix_sort(J.pairs, J.sort_scratch, J.sort_hist, std::min(J.n, (uint32_t)J.pairs_count),
                                      0xFFFFFFFFu, S));
            int k = std::min(desired, std::min(J.cap - J.actual, 65536));
            if (k > 0) {
                SAD_TRY(launch_write_indices(J.pairs, J.split_idx, (uint32_t)k, S));
                SAD_TRY(launch_split(J.d_sites, J.d_adam, J.split_idx, (uint32_t)k, J.st, (uint32_t)J.actual,
                                     J.d_target, J.width, J.height, S));
                J.actual += k;
                J.n = (uint32_t)J.actual;
                J.active_est += k;
                upd_idx.push_back((uint32_t)j);
                upd_n.push_back(J.n);
                upd_flip.push_back(0u);
            }
        }
        if (!upd_idx.empty())
            SAD_TRY(launch_update_jobs(d_jobs, upd_idx.data(), upd_n.data(), upd_flip.data(), (uint32_t)upd_idx.size(),
                                       S));
        if (any_prune)
            for (int j = 0; j < n_jobs; ++j) SAD_TRY(launch_clear(&jobs[j]->removal, 1, jobs[j]->n, S));

        if (o.profile_gradients) {
            hipEvent_t a, b;
            HIP_TRY(hipEventCreate(&a));
            gev.push_back(a);
            HIP_TRY(hipEventCreate(&b));
            gev.push_back(b);
            HIP_TRY(hipEventRecord(a, S));
        }
        SAD_TRY(launch_gradients_batch(d_jobs, (uint32_t)n_jobs, total_tiles, any_prune, S));
        if (o.profile_gradients) {
            HIP_TRY(hipEventRecord(gev.back(), S));
            double b = 0.0;
            for (int j = 0; j < n_jobs; ++j)
                b += 48.0 * jobs[j]->np + 120.0 * jobs[j]->n + (any_prune ? 8.0 * jobs[j]->n : 0.0);
            gbytes.push_back(b);
        }
        if (pending) HIP_TRY(hipStreamWaitEvent(S, sp.cand_done, 0));
        if (!late_flip.empty()) {
            SAD_TRY(flip_jobs(d_jobs, late_flip, S));
            for (uint32_t j : late_flip) jobs[j]->cur ^= 1u;
        }
        BatchStep st = {};
        const float blend = (float)iter / (float)std::max(1, o.iterations);
        st.lambda = 0.05f * (0.1f + 0.9f * blend);
        st.lr_pos = lr_pos; st.lr_tau = lr_tau; st.lr_radius = lr_rad; st.lr_color = lr_col;
        st.lr_dir = lr_dir; st.lr_aniso = lr_an;
        st.beta1 = o.beta1; st.beta2 = o.beta2; st.eps = o.eps;
        const uint32_t t = (uint32_t)(iter + 1);
        st.b1c = 1.0f / (1.0f - powf(o.beta1, (float)t));
        st.b2c = 1.0f / (1.0f - powf(o.beta2, (float)t));
        SAD_TRY(launch_tau_adam_batch(d_jobs, (uint32_t)n_jobs, total_sites, st, S));

        for (int j = 0; j < n_jobs; ++j) {
            JobRun& J = *jobs[j];
            if (!prune[j]) continue;
            J.n_prune++;
            SAD_TRY(launch_prune_pairs(J.d_sites, J.removal, J.pairs, J.n, J.delta_norm, (uint32_t)J.pairs_count, S));
            SAD_TRY(launch_radix_sort(J.pairs, J.sort_scratch, J.sort_hist, std::min(J.n, (uint32_t)J.pairs_count),
                                      0xFFFFFFFFu, S));
            int k = std::min(std::max(0, (int)((float)J.active_est * J.p_pct)), 65536);
            if (k > 0) {
                SAD_TRY(launch_write_indices(J.pairs, J.prune_idx, (uint32_t)k, S));
                SAD_TRY(launch_prune_sites(J.d_sites, J.prune_idx, (uint32_t)k, S));
                J.active_est = std::max(0, J.active_est - k);
            }
        }
        HIP_TRY(hipEventRecord(sp.sites_ready, S));

        if (iter % std::max(1, o.log_freq) == 0 || iter == o.iterations - 1) {
            std::vector<float> sse(n_jobs);
            std::vector<uint32_t> cnt(n_jobs);
            for (int j = 0; j < n_jobs; ++j) {
                JobRun& J = *jobs[j];
                SAD_TRY(launch_render(J.c0(), J.c1(), J.d_render, J.d_sites, J.iss, J.n, J.width, J.height, J.cw, J.ch,
                                      S));
                SAD_TRY(launch_sse(J.d_render, J.d_target, J.d_mask, J.d_mse, J.width, J.height, S));
                SAD_TRY(launch_count_active(J.d_sites, J.d_count, J.n, S));
                HIP_TRY(hipMemcpyAsync(&sse[j], J.d_mse, sizeof(float), hipMemcpyDeviceToHost, S));
                HIP_TRY(hipMemcpyAsync(&cnt[j], J.d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, S));
            }
            HIP_TRY(hipStreamSynchronize(S));
            for (int j = 0; j < n_jobs; ++j) {
                JobRun& J = *jobs[j];
                float mse = sse[j] / (std::max(1.0f, J.mask_sum) * 3.0f);
                float psnr = mse > 0.0f ? 20.0f * log10f(1.0f / std::sqrt(mse)) : 100.0f;
                if (J.io->log_psnr && J.n_logs < J.io->log_cap) J.io->log_psnr[J.n_logs] = psnr;
                J.n_logs++;
                J.final_psnr = psnr;
                J.best = std::max(J.best, psnr);
                J.active_count = cnt[j];
            }
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
    }
    for (double b : gbytes) gb += b;
    const double secs = std::chrono::duration<double>(t1 - t0).count();
    for (int j = 0; j < n_jobs; ++j) {
        JobRun& J = *jobs[j];
        HIP_TRY(hipMemcpy(J.io->sites_io, J.d_sites, sizeof(SadSite) * J.cap, hipMemcpyDeviceToHost));
        SadhFitResult* r = J.io->result;
        if (!r) continue;
        r->final_psnr = J.final_psnr;
        r->best_psnr = J.best;
        r->actual_sites = J.actual;
        r->active_estimate = J.active_est;
        r->active_count = (int)J.active_count;
        r->densify_percentile = J.d_pct;
        r->prune_percentile = J.p_pct;
        r->capacity = J.cap;
        r->n_logs = J.n_logs;
        r->n_densify = J.n_dens;
        r->n_prune = J.n_prune;
        r->n_cand_updates = J.n_upd;
        r->train_seconds = secs;
        r->grad_kernel_ms = 0.0;
        r->grad_launches = 0;
        r->grad_bytes = 0.0;
    }
    if (bres) {
        bres->train_seconds = secs;
        bres->grad_kernel_ms = gms;
        bres->grad_launches = (int)(gev.size() / 2);
        bres->grad_bytes = gb;
        bres->total_pixels = 0.0;
        for (int j = 0; j < n_jobs; ++j) bres->total_pixels += (double)jobs[j]->np;
    }
    return SADH_OK;
}

}  // namespace sadh

using namespace sadh;

extern "C" int sadh_fit_batch(const SadhFitOptions* opts, SadhBatchJob* jobs, int n_jobs, SadhBatchResult* result)
{
    SADH_REQUIRE(opts && jobs, "fit_batch: null pointer");
    return fit_batch_impl(*opts, jobs, n_jobs, result);
}
