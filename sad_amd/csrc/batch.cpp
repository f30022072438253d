/*
 * batch.cpp - the batching engine: many fits on one device, each at its own
 * iteration, advanced together one "tick" (one iteration of every active fit)
 * at a time; and the continuous-batching server built on it.
 *
 * A tick issues one launch per kernel for all active fits: the VPT pass (with
 * its search-site preparation) for the fits that update candidates this tick,
 * the tiled gradient, and the four tau passes (the last fused with Adam +
 * clear).  Each launch runs over the concatenation of the fits' tiles or sites
 * and finds its fit through a device table of descriptors (batch.h); the
 * per-fit iteration scalars (tau weight, Adam bias corrections, VPT jump index,
 * "update now") come from a small per-tick table.  Rare per-fit work keeps its
 * own launches: densify (site statistics, score pairs, sort, split), prune,
 * Hilbert tables, and the log render + PSNR.
 *
 * Every fit follows exactly the schedule of the single-image fit (fit.cpp,
 * the reference's trainVoronoi, sad_cuda.cu:1860-2533), including the overlap
 * semantics: on a candidate-update iteration a fit that does not densify
 * computes its gradient from the field before the update (the VPT runs on a
 * second stream), a fit that densifies updates its field first.  Which field
 * is current is a per-fit flag in the descriptor, flipped in stream order.
 *
 * sadh_fit_batch adds all fits at once and ticks until they finish.  The
 * server (sadh_server_*) admits queued fit and render requests in arrival
 * order while their projected device memory fits the budget (and a fit slot
 * is free), ticks the running fits, runs render requests on a third stream so
 * they overlap the fits, and retires finished requests with their times.
 * The reference has neither: it fits or renders one image per process.
 */
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>

#include "fit_common.h"

namespace sadh {

namespace {

using Clock = std::chrono::steady_clock;

struct FitRun {
    SadhBatchJob io = {};
    long long id = 0;
    int iterations = 0, iter = 0;
    uint32_t jump_index = 0;
    float d_pct = 0.0f, p_pct = 0.0f;
    int cap = 0, pairs_count = 0;
    int width = 0, height = 0, np = 0, ds = 1, cw = 0, ch = 0;
    size_t cells = 0;
    float mask_sum = 0.0f, iss = 0.0f, delta_norm = 0.0f;
    DevArena mem;
    float *d_target = nullptr, *d_mask = nullptr, *d_render = nullptr;
    long long* d_mse = nullptr;        /* fixed-point (lo, hi) sums */
    long long* d_ssim = nullptr;       /* ssim_metric: sum of the per-pixel SSIM */
    float ssim_weight = 0.0f;          /* SSIM loss weight */
    SadInferenceSite* d_packed = nullptr;   /* packed_psnr: quantised sites, their quant, decoded sites */
    SadInferenceQuant* d_quant = nullptr;
    SadSite* d_dec = nullptr;
    float *tau_raw = nullptr, *tau_tmp = nullptr, *tau_tmp2 = nullptr;
    long long* removal = nullptr;      /* fixed-point (lo, hi) pair per site */
    uint32_t* tau_nbr = nullptr;
    SadSite* d_sites = nullptr;
    SadAdamState* d_adam = nullptr;
    long long* gacc = nullptr;         /* fixed-point gradient lines, FIX_GRAD_STRIDE words per site */
    long long* st = nullptr;           /* fixed-point statistics lines, FIX_STAT_STRIDE words per site */
    SadPair *pairs = nullptr, *sort_scratch = nullptr;
    uint32_t *sort_hist = nullptr, *split_idx = nullptr, *prune_idx = nullptr, *d_count = nullptr;
    uint32_t *fa0 = nullptr, *fa1 = nullptr, *fb0 = nullptr, *fb1 = nullptr;
    SadSearchSite* search = nullptr;
    uint32_t *h_order = nullptr, *h_pos = nullptr, *h_hist = nullptr;
    SadPair *h_pairs = nullptr, *h_scratch = nullptr;
    uint32_t hil_padded = 0;
    int actual = 0, active_est = 0;
    uint32_t n = 0, cur = 0, hil_count = 0;
    bool hil_ready = false;
    float best = 0.0f, final_psnr = 0.0f;
    float final_ssim = 0.0f, best_ssim = 0.0f, packed_psnr = 0.0f;
    long long sse[2] = {0, 0}, ssim_sum[2] = {0, 0};
    int n_logs = 0, n_dens = 0, n_prune = 0, n_upd = 0;
    uint32_t active_count = 0;
    double submit_s = 0.0, admit_s = 0.0;
    uint32_t site_begin = 0;           /* offset of the job's sites in the flattened site range */
    /* this tick */
    bool upd = false, dens = false, prune = false, log = false;

    uint32_t* c0() const { return cur ? fb0 : fa0; }
    uint32_t* c1() const { return cur ? fb1 : fa1; }
};

/* device bytes a fit allocates (the admission projection) */
size_t fit_bytes(const SadhFitOptions& o, int width, int height, int cap, int pairs_count, bool use_hil)
{
    const size_t np = (size_t)width * height;
    const int ds = std::max(1, o.cand_downscale);
    const size_t cells = (size_t)((width + ds - 1) / ds) * ((height + ds - 1) / ds);
    size_t b = cells * 64 + np * (12 + 4 + 12);
    /* site, Adam state, fixed-point gradient line, 3 tau buffers, removal pair, tau neighbour list, search site */
    b += (size_t)cap * (sizeof(SadSite) + sizeof(SadAdamState) + 8 * FIX_GRAD_STRIDE + 12 + 16 + 32 +
                        sizeof(SadSearchSite));
    if (o.densify_enabled) b += (size_t)cap * 8 * FIX_STAT_STRIDE;
    if (o.packed_psnr) b += (size_t)cap * (sizeof(SadInferenceSite) + sizeof(SadSite)) + sizeof(SadInferenceQuant);
    if (pairs_count > 0) b += (size_t)pairs_count * 16 + 256 * 4 * (pairs_count / 1024 + 1) + 2 * 65536 * 4;
    if (use_hil) b += (size_t)cap * 8 + 16 * (size_t)(cap / 1024 + 1) * 1024 * 2;
    return b;
}

size_t render_bytes(uint32_t n, int width, int height, bool use_hil)
{
    size_t b = (size_t)width * height * (64 + 12) + (size_t)n * (sizeof(SadSite) + sizeof(SadSearchSite));
    if (use_hil) b += (size_t)n * 8 + 16 * (size_t)(n / 1024 + 1) * 1024 * 2;
    return b;
}

int fit_plan_for(const SadhFitOptions& base, const SadhBatchJob& io, float* d_pct, float* p_pct, int* cap,
                 int* pairs)
{
    SadhFitOptions o = base;
    o.n_sites = io.n_sites;
    o.init_mode = 0;
    return plan(o, io.width, io.height, d_pct, p_pct, cap, pairs);
}

int setup_fit(const SadhFitOptions& base, FitRun& J, hipStream_t S, bool use_hil)
{
    SadhBatchJob& io = J.io;
    J.iterations = base.iterations;
    J.width = io.width;
    J.height = io.height;
    SAD_TRY(fit_plan_for(base, io, &J.d_pct, &J.p_pct, &J.cap, &J.pairs_count));
    if (!io.target || !io.sites_io) return fail(SADH_ERR_INVALID, "fit_batch: null target or sites");
    if (io.sites_cap < J.cap)
        return fail(SADH_ERR_CAPACITY, "fit_batch: a job's sites_io holds " + std::to_string(io.sites_cap) +
                                           " sites, its plan needs " + std::to_string(J.cap));
    J.np = J.width * J.height;
    J.ds = std::max(1, base.cand_downscale);
    J.cw = std::max(1, (J.width + J.ds - 1) / J.ds);
    J.ch = std::max(1, (J.height + J.ds - 1) / J.ds);
    J.cells = (size_t)J.cw * J.ch;
    const int md = std::max(J.width, J.height);
    J.iss = 1.0f / ((float)md * (float)md);

    std::vector<float> mask(J.np, 1.0f);
    J.mask_sum = (float)J.np;
    if (io.mask) {
        J.mask_sum = 0.0f;
        /* loadMask, sad_cuda.cu:207-213: the normaliser counts pixels with a positive mask */
        for (int i = 0; i < J.np; ++i) { mask[i] = io.mask[i]; J.mask_sum += io.mask[i] > 0.0f ? 1.0f : 0.0f; }
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
    SAD_TRY(J.mem.alloc(&J.d_mse, 2));
    SAD_TRY(J.mem.alloc(&J.d_ssim, 2));
    J.ssim_weight = base.ssim_weight > 0.0f ? base.ssim_weight : 0.0f;
    if (base.packed_psnr) {
        SAD_TRY(J.mem.alloc(&J.d_packed, (size_t)cap));
        SAD_TRY(J.mem.alloc(&J.d_quant, 1));
        SAD_TRY(J.mem.alloc(&J.d_dec, (size_t)cap));
    }
    SAD_TRY(J.mem.alloc(&J.d_count, 1));
    SAD_TRY(J.mem.alloc(&J.d_sites, cap));
    SAD_TRY(J.mem.alloc(&J.d_adam, cap));
    SAD_TRY(J.mem.alloc(&J.gacc, (size_t)cap * FIX_GRAD_STRIDE));
    SAD_TRY(J.mem.alloc(&J.tau_raw, cap));
    SAD_TRY(J.mem.alloc(&J.tau_tmp, cap));
    SAD_TRY(J.mem.alloc(&J.tau_tmp2, cap));
    SAD_TRY(J.mem.alloc(&J.tau_nbr, (size_t)cap * 8));
    SAD_TRY(J.mem.alloc(&J.removal, (size_t)cap * 2));
    if (base.densify_enabled) SAD_TRY(J.mem.alloc(&J.st, (size_t)cap * FIX_STAT_STRIDE));
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
    HIP_TRY(hipMemsetAsync(J.gacc, 0, sizeof(long long) * FIX_GRAD_STRIDE * cap, S));
    HIP_TRY(hipMemsetAsync(J.removal, 0, sizeof(long long) * 2 * cap, S));
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

void describe(const FitRun& J, BatchJob& d, uint32_t tile_begin, uint32_t ctile_begin, uint32_t site_begin,
              uint32_t pblock_begin, uint32_t sblock_begin)
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
    d.tau_nbr = J.tau_nbr;
    d.removal = J.removal;
    d.rendered = J.d_render;
    d.render = J.d_render;
    d.ssim_weight = J.ssim_weight;
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
    d.pblock_begin = pblock_begin;
    d.sblock_begin = sblock_begin;
}

/* The engine: running fits, their device table, streams and per-tick tables. */
struct Engine {
    SadhFitOptions o = {};
    bool use_hil = false;
    StreamPair sp;
    hipEvent_t steps_ready = nullptr;
    std::vector<std::unique_ptr<FitRun>> fits;
    bool table_dirty = true;
    DevArena tmem;                      /* device tables (a new one when the count grows) */
    BatchJob* d_jobs = nullptr;
    int d_jobs_cap = 0;
    uint32_t total_tiles = 0, total_ctiles = 0, total_sites = 0, total_pblocks = 0, total_sblocks = 0;
    long long* d_logv = nullptr;        /* log ticks: (lo, hi, count, 0) per job, d_jobs_cap entries */
    long long* h_logv = nullptr;        /* pinned host copy */
    /* batched Hilbert tables: the tick's pair array, its sort scratch and histogram (hb_cap pairs) */
    SadPair *hb_pairs = nullptr, *hb_scratch = nullptr;
    uint32_t* hb_hist = nullptr;
    uint32_t hb_cap = 0, hil_total = 0, hil_jobs = 0, hil_key_bits = 0;
    /* retirement: the finishing jobs' sites gathered on the device and copied into pinned memory */
    SadSite* d_stage = nullptr;
    SadSite* h_stage = nullptr;
    uint32_t stage_cap = 0;
    static constexpr int RING = 4;      /* per-tick step tables: pinned host + device slots */
    JobStep* h_steps = nullptr;
    JobStep* d_steps = nullptr;
    int steps_cap = 0;
    hipEvent_t slot_done[RING] = {};
    uint64_t tick_no = 0;
    std::vector<hipEvent_t> gev;
    std::vector<double> gbytes;
    Clock::time_point t_origin = Clock::now();
    /* graph_ticks: one executable graph per ring slot, valid while its epoch is current.  The epoch
     * moves whenever a launch argument of a steady iteration would change: a new job table, a site
     * count changed by densify, or (with the SSIM render) a candidate field flipped on the host. */
    hipGraphExec_t tick_exec[RING] = {};
    uint64_t tick_epoch[RING] = {};
    int tick_sched[RING] = {};          /* the schedule a slot's graph was captured with (index + 1) */
    uint64_t epoch = 1;
    int graph_replays = 0, graph_captures = 0;
    /* The schedule of a tick's launches: the tile-hash size of the batched gradient and the block size of the
     * batched tau passes.  Neither changes a result.  With autotune, steady ticks take the candidates in turn,
     * each timed by the stream period it takes, and the fastest is kept.  (Graph replays are not a candidate:
     * they measured slower in every configuration, and trying one costs a capture per ring slot.) */
    struct Schedule {
        int hash_slots;
        uint32_t tau_block;
    };
    static constexpr int N_SCHEDULES = 3;
    static constexpr int TUNE_SAMPLES = 5;              /* ticks per candidate; the first is not counted */
    /* Tuning starts after this many ticks: a fit's first iterations, over the random initial field, cost several
     * times a converged one and fall quickly, which would favour whichever candidate runs later.  The rounds
     * alternate their order (ABC, CBA, ...), so a steady trend weighs on every candidate alike. */
    static constexpr uint64_t TUNE_START = 40;
    static constexpr Schedule CANDIDATES[N_SCHEDULES] = {{128, 256u}, {64, 256u}, {128, 64u}};
    int sched_id = 0;                   /* index into CANDIDATES */
    bool tuning = false;
    int tune_ticks = 0;
    int tune_open = -1;                 /* the tuning tick whose end is the next tick's start */
    /* Per tuning tick, an event where its stream work starts and one where the next tick's starts: the whole
     * period the schedule costs the stream, its launch overheads included.  Nothing waits on them until the
     * last, so stream launches stay queued ahead of the device as in untimed ticks. */
    hipEvent_t tev[2 * N_SCHEDULES * TUNE_SAMPLES] = {};
    int tuned_hash_slots = 0, tuned_tau_block = 0;

    ~Engine()
    {
        for (auto& g : tick_exec) if (g) (void)hipGraphExecDestroy(g);
        for (auto e : gev) (void)hipEventDestroy(e);
        for (auto& e : slot_done) if (e) (void)hipEventDestroy(e);
        for (auto& e : tev) if (e) (void)hipEventDestroy(e);
        if (steps_ready) (void)hipEventDestroy(steps_ready);
        if (h_steps) (void)hipHostFree(h_steps);
        if (d_steps) (void)hipFree(d_steps);
        if (h_logv) (void)hipHostFree(h_logv);
        if (h_stage) (void)hipHostFree(h_stage);
    }

    double now() const { return std::chrono::duration<double>(Clock::now() - t_origin).count(); }

    int init(const SadhFitOptions& opts)
    {
        o = opts;
        if (o.init_mode != 0) return fail(SADH_ERR_UNSUPPORTED, "fit_batch: only caller-provided initial sites");
        use_hil = o.cand_hilbert_probes > 0 && o.cand_hilbert_window > 0;
        HIP_TRY(hipStreamCreateWithFlags(&sp.main, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&sp.cand, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&sp.cand_done, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&sp.sites_ready, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&steps_ready, hipEventDisableTiming));
        for (auto& e : slot_done) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        sched_id = 0;
        tuning = o.autotune != 0;
        if (tuning)
            for (auto& e : tev) HIP_TRY(hipEventCreate(&e));
        return SADH_OK;
    }

    int add_fit(const SadhBatchJob& io, long long id, double submit_s)
    {
        std::unique_ptr<FitRun> f(new FitRun());
        f->io = io;
        f->id = id;
        f->submit_s = submit_s;
        /* the device table is read by in-flight kernels: finish them before it changes */
        HIP_TRY(hipStreamSynchronize(sp.main));
        HIP_TRY(hipStreamSynchronize(sp.cand));
        SAD_TRY(setup_fit(o, *f, sp.main, use_hil));
        f->admit_s = now();
        fits.push_back(std::move(f));
        table_dirty = true;
        return SADH_OK;
    }

    int rebuild_table()
    {
        HIP_TRY(hipStreamSynchronize(sp.main));
        HIP_TRY(hipStreamSynchronize(sp.cand));
        const int nj = (int)fits.size();
        std::vector<BatchJob> desc(nj);
        total_tiles = total_ctiles = total_sites = total_pblocks = total_sblocks = 0;
        for (int j = 0; j < nj; ++j) {
            describe(*fits[j], desc[j], total_tiles, total_ctiles, total_sites, total_pblocks, total_sblocks);
            total_tiles += desc[j].ntiles;
            total_ctiles += desc[j].cntiles;
            total_sites += desc[j].cap;
            total_pblocks += div_up((uint32_t)fits[j]->np, 256u);
            total_sblocks += div_up(desc[j].cap, 256u);
            fits[j]->site_begin = desc[j].site_begin;
        }
        if (nj > d_jobs_cap) {
            d_jobs_cap = std::max(nj, 2 * d_jobs_cap);
            SAD_TRY(tmem.alloc(&d_jobs, (size_t)d_jobs_cap));
            SAD_TRY(tmem.alloc(&d_logv, 4u * (size_t)d_jobs_cap));
            if (h_logv) (void)hipHostFree(h_logv);
            h_logv = nullptr;
            HIP_TRY(hipHostMalloc((void**)&h_logv, sizeof(long long) * 4u * (size_t)d_jobs_cap, hipHostMallocDefault));
        }
        if (nj > steps_cap) {
            if (h_steps) (void)hipHostFree(h_steps);
            if (d_steps) (void)hipFree(d_steps);
            h_steps = nullptr;
            d_steps = nullptr;
            steps_cap = std::max(nj, 2 * steps_cap);
            /* per ring slot: the tick's table, then a copy whose vpt flags select the updates that run after
             * the gradient (multi-pass updates of fits that do not densify) */
            HIP_TRY(hipHostMalloc((void**)&h_steps, sizeof(JobStep) * RING * 2 * steps_cap, hipHostMallocDefault));
            HIP_TRY(hipMalloc((void**)&d_steps, sizeof(JobStep) * RING * 2 * steps_cap));
        }
        if (nj > 0) HIP_TRY(hipMemcpy(d_jobs, desc.data(), sizeof(BatchJob) * nj, hipMemcpyHostToDevice));
        table_dirty = false;
        epoch++;
        return SADH_OK;
    }

    /* the launches of a steady iteration: the SSIM renders (when the loss has that term), the batched
     * gradient and the batched tau diffusion + Adam; every per-iteration value comes from the device
     * step table `ds` and the job table, so the same launches replay as a graph */
    int steady_launches(int nj, const JobStep* ds, hipStream_t S, const Schedule& k)
    {
        if (o.ssim_weight > 0.0f)
            for (int j = 0; j < nj; ++j) {
                FitRun& J = *fits[j];
                SAD_TRY(launch_render(J.c0(), J.c1(), J.d_render, J.d_sites, J.iss, J.n, J.width, J.height, J.cw,
                                      J.ch, S));
            }
        SAD_TRY(launch_gradients_batch(d_jobs, (uint32_t)nj, total_tiles, false, S, k.hash_slots));
        BatchStep st = adam_step();
        return launch_tau_adam_batch(d_jobs, (uint32_t)nj, ds, total_sites, st, S, k.tau_block);
    }

    BatchStep adam_step() const
    {
        BatchStep st = {};
        st.lr_pos = o.lr_pos * o.lr_scale; st.lr_tau = o.lr_tau * o.lr_scale; st.lr_radius = o.lr_radius * o.lr_scale;
        st.lr_color = o.lr_color * o.lr_scale; st.lr_dir = o.lr_dir * o.lr_scale; st.lr_aniso = o.lr_aniso * o.lr_scale;
        st.beta1 = o.beta1; st.beta2 = o.beta2; st.eps = o.eps;
        return st;
    }

    /* replay slot `slot`'s graph of the steady launches, capturing it first when its epoch or schedule is old */
    int graph_tick(int slot, int nj, const JobStep* ds)
    {
        hipStream_t S = sp.main;
        if (!tick_exec[slot] || tick_epoch[slot] != epoch || tick_sched[slot] != sched_id + 1) {
            if (tick_exec[slot]) {
                HIP_TRY(hipGraphExecDestroy(tick_exec[slot]));
                tick_exec[slot] = nullptr;
            }
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(S, hipStreamCaptureModeRelaxed));
            const int rc = steady_launches(nj, ds, S, CANDIDATES[sched_id]);
            const hipError_t ec = hipStreamEndCapture(S, &g);
            if (rc != SADH_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            HIP_TRY(ec);
            const hipError_t ei = hipGraphInstantiate(&tick_exec[slot], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_TRY(ei);
            tick_epoch[slot] = epoch;
            tick_sched[slot] = sched_id + 1;
            graph_captures++;
        }
        HIP_TRY(hipGraphLaunch(tick_exec[slot], S));
        graph_replays++;
        return SADH_OK;
    }

    int flip(const std::vector<uint32_t>& which, hipStream_t s)
    {
        if (which.empty()) return SADH_OK;
        std::vector<uint32_t> keep(which.size(), 0xFFFFFFFFu), one(which.size(), 1u);
        return launch_update_jobs(d_jobs, which.data(), keep.data(), one.data(), (uint32_t)which.size(), s);
    }

    /* one iteration of every running fit; `done` receives the fits that finished */
    int tick(std::vector<std::unique_ptr<FitRun>>& done)
    {
        if (fits.empty()) return SADH_OK;
        if (table_dirty) SAD_TRY(rebuild_table());
        hipStream_t S = sp.main;
        const int nj = (int)fits.size();
        int eff_ps = o.prune_start;
        if (o.densify_enabled && !o.prune_during_densify && o.prune_start < o.densify_end) eff_ps = o.densify_end;
        /* per-fit flags and step table (a ring slot whose previous tick has finished) */
        const int slot = (int)(tick_no % RING);
        HIP_TRY(hipEventSynchronize(slot_done[slot]));
        JobStep* hs = h_steps + (size_t)slot * 2 * steps_cap;
        JobStep* ds = d_steps + (size_t)slot * 2 * steps_cap;
        /* An update of several passes ping-pongs through the field its fit's gradient reads, so it never runs
         * beside that gradient: a densifying fit's update runs before its gradient (which reads the new
         * field), the others' after it. */
        const bool split_upd = std::max(1, o.cand_update_passes) > 1;
        bool any_upd = false, any_dens_upd = false, any_prune = false, any_log = false, any_dens = false;
        hil_total = hil_jobs = hil_key_bits = 0;
        for (int j = 0; j < nj; ++j) {
            FitRun& J = *fits[j];
            const int it = J.iter;
            const int eff_pe = o.prune_end > 0 ? o.prune_end : J.iterations - 1;
            const int desired = std::max(0, (int)((float)J.active_est * J.d_pct));
            J.upd = o.cand_update_freq > 0 && it % o.cand_update_freq == 0;
            J.dens = o.densify_enabled && it >= o.densify_start && it <= o.densify_end &&
                     it % std::max(1, o.densify_freq) == 0 && J.actual < J.cap && desired > 0;
            J.prune = J.p_pct > 0.0f && J.pairs_count > 0 && it >= eff_ps && it < eff_pe &&
                      it % std::max(1, o.prune_freq) == 0;
            J.log = it % std::max(1, o.log_freq) == 0 || it == J.iterations - 1;
            any_upd = any_upd || J.upd;
            any_dens_upd = any_dens_upd || (J.upd && J.dens);
            any_dens = any_dens || (J.dens && J.pairs_count > 0);
            any_prune = any_prune || J.prune;
            any_log = any_log || J.log;
            const float blend = (float)it / (float)std::max(1, J.iterations);
            const uint32_t t = (uint32_t)(it + 1);
            hs[j].lambda = 0.05f * (0.1f + 0.9f * blend);
            hs[j].b1c = 1.0f / (1.0f - sad_powi_f(o.beta1, t));
            hs[j].b2c = 1.0f / (1.0f - sad_powi_f(o.beta2, t));
            hs[j].jump_index = J.jump_index;
            hs[j].vpt = (J.upd && (!split_upd || J.dens)) ? 1u : 0u;
            hs[j].log = J.log ? 1u : 0u;
            hs[j].retire = it + 1 >= J.iterations ? 1u : 0u;
            /* a Hilbert table is (re)built before the job's first update and after its site count changes */
            const bool hil = use_hil && J.upd && (!J.hil_ready || J.hil_count != J.n);
            hs[j].hil_begin = hil_total;
            hs[j].hil_n = hil ? J.n : 0u;
            hs[j].hil_bits = (uint32_t)sad_hilbert_bits(J.width, J.height);
            if (hil) {
                hil_total += J.n;
                hil_jobs++;
                hil_key_bits = std::max(hil_key_bits, 2u * hs[j].hil_bits);
            }
        }
        for (int j = 0; j < nj; ++j) {
            hs[nj + j] = hs[j];
            hs[nj + j].vpt = (split_upd && fits[j]->upd && !fits[j]->dens) ? 1u : 0u;
        }
        const bool steady = !any_upd && !any_dens && !any_prune && !any_log && !o.profile_gradients;
        if (tune_open >= 0) {
            HIP_TRY(hipEventRecord(tev[2 * tune_open + 1], S));
            tune_open = -1;
            if (tune_ticks == N_SCHEDULES * TUNE_SAMPLES) SAD_TRY(choose_schedule());
        }
        const bool tune_now = steady && tuning && tick_no >= TUNE_START;
        if (tune_now) HIP_TRY(hipEventRecord(tev[2 * tune_ticks], S));
        HIP_TRY(hipMemcpyAsync(ds, hs, sizeof(JobStep) * 2 * nj, hipMemcpyHostToDevice, S));
        if (any_upd) HIP_TRY(hipEventRecord(steps_ready, S));

        if (tune_now) {
            SAD_TRY(steady_launches(nj, ds, S, CANDIDATES[tune_candidate(tune_ticks)]));
            tune_open = tune_ticks++;
        } else if (steady && o.graph_ticks)
            SAD_TRY(graph_tick(slot, nj, ds));
        else
            SAD_TRY(eager_tick(nj, ds, any_upd, any_dens_upd, any_prune));
        HIP_TRY(hipEventRecord(slot_done[slot], S));
        tick_no++;
        return finish_tick(any_log, ds, done);
    }

    /* the candidate of tuning tick i: round i / N in order, then reversed, alternately */
    static int tune_candidate(int i)
    {
        const int r = i / N_SCHEDULES, k = i % N_SCHEDULES;
        return (r & 1) ? N_SCHEDULES - 1 - k : k;
    }

    /* after TUNE_SAMPLES tuning ticks of each candidate: the schedule with the least total period over its
     * ticks after the first (which may include a graph capture or a first launch) */
    int choose_schedule()
    {
        const int last = 2 * (N_SCHEDULES * TUNE_SAMPLES) - 1;
        HIP_TRY(hipEventSynchronize(tev[last]));
        double total[N_SCHEDULES] = {};
        for (int i = N_SCHEDULES; i < N_SCHEDULES * TUNE_SAMPLES; ++i) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, tev[2 * i], tev[2 * i + 1]));
            total[tune_candidate(i)] += ms;
        }
        int best = 0;
        for (int i = 1; i < N_SCHEDULES; ++i)
            if (total[i] < total[best]) best = i;
        sched_id = best;
        tuning = false;
        tuned_hash_slots = CANDIDATES[best].hash_slots;
        tuned_tau_block = (int)CANDIDATES[best].tau_block;
        return SADH_OK;
    }

    /* the VPT passes of this tick's update for the jobs whose flag is set in `steps`, on stream U */
    int vpt_passes(int nj, const JobStep* steps, hipStream_t U)
    {
        const int passes = std::max(1, o.cand_update_passes);
        for (int p = 0; p < passes; ++p) {
            BatchStep stp = {};
            stp.vpt_pass = (uint32_t)p;
            stp.radius_scale = o.cand_radius_scale;
            stp.radius_probes = o.cand_radius_probes;
            stp.inject_count = o.cand_inject_count;
            stp.hil_probes = use_hil ? o.cand_hilbert_probes : 0u;
            stp.hil_window = use_hil ? o.cand_hilbert_window : 0u;
            SAD_TRY(launch_vpt_batch(d_jobs, (uint32_t)nj, steps, total_ctiles, total_sites, stp, U));
        }
        return SADH_OK;
    }

    /* a tick's launches when some fit updates candidates, densifies, prunes or logs (or graphs are off) */
    int eager_tick(int nj, const JobStep* ds, bool any_upd, bool any_dens_upd, bool any_prune)
    {
        hipStream_t S = sp.main;
        const int passes = std::max(1, o.cand_update_passes);
        const bool split_upd = passes > 1;
        std::vector<uint32_t> late;
        bool pending = false, deferred = false;
        if (any_upd) {
            /* an updating fit that densifies needs its new field first: then all single-pass updates run on
             * S; multi-pass updates run on S as well, those of non-densifying fits after the gradient */
            hipStream_t U = (any_dens_upd || split_upd) ? S : sp.cand;
            if (U != S) HIP_TRY(hipStreamWaitEvent(sp.cand, steps_ready, 0));
            /* Two or more tables to build: one launch set for all of them, keyed (job << shift) | Hilbert
             * index, when the job index and the widest Hilbert index fit 32 bits together */
            uint32_t job_bits = 0;
            while ((1u << job_bits) < (uint32_t)nj) job_bits++;
            const bool hil_batched = hil_jobs >= 2 && hil_key_bits < 32u && job_bits + hil_key_bits <= 32u;
            if (hil_batched) {
                const uint32_t padded = div_up(hil_total, 1024u) * 1024u;
                if (padded > hb_cap) {
                    hb_cap = std::max(padded, 2u * hb_cap);
                    SAD_TRY(tmem.alloc(&hb_pairs, hb_cap));
                    SAD_TRY(tmem.alloc(&hb_scratch, hb_cap));
                    SAD_TRY(tmem.alloc(&hb_hist, 256u * (hb_cap / 1024u)));
                }
                const uint32_t key_bits = job_bits + hil_key_bits;
                const uint32_t mk = key_bits >= 32u ? 0xFFFFFFFFu : (1u << key_bits);
                SAD_TRY(launch_hilbert_batch(d_jobs, (uint32_t)nj, ds, hil_total, hil_key_bits, mk, hb_pairs, hb_scratch,
                                             hb_hist, U));
            }
            for (int j = 0; j < nj; ++j) {
                FitRun& J = *fits[j];
                if (!J.upd) continue;
                J.n_upd++;
                if (hil_batched && use_hil && (!J.hil_ready || J.hil_count != J.n)) {
                    J.hil_ready = true;
                    J.hil_count = J.n;
                } else if (use_hil && (!J.hil_ready || J.hil_count != J.n)) {
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
            SAD_TRY(vpt_passes(nj, ds, U));
            /* an odd pass count leaves the result in the other field: densifying fits switch now,
             * the others after their gradient, which reads the field from before the update */
            std::vector<uint32_t> now_flip;
            for (int j = 0; j < nj; ++j) {
                FitRun& J = *fits[j];
                if (!J.upd) continue;
                J.jump_index += (uint32_t)passes;
                if (passes & 1) (J.dens ? now_flip : late).push_back((uint32_t)j);
                if (split_upd && !J.dens) deferred = true;
            }
            SAD_TRY(flip(now_flip, S));
            for (uint32_t j : now_flip) fits[j]->cur ^= 1u;
            if (o.ssim_weight > 0.0f && (!now_flip.empty() || !late.empty())) epoch++;
            if (U != S) {
                HIP_TRY(hipEventRecord(sp.cand_done, sp.cand));
                pending = true;
            }
        }
        /* densify (per fit) */
        std::vector<uint32_t> upd_idx, upd_n, upd_flip;
        for (int j = 0; j < nj; ++j) {
            FitRun& J = *fits[j];
            if (!J.dens || J.pairs_count <= 0) continue;
            J.n_dens++;
            const int desired = std::max(0, (int)((float)J.active_est * J.d_pct));
            SAD_TRY(launch_clear_fix(J.st, (size_t)J.n * FIX_STAT_STRIDE, S));
            SAD_TRY(launch_site_stats_fix(J.c0(), J.c1(), J.d_target, J.d_mask, J.d_sites, J.iss, J.n, J.st, J.width,
                                          J.height, J.cw, J.ch, S));
            SAD_TRY(launch_densify_pairs_fix(J.d_sites, J.st, J.pairs, J.n, 1.0f, o.densify_score_alpha,
                                             (uint32_t)J.pairs_count, S));
            SAD_TRY(launch_radix_sort(J.pairs, J.sort_scratch, J.sort_hist, std::min(J.n, (uint32_t)J.pairs_count),
                                      0xFFFFFFFFu, S));
            int k = std::min(desired, std::min(J.cap - J.actual, 65536));
            if (k > 0) {
                SAD_TRY(launch_write_indices(J.pairs, J.split_idx, (uint32_t)k, S));
                SAD_TRY(launch_split_fix(J.d_sites, J.d_adam, J.split_idx, (uint32_t)k, J.st, (uint32_t)J.actual,
                                         J.d_target, J.width, J.height, S));
                J.actual += k;
                J.n = (uint32_t)J.actual;
                J.active_est += k;
                upd_idx.push_back((uint32_t)j);
                upd_n.push_back(J.n);
                upd_flip.push_back(0u);
            }
        }
        if (!upd_idx.empty()) {
            SAD_TRY(launch_update_jobs(d_jobs, upd_idx.data(), upd_n.data(), upd_flip.data(), (uint32_t)upd_idx.size(),
                                       S));
            epoch++;
        }
        if (any_prune)
            for (int j = 0; j < nj; ++j) SAD_TRY(launch_clear_fix(fits[j]->removal, 2u * (size_t)fits[j]->n, S));

        if (o.profile_gradients) {
            hipEvent_t a, b;
            HIP_TRY(hipEventCreate(&a));
            gev.push_back(a);
            HIP_TRY(hipEventCreate(&b));
            gev.push_back(b);
            HIP_TRY(hipEventRecord(a, S));
        }
        /* the SSIM loss term reads a render of each job's current state (sad_cuda.cu:2307-2318) */
        if (o.ssim_weight > 0.0f)
            for (int j = 0; j < nj; ++j) {
                FitRun& J = *fits[j];
                SAD_TRY(launch_render(J.c0(), J.c1(), J.d_render, J.d_sites, J.iss, J.n, J.width, J.height, J.cw,
                                      J.ch, S));
            }
        SAD_TRY(launch_gradients_batch(d_jobs, (uint32_t)nj, total_tiles, any_prune, S,
                                       CANDIDATES[sched_id].hash_slots));
        if (o.profile_gradients) {
            HIP_TRY(hipEventRecord(gev.back(), S));
            double b = 0.0;
            for (int j = 0; j < nj; ++j)
                b += 48.0 * fits[j]->np + 120.0 * fits[j]->n + (any_prune ? 8.0 * fits[j]->n : 0.0);
            gbytes.push_back(b);
        }
        if (pending) HIP_TRY(hipStreamWaitEvent(S, sp.cand_done, 0));
        if (deferred) SAD_TRY(vpt_passes(nj, ds + nj, S));
        SAD_TRY(flip(late, S));
        for (uint32_t j : late) fits[j]->cur ^= 1u;

        SAD_TRY(launch_tau_adam_batch(d_jobs, (uint32_t)nj, ds, total_sites, adam_step(), S,
                                      CANDIDATES[sched_id].tau_block));

        for (int j = 0; j < nj; ++j) {
            FitRun& J = *fits[j];
            if (!J.prune) continue;
            J.n_prune++;
            SAD_TRY(launch_prune_pairs_fix(J.d_sites, J.removal, J.pairs, J.n, J.delta_norm, (uint32_t)J.pairs_count,
                                           S));
            SAD_TRY(launch_radix_sort(J.pairs, J.sort_scratch, J.sort_hist, std::min(J.n, (uint32_t)J.pairs_count),
                                      0xFFFFFFFFu, S));
            int k = std::min(std::max(0, (int)((float)J.active_est * J.p_pct)), 65536);
            if (k > 0) {
                SAD_TRY(launch_write_indices(J.pairs, J.prune_idx, (uint32_t)k, S));
                SAD_TRY(launch_prune_sites(J.d_sites, J.prune_idx, (uint32_t)k, S));
                J.active_est = std::max(0, J.active_est - k);
            }
        }
        return SADH_OK;
    }

    /* the log renders and metrics, then the iteration count; retires finished fits */
    int finish_tick(bool any_log, const JobStep* ds, std::vector<std::unique_ptr<FitRun>>& done)
    {
        hipStream_t S = sp.main;
        const int nj = (int)fits.size();
        if (any_log) {
            int n_log = 0;
            for (auto& f : fits) n_log += f->log ? 1 : 0;
            /* Two or more logging jobs (a batch's jobs log on the same ticks): one render launch and one sum
             * launch for all of them and one copy back, the same values as the per-job launches below (the
             * same tiles and the same 256-pixel partials into an exact sum).  The SSIM metric keeps the
             * per-job launches. */
            const bool batched = n_log >= 2 && !o.ssim_metric;
            if (batched) {
                SAD_TRY(launch_render_batch(d_jobs, (uint32_t)nj, ds, total_tiles, S));
                SAD_TRY(launch_log_sums_batch(d_jobs, (uint32_t)nj, ds, total_pblocks, total_sblocks, d_logv, S));
                HIP_TRY(hipMemcpyAsync(h_logv, d_logv, sizeof(long long) * 4u * (size_t)nj, hipMemcpyDeviceToHost, S));
            }
            for (int j = 0; j < nj && !batched; ++j) {
                FitRun& J = *fits[j];
                if (!J.log) continue;
                SAD_TRY(launch_render(J.c0(), J.c1(), J.d_render, J.d_sites, J.iss, J.n, J.width, J.height, J.cw,
                                      J.ch, S));
                SAD_TRY(launch_sse_fix(J.d_render, J.d_target, J.d_mask, J.d_mse, J.width, J.height, S));
                if (o.ssim_metric) {
                    SAD_TRY(launch_ssim_fix(J.d_render, J.d_target, J.d_mask, J.d_ssim, J.width, J.height, S));
                    HIP_TRY(hipMemcpyAsync(J.ssim_sum, J.d_ssim, sizeof(J.ssim_sum), hipMemcpyDeviceToHost, S));
                }
                SAD_TRY(launch_count_active(J.d_sites, J.d_count, J.n, S));
                HIP_TRY(hipMemcpyAsync(J.sse, J.d_mse, sizeof(J.sse), hipMemcpyDeviceToHost, S));
                HIP_TRY(hipMemcpyAsync(&J.active_count, J.d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, S));
            }
            HIP_TRY(hipStreamSynchronize(S));
            for (int j = 0; j < nj; ++j) {
                FitRun& J = *fits[j];
                if (!J.log) continue;
                if (batched) {
                    J.sse[0] = h_logv[4 * j];
                    J.sse[1] = h_logv[4 * j + 1];
                    J.active_count = (uint32_t)h_logv[4 * j + 2];
                }
                float mse = fix_value_host(J.sse) / (std::max(1.0f, J.mask_sum) * 3.0f);
                float psnr = mse > 0.0f ? 20.0f * log10f(1.0f / std::sqrt(mse)) : 100.0f;
                if (J.io.log_psnr && J.n_logs < J.io.log_cap) J.io.log_psnr[J.n_logs] = psnr;
                if (o.ssim_metric) {
                    const float ssim = fix_value_host(J.ssim_sum) / std::max(1.0f, J.mask_sum);   /* sad_cuda.cu:2443-2448 */
                    SadhFitResult* r = J.io.result;
                    if (r && r->log_ssim && J.n_logs < J.io.log_cap) r->log_ssim[J.n_logs] = ssim;
                    J.final_ssim = ssim;
                    J.best_ssim = std::max(J.best_ssim, ssim);
                }
                J.n_logs++;
                J.final_psnr = psnr;
                J.best = std::max(J.best, psnr);
            }
        }
        /* advance; retire finished fits */
        bool finished = false;
        for (auto& f : fits) {
            f->iter++;
            finished = finished || f->iter >= f->iterations;
        }
        if (finished) {
            /* Two or more fits finishing: one gather of their sites, one copy per run of consecutive finishing
             * jobs into pinned memory and one synchronisation, instead of a blocking copy per fit */
            int n_fin = 0;
            for (auto& f : fits) n_fin += f->iter >= f->iterations ? 1 : 0;
            const bool gather = n_fin >= 2;
            if (gather) {
                if (total_sites > stage_cap) {
                    stage_cap = std::max(total_sites, 2u * stage_cap);
                    SAD_TRY(tmem.alloc(&d_stage, stage_cap));
                    if (h_stage) (void)hipHostFree(h_stage);
                    h_stage = nullptr;
                    HIP_TRY(hipHostMalloc((void**)&h_stage, sizeof(SadSite) * stage_cap, hipHostMallocDefault));
                }
                SAD_TRY(launch_gather_sites_batch(d_jobs, (uint32_t)nj, ds, total_sites, d_stage, S));
                for (int j = 0; j < nj;) {
                    if (fits[j]->iter < fits[j]->iterations) { ++j; continue; }
                    int e = j + 1;
                    while (e < nj && fits[e]->iter >= fits[e]->iterations) ++e;
                    const uint32_t b = fits[j]->site_begin, end = fits[e - 1]->site_begin + (uint32_t)fits[e - 1]->cap;
                    HIP_TRY(hipMemcpyAsync(h_stage + b, d_stage + b, sizeof(SadSite) * (end - b), hipMemcpyDeviceToHost,
                                           S));
                    j = e;
                }
            }
            HIP_TRY(hipStreamSynchronize(S));
            HIP_TRY(hipStreamSynchronize(sp.cand));
            std::vector<std::unique_ptr<FitRun>> keep;
            for (auto& f : fits) {
                if (f->iter < f->iterations) { keep.push_back(std::move(f)); continue; }
                if (gather)
                    std::memcpy(f->io.sites_io, h_stage + f->site_begin, sizeof(SadSite) * f->cap);
                else
                    HIP_TRY(hipMemcpy(f->io.sites_io, f->d_sites, sizeof(SadSite) * f->cap, hipMemcpyDeviceToHost));
                SAD_TRY(finish_renders(*f, S));
                done.push_back(std::move(f));
            }
            fits.swap(keep);
            table_dirty = true;
        }
        return SADH_OK;
    }

    /* a finished fit: the packed-site PSNR and render (training.swift:676-703) and the final render, as
     * sadh_fit produces them */
    int finish_renders(FitRun& J, hipStream_t S)
    {
        SadhFitResult* r = J.io.result;
        const size_t img = sizeof(float) * 3 * (size_t)J.np;
        if (o.packed_psnr) {
            SAD_TRY(launch_pack_inference(J.d_sites, J.n, J.width, J.height, J.d_packed, J.d_quant, S));
            SAD_TRY(launch_unpack_inference(J.d_packed, J.d_quant, J.n, J.width, J.height, J.d_dec, S));
            SAD_TRY(launch_render(J.c0(), J.c1(), J.d_render, J.d_dec, J.iss, J.n, J.width, J.height, J.cw, J.ch, S));
            SAD_TRY(launch_sse_fix(J.d_render, J.d_target, J.d_mask, J.d_mse, J.width, J.height, S));
            long long sse[2] = {0, 0};
            HIP_TRY(hipMemcpyAsync(sse, J.d_mse, sizeof(sse), hipMemcpyDeviceToHost, S));
            if (r && r->packed_render) HIP_TRY(hipMemcpyAsync(r->packed_render, J.d_render, img, hipMemcpyDeviceToHost, S));
            HIP_TRY(hipStreamSynchronize(S));
            const float mse = fix_value_host(sse) / (std::max(1.0f, J.mask_sum) * 3.0f);
            J.packed_psnr = mse > 0.0f ? 20.0f * log10f(1.0f / std::sqrt(mse)) : 100.0f;
        }
        if (r && r->final_render) {
            SAD_TRY(launch_render(J.c0(), J.c1(), J.d_render, J.d_sites, J.iss, J.n, J.width, J.height, J.cw, J.ch, S));
            HIP_TRY(hipMemcpyAsync(r->final_render, J.d_render, img, hipMemcpyDeviceToHost, S));
            HIP_TRY(hipStreamSynchronize(S));
        }
        /* the host copy of the sites is in sites_io already */
        return debug_renders(J.c0(), J.c1(), J.d_sites, J.io.sites_io, J.n, J.iss, J.width, J.height, J.cw, J.ch,
                             J.d_render, r, S);
    }

    void fill_result(const FitRun& J, double secs) const
    {
        SadhFitResult* r = J.io.result;
        if (!r) return;
        r->final_psnr = J.final_psnr;
        r->best_psnr = J.best;
        r->final_ssim = J.final_ssim;
        r->best_ssim = J.best_ssim;
        r->packed_psnr = J.packed_psnr;
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

    int grad_totals(double* ms, int* launches, double* bytes)
    {
        double g = 0.0, b = 0.0;
        for (size_t i = 0; i + 1 < gev.size(); i += 2) {
            float t = 0.0f;
            HIP_TRY(hipEventElapsedTime(&t, gev[i], gev[i + 1]));
            g += t;
        }
        for (double x : gbytes) b += x;
        *ms = g;
        *launches = (int)(gev.size() / 2);
        *bytes = b;
        return SADH_OK;
    }
};

}  // namespace

static int fit_batch_impl(const SadhFitOptions& o, SadhBatchJob* ios, int n_jobs, SadhBatchResult* bres)
{
    if (n_jobs <= 0 || !ios) return fail(SADH_ERR_INVALID, "fit_batch: no jobs");
    Engine E;
    SAD_TRY(E.init(o));
    for (int j = 0; j < n_jobs; ++j) SAD_TRY(E.add_fit(ios[j], j, 0.0));
    std::vector<std::unique_ptr<FitRun>> done;
    auto t0 = Clock::now();
    while (!E.fits.empty()) SAD_TRY(E.tick(done));
    HIP_TRY(hipStreamSynchronize(E.sp.main));
    HIP_TRY(hipStreamSynchronize(E.sp.cand));
    const double secs = std::chrono::duration<double>(Clock::now() - t0).count();
    double px = 0.0;
    for (auto& f : done) {
        E.fill_result(*f, secs);
        px += (double)f->np;
    }
    if (bres) {
        bres->train_seconds = secs;
        SAD_TRY(E.grad_totals(&bres->grad_kernel_ms, &bres->grad_launches, &bres->grad_bytes));
        bres->total_pixels = px;
        bres->graph_ticks = E.graph_replays;
        bres->graph_captures = E.graph_captures;
        bres->tuned_hash_slots = E.tuned_hash_slots;
        bres->tuned_tau_block = E.tuned_tau_block;
    }
    return SADH_OK;
}

/* ---------------------------------------------------------------- server -- */

namespace {

struct RenderRun {
    long long id = 0;
    const SadSite* sites_h = nullptr;
    uint32_t n = 0;
    int width = 0, height = 0;
    float* out_h = nullptr;
    size_t bytes = 0;
    double submit_s = 0.0, admit_s = 0.0;
    DevArena mem;
    hipEvent_t done = nullptr;
    ~RenderRun()
    {
        if (done) (void)hipEventDestroy(done);
    }
};

struct Request {
    int kind = 0;            /* 0 fit, 1 render */
    long long id = 0;
    double submit_s = 0.0;
    size_t bytes = 0;
    SadhBatchJob fit = {};
    const SadSite* sites = nullptr;
    uint32_t n = 0;
    int width = 0, height = 0;
    float* out = nullptr;
};

}  // namespace

struct SadhServerState {
    SadhServerOptions so = {};
    Engine E;
    hipStream_t rstream = nullptr;
    std::deque<Request> queue;
    std::vector<std::unique_ptr<RenderRun>> renders;
    std::vector<SadhServerEvent> events;
    size_t budget = 0, in_use = 0;
    std::vector<size_t> fit_bytes_of;   /* parallel to E.fits */
    long long next_id = 1;
    bool use_hil_render = true;

    ~SadhServerState()
    {
        if (E.sp.main) (void)hipStreamSynchronize(E.sp.main);
        if (E.sp.cand) (void)hipStreamSynchronize(E.sp.cand);
        if (rstream) (void)hipStreamSynchronize(rstream);
        renders.clear();
        if (rstream) (void)hipStreamDestroy(rstream);
    }

    int render_start(RenderRun& R)
    {
        const int w = R.width, h = R.height;
        const uint32_t n = R.n;
        const int passes = so.render_passes > 0 ? so.render_passes : 16;
        const bool use_hil = use_hil_render;
        const size_t cells = (size_t)w * h;
        SadSite* d_sites;
        SadSearchSite* search;
        uint32_t *a0, *a1, *b0, *b1;
        float* d_out;
        SAD_TRY(R.mem.alloc(&d_sites, n));
        SAD_TRY(R.mem.alloc(&search, n));
        SAD_TRY(R.mem.alloc(&a0, cells * 4));
        SAD_TRY(R.mem.alloc(&a1, cells * 4));
        SAD_TRY(R.mem.alloc(&b0, cells * 4));
        SAD_TRY(R.mem.alloc(&b1, cells * 4));
        SAD_TRY(R.mem.alloc(&d_out, cells * 3));
        const uint32_t padded = div_up(n, 1024u) * 1024u;
        uint32_t *h_order = nullptr, *h_pos = nullptr, *h_hist = nullptr;
        SadPair *h_pairs = nullptr, *h_scratch = nullptr;
        if (use_hil) {
            SAD_TRY(R.mem.alloc(&h_order, n));
            SAD_TRY(R.mem.alloc(&h_pos, n));
            SAD_TRY(R.mem.alloc(&h_pairs, padded));
            SAD_TRY(R.mem.alloc(&h_scratch, padded));
            SAD_TRY(R.mem.alloc(&h_hist, 256u * (padded / 1024u)));
        }
        hipStream_t S = rstream;
        HIP_TRY(hipMemcpyAsync(d_sites, R.sites_h, sizeof(SadSite) * n, hipMemcpyHostToDevice, S));
        const int md = std::max(w, h);
        const float iss = 1.0f / ((float)md * (float)md);
        if (use_hil) {
            int bits = sad_hilbert_bits(w, h);
            uint32_t mk = bits >= 16 ? 0xFFFFFFFFu : (1u << (2 * bits));
            SAD_TRY(launch_hilbert_pairs(d_sites, h_pairs, n, padded, w, h, bits, S));
            SAD_TRY(launch_radix_sort(h_pairs, h_scratch, h_hist, padded, mk, S));
            SAD_TRY(launch_hilbert_order(h_pairs, h_order, h_pos, n, S));
        }
        /* render-job recipe (sad_cuda.cu:1616-1655, 1678-1858): clear, seed, JFA flood, VPT passes, render;
         * the same launches as sadh_render_sites with one JFA round */
        SAD_TRY(sadh_clear_candidates(a0, a1, w, h, (sadh_stream_t)S));
        SAD_TRY(sadh_jfa_seed(a0, d_sites, n, w, h, 1, (sadh_stream_t)S));
        SAD_TRY(launch_search_sites(d_sites, search, n, iss, false, S));
        int step = 1;
        while (step < std::max(w, h)) step <<= 1;
        uint32_t *in = a0, *outp = a1;
        for (int s = step / 2; s >= 1; s /= 2) {
            SAD_TRY(launch_jfa_flood_search(in, outp, search, n, (uint32_t)s, iss, w, h, 1, w, h, S));
            std::swap(in, outp);
        }
        if (in != a0) HIP_TRY(hipMemcpyAsync(a0, in, cells * 16, hipMemcpyDeviceToDevice, S));
        SAD_TRY(launch_search_sites(d_sites, search, n, iss, true, S));
        uint32_t jump_index = 0;
        for (int p = 0; p < passes; ++p) {
            uint32_t stp = sad_pack_jump_step(jump_index, w, h);
            SAD_TRY(launch_vpt_search(a0, a1, b0, b1, search, n, stp, iss, jump_index >> 16, so.fit.cand_radius_scale,
                                      so.fit.cand_radius_probes, so.fit.cand_inject_count, h_order, h_pos,
                                      use_hil ? so.fit.cand_hilbert_probes : 0u,
                                      use_hil ? so.fit.cand_hilbert_window : 0u, 1, w, h, w, h, S));
            jump_index++;
            std::swap(a0, b0);
            std::swap(a1, b1);
        }
        SAD_TRY(launch_render(a0, a1, d_out, d_sites, iss, n, w, h, w, h, S));
        HIP_TRY(hipMemcpyAsync(R.out_h, d_out, sizeof(float) * 3 * cells, hipMemcpyDeviceToHost, S));
        HIP_TRY(hipEventCreateWithFlags(&R.done, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(R.done, S));
        return SADH_OK;
    }

    int admit()
    {
        const int max_fits = so.max_fit_jobs > 0 ? so.max_fit_jobs : 64;
        while (!queue.empty()) {
            const Request& q = queue.front();
            if (in_use + q.bytes > budget) break;   /* arrival order: a request waits for the one before it */
            if (q.kind == 0 && (int)E.fits.size() >= max_fits) break;
            Request r = q;
            queue.pop_front();
            if (r.kind == 0) {
                SAD_TRY(E.add_fit(r.fit, r.id, r.submit_s));
                fit_bytes_of.push_back(r.bytes);
            } else {
                std::unique_ptr<RenderRun> R(new RenderRun());
                R->id = r.id;
                R->sites_h = r.sites;
                R->n = r.n;
                R->width = r.width;
                R->height = r.height;
                R->out_h = r.out;
                R->bytes = r.bytes;
                R->submit_s = r.submit_s;
                R->admit_s = E.now();
                SAD_TRY(render_start(*R));
                renders.push_back(std::move(R));
            }
            in_use += r.bytes;
        }
        return SADH_OK;
    }

    int retire_renders()
    {
        std::vector<std::unique_ptr<RenderRun>> keep;
        for (auto& R : renders) {
            hipError_t e = hipEventQuery(R->done);
            if (e == hipErrorNotReady) { keep.push_back(std::move(R)); continue; }
            if (e != hipSuccess) return fail((int)e, std::string("server render: ") + hipGetErrorString(e));
            SadhServerEvent ev = {};
            ev.id = R->id;
            ev.kind = 1;
            ev.submit_s = R->submit_s;
            ev.admit_s = R->admit_s;
            ev.done_s = E.now();
            events.push_back(ev);
            in_use -= R->bytes;
        }
        renders.swap(keep);
        return SADH_OK;
    }

    int step(int* outstanding)
    {
        SAD_TRY(retire_renders());
        SAD_TRY(admit());
        if (!E.fits.empty()) {
            std::vector<long long> ids_before;
            for (auto& f : E.fits) ids_before.push_back(f->id);
            std::vector<std::unique_ptr<FitRun>> done;
            SAD_TRY(E.tick(done));
            if (!done.empty()) {
                const double t = E.now();
                std::vector<size_t> kept;
                for (size_t i = 0; i < ids_before.size(); ++i) {
                    bool gone = false;
                    for (auto& f : done) gone = gone || f->id == ids_before[i];
                    if (gone) in_use -= fit_bytes_of[i];
                    else kept.push_back(fit_bytes_of[i]);
                }
                fit_bytes_of.swap(kept);
                for (auto& f : done) {
                    E.fill_result(*f, t - f->admit_s);
                    SadhServerEvent ev = {};
                    ev.id = f->id;
                    ev.kind = 0;
                    ev.submit_s = f->submit_s;
                    ev.admit_s = f->admit_s;
                    ev.done_s = t;
                    events.push_back(ev);
                }
            }
        } else if (!renders.empty()) {
            /* no fit to advance: wait for the oldest render rather than spin */
            HIP_TRY(hipEventSynchronize(renders.front()->done));
        }
        SAD_TRY(retire_renders());
        if (outstanding) *outstanding = (int)(E.fits.size() + renders.size() + queue.size());
        return SADH_OK;
    }
};

}  // namespace sadh

using namespace sadh;

struct SadhServer : SadhServerState {};

extern "C" {

int sadh_fit_batch(const SadhFitOptions* opts, SadhBatchJob* jobs, int n_jobs, SadhBatchResult* result)
{
    SADH_REQUIRE(opts && jobs, "fit_batch: null pointer");
    return fit_batch_impl(*opts, jobs, n_jobs, result);
}

int sadh_server_create(const SadhServerOptions* opts, sadh_server_t* out)
{
    SADH_REQUIRE(opts && out, "server_create: null pointer");
    std::unique_ptr<SadhServer> S(new SadhServer());
    S->so = *opts;
    SAD_TRY(S->E.init(opts->fit));
    S->use_hil_render = opts->fit.cand_hilbert_probes > 0 && opts->fit.cand_hilbert_window > 0;
    HIP_TRY(hipStreamCreateWithFlags(&S->rstream, hipStreamNonBlocking));
    if (opts->memory_budget > 0.0) {
        S->budget = (size_t)opts->memory_budget;
    } else {
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        S->budget = (size_t)(0.8 * (double)total_b);
    }
    *out = S.release();
    return SADH_OK;
}

int sadh_server_submit_fit(sadh_server_t server, const SadhBatchJob* job, long long* id)
{
    SADH_REQUIRE(server && job && id, "server_submit_fit: null pointer");
    SADH_REQUIRE(job->target && job->sites_io && job->width > 0 && job->height > 0 && job->n_sites > 0,
                 "server_submit_fit: bad job");
    float dp, pp;
    int cap, pairs;
    SAD_TRY(fit_plan_for(server->so.fit, *job, &dp, &pp, &cap, &pairs));
    if (job->sites_cap < cap)
        return fail(SADH_ERR_CAPACITY, "server_submit_fit: sites_io holds " + std::to_string(job->sites_cap) +
                                           " sites, the plan needs " + std::to_string(cap));
    Request r;
    r.kind = 0;
    r.id = server->next_id++;
    r.submit_s = server->E.now();
    r.fit = *job;
    r.bytes = fit_bytes(server->so.fit, job->width, job->height, cap, pairs, server->E.use_hil);
    if (r.bytes > server->budget) return fail(SADH_ERR_CAPACITY, "server_submit_fit: job exceeds the memory budget");
    server->queue.push_back(r);
    *id = r.id;
    return SADH_OK;
}

int sadh_server_submit_render(sadh_server_t server, const SadSite* sites, uint32_t site_count, int width, int height,
                              float* output, long long* id)
{
    SADH_REQUIRE(server && sites && output && id && site_count > 0 && width > 0 && height > 0,
                 "server_submit_render: bad arguments");
    Request r;
    r.kind = 1;
    r.id = server->next_id++;
    r.submit_s = server->E.now();
    r.sites = sites;
    r.n = site_count;
    r.width = width;
    r.height = height;
    r.out = output;
    r.bytes = render_bytes(site_count, width, height, server->use_hil_render);
    if (r.bytes > server->budget) return fail(SADH_ERR_CAPACITY, "server_submit_render: job exceeds the memory budget");
    server->queue.push_back(r);
    *id = r.id;
    return SADH_OK;
}

int sadh_server_step(sadh_server_t server, int* outstanding)
{
    SADH_REQUIRE(server, "server_step: null server");
    return server->step(outstanding);
}

int sadh_server_poll(sadh_server_t server, SadhServerEvent* events, int cap, int* count)
{
    SADH_REQUIRE(server && count && (cap == 0 || events), "server_poll: null pointer");
    int k = std::min(cap, (int)server->events.size());
    for (int i = 0; i < k; ++i) events[i] = server->events[i];
    server->events.erase(server->events.begin(), server->events.begin() + k);
    *count = k;
    return SADH_OK;
}

double sadh_server_time(sadh_server_t server) { return server ? server->E.now() : 0.0; }

int sadh_server_destroy(sadh_server_t server)
{
    delete server;
    return SADH_OK;
}

}  // extern "C"
