/*
 * fit_common.h - host-side pieces shared by the single-image and batched fit
 * drivers: kernel launcher declarations, status macros, RAII device arena and
 * stream pair, and the capacity/percentile plan.
 */
#ifndef SAD_FIT_COMMON_H
#define SAD_FIT_COMMON_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "batch.h"

namespace sadh {

int launch_render_cells(const uint32_t*, const uint32_t*, float*, const SadSite*, float, uint32_t, int, int, int, int,
                        hipStream_t);
int launch_tau_heatmap(const uint32_t*, const uint32_t*, float*, const SadSite*, uint32_t, float, float, float, float,
                       int, int, int, int, hipStream_t);
/* exportActiveSites, training_report.swift:21-55: min, mean and max of exp(log_tau) over the active sites
 * (host, index order; all zero when none is active) */
void tau_range(const SadSite* sites, uint32_t n, float* tmin, float* tmean, float* tmax);
/* the debug pictures of a finished fit into r->cells_render and r->tau_heatmap (each optional) */
int debug_renders(const uint32_t* c0, const uint32_t* c1, const SadSite* d_sites, const SadSite* host_sites,
                  uint32_t n, float iss, int width, int height, int cw, int ch, float* d_scratch, SadhFitResult* r,
                  hipStream_t stream);
int launch_render(const uint32_t*, const uint32_t*, float*, const SadSite*, float, uint32_t, int, int, int, int,
                  hipStream_t);
int launch_gradients(const uint32_t*, const uint32_t*, const float*, const float*, float* const[10], const SadSite*,
                     float, uint32_t, float*, bool, int, int, int, int, hipStream_t, const float* rendered = nullptr,
                     float ssim_weight = 0.0f);
int launch_site_stats(const uint32_t*, const uint32_t*, const float*, const float*, const SadSite*, float, uint32_t,
                      float* const[8], int, int, int, int, hipStream_t);
int launch_adam(SadSite*, SadAdamState*, float* const[10], float, float, float, float, float, float, float, float,
                float, uint32_t, uint32_t, int, int, bool, hipStream_t);
int launch_clear(float* const*, int, uint32_t, hipStream_t);
int launch_clear_fix(long long*, size_t, hipStream_t);
int launch_tau_adam(const uint32_t*, const uint32_t*, float*, float*, uint32_t*, float, int, int, int, SadSite*,
                    SadAdamState*, float* const[10], float, float, float, float, float, float, float, float, float,
                    uint32_t, uint32_t, int, int, hipStream_t);
int launch_tau_diffuse(const uint32_t*, const uint32_t*, const SadSite*, const float*, const float*, float*, float*,
                       uint32_t, float, int, int, int, hipStream_t);
int launch_split(SadSite*, SadAdamState*, const uint32_t*, uint32_t, const float* const[8], uint32_t, const float*,
                 int, int, hipStream_t);
int launch_sse(const float*, const float*, const float*, float*, int, int, hipStream_t);
int launch_sse_fix(const float*, const float*, const float*, long long*, int, int, hipStream_t);
int launch_count_active(const SadSite*, uint32_t*, uint32_t, hipStream_t);
int launch_init_gradient(SadSite*, uint32_t, uint32_t*, const float*, const float*, float, uint32_t, float, float, int,
                         int, hipStream_t);
int launch_densify_pairs(const SadSite*, const float*, const float*, SadPair*, uint32_t, float, float, uint32_t,
                         hipStream_t);
int launch_prune_pairs(const SadSite*, const float*, SadPair*, uint32_t, float, uint32_t, hipStream_t);
int launch_write_indices(const SadPair*, uint32_t*, uint32_t, hipStream_t);
int launch_prune_sites(SadSite*, const uint32_t*, uint32_t, hipStream_t);
int launch_hilbert_pairs(const SadSite*, SadPair*, uint32_t, uint32_t, int, int, int, hipStream_t);
int launch_hilbert_order(const SadPair*, uint32_t*, uint32_t*, uint32_t, hipStream_t);
int launch_radix_sort(SadPair*, SadPair*, uint32_t*, uint32_t, uint32_t, hipStream_t);
int launch_vpt_search(const uint32_t*, const uint32_t*, uint32_t*, uint32_t*, const SadSearchSite*, uint32_t,
                      uint32_t, float, uint32_t, float, uint32_t, uint32_t, const uint32_t*, const uint32_t*, uint32_t,
                      uint32_t, int, int, int, int, int, hipStream_t);
int launch_search_sites(const SadSite*, SadSearchSite*, uint32_t, float, bool, hipStream_t);
int launch_jfa_flood_search(const uint32_t*, uint32_t*, const SadSearchSite*, uint32_t, uint32_t, float, int, int,
                            int, int, int, hipStream_t);

/* ------------------------------------------------------------ helpers --- */

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) return fail((int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define SAD_TRY(expr)                  \
    do {                               \
        int r_ = (expr);               \
        if (r_ != SADH_OK) return r_;  \
    } while (0)

/* Device allocations released at scope exit. */
struct DevArena {
    std::vector<void*> ptrs;
    ~DevArena()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T** out, size_t count)
    {
        void* p = nullptr;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return fail((int)e, "hipMalloc(" + std::to_string(bytes) + " B) failed");
        ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return SADH_OK;
    }
};

struct StreamPair {
    hipStream_t main = nullptr, cand = nullptr;
    hipEvent_t cand_done = nullptr, sites_ready = nullptr;
    ~StreamPair()
    {
        if (cand_done) (void)hipEventDestroy(cand_done);
        if (sites_ready) (void)hipEventDestroy(sites_ready);
        if (cand) (void)hipStreamDestroy(cand);
        if (main) (void)hipStreamDestroy(main);
    }
};

/* AoS [N][16] gradient lines (the fit's accumulation layout) */
int launch_gradients_aos(const uint32_t*, const uint32_t*, const float*, const float*, float*, const SadSite*, float,
                         uint32_t, float*, bool, int, int, int, int, hipStream_t, const float* rendered = nullptr,
                         float ssim_weight = 0.0f);
int launch_ssim(const float* rendered, const float* target, const float* mask, float* acc, int width, int height,
                hipStream_t stream);
int launch_ssim_fix(const float* rendered, const float* target, const float* mask, long long* acc, int width,
                    int height, hipStream_t stream);
int launch_pack_inference(const SadSite* sites, uint32_t n, int width, int height, SadInferenceSite* out,
                          SadInferenceQuant* quant, hipStream_t stream);
int launch_unpack_inference(const SadInferenceSite* in, const SadInferenceQuant* quant, uint32_t n, int width,
                            int height, SadSite* out, hipStream_t stream);
/* The fit's exact cross-tile sums: gradients as fixed-point lines of FIX_GRAD_STRIDE int64 words per site
 * ((lo, hi) of values 0-9 in Site field order), removal deltas as one (lo, hi) pair per site, densify
 * statistics as lines of FIX_STAT_STRIDE words (fix_split, sad_device.h). */
/* the fit's tail (4 tau passes, the last fused with Adam + clear of the lines);
 * nbr_lists: 8 uint32 per site of scratch for the neighbour lists the passes share */
int launch_tau_adam_fix(const uint32_t*, const uint32_t*, long long*, float*, float*, float*, float, int, int, int,
                        SadSite*, SadAdamState*, float, float, float, float, float, float, float, float, float,
                        uint32_t, uint32_t, int, int, hipStream_t, uint32_t* nbr_lists, bool lists_built);
/* nbr_lists (nullable): the same launch also writes the tau passes' neighbour lists on this field */
int launch_gradients_fix(const uint32_t*, const uint32_t*, const float*, const float*, long long*, const SadSite*,
                         float, uint32_t, long long*, bool, int, int, int, int, hipStream_t,
                         const float* rendered = nullptr, float ssim_weight = 0.0f, uint32_t* nbr_lists = nullptr,
                         float nbr_lambda = 0.0f, int ds = 1);
int launch_site_stats_fix(const uint32_t*, const uint32_t*, const float*, const float*, const SadSite*, float,
                          uint32_t, long long*, int, int, int, int, hipStream_t);
int launch_split_fix(SadSite*, SadAdamState*, const uint32_t*, uint32_t, const long long*, uint32_t, const float*,
                     int, int, hipStream_t);
int launch_densify_pairs_fix(const SadSite*, const long long*, SadPair*, uint32_t, float, float, uint32_t,
                             hipStream_t);
int launch_prune_pairs_fix(const SadSite*, const long long*, SadPair*, uint32_t, float, uint32_t, hipStream_t);

/* batched fit (one launch over every job) */
int launch_gradients_batch(const BatchJob*, uint32_t, uint32_t, bool, hipStream_t, int hash_slots = 128);
int launch_vpt_batch(const BatchJob*, uint32_t, const JobStep*, uint32_t, uint32_t, const BatchStep&, hipStream_t);
int launch_tau_adam_batch(const BatchJob*, uint32_t, const JobStep*, uint32_t, const BatchStep&, hipStream_t,
                          uint32_t block = 256);
int launch_update_jobs(BatchJob*, const uint32_t*, const uint32_t*, const uint32_t*, uint32_t, hipStream_t);
/* log ticks: the renders of the jobs whose step flag `log` is set, then their squared-error sums and active-site
 * counts as (lo, hi, count, 0) words per job */
int launch_render_batch(const BatchJob*, uint32_t, const JobStep*, uint32_t, hipStream_t);
int launch_log_sums_batch(const BatchJob*, uint32_t, const JobStep*, uint32_t, uint32_t, long long*, hipStream_t);
/* the sites of the jobs whose step flag `retire` is set, at their site_begin offsets of `out` */
int launch_gather_sites_batch(const BatchJob*, uint32_t, const JobStep*, uint32_t, SadSite*, hipStream_t);
/* the Hilbert tables of the jobs with hil_n > 0 (pairs, scratch: 1024-padded total; hist: 256 per 1024 pairs) */
int launch_hilbert_batch(const BatchJob*, uint32_t, const JobStep*, uint32_t, uint32_t, uint32_t, SadPair*, SadPair*,
                         uint32_t*, hipStream_t);

/* a fixed-point (lo, hi) sum read back on the host, with the device's rounding (fix_value) */
inline float fix_value_host(const long long w[2]) { return (float)((double)w[1] + (double)w[0] * 0x1p-40); }

/* capacity and densify/prune percentiles of a fit (sad_cuda.cu:904-1098) */
int plan(const SadhFitOptions& o, int width, int height, float* d_pct, float* p_pct, int* cap, int* pairs);

}  // namespace sadh

#endif /* SAD_FIT_COMMON_H */
