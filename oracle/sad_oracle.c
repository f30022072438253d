/*
 * sad_oracle.c - CPU restatement of the SAD fit/render path (TEST INFRASTRUCTURE).
 *
 * This file is the checker for the HIP library, nothing else: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product path never calls into it.
 *
 * Every function restates one kernel or host routine of the reference and cites
 * the file:line it follows (paths relative to the reference checkout).  Loops run
 * over pixels, cells or sites exactly as the reference's CUDA threads do; where a
 * reference kernel has a data race (jfaSeedKernel: several sites writing one
 * cell) the deterministic rule the HIP library implements is stated and used.
 * Gradient-type sums (fp32 atomics in the reference, order-nondeterministic) come
 * in two forms: accumulated in double, with the sum of absolute contributions
 * alongside so tests can bound the float-output entry points' fp32 summation
 * error; and, for the fit, in the library's own order (per 16x16 tile and site,
 * chunks of ORC_CHUNK entries, partials added in exact fixed point), so the
 * oracle's fit matches the GPU's bit for bit.
 *
 * Floating-point primitives come from sad_amd/csrc/sad_math.h, compiled with
 * -ffp-contract=off, so candidate lists and score keys are bit-comparable with
 * the GPU.  Because the primitives are shared, they and the functions built on
 * them are pinned independently of the GPU: exp/log/pow/fp16 against libm and
 * numpy error bounds, render and gradients against closed forms and float64
 * central differences, site statistics, removal deltas, the fixed-point sums and
 * the bpp solver against brute force, conservation laws and hand-counted
 * schedules (tests/test_oracle.py, tests/test_oracle_sums.py).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#include "../sad_amd/csrc/sad_math.h"

#define ORC_API __attribute__((visibility("default")))

int orc_num_threads(void) { return omp_get_max_threads(); }
ORC_API void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
ORC_API int orc_get_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------ primitives (exported) -- */
ORC_API void orc_expf_array(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = sad_expf(x[i]); }
ORC_API void orc_logf_array(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = sad_logf(x[i]); }
ORC_API void orc_f32_to_f16_array(const float* x, uint16_t* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = sad_f32_to_f16(x[i]); }
ORC_API void orc_f16_to_f32_array(const uint16_t* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = sad_f16_to_f32(x[i]); }
ORC_API uint32_t orc_hilbert_index(uint32_t x, uint32_t y, int bits) { return sad_hilbert_index(x, y, bits); }
ORC_API uint32_t orc_jump_step(uint32_t i, int w, int h) { return sad_jump_step(i, w, h); }

/* ------------------------------------------------------ forward weights --
 * Shared by render, gradients and stats: logits -exp(log_tau)*dmix over the 8
 * candidates, softmax with max subtraction, blended colour
 * (render.cu:34-91, gradients.cu:200-247, utils.cu:481-523).
 * Returns 0 when every candidate is invalid (max logit == -inf). */
static int orc_forward_px(const uint32_t ids[SAD_K], const SadSite* sites, uint32_t n,
                          float ux, float uy, float iss, float w[SAD_K], float pred[3])
{
    float logits[SAD_K];
    float mx = -INFINITY;
    for (int i = 0; i < SAD_K; ++i) {
        uint32_t id = ids[i];
        if (id >= n || sites[id].pos_x < 0.0f) { logits[i] = -INFINITY; continue; }
        const SadSite* s = &sites[id];
        float tau = sad_expf(s->log_tau);
        float d = sad_dmix2(s->pos_x, s->pos_y, s->dir_x, s->dir_y, s->log_aniso, s->radius, ux, uy, iss);
        logits[i] = -tau * d;
        mx = fmaxf(mx, logits[i]);
    }
    if (isinf(mx) && mx < 0.0f) return 0;
    float sum = 0.0f;
    for (int i = 0; i < SAD_K; ++i) { w[i] = sad_expf(logits[i] - mx); sum += w[i]; }
    float inv = 1.0f / fmaxf(sum, 1e-8f);
    pred[0] = pred[1] = pred[2] = 0.0f;
    for (int i = 0; i < SAD_K; ++i) {
        w[i] = w[i] * inv;
        uint32_t id = ids[i];
        if (id < n && sites[id].pos_x >= 0.0f) {
            pred[0] += w[i] * sites[id].col_r;
            pred[1] += w[i] * sites[id].col_g;
            pred[2] += w[i] * sites[id].col_b;
        }
    }
    return 1;
}

static void orc_load_ids(const uint32_t* c0, const uint32_t* c1, int cell, uint32_t ids[SAD_K])
{
    for (int k = 0; k < 4; ++k) { ids[k] = c0[4 * (size_t)cell + k]; ids[4 + k] = c1[4 * (size_t)cell + k]; }
}

/* renderVoronoiKernel, render.cu:5-92.  The colour blend additionally skips
 * NaN/Inf weights (render.cu:83), which the gradient kernel does not. */
/* cells != 0: the site colours replaced by sad_hash_color of the id
 * (renderVoronoiColoring, shaders/encoders/render_encoder.metal:73-134) */
static void render_impl(const uint32_t* cand0, const uint32_t* cand1, float* out, const SadSite* sites, float iss,
                        uint32_t n, int width, int height, int cand_w, int cand_h, int cells)
{
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            size_t p = (size_t)y * width + x;
            uint32_t ids[SAD_K];
            orc_load_ids(cand0, cand1, sad_cand_index(x, y, width, height, cand_w, cand_h), ids);
            float logits[SAD_K];
            float mx = -INFINITY;
            for (int i = 0; i < SAD_K; ++i) {
                uint32_t id = ids[i];
                if (id >= n || sites[id].pos_x < 0.0f) { logits[i] = -INFINITY; continue; }
                const SadSite* s = &sites[id];
                float tau = sad_expf(s->log_tau);
                float d = sad_dmix2(s->pos_x, s->pos_y, s->dir_x, s->dir_y, s->log_aniso, s->radius,
                                    (float)x, (float)y, iss);
                logits[i] = -tau * d;
                mx = fmaxf(mx, logits[i]);
            }
            if (isinf(mx) && mx < 0.0f) { out[3 * p] = out[3 * p + 1] = out[3 * p + 2] = 0.0f; continue; }
            float w[SAD_K], sum = 0.0f;
            for (int i = 0; i < SAD_K; ++i) { w[i] = sad_expf(logits[i] - mx); sum += w[i]; }
            float inv = 1.0f / fmaxf(sum, 1e-8f);
            float c[3] = {0.0f, 0.0f, 0.0f};
            for (int i = 0; i < SAD_K; ++i) {
                float wi = w[i] * inv;
                uint32_t id = ids[i];
                if (!isnan(wi) && !isinf(wi) && id < n && sites[id].pos_x >= 0.0f) {
                    float col[3] = {sites[id].col_r, sites[id].col_g, sites[id].col_b};
                    if (cells) sad_hash_color(id, col);
                    c[0] += wi * col[0];
                    c[1] += wi * col[1];
                    c[2] += wi * col[2];
                }
            }
            out[3 * p] = c[0]; out[3 * p + 1] = c[1]; out[3 * p + 2] = c[2];
        }
    }
}

ORC_API void orc_render(const uint32_t* cand0, const uint32_t* cand1, float* out,
                        const SadSite* sites, float iss, uint32_t n,
                        int width, int height, int cand_w, int cand_h)
{
    render_impl(cand0, cand1, out, sites, iss, n, width, height, cand_w, cand_h, 0);
}

ORC_API void orc_render_cells(const uint32_t* cand0, const uint32_t* cand1, float* out,
                              const SadSite* sites, float iss, uint32_t n,
                              int width, int height, int cand_w, int cand_h)
{
    render_impl(cand0, cand1, out, sites, iss, n, width, height, cand_w, cand_h, 1);
}

/* hashColor (shaders/sad_common.metal:542-565) of each id */
ORC_API void orc_hash_colors(const uint32_t* ids, uint32_t n, float* rgb)
{
    for (uint32_t i = 0; i < n; ++i) sad_hash_color(ids[i], rgb + 3 * (size_t)i);
}

/* renderCentroidsTauHeatmap, shaders/encoders/render_encoder.metal:136-190 */
ORC_API void orc_tau_heatmap(const uint32_t* cand0, const uint32_t* cand1, float* out, const SadSite* sites,
                             uint32_t n, float tmin, float tmean, float tmax, float dot, int width, int height,
                             int cand_w, int cand_h)
{
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            size_t p = (size_t)y * width + x;
            uint32_t ids[SAD_K];
            orc_load_ids(cand0, cand1, sad_cand_index(x, y, width, height, cand_w, cand_h), ids);
            float acc[3] = {0.0f, 0.0f, 0.0f}, acc_a = 0.0f;
            for (int i = 0; i < SAD_K; ++i) {
                uint32_t id = ids[i];
                if (id >= n || sites[id].pos_x < 0.0f) continue;
                const SadSite* s = &sites[id];
                float a = sad_dot_alpha((float)x - s->pos_x, (float)y - s->pos_y, dot);
                if (a <= 0.0f) continue;
                float col[3];
                sad_heat_color(sad_expf(s->log_tau), tmin, tmean, tmax, col);
                float blend = a * (1.0f - acc_a);
                acc[0] += col[0] * blend;
                acc[1] += col[1] * blend;
                acc[2] += col[2] * blend;
                acc_a += blend;
            }
            for (int c = 0; c < 3; ++c) out[3 * p + c] = acc[c] + 0.05f * (1.0f - acc_a);
        }
    }
}

/* initCandidatesKernel, render.cu:95-136 */
ORC_API void orc_init_candidates(uint32_t* cand0, uint32_t* cand1, uint32_t n, uint32_t seed,
                                 int per_pixel, int width, int height)
{
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            size_t p = (size_t)y * width + x;
            uint32_t idx[SAD_K];
            if (per_pixel && n == (uint32_t)(width * height)) {
                static const int nb[7][2] = {{-1,0},{1,0},{0,-1},{0,1},{-1,-1},{1,-1},{-1,1}};
                idx[0] = (uint32_t)p;
                for (int i = 0; i < 7; ++i) {
                    int nx = sad_clampi(x + nb[i][0], 0, width - 1);
                    int ny = sad_clampi(y + nb[i][1], 0, height - 1);
                    idx[i + 1] = (uint32_t)(ny * width + nx);
                }
            } else {
                uint32_t state = seed ^ ((uint32_t)x * 1973u) ^ ((uint32_t)y * 9277u);
                for (int i = 0; i < SAD_K; ++i) {
                    state = sad_xorshift32(state + (uint32_t)i);
                    idx[i] = state % n;
                }
            }
            for (int k = 0; k < 4; ++k) { cand0[4 * p + k] = idx[k]; cand1[4 * p + k] = idx[4 + k]; }
        }
    }
}

/* clearCandidatesKernel, render.cu:139-154 */
ORC_API void orc_clear_candidates(uint32_t* cand0, uint32_t* cand1, int width, int height)
{
    size_t m = (size_t)width * height * 4;
    for (size_t i = 0; i < m; ++i) { cand0[i] = SAD_EMPTY; cand1[i] = SAD_EMPTY; }
}

/* jfaSeedKernel, jfa.cu:66-92: every active site writes its id into slot .x of
 * its home cell int(pos / downscale), clamped.  When several sites share a cell
 * the reference's outcome depends on thread scheduling; this build defines it as
 * the largest site id (what sequential execution in id order leaves behind). */
ORC_API void orc_jfa_seed(uint32_t* cand0, const SadSite* sites, uint32_t n,
                          int width, int height, int downscale)
{
    float inv_ds = 1.0f / fmaxf((float)downscale, 1.0f);
    for (uint32_t g = 0; g < n; ++g) {
        const SadSite* s = &sites[g];
        if (s->pos_x < 0.0f) continue;
        int x = sad_clampi((int)(s->pos_x * inv_ds), 0, width - 1);
        int y = sad_clampi((int)(s->pos_y * inv_ds), 0, height - 1);
        cand0[4 * ((size_t)y * width + x)] = g;
    }
}

/* insertClosest4, jfa.cu:14-63 */
static void orc_insert4(uint32_t bi[4], float bd[4], uint32_t c, float ux, float uy,
                        const SadSite* sites, uint32_t n, float iss)
{
    if (c >= n) return;
    for (int i = 0; i < 4; ++i) if (bi[i] == c) return;
    if (sites[c].pos_x < 0.0f) return;
    float d2 = sad_site_key(&sites[c], ux, uy, iss);
    if (!(d2 < bd[3])) return;
    int pos = 3;
    while (pos > 0 && d2 < bd[pos - 1]) pos--;
    for (int i = 3; i > pos; --i) { bd[i] = bd[i - 1]; bi[i] = bi[i - 1]; }
    bd[pos] = d2; bi[pos] = c;
}

/* jfaFloodKernel, jfa.cu:95-144 */
ORC_API void orc_jfa_flood(const uint32_t* in0, uint32_t* out0, const SadSite* sites, uint32_t n,
                           uint32_t step, float iss, int width, int height, int downscale,
                           int target_w, int target_h)
{
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            float ux = fminf((float)x * (float)downscale, (float)(target_w - 1));
            float uy = fminf((float)y * (float)downscale, (float)(target_h - 1));
            uint32_t bi[4] = {SAD_EMPTY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY};
            float bd[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
            for (int dy = -1; dy <= 1; ++dy) {
                for (int dx = -1; dx <= 1; ++dx) {
                    int sx = sad_clampi(x + dx * (int)step, 0, width - 1);
                    int sy = sad_clampi(y + dy * (int)step, 0, height - 1);
                    const uint32_t* c = &in0[4 * ((size_t)sy * width + sx)];
                    for (int k = 0; k < 4; ++k) orc_insert4(bi, bd, c[k], ux, uy, sites, n, iss);
                }
            }
            uint32_t* o = &out0[4 * ((size_t)y * width + x)];
            for (int k = 0; k < 4; ++k) o[k] = bi[k];
        }
    }
}

/* packCandidateSitesKernel, candidate_pack.cu:3-27 */
ORC_API void orc_pack_sites(const SadSite* sites, SadPackedSite* packed, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i) packed[i] = sad_pack_site(&sites[i]);
}

/* insertClosest8Packed, jfa.cu:176-248: sorted top-8 with strict '<' insertion */
static void orc_insert8(uint32_t bi[8], float bd[8], uint32_t c, float ux, float uy,
                        const SadPackedSite* ps, uint32_t n, float iss)
{
    if (c >= n) return;
    for (int i = 0; i < 8; ++i) if (bi[i] == c) return;
    float d2 = sad_packed_key(ps[c], ux, uy, iss);
    if (!(d2 < bd[7])) return;
    int pos = 7;
    while (pos > 0 && d2 < bd[pos - 1]) pos--;
    for (int i = 7; i > pos; --i) { bd[i] = bd[i - 1]; bi[i] = bi[i - 1]; }
    bd[pos] = d2; bi[pos] = c;
}

/* updateCandidatesKernel (VPT pass), jfa.cu:271-386 */
ORC_API void orc_update_candidates(const uint32_t* in0, const uint32_t* in1,
                                   uint32_t* out0, uint32_t* out1,
                                   const SadPackedSite* ps, uint32_t n, uint32_t step, float iss,
                                   uint32_t step_high, float radius_scale, uint32_t radius_probes,
                                   uint32_t inject_count, const uint32_t* hil_order,
                                   const uint32_t* hil_pos, uint32_t hil_probes, uint32_t hil_window,
                                   int downscale, int target_w, int target_h, int width, int height)
{
    uint32_t step_index = step & 0xFFFFu;
    uint32_t jump = step >> 16;
    if (jump == 0u) jump = 1u;
    uint32_t full_index = (step_high << 16) | step_index;
    float rad = radius_scale * ((float)width / 1024.0f);
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            size_t p = (size_t)y * width + x;
            float ux = fminf((float)x * (float)downscale, (float)(target_w - 1));
            float uy = fminf((float)y * (float)downscale, (float)(target_h - 1));
            uint32_t bi[8];
            float bd[8];
            for (int i = 0; i < 8; ++i) { bi[i] = SAD_EMPTY; bd[i] = INFINITY; }
            for (int k = 0; k < 4; ++k) orc_insert8(bi, bd, in0[4 * p + k], ux, uy, ps, n, iss);
            for (int k = 0; k < 4; ++k) orc_insert8(bi, bd, in1[4 * p + k], ux, uy, ps, n, iss);
            static const int off[4][2] = {{-1,0},{1,0},{0,-1},{0,1}};
            for (int j = 0; j < 4; ++j) {
                int nx = sad_clampi(x + off[j][0] * (int)jump, 0, width - 1);
                int ny = sad_clampi(y + off[j][1] * (int)jump, 0, height - 1);
                size_t q = (size_t)ny * width + nx;
                for (int k = 0; k < 4; ++k) orc_insert8(bi, bd, in0[4 * q + k], ux, uy, ps, n, iss);
                for (int k = 0; k < 4; ++k) orc_insert8(bi, bd, in1[4 * q + k], ux, uy, ps, n, iss);
            }
            uint32_t state = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((step_index + jump) * 83492791u);
            for (uint32_t r = 0; r < radius_probes; ++r) {
                state = sad_xorshift32(state);
                float a = ((float)state * (1.0f / 4294967296.0f)) * 6.2831853f;
                int dx = (int)(sad_cosf(a) * rad);
                int dy = (int)(sad_sinf(a) * rad);
                int px = sad_clampi(x + dx, 0, width - 1);
                int py = sad_clampi(y + dy, 0, height - 1);
                if (px == x && py == y) continue;
                size_t q = (size_t)py * width + px;
                for (int k = 0; k < 4; ++k) orc_insert8(bi, bd, in0[4 * q + k], ux, uy, ps, n, iss);
                for (int k = 0; k < 4; ++k) orc_insert8(bi, bd, in1[4 * q + k], ux, uy, ps, n, iss);
            }
            if (hil_probes > 0 && hil_window > 0) {
                uint32_t best = bi[0];
                if (best < n) {
                    uint32_t pos = hil_pos[best];
                    uint32_t span = hil_window * 2u + 1u;
                    uint32_t hs = ((uint32_t)x * 2654435761u) ^ ((uint32_t)y * 1597334677u) ^ (full_index * 374761393u);
                    for (uint32_t i = 0; i < hil_probes; ++i) {
                        hs = sad_xorshift32(hs + i);
                        int o = (int)(hs % span) - (int)hil_window;
                        int idx = sad_clampi((int)pos + o, 0, (int)(n - 1u));
                        orc_insert8(bi, bd, hil_order[idx], ux, uy, ps, n, iss);
                    }
                }
            }
            uint32_t is = ((uint32_t)x * 1664525u) ^ ((uint32_t)y * 1013904223u) ^ (full_index * 374761393u);
            for (uint32_t i = 0; i < inject_count; ++i) {
                is = sad_xorshift32(is + i);
                orc_insert8(bi, bd, is % n, ux, uy, ps, n, iss);
            }
            for (int k = 0; k < 4; ++k) { out0[4 * p + k] = bi[k]; out1[4 * p + k] = bi[4 + k]; }
        }
    }
}

/* -------------------------------------------------------------- gradients --
 * computeGradientsTiledKernel, gradients.cu:103-486: L2 loss, plus
 * -ssim_weight * SSIM when ssim_weight > 0 (gradients.cu:236-266: each pixel's
 * own 3x3 window SSIM differentiated through its own rendered luma, read from
 * `rendered`).  The per-candidate analytic
 * derivative of gradients.cu:352-420 is sad_grad_terms (sad_math.h), which is
 * pinned separately against central finite differences of the render; the
 * tile hash is an accumulation device only, so the oracle sums per site directly.
 * acc[10*id + j] (double) receives the sum of contributions for parameter j in
 * Site field order {pos.x, pos.y, log_tau, radius, r, g, b, dir.x, dir.y,
 * log_aniso}; abs (nullable) receives the sum of |contribution|.  rdelta/rabs
 * (nullable) receive removal deltas (gradients.cu:294-350) when compute_removal. */
static void orc_grad_candidate(const SadSite* s, float w, const float g[3], const float pred[3],
                               float ux, float uy, float iss, float out[10])
{
    float l1 = sad_expf(s->log_aniso);
    float l2 = 1.0f / fmaxf(l1, 1e-8f);
    sad_grad_terms(s->pos_x, s->pos_y, sad_expf(s->log_tau), s->radius * sqrtf(iss), s->dir_x, s->dir_y, l1, l2,
                   s->col_r, s->col_g, s->col_b, ux, uy, w, g[0], g[1], g[2], pred[0], pred[1], pred[2], iss,
                   sqrtf(iss), out);
}

static float orc_removal_delta(const float pred[3], const float tgt[3], const SadSite* s, float wt)
{
    return sad_removal_delta(pred[0], pred[1], pred[2], tgt[0], tgt[1], tgt[2], s->col_r, s->col_g, s->col_b, wt);
}

/* Generic per-pixel gradient driver.  mode 0: loss gradient 2(pred - tgt) with
 * mask (tiled kernel); mode 1: upstream grad_output, no mask (autograd kernel,
 * sad_bindings.cu:100-207, which also skips pixels whose max logit is +inf). */
static void orc_gradients_impl(int mode, const uint32_t* cand0, const uint32_t* cand1,
                               const float* target, const float* mask, const float* grad_out,
                               const SadSite* sites, float iss, uint32_t n,
                               int compute_removal, int width, int height, int cand_w, int cand_h,
                               double* acc, double* abs_acc, double* rdelta, double* rabs,
                               const float* rendered, float ssim_weight)
{
    int T = omp_get_max_threads();
    size_t nn = (size_t)n;
    double* priv = (double*)calloc((size_t)T * nn * 10, sizeof(double));
    double* privabs = abs_acc ? (double*)calloc((size_t)T * nn * 10, sizeof(double)) : NULL;
    double* privr = rdelta ? (double*)calloc((size_t)T * nn, sizeof(double)) : NULL;
    double* privrabs = rabs ? (double*)calloc((size_t)T * nn, sizeof(double)) : NULL;
    #pragma omp parallel
    {
        int t = omp_get_thread_num();
        double* my = priv + (size_t)t * nn * 10;
        double* myabs = privabs ? privabs + (size_t)t * nn * 10 : NULL;
        double* myr = privr ? privr + (size_t)t * nn : NULL;
        double* myrabs = privrabs ? privrabs + (size_t)t * nn : NULL;
        #pragma omp for schedule(static)
        for (int y = 0; y < height; ++y) {
            for (int x = 0; x < width; ++x) {
                size_t p = (size_t)y * width + x;
                if (mode == 0 && mask[p] <= 0.0f) continue;
                uint32_t ids[SAD_K];
                orc_load_ids(cand0, cand1, sad_cand_index(x, y, width, height, cand_w, cand_h), ids);
                float ux = (float)x, uy = (float)y;
                float w[SAD_K], pred[3];
                if (mode == 1) {
                    /* compute_candidate_weights_raw returns false on !any or isinf(max) */
                    float mx = -INFINITY; int any = 0;
                    for (int i = 0; i < SAD_K; ++i) {
                        uint32_t id = ids[i];
                        if (id >= n || sites[id].pos_x < 0.0f) continue;
                        const SadSite* s = &sites[id];
                        float l = -sad_expf(s->log_tau) * sad_dmix2(s->pos_x, s->pos_y, s->dir_x, s->dir_y,
                                                                    s->log_aniso, s->radius, ux, uy, iss);
                        mx = fmaxf(mx, l); any = 1;
                    }
                    if (!any || isinf(mx)) continue;
                }
                if (!orc_forward_px(ids, sites, n, ux, uy, iss, w, pred)) continue;
                float g[3], tgt[3] = {0, 0, 0};
                if (mode == 0) {
                    tgt[0] = target[3 * p]; tgt[1] = target[3 * p + 1]; tgt[2] = target[3 * p + 2];
                    g[0] = 2.0f * (pred[0] - tgt[0]);
                    g[1] = 2.0f * (pred[1] - tgt[1]);
                    g[2] = 2.0f * (pred[2] - tgt[2]);
                    if (ssim_weight > 0.0f) {
                        SadSsimStats st = sad_ssim_window(rendered, target, width, height, x, y);
                        float f = -ssim_weight * sad_ssim_dcentre(st, sad_luma3(rendered + 3 * p), sad_luma3(tgt));
                        g[0] += f * 0.299f;
                        g[1] += f * 0.587f;
                        g[2] += f * 0.114f;
                    }
                } else {
                    g[0] = grad_out[3 * p]; g[1] = grad_out[3 * p + 1]; g[2] = grad_out[3 * p + 2];
                }
                if (mode == 0 && compute_removal && myr) {
                    uint32_t us[SAD_K]; float uw[SAD_K]; int nu = 0;
                    for (int i = 0; i < SAD_K; ++i) {
                        uint32_t id = ids[i];
                        if (id >= n || sites[id].pos_x < 0.0f) continue;
                        int found = 0;
                        for (int j = 0; j < nu; ++j) if (us[j] == id) { uw[j] += w[i]; found = 1; break; }
                        if (!found) { us[nu] = id; uw[nu] = w[i]; nu++; }
                    }
                    for (int j = 0; j < nu; ++j) {
                        if (uw[j] > 0.0f) {
                            float dl = orc_removal_delta(pred, tgt, &sites[us[j]], uw[j]);
                            myr[us[j]] += dl;
                            if (myrabs) myrabs[us[j]] += fabs((double)dl);
                        }
                    }
                }
                for (int i = 0; i < SAD_K; ++i) {
                    uint32_t id = ids[i];
                    if (id >= n || sites[id].pos_x < 0.0f) continue;
                    float c[10];
                    orc_grad_candidate(&sites[id], w[i], g, pred, ux, uy, iss, c);
                    for (int j = 0; j < 10; ++j) {
                        my[10 * (size_t)id + j] += c[j];
                        if (myabs) myabs[10 * (size_t)id + j] += fabs((double)c[j]);
                    }
                }
            }
        }
    }
    for (int t = 0; t < T; ++t) {
        for (size_t i = 0; i < nn * 10; ++i) {
            acc[i] += priv[(size_t)t * nn * 10 + i];
            if (abs_acc) abs_acc[i] += privabs[(size_t)t * nn * 10 + i];
        }
        if (rdelta) for (size_t i = 0; i < nn; ++i) rdelta[i] += privr[(size_t)t * nn + i];
        if (rabs) for (size_t i = 0; i < nn; ++i) rabs[i] += privrabs[(size_t)t * nn + i];
    }
    free(priv); free(privabs); free(privr); free(privrabs);
}

ORC_API void orc_gradients(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                           const float* mask, const SadSite* sites, float iss, uint32_t n,
                           int compute_removal, int width, int height, int cand_w, int cand_h,
                           double* acc, double* abs_acc, double* rdelta, double* rabs)
{
    orc_gradients_impl(0, cand0, cand1, target, mask, NULL, sites, iss, n, compute_removal,
                       width, height, cand_w, cand_h, acc, abs_acc, rdelta, rabs, NULL, 0.0f);
}

/* the tiled gradient with the SSIM loss term (rendered = the current render) */
ORC_API void orc_gradients_ssim(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                                const float* rendered, const float* mask, const SadSite* sites, float iss,
                                uint32_t n, int compute_removal, float ssim_weight, int width, int height,
                                int cand_w, int cand_h, double* acc, double* abs_acc, double* rdelta, double* rabs)
{
    orc_gradients_impl(0, cand0, cand1, target, mask, NULL, sites, iss, n, compute_removal,
                       width, height, cand_w, cand_h, acc, abs_acc, rdelta, rabs, rendered, ssim_weight);
}

/* render_sad_backward_kernel, sad_bindings.cu:100-207: grad_sites [N,10] AoS */
ORC_API void orc_render_backward(const uint32_t* cand0, const uint32_t* cand1, const float* grad_out,
                                 const SadSite* sites, float iss, uint32_t n, int width, int height,
                                 int cand_w, int cand_h, double* acc, double* abs_acc)
{
    orc_gradients_impl(1, cand0, cand1, NULL, NULL, grad_out, sites, iss, n, 0,
                       width, height, cand_w, cand_h, acc, abs_acc, NULL, NULL, NULL, 0.0f);
}

/* ------------------------------------------------ exact tile-ordered sums --
 * The HIP library sums every per-(pixel, candidate) contribution of a 16x16
 * tile into one float partial per (tile, site, value), in a fixed order, and
 * adds the partials of different tiles as exact fixed-point integers.  The
 * functions below form the same partials in the same order, so the library's
 * sums are reproduced bit for bit, and the order of the reference's float
 * atomics (gradients.cu:422-459, utils.cu:106-231), which no run repeats, is
 * replaced by this one.  The fp64 per-site sums above remain the independent
 * check of these: both are sums of the same per-pair terms.
 *
 * Order of one (tile, site) sum: the site's entries (pixel p of the tile in
 * raster order, candidate rank k) sorted by (p / 64, k, p % 64) - the wave,
 * then the rank, then the lane - cut into chunks of ORC_CHUNK entries; in a
 * chunk, lane l (of ORC_LANES) keeps a running float sum from 0 of entries
 * l, l + 4, ...; the chunk's value is (s0 + s1) + (s2 + s3); the site's value
 * is the running float sum from 0 of its chunks in order.
 *
 * Fixed point: a partial v becomes hi = trunc(v), lo = round-half-even of
 * (v - trunc(v)) * 2^40 (|v| saturated at 2^62, NaN taken as 0); sums of hi and
 * of lo are int64; the value is (float)((double)hi + (double)lo * 2^-40). */
#define ORC_CHUNK 20
#define ORC_LANES 4
#define ORC_TILE_ENTRIES (256 * SAD_K)

ORC_API void orc_fix_split(float v, int64_t* lo, int64_t* hi)
{
    if (isnan(v)) v = 0.0f;
    if (v > 0x1p62f) v = 0x1p62f;
    if (v < -0x1p62f) v = -0x1p62f;
    float t = truncf(v);
    *hi = (int64_t)t;
    *lo = (int64_t)nearbyintf((v - t) * 0x1p40f);
}

ORC_API float orc_fix_value(int64_t lo, int64_t hi) { return (float)((double)hi + (double)lo * 0x1p-40); }

static void orc_fix_add(int64_t* pair, float v)
{
    int64_t lo, hi;
    orc_fix_split(v, &lo, &hi);
    if (lo) __atomic_fetch_add(&pair[0], lo, __ATOMIC_RELAXED);
    if (hi) __atomic_fetch_add(&pair[1], hi, __ATOMIC_RELAXED);
}

/* fixed-point pairs [count][2] (lo, hi) -> floats */
ORC_API void orc_fix_values(const int64_t* pairs, float* out, int64_t count)
{
    for (int64_t i = 0; i < count; ++i) out[i] = orc_fix_value(pairs[2 * i], pairs[2 * i + 1]);
}

typedef struct OrcTilePixel {
    uint32_t ids[SAD_K];
    float w[SAD_K];
    float pred[3], g[3], tgt[3];
    float err;
    uint8_t use;               /* bit k: the pair (pixel, k) is summed */
} OrcTilePixel;

/* mode 0: loss gradient (mask, 2(pred - tgt) + optional SSIM term, optional removal);
 * mode 1: upstream grad_output (no mask; a +inf max logit drops the pixel, sad_bindings.cu:72-74);
 * mode 2: densify statistics */
typedef struct OrcTileJob {
    int mode, removal;
    const uint32_t *cand0, *cand1;
    const float *target, *mask, *grad_out, *rendered;
    float ssim_weight;
    const SadSite* sites;
    float iss;
    uint32_t n;
    int width, height, cand_w, cand_h;
    int64_t* out;              /* [n][stride] (lo, hi) pairs of value j at 2 j */
    int stride;
    int64_t* rout;             /* removal: [n][2] */
    int32_t* parts;            /* nullable: partials (tiles) added per site */
} OrcTileJob;

/* the pixels of one tile: forward weights, dL/dpred, the summed pairs */
static void orc_tile_pixels(const OrcTileJob* J, int tx0, int ty0, OrcTilePixel* px)
{
    for (int p = 0; p < 256; ++p) {
        OrcTilePixel* q = &px[p];
        memset(q, 0, sizeof(*q));
        const int x = tx0 + (p & 15), y = ty0 + (p >> 4);
        if (x >= J->width || y >= J->height) continue;
        const size_t pix = (size_t)y * J->width + x;
        if (J->mode != 1 && J->mask[pix] <= 0.0f) continue;
        orc_load_ids(J->cand0, J->cand1, sad_cand_index(x, y, J->width, J->height, J->cand_w, J->cand_h), q->ids);
        const float ux = (float)x, uy = (float)y;
        if (J->mode == 1) {
            float mx = -INFINITY;
            int any = 0;
            for (int i = 0; i < SAD_K; ++i) {
                uint32_t id = q->ids[i];
                if (id >= J->n || J->sites[id].pos_x < 0.0f) continue;
                const SadSite* s = &J->sites[id];
                float l = -sad_expf(s->log_tau) * sad_dmix2(s->pos_x, s->pos_y, s->dir_x, s->dir_y, s->log_aniso,
                                                            s->radius, ux, uy, J->iss);
                mx = fmaxf(mx, l);
                any = 1;
            }
            if (!any || isinf(mx)) continue;
        }
        if (!orc_forward_px(q->ids, J->sites, J->n, ux, uy, J->iss, q->w, q->pred)) continue;
        if (J->mode == 1) {
            q->g[0] = J->grad_out[3 * pix]; q->g[1] = J->grad_out[3 * pix + 1]; q->g[2] = J->grad_out[3 * pix + 2];
        } else {
            q->tgt[0] = J->target[3 * pix]; q->tgt[1] = J->target[3 * pix + 1]; q->tgt[2] = J->target[3 * pix + 2];
            float d0 = q->pred[0] - q->tgt[0], d1 = q->pred[1] - q->tgt[1], d2 = q->pred[2] - q->tgt[2];
            q->g[0] = 2.0f * d0; q->g[1] = 2.0f * d1; q->g[2] = 2.0f * d2;
            q->err = d0 * d0 + d1 * d1 + d2 * d2;
            if (J->mode == 0 && J->ssim_weight > 0.0f) {
                SadSsimStats st = sad_ssim_window(J->rendered, J->target, J->width, J->height, x, y);
                float f = -J->ssim_weight * sad_ssim_dcentre(st, sad_luma3(J->rendered + 3 * pix), sad_luma3(q->tgt));
                q->g[0] += f * 0.299f;
                q->g[1] += f * 0.587f;
                q->g[2] += f * 0.114f;
            }
        }
        for (int k = 0; k < SAD_K; ++k) {
            uint32_t id = q->ids[k];
            if (id < J->n && J->sites[id].pos_x >= 0.0f) q->use |= (uint8_t)(1u << k);
        }
    }
}

/* the values one entry adds to its lane's running sums (nv of them); returns 0 when it adds none */
static int orc_entry_values(const OrcTileJob* J, const OrcTilePixel* px, int p, int k, int first, int tx0, int ty0,
                            float* v)
{
    const OrcTilePixel* q = &px[p];
    const uint32_t id = q->ids[k];
    const SadSite* s = &J->sites[id];
    const float ex = (float)(tx0 + (p & 15)), ey = (float)(ty0 + (p >> 4));
    int any = 0;
    float wt = 0.0f;
    if (first && (J->mode == 2 || J->removal))
        for (int j = 0; j < SAD_K; ++j) wt += (q->ids[j] == id && (q->use >> j & 1u)) ? q->w[j] : 0.0f;
    if (J->mode == 2) {
        for (int j = 0; j < 8; ++j) v[j] = 0.0f;
        if (first && wt > 0.0f) {
            const float er = q->err, we = wt * fminf(er, 1.0f);
            v[0] = wt; v[1] = wt * er; v[2] = we; v[3] = we * ex; v[4] = we * ey;
            v[5] = we * ex * ex; v[6] = we * ex * ey; v[7] = we * ey * ey;
            any = 1;
        }
        return any;
    }
    orc_grad_candidate(s, q->w[k], q->g, q->pred, ex, ey, J->iss, v);
    if (J->removal) {
        v[10] = 0.0f;
        if (first && wt > 0.0f) v[10] = orc_removal_delta(q->pred, q->tgt, s, wt);
    }
    return 1;
}

/* one tile: every site's partial in the library's order, added as fixed point */
static void orc_tile(const OrcTileJob* J, int tile, int tiles_x, OrcTilePixel* px, uint32_t* order_id,
                     uint16_t* order_ent, uint32_t* site_ids, uint32_t* site_count, uint32_t* site_start,
                     uint32_t* hkey, uint32_t* hval)
{
    const int tx0 = (tile % tiles_x) * 16, ty0 = (tile / tiles_x) * 16;
    const int nv = J->mode == 2 ? 8 : (J->removal ? 11 : 10);
    orc_tile_pixels(J, tx0, ty0, px);
    /* the tile's sites (any order) via a small open-addressing map */
    const uint32_t HM = 4096;
    for (uint32_t i = 0; i < HM; ++i) hkey[i] = SAD_EMPTY;
    uint32_t ns = 0, ne = 0;
    for (int wv = 0; wv < 4; ++wv)
        for (int k = 0; k < SAD_K; ++k)
            for (int lane = 0; lane < 64; ++lane) {
                const int p = wv * 64 + lane;
                if (!(px[p].use >> k & 1u)) continue;
                const uint32_t id = px[p].ids[k];
                uint32_t h = (id * 2654435761u) & (HM - 1);
                while (hkey[h] != SAD_EMPTY && hkey[h] != id) h = (h + 1) & (HM - 1);
                if (hkey[h] == SAD_EMPTY) { hkey[h] = id; hval[h] = ns; site_ids[ns] = id; site_count[ns] = 0; ns++; }
                order_id[ne] = hval[h];
                order_ent[ne] = (uint16_t)(p * SAD_K + k);
                site_count[hval[h]]++;
                ne++;
            }
    /* bucket the entries per site, keeping their order */
    uint32_t run = 0;
    for (uint32_t u = 0; u < ns; ++u) { site_start[u] = run; run += site_count[u]; site_count[u] = 0; }
    uint16_t* lists = (uint16_t*)(order_ent + ORC_TILE_ENTRIES);
    for (uint32_t e = 0; e < ne; ++e) {
        const uint32_t u = order_id[e];
        lists[site_start[u] + site_count[u]++] = order_ent[e];
    }
    for (uint32_t u = 0; u < ns; ++u) {
        const uint32_t id = site_ids[u], beg0 = site_start[u], cnt = site_count[u];
        float total[11] = {0};
        for (uint32_t c0 = 0; c0 < cnt; c0 += ORC_CHUNK) {
            const uint32_t cend = c0 + ORC_CHUNK < cnt ? c0 + ORC_CHUNK : cnt;
            float lane[ORC_LANES][11];
            memset(lane, 0, sizeof(lane));
            for (int l = 0; l < ORC_LANES; ++l) {
                for (uint32_t e = c0 + (uint32_t)l; e < cend; e += ORC_LANES) {
                    const int ent = lists[beg0 + e];
                    const int p = ent / SAD_K, k = ent % SAD_K;
                    int first = 1;
                    for (int j = 0; j < k; ++j)
                        if (px[p].ids[j] == id && (px[p].use >> j & 1u)) first = 0;
                    float v[11];
                    if (!orc_entry_values(J, px, p, k, first, tx0, ty0, v)) continue;
                    for (int j = 0; j < nv; ++j) lane[l][j] += v[j];
                }
            }
            for (int j = 0; j < nv; ++j) total[j] += (lane[0][j] + lane[1][j]) + (lane[2][j] + lane[3][j]);
        }
        for (int j = 0; j < nv; ++j) {
            if (J->removal && j == 10) orc_fix_add(&J->rout[2 * (size_t)id], total[j]);
            else orc_fix_add(&J->out[(size_t)id * J->stride + 2 * j], total[j]);
        }
        if (J->parts) __atomic_fetch_add(&J->parts[id], 1, __ATOMIC_RELAXED);
    }
}

static void orc_tiles_run(const OrcTileJob* J)
{
    const int tiles_x = (J->width + 15) / 16, tiles_y = (J->height + 15) / 16;
    const int ntiles = tiles_x * tiles_y;
    #pragma omp parallel
    {
        OrcTilePixel* px = (OrcTilePixel*)malloc(sizeof(OrcTilePixel) * 256);
        uint32_t* order_id = (uint32_t*)malloc(sizeof(uint32_t) * ORC_TILE_ENTRIES);
        uint16_t* order_ent = (uint16_t*)malloc(sizeof(uint16_t) * 2 * ORC_TILE_ENTRIES);
        uint32_t* site_ids = (uint32_t*)malloc(sizeof(uint32_t) * ORC_TILE_ENTRIES);
        uint32_t* site_count = (uint32_t*)malloc(sizeof(uint32_t) * ORC_TILE_ENTRIES);
        uint32_t* site_start = (uint32_t*)malloc(sizeof(uint32_t) * ORC_TILE_ENTRIES);
        uint32_t* hkey = (uint32_t*)malloc(sizeof(uint32_t) * 4096);
        uint32_t* hval = (uint32_t*)malloc(sizeof(uint32_t) * 4096);
        #pragma omp for schedule(dynamic, 4)
        for (int t = 0; t < ntiles; ++t)
            orc_tile(J, t, tiles_x, px, order_id, order_ent, site_ids, site_count, site_start, hkey, hval);
        free(px); free(order_id); free(order_ent); free(site_ids); free(site_count); free(site_start);
        free(hkey); free(hval);
    }
}

/* computeGradientsTiledKernel (gradients.cu:103-486) with the library's sums:
 * gfix [n][20] receives the (lo, hi) pairs of the 10 values in Site field order,
 * rfix [n][2] (when compute_removal) the removal deltas; both are added to. */
ORC_API void orc_gradients_fix(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                               const SadSite* sites, float iss, uint32_t n, int compute_removal, int width,
                               int height, int cand_w, int cand_h, const float* rendered, float ssim_weight,
                               int64_t* gfix, int64_t* rfix, int32_t* parts)
{
    OrcTileJob J = {0, compute_removal ? 1 : 0, cand0, cand1, target, mask, NULL, rendered,
                    ssim_weight > 0.0f ? ssim_weight : 0.0f, sites, iss, n, width, height, cand_w, cand_h,
                    gfix, 20, rfix, parts};
    orc_tiles_run(&J);
}

/* render_sad_backward_kernel (sad_bindings.cu:100-207) with the library's sums, into gfix [n][20] */
ORC_API void orc_render_backward_fix(const uint32_t* cand0, const uint32_t* cand1, const float* grad_out,
                                     const SadSite* sites, float iss, uint32_t n, int width, int height,
                                     int cand_w, int cand_h, int64_t* gfix, int32_t* parts)
{
    OrcTileJob J = {1, 0, cand0, cand1, NULL, NULL, grad_out, NULL, 0.0f, sites, iss, n, width, height, cand_w,
                    cand_h, gfix, 20, NULL, parts};
    orc_tiles_run(&J);
}

/* computeSiteStatsSimpleKernel (utils.cu:106-231) with the library's sums, into sfix [n][16]:
 * (lo, hi) of mass, energy, w.err, w.err.x, w.err.y, w.err.xx, w.err.xy, w.err.yy */
ORC_API void orc_site_stats_fix(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                                const SadSite* sites, float iss, uint32_t n, int width, int height, int cand_w,
                                int cand_h, int64_t* sfix, int32_t* parts)
{
    OrcTileJob J = {2, 0, cand0, cand1, target, mask, NULL, NULL, 0.0f, sites, iss, n, width, height, cand_w, cand_h,
                    sfix, 16, NULL, parts};
    orc_tiles_run(&J);
}

/* computePSNRKernel (utils.cu:5-26) with the library's block sums: per block of
 * 256 pixels (index order) a shuffle-down tree per 64-pixel wave, the waves
 * combined as (w0 + w1) + (w2 + w3), the blocks added as fixed point. */
ORC_API float orc_sse_fix(const float* rendered, const float* target, const float* mask, int width, int height)
{
    const size_t np = (size_t)width * height;
    int64_t acc[2] = {0, 0};
    for (size_t b0 = 0; b0 < np; b0 += 256) {
        float wsum[4];
        for (int w = 0; w < 4; ++w) {
            float v[64];
            for (int l = 0; l < 64; ++l) {
                const size_t i = b0 + (size_t)w * 64 + l;
                v[l] = 0.0f;
                if (i < np && mask[i] > 0.0f) {
                    float d0 = rendered[3 * i] - target[3 * i], d1 = rendered[3 * i + 1] - target[3 * i + 1];
                    float d2 = rendered[3 * i + 2] - target[3 * i + 2];
                    v[l] = d0 * d0 + d1 * d1 + d2 * d2;
                }
            }
            for (int off = 32; off > 0; off >>= 1)
                for (int l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
            wsum[w] = v[0];
        }
        orc_fix_add(acc, (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]));
    }
    return orc_fix_value(acc[0], acc[1]);
}

/* tau_diffuse_kernel, tau_diffusion.cu:27-89 */
ORC_API void orc_tau_diffuse(const uint32_t* cand0, const uint32_t* cand1, const SadSite* sites,
                             const float* graw, const float* gin, float* gout, uint32_t n,
                             float lambda, int width, int height, int downscale)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (lambda <= 0.0f || sites[i].pos_x < 0.0f) { gout[i] = gin[i]; continue; }
        float inv_ds = 1.0f / fmaxf((float)downscale, 1.0f);
        int px = sad_clampi((int)(sites[i].pos_x * inv_ds + 0.5f), 0, width - 1);
        int py = sad_clampi((int)(sites[i].pos_y * inv_ds + 0.5f), 0, height - 1);
        uint32_t ids[SAD_K];
        orc_load_ids(cand0, cand1, py * width + px, ids);
        float sum = 0.0f;
        uint32_t cnt = 0;
        for (int k = 0; k < SAD_K; ++k) {
            uint32_t nb = ids[k];
            if (nb >= n || nb == (uint32_t)i) continue;
            int seen = 0;
            for (int j = 0; j < k; ++j) if (ids[j] == nb) { seen = 1; break; }
            if (seen || sites[nb].pos_x < 0.0f) continue;
            sum += gin[nb];
            cnt += 1;
        }
        float diag = 1.0f + lambda * (float)cnt;
        gout[i] = (graw[i] + lambda * sum) / diag;
    }
}

/* adamUpdateKernel, adam.cu:5-180.  b1c = 1/(1-beta1^t), b2c = 1/(1-beta2^t) are
 * computed once per step with sad_powi_f, an integer power in double (the HIP
 * launcher does the same), in place of the reference's powf. */
ORC_API void orc_adam_bias(float beta1, float beta2, uint32_t t, float* b1c, float* b2c)
{
    *b1c = 1.0f / (1.0f - sad_powi_f(beta1, t));
    *b2c = 1.0f / (1.0f - sad_powi_f(beta2, t));
}

ORC_API void orc_adam(SadSite* sites, SadAdamState* st, float* const g[10],
                      float lr_pos, float lr_tau, float lr_radius, float lr_color, float lr_dir,
                      float lr_aniso, float beta1, float beta2, float eps, uint32_t t, uint32_t n,
                      int width, int height)
{
    float b1c, b2c;
    orc_adam_bias(beta1, beta2, t, &b1c, &b2c);
    float norm = 1.0f / (float)(width * height);
    float ob1 = 1.0f - beta1, ob2 = 1.0f - beta2;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float gpx = g[0][i] * norm, gpy = g[1][i] * norm, glt = g[2][i] * norm, grd = g[3][i] * norm;
        float gcr = g[4][i] * norm, gcg = g[5][i] * norm, gcb = g[6][i] * norm;
        float gdx = g[7][i] * norm, gdy = g[8][i] * norm, gla = g[9][i] * norm;
        SadSite* s = &sites[i];
        SadAdamState* a = &st[i];
        if (s->pos_x < 0.0f) continue;
        /* position: isotropic Adam (one v for both axes) */
        a->m_pos_x = beta1 * a->m_pos_x + ob1 * gpx;
        a->m_pos_y = beta1 * a->m_pos_y + ob1 * gpy;
        a->v_pos = beta2 * a->v_pos + ob2 * (gpx * gpx + gpy * gpy);
        float inv_den = 1.0f / (sqrtf(a->v_pos * b2c) + eps);
        s->pos_x -= lr_pos * (a->m_pos_x * b1c) * inv_den;
        s->pos_y -= lr_pos * (a->m_pos_y * b1c) * inv_den;
        s->pos_x = sad_clampf(s->pos_x, 0.0f, (float)(width - 1));
        s->pos_y = sad_clampf(s->pos_y, 0.0f, (float)(height - 1));
        /* log_tau */
        a->m_log_tau = beta1 * a->m_log_tau + ob1 * glt;
        a->v_log_tau = beta2 * a->v_log_tau + ob2 * (glt * glt);
        s->log_tau -= lr_tau * (a->m_log_tau * b1c) / (sqrtf(a->v_log_tau * b2c) + eps);
        /* radius, clamped to [1, 512] */
        a->m_radius = beta1 * a->m_radius + ob1 * grd;
        a->v_radius = beta2 * a->v_radius + ob2 * (grd * grd);
        s->radius -= lr_radius * (a->m_radius * b1c) / (sqrtf(a->v_radius * b2c) + eps);
        s->radius = sad_clampf(s->radius, 1.0f, 512.0f);
        /* colour: per-channel Adam, clamped to [0, 1] */
        a->m_col_r = beta1 * a->m_col_r + ob1 * gcr;
        a->m_col_g = beta1 * a->m_col_g + ob1 * gcg;
        a->m_col_b = beta1 * a->m_col_b + ob1 * gcb;
        a->v_col_r = beta2 * a->v_col_r + ob2 * (gcr * gcr);
        a->v_col_g = beta2 * a->v_col_g + ob2 * (gcg * gcg);
        a->v_col_b = beta2 * a->v_col_b + ob2 * (gcb * gcb);
        float cr = s->col_r - lr_color * (a->m_col_r * b1c) / (sqrtf(a->v_col_r * b2c) + eps);
        float cg = s->col_g - lr_color * (a->m_col_g * b1c) / (sqrtf(a->v_col_g * b2c) + eps);
        float cb = s->col_b - lr_color * (a->m_col_b * b1c) / (sqrtf(a->v_col_b * b2c) + eps);
        s->col_r = sad_clampf(cr, 0.0f, 1.0f);
        s->col_g = sad_clampf(cg, 0.0f, 1.0f);
        s->col_b = sad_clampf(cb, 0.0f, 1.0f);
        /* direction: tangent projection + isotropic Adam + renormalise */
        float ddot = gdx * s->dir_x + gdy * s->dir_y;
        float tx = gdx - s->dir_x * ddot, ty = gdy - s->dir_y * ddot;
        a->m_dir_x = beta1 * a->m_dir_x + ob1 * tx;
        a->m_dir_y = beta1 * a->m_dir_y + ob1 * ty;
        a->v_dir = beta2 * a->v_dir + ob2 * (tx * tx + ty * ty);
        float inv_dd = 1.0f / (sqrtf(a->v_dir * b2c) + eps);
        float nx = s->dir_x - lr_dir * (a->m_dir_x * b1c) * inv_dd;
        float ny = s->dir_y - lr_dir * (a->m_dir_y * b1c) * inv_dd;
        float len2 = nx * nx + ny * ny;
        if (len2 < 1e-12f) { nx = 1.0f; ny = 0.0f; }
        else { float il = 1.0f / sqrtf(len2); nx *= il; ny *= il; }
        s->dir_x = nx; s->dir_y = ny;
        /* log_aniso, clamped to [-4, 4] */
        a->m_log_aniso = beta1 * a->m_log_aniso + ob1 * gla;
        a->v_log_aniso = beta2 * a->v_log_aniso + ob2 * (gla * gla);
        s->log_aniso -= lr_aniso * (a->m_log_aniso * b1c) / (sqrtf(a->v_log_aniso * b2c) + eps);
        s->log_aniso = sad_clampf(s->log_aniso, -4.0f, 4.0f);
    }
}

/* computeDensifyScorePairsKernel, densify_prune.cu:223-254 */
ORC_API void orc_densify_pairs(const SadSite* sites, const float* mass, const float* energy,
                               SadPair* pairs, uint32_t n, float min_mass, float alpha, uint32_t count)
{
    for (uint32_t i = 0; i < count; ++i) {
        SadPair out = {SAD_EMPTY, SAD_EMPTY};
        if (i < n) {
            float m = mass[i], e = energy[i];
            if (sites[i].pos_x >= 0.0f && m > min_mass && isfinite(e)) {
                float den = sad_powf_pos(fmaxf(m, 1e-8f), alpha);
                float score = fmaxf(e, 0.0f) / den;
                out.key = 0xFFFFFFFFu - sad_f2u(score);
                out.val = i;
            }
        }
        pairs[i] = out;
    }
}

/* computePruneScorePairsKernel, densify_prune.cu:257-281 */
ORC_API void orc_prune_pairs(const SadSite* sites, const float* rdelta, SadPair* pairs,
                             uint32_t n, float delta_norm, uint32_t count)
{
    for (uint32_t i = 0; i < count; ++i) {
        SadPair out = {SAD_EMPTY, SAD_EMPTY};
        if (i < n) {
            float d = rdelta[i] * delta_norm;
            if (sites[i].pos_x >= 0.0f && isfinite(d)) { out.key = sad_f2u(fmaxf(d, 0.0f)); out.val = i; }
        }
        pairs[i] = out;
    }
}

/* launchRadixSortUInt2, radix_sort.cu:279-336: stable LSD sort by key with 8-bit
 * digits, 2 passes (low 16 bits) when max_key_exclusive <= 2^16, else 4. */
ORC_API void orc_radix_sort(SadPair* data, uint32_t count, uint32_t max_key_exclusive)
{
    if (count == 0) return;
    int passes = (max_key_exclusive <= (1u << 16)) ? 2 : 4;
    SadPair* tmp = (SadPair*)malloc((size_t)count * sizeof(SadPair));
    SadPair* in = data;
    SadPair* out = tmp;
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t shift = 8u * (uint32_t)pass;
        size_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (uint32_t i = 0; i < count; ++i) cnt[((in[i].key >> shift) & 0xFFu) + 1]++;
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (uint32_t i = 0; i < count; ++i) out[cnt[(in[i].key >> shift) & 0xFFu]++] = in[i];
        SadPair* t = in; in = out; out = t;
    }
    free(tmp);
}

/* writeSplitIndicesFromSortedKernel, densify_prune.cu:284-292 */
ORC_API void orc_write_indices(const SadPair* sorted, uint32_t* idx, uint32_t k)
{
    for (uint32_t i = 0; i < k; ++i) idx[i] = sorted[i].val;
}

/* pruneSitesByIndexKernel, densify_prune.cu:295-307 */
ORC_API void orc_prune_sites(SadSite* sites, const uint32_t* idx, uint32_t k)
{
    for (uint32_t i = 0; i < k; ++i) {
        if (idx[i] == SAD_EMPTY) continue;
        sites[idx[i]].pos_x = -1.0f;
        sites[idx[i]].pos_y = -1.0f;
    }
}

/* splitSitesKernel, densify_prune.cu:86-220.  A parent index >= current site
 * count (a sentinel pair that sorted into the first k) is skipped; the
 * reference would read sites[0xFFFFFFFF] there (densify_prune.cu:106-107). */
ORC_API void orc_split_sites(SadSite* sites, SadAdamState* st, const uint32_t* idx, uint32_t k,
                             const float* mass, const float* ew, const float* ewx, const float* ewy,
                             const float* ewxx, const float* ewxy, const float* ewyy,
                             uint32_t cur, const float* target, int width, int height)
{
    for (uint32_t gid = 0; gid < k; ++gid) {
        uint32_t pi = idx[gid];
        if (pi >= cur) continue;
        SadSite parent = sites[pi];
        if (parent.pos_x < 0.0f) continue;
        float w = ew[pi];
        float ax = parent.dir_x, ay = parent.dir_y;
        float cx = parent.pos_x, cy = parent.pos_y;
        float la = parent.log_aniso * 0.8f;
        if (w > 1e-3f) {
            float mx = ewx[pi] / w, my = ewy[pi] / w;
            cx = 0.4f * parent.pos_x + 0.6f * mx;
            cy = 0.4f * parent.pos_y + 0.6f * my;
            float exx = ewxx[pi] / w - mx * mx;
            float exy = ewxy[pi] / w - mx * my;
            float eyy = ewyy[pi] / w - my * my;
            exx = fmaxf(exx, 1e-4f);
            eyy = fmaxf(eyy, 1e-4f);
            /* the library's own atan2 and sincos: CUDA's libdevice is not available to pin against */
            float th = 0.5f * sad_atan2f(2.0f * exy, exx - eyy);
            sad_sincosf(th, &ay, &ax);
            float tr = exx + eyy;
            float disc = sqrtf(fmaxf(0.0f, 0.25f * (exx - eyy) * (exx - eyy) + exy * exy));
            float l1 = fmaxf(1e-4f, 0.5f * tr + disc);
            float l2 = fmaxf(1e-4f, 0.5f * tr - disc);
            la = sad_clampf(-0.5f * sad_logf(l1 / l2), -2.0f, 2.0f);
        } else {
            int ix = sad_clampi((int)parent.pos_x, 0, width - 1);
            int iy = sad_clampi((int)parent.pos_y, 0, height - 1);
            float lum[3][3];
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int sx = sad_clampi(ix + dx, 0, width - 1), sy = sad_clampi(iy + dy, 0, height - 1);
                    const float* c = &target[3 * ((size_t)sy * width + sx)];
                    lum[dy + 1][dx + 1] = 0.299f * c[0] + 0.587f * c[1] + 0.114f * c[2];
                }
            float TL = lum[0][0], T = lum[0][1], TR = lum[0][2], L = lum[1][0], R = lum[1][2];
            float BL = lum[2][0], B = lum[2][1], BR = lum[2][2];
            float gx = (-TL + TR - 2.0f * L + 2.0f * R - BL + BR) / 4.0f;
            float gy = (-TL - 2.0f * T - TR + BL + 2.0f * B + BR) / 4.0f;
            float g2 = gx * gx + gy * gy;
            if (g2 > 1e-8f) {
                float len = sqrtf(g2);
                if (len > 1e-8f) { ax = gx / len; ay = gy / len; } else { ax = 1.0f; ay = 0.0f; }
            }
        }
        float m = fmaxf(mass[pi], 1.0f);
        float od = sad_clampf(sqrtf(m) * 0.5f, 1.5f, 48.0f);
        float ox = ax * od, oy = ay * od;
        float pax = sad_clampf(cx + ox, 0.0f, (float)(width - 1));
        float pay = sad_clampf(cy + oy, 0.0f, (float)(height - 1));
        float pbx = sad_clampf(cx - ox, 0.0f, (float)(width - 1));
        float pby = sad_clampf(cy - oy, 0.0f, (float)(height - 1));
        const float* ca = &target[3 * ((size_t)(int)pay * width + (int)pax)];
        const float* cb = &target[3 * ((size_t)(int)pby * width + (int)pbx)];
        uint32_t child = cur + gid;
        if (child == pi) continue;
        SadSite c0 = parent, c1 = parent;
        c0.pos_x = pax; c0.pos_y = pay; c1.pos_x = pbx; c1.pos_y = pby;
        c0.col_r = ca[0]; c0.col_g = ca[1]; c0.col_b = ca[2];
        c1.col_r = cb[0]; c1.col_g = cb[1]; c1.col_b = cb[2];
        c0.log_tau = c1.log_tau = parent.log_tau - 0.25f;
        c0.radius = c1.radius = parent.radius * 0.85f;
        c0.dir_x = c1.dir_x = ax; c0.dir_y = c1.dir_y = ay;
        c0.log_aniso = c1.log_aniso = la;
        sites[pi] = c0;
        sites[child] = c1;
        memset(&st[pi], 0, sizeof(SadAdamState));
        memset(&st[child], 0, sizeof(SadAdamState));
    }
}

/* computeHilbertPairsKernel / writeHilbertOrderKernel, sad_cuda.cu:1547-1584 */
ORC_API void orc_hilbert_pairs(const SadSite* sites, SadPair* pairs, uint32_t n, uint32_t padded,
                               int width, int height, int bits)
{
    for (uint32_t i = 0; i < padded; ++i) {
        if (i < n) {
            int px = (int)sites[i].pos_x, py = (int)sites[i].pos_y;
            px = px < 0 ? 0 : (px >= width ? width - 1 : px);
            py = py < 0 ? 0 : (py >= height ? height - 1 : py);
            pairs[i].key = sad_hilbert_index((uint32_t)px, (uint32_t)py, bits);
            pairs[i].val = i;
        } else {
            pairs[i].key = SAD_EMPTY;
            pairs[i].val = 0u;
        }
    }
}

ORC_API void orc_hilbert_order(const SadPair* sorted, uint32_t* order, uint32_t* pos, uint32_t n)
{
    for (uint32_t i = 0; i < n; ++i) { order[i] = sorted[i].val; pos[sorted[i].val] = i; }
}

/* computeSiteStatsSimpleKernel, utils.cu:106-231: per unique candidate site,
 * mass += w, energy += w*err and err moments with werr = w*min(err,1).
 * out[8*id + j] (double) for j in {mass, energy, w, wx, wy, wxx, wxy, wyy}. */
ORC_API void orc_site_stats(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                            const float* mask, const SadSite* sites, float iss, uint32_t n,
                            int width, int height, int cand_w, int cand_h, double* out)
{
    int T = omp_get_max_threads();
    double* priv = (double*)calloc((size_t)T * n * 8, sizeof(double));
    #pragma omp parallel
    {
        double* my = priv + (size_t)omp_get_thread_num() * n * 8;
        #pragma omp for schedule(static)
        for (int y = 0; y < height; ++y) {
            for (int x = 0; x < width; ++x) {
                size_t p = (size_t)y * width + x;
                if (mask[p] <= 0.0f) continue;
                uint32_t ids[SAD_K];
                orc_load_ids(cand0, cand1, sad_cand_index(x, y, width, height, cand_w, cand_h), ids);
                float w[SAD_K], pred[3];
                float ux = (float)x, uy = (float)y;
                if (!orc_forward_px(ids, sites, n, ux, uy, iss, w, pred)) continue;
                float d0 = pred[0] - target[3 * p], d1 = pred[1] - target[3 * p + 1], d2 = pred[2] - target[3 * p + 2];
                float err = d0 * d0 + d1 * d1 + d2 * d2;
                int seen[SAD_K] = {0};
                for (int i = 0; i < SAD_K; ++i) {
                    uint32_t id = ids[i];
                    if (id >= n || sites[id].pos_x < 0.0f || seen[i]) continue;
                    int dup = 0;
                    for (int j = 0; j < i; ++j) if (ids[j] == id) { dup = 1; seen[j] = 1; break; }
                    if (dup) continue;
                    float wt = 0.0f;
                    for (int j = i; j < SAD_K; ++j) if (ids[j] == id) { wt += w[j]; seen[j] = 1; }
                    if (wt > 0.0f) {
                        double* o = &my[8 * (size_t)id];
                        float we = wt * fminf(err, 1.0f);
                        o[0] += wt; o[1] += wt * err; o[2] += we;
                        o[3] += we * ux; o[4] += we * uy;
                        o[5] += we * ux * ux; o[6] += we * ux * uy; o[7] += we * uy * uy;
                    }
                }
            }
        }
    }
    for (int t = 0; t < T; ++t)
        for (size_t i = 0; i < (size_t)n * 8; ++i) out[i] += priv[(size_t)t * n * 8 + i];
    free(priv);
}

/* computePSNRKernel, utils.cu:5-26 (sum of squared error over masked pixels) */
ORC_API double orc_sse(const float* rendered, const float* target, const float* mask, int width, int height)
{
    double s = 0.0;
    for (size_t p = 0; p < (size_t)width * height; ++p) {
        if (mask[p] <= 0.0f) continue;
        float d0 = rendered[3 * p] - target[3 * p], d1 = rendered[3 * p + 1] - target[3 * p + 1];
        float d2 = rendered[3 * p + 2] - target[3 * p + 2];
        s += (double)(d0 * d0 + d1 * d1 + d2 * d2);
    }
    return s;
}

/* computeSSIMKernel, utils.cu:29-87: sum over masked pixels of the window SSIM
 * (double accumulation); per-pixel values into per_px when non-null */
ORC_API double orc_ssim_sum(const float* rendered, const float* target, const float* mask, int width, int height,
                            float* per_px)
{
    double s = 0.0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            size_t p = (size_t)y * width + x;
            float v = 0.0f;
            if (mask[p] > 0.0f) {
                float A, B, C, D;
                v = sad_ssim_terms(sad_ssim_window(rendered, target, width, height, x, y), &A, &B, &C, &D);
                s += (double)v;
            }
            if (per_px) per_px[p] = v;
        }
    return s;
}

/* the SSIM loss factor -w dSSIM/dx_k of every pixel (0 outside the mask) */
ORC_API void orc_ssim_dcentre(const float* rendered, const float* target, const float* mask, int width, int height,
                              float* out)
{
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            size_t p = (size_t)y * width + x;
            out[p] = 0.0f;
            if (mask[p] > 0.0f)
                out[p] = sad_ssim_dcentre(sad_ssim_window(rendered, target, width, height, x, y),
                                          sad_luma3(rendered + 3 * p), sad_luma3(target + 3 * p));
        }
}

/* initGradientWeightedKernel, densify_prune.cu:5-83.  The reference's site
 * threads race for seeds on one atomic counter; here the draws are serialised
 * in site order (site i draws from the seed after site i-1's last draw), the
 * rule the HIP library implements.  Returns one past the last seed drawn. */
ORC_API uint32_t orc_init_gradient_weighted(SadSite* sites, uint32_t n, uint32_t seed0, const float* target,
                                            const float* mask, float thr, uint32_t attempts, float log_tau0,
                                            float radius0, int width, int height)
{
    uint32_t seed = seed0;
    for (uint32_t i = 0; i < n; ++i) {
        float px = 0.0f, py = 0.0f, g = 0.0f;
        int pix = -1;
        for (uint32_t a = 0; a < attempts; ++a) {
            float x, y, gg;
            int p;
            if (sad_init_grad_eval(target, mask, width, height, seed++, thr, &x, &y, &p, &gg)) {
                px = x; py = y; pix = p; g = gg;
                break;
            }
        }
        sites[i] = sad_init_grad_site(target, px, py, pix, g, log_tau0, radius0);
    }
    return seed;
}

/* one R2 seed: acceptance, position, pixel and gradient magnitude */
ORC_API int orc_init_grad_eval(const float* target, const float* mask, int width, int height, uint32_t seed,
                               float thr, float* out4)
{
    float x, y, g;
    int p;
    int ok = sad_init_grad_eval(target, mask, width, height, seed, thr, &x, &y, &p, &g);
    out4[0] = x; out4[1] = y; out4[2] = (float)p; out4[3] = g;
    return ok;
}

/* makePackedInferenceSites, backends/metal/sources/utils/packed_inference.swift:86-199:
 * ranges over active sites (both coordinates >= 0, finite values), then each
 * site packed with sad_pack_inference (sad_math.h). */
ORC_API void orc_pack_inference(const SadSite* sites, uint32_t n, int width, int height, SadInferenceSite* out,
                                SadInferenceQuant* quant)
{
    float mn[5], mx[5];
    for (int k = 0; k < 5; ++k) { mn[k] = INFINITY; mx[k] = -INFINITY; }
    for (uint32_t i = 0; i < n; ++i) {
        const SadSite* s = &sites[i];
        if (!(s->pos_x >= 0.0f && s->pos_y >= 0.0f)) continue;
        const float v[5] = {s->log_tau, s->radius, s->col_r, s->col_g, s->col_b};
        for (int k = 0; k < 5; ++k)
            if (isfinite(v[k])) { mn[k] = fminf(mn[k], v[k]); mx[k] = fmaxf(mx[k], v[k]); }
    }
    sad_quant_finish(mn, mx, quant);
    const float maxx = (float)(width > 1 ? width - 1 : 0), maxy = (float)(height > 1 ? height - 1 : 0);
    for (uint32_t i = 0; i < n; ++i) out[i] = sad_pack_inference(&sites[i], quant, maxx, maxy);
}

/* decodePackedSite, backends/metal/shaders/sad_common.metal:71-120 */
ORC_API void orc_unpack_inference(const SadInferenceSite* in, const SadInferenceQuant* quant, uint32_t n, int width,
                                  int height, SadSite* out)
{
    const float maxx = (float)(width > 1 ? width - 1 : 0), maxy = (float)(height > 1 ? height - 1 : 0);
    for (uint32_t i = 0; i < n; ++i) out[i] = sad_unpack_inference(in[i], quant, maxx, maxy);
}

ORC_API void orc_atan2_array(const float* y, const float* x, float* out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = sad_atan2f(y[i], x[i]);
}
