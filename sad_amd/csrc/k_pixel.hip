/*
 * k_pixel.hip - per-pixel kernels: softmax render forward, tiled hashed gradient
 * backward (forward recomputed), autograd backward into [N][10], and the
 * densify site statistics.
 *
 * Design (gfx950):
 *   - One 256-thread workgroup per 16x16-pixel tile.  Tiles are mapped to
 *     workgroups so that consecutive tiles share an XCD (and its L2).
 *   - Stage 1: every (pixel, candidate) pair inserts its site id into an LDS
 *     hash of H slots (multiplicative hash, linear probing, CAS only on empty
 *     slots).  Stage 2: one thread per occupied slot loads the 40-byte site once
 *     and stores the per-site terms of the forward (tau, e^a, e^-a, r/S, colour)
 *     in LDS, so the exp/divisions run once per site per tile instead of once
 *     per pixel-candidate.  Ids that do not fit (H slots, 8 probes) are evaluated
 *     from global memory and accumulate with global atomics.
 *   - Stage 3: a wave64 covers 8 pixels x 8 candidates: lane = 8 * pixel +
 *     candidate.  The softmax max/sum and the colour blend are 8-lane DPP
 *     reductions, and each lane then owns one candidate's gradient.  Because a
 *     wave's atomics now target the candidates of 8 pixels (not the same
 *     candidate rank of 64 neighbouring pixels, which is mostly one site), LDS
 *     atomic conflicts drop roughly 8x compared with thread-per-pixel.
 *   - Stage 4: occupied slots are flushed to the SoA gradient arrays with one
 *     global float atomic per parameter.
 * Numerics: per-site terms use the exact primitives; the per-pixel distance and
 * weights use the hardware sqrt/exp (~1 ulp).  Parity with the oracle is a
 * tolerance for these kernels.
 */
#include "sad_device.h"

namespace sadh {

enum PixelMode { MODE_RENDER = 0, MODE_GRAD = 1, MODE_BWD = 2, MODE_STATS = 3 };

constexpr uint16_t SLOT_INVALID = 0xFFFFu;   /* id >= site count or pixel not processed */
constexpr uint16_t SLOT_OVERFLOW = 0xFFFEu;  /* valid id without a hash slot */
constexpr int NACC = 10;

template <int H>
struct TileLds {
    uint32_t key[H];
    float4 p0[H];              /* pos.x, pos.y, tau, r/S      (pos.x < 0: inactive) */
    float4 p1[H];              /* dir.x, dir.y, e^a, 1/e^a */
    float4 p2[H];              /* r, g, b, 1/max(e^a, 1e-8) */
    float acc[NACC][H];
    float delta[H];
    uint16_t slot[256 * 8];
};

struct SiteTerms {
    float4 p0, p1, p2;
};

__device__ __forceinline__ SiteTerms site_terms(const SadSite* __restrict__ sites, uint32_t id, float inv_scale)
{
    const float2* s2 = reinterpret_cast<const float2*>(sites + id);
    float2 a = s2[0], b = s2[1], c = s2[2], d = s2[3], e = s2[4];
    /* a = pos, b = (log_tau, radius), c = (r, g), d = (b, dir.x), e = (dir.y, log_aniso) */
    SiteTerms t;
    float l1 = sad_expf(e.y);
    t.p0 = make_float4(a.x, a.y, sad_expf(b.x), b.y * inv_scale);
    t.p1 = make_float4(d.y, e.x, l1, 1.0f / l1);
    t.p2 = make_float4(c.x, c.y, d.x, 1.0f / fmaxf(l1, 1e-8f));
    return t;
}

template <int H>
__device__ __forceinline__ uint16_t hash_insert(TileLds<H>& L, uint32_t id)
{
    uint32_t h = (id * 2654435761u) & (uint32_t)(H - 1);
#pragma unroll 1
    for (int probe = 0; probe < 8; ++probe) {
        uint32_t k = L.key[h];
        if (k == id) return (uint16_t)h;
        if (k == SAD_EMPTY) {
            uint32_t old = atomicCAS(&L.key[h], SAD_EMPTY, id);
            if (old == SAD_EMPTY || old == id) return (uint16_t)h;
        }
        h = (h + 1u) & (uint32_t)(H - 1);
    }
    return SLOT_OVERFLOW;
}

struct PixelArgs {
    const uint32_t* cand0;
    const uint32_t* cand1;
    const SadSite* sites;
    const float* target;       /* GRAD, STATS */
    const float* mask;         /* GRAD, STATS */
    const float* grad_out;     /* BWD */
    float* out;                /* RENDER: [H][W][3]; BWD: [N][10] */
    float* g[NACC];            /* GRAD: 10 SoA arrays; STATS: 8 arrays */
    float* removal;            /* GRAD with removal */
    float iss, inv_scale;
    uint32_t n;
    int width, height, cw, ch;
    uint32_t tiles_x, ntiles;
};

template <int MODE, bool REMOVAL, int H>
__global__ __launch_bounds__(256) void k_pixel(PixelArgs A)
{
    __shared__ TileLds<H> L;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t k = tid & 7u;
    const uint32_t tile = xcd_tile(blockIdx.x, A.ntiles);
    const int tx0 = (int)(tile % A.tiles_x) * 16;
    const int ty0 = (int)(tile / A.tiles_x) * 16;
    const uint32_t* cplane = (k < 4u) ? A.cand0 : A.cand1;
    const uint32_t kk = k & 3u;

    for (uint32_t s = tid; s < (uint32_t)H; s += 256u) {
        L.key[s] = SAD_EMPTY;
        if (MODE != MODE_RENDER) {
#pragma unroll
            for (int j = 0; j < NACC; ++j) L.acc[j][s] = 0.0f;
            L.delta[s] = 0.0f;
        }
    }
    __syncthreads();

    /* stage 1: hash the tile's candidate ids */
    for (int r = 0; r < 8; ++r) {
        uint32_t p = (uint32_t)r * 32u + (tid >> 3);
        int x = tx0 + (int)(p & 15u), y = ty0 + (int)(p >> 4);
        uint16_t slot = SLOT_INVALID;
        bool act = x < A.width && y < A.height;
        if (act && (MODE == MODE_GRAD || MODE == MODE_STATS)) act = A.mask[y * A.width + x] > 0.0f;
        if (act) {
            int cell = sad_cand_index(x, y, A.width, A.height, A.cw, A.ch);
            uint32_t id = cplane[4u * (uint32_t)cell + kk];
            if (id < A.n) slot = hash_insert<H>(L, id);
        }
        L.slot[p * 8u + k] = slot;
    }
    __syncthreads();

    /* stage 2: per-site terms of the occupied slots */
    for (uint32_t s = tid; s < (uint32_t)H; s += 256u) {
        uint32_t id = L.key[s];
        if (id == SAD_EMPTY) continue;
        SiteTerms t = site_terms(A.sites, id, A.inv_scale);
        L.p0[s] = t.p0;
        L.p1[s] = t.p1;
        L.p2[s] = t.p2;
    }
    __syncthreads();

    /* stage 3: 8 lanes per pixel */
    for (int r = 0; r < 8; ++r) {
        uint32_t p = (uint32_t)r * 32u + (tid >> 3);
        int x = tx0 + (int)(p & 15u), y = ty0 + (int)(p >> 4);
        bool inb = x < A.width && y < A.height;
        uint16_t slot = L.slot[p * 8u + k];
        uint32_t pix = inb ? (uint32_t)(y * A.width + x) : 0u;
        SiteTerms t;
        t.p0 = make_float4(-1.0f, -1.0f, 0.0f, 0.0f);
        t.p1 = t.p2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint32_t id = SAD_EMPTY;
        if (slot < (uint16_t)H) {
            t.p0 = L.p0[slot];
            t.p1 = L.p1[slot];
            t.p2 = L.p2[slot];
            id = L.key[slot];
        } else if (slot == SLOT_OVERFLOW) {
            int cell = sad_cand_index(x, y, A.width, A.height, A.cw, A.ch);
            id = cplane[4u * (uint32_t)cell + kk];
            t = site_terms(A.sites, id, A.inv_scale);
        }
        const bool valid = slot != SLOT_INVALID && t.p0.x >= 0.0f;
        const float ux = (float)x, uy = (float)y;

        /* forward: logit = -tau * dmix (render.cu:52-55) */
        float dx = ux - t.p0.x, dy = uy - t.p0.y;
        float diff2 = fmaf(dx, dx, dy * dy);
        float proj = fmaf(t.p1.x, dx, t.p1.y * dy);
        float proj2 = proj * proj;
        float perp2 = fmaxf(diff2 - proj2, 0.0f);
        float d2s_f = fmaxf(fmaf(t.p1.z, proj2, t.p1.w * perp2) * A.iss, 1e-8f);
        float logit = valid ? -(t.p0.z * (fast_sqrt(d2s_f) - t.p0.w)) : -INFINITY;
        float mx = group8_max(logit);
        bool has = mx > -INFINITY;
        /* the autograd kernel also skips a +inf maximum (sad_bindings.cu:72-74) */
        if (MODE == MODE_BWD) has = has && mx < INFINITY;
        float e = (valid && has) ? fast_exp(logit - mx) : 0.0f;
        float sum = group8_sum(e);
        float w = e * (1.0f / fmaxf(sum, 1e-8f));
        float wc = w;
        if (MODE == MODE_RENDER) wc = (valid && !isnan(w) && !isinf(w)) ? w : 0.0f;
        else wc = valid ? w : 0.0f;
        float pr = group8_sum(wc * t.p2.x);
        float pg = group8_sum(wc * t.p2.y);
        float pb = group8_sum(wc * t.p2.z);

        if (MODE == MODE_RENDER) {
            if (inb && k == 0u) {
                float* o = A.out + 3u * pix;
                o[0] = has ? pr : 0.0f;
                o[1] = has ? pg : 0.0f;
                o[2] = has ? pb : 0.0f;
            }
            continue;
        }

        float tr = 0.0f, tg = 0.0f, tb = 0.0f;
        if ((MODE == MODE_GRAD || MODE == MODE_STATS) && inb) {
            const float* tp = A.target + 3u * pix;
            tr = tp[0]; tg = tp[1]; tb = tp[2];
        }

        if (MODE == MODE_STATS || (MODE == MODE_GRAD && REMOVAL)) {
            /* per unique site of the pixel: summed weight, counted at its first rank
             * (utils.cu:529-569, gradients.cu:295-321) */
            uint32_t myid = valid ? id : (0xFFFFFF00u | lane);
            float wt = 0.0f;
            bool first = true;
            const int base = (int)(lane & ~7u);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                uint32_t oj = (uint32_t)__shfl((int)myid, base + j);
                float wj = __shfl(w, base + j);
                if (oj == myid) {
                    wt += wj;
                    if (j < (int)k) first = false;
                }
            }
            if (valid && has && first && wt > 0.0f) {
                if (MODE == MODE_STATS) {
                    float er = pr - tr, eg = pg - tg, eb = pb - tb;
                    float err = er * er + eg * eg + eb * eb;
                    float we = wt * fminf(err, 1.0f);
                    float v[8] = {wt, wt * err, we, we * ux, we * uy, we * ux * ux, we * ux * uy, we * uy * uy};
                    if (slot < (uint16_t)H) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) atomicAdd(&L.acc[j][slot], v[j]);
                    } else {
#pragma unroll
                        for (int j = 0; j < 8; ++j) unsafeAtomicAdd(A.g[j] + id, v[j]);
                    }
                } else {
                    /* removalDeltaForSite, gradients.cu:82-100 */
                    float keep = 1.0f - wt;
                    float dl;
                    if (keep < 1e-4f) {
                        dl = 1.0e6f;
                    } else {
                        float ik = 1.0f / keep;
                        float k0 = (pr - wt * t.p2.x) * ik - tr;
                        float k1 = (pg - wt * t.p2.y) * ik - tg;
                        float k2 = (pb - wt * t.p2.z) * ik - tb;
                        float d0 = pr - tr, d1 = pg - tg, d2 = pb - tb;
                        dl = (k0 * k0 + k1 * k1 + k2 * k2) - (d0 * d0 + d1 * d1 + d2 * d2);
                    }
                    if (slot < (uint16_t)H) atomicAdd(&L.delta[slot], dl);
                    else unsafeAtomicAdd(A.removal + id, dl);
                }
            }
        }
        if (MODE == MODE_STATS) continue;

        /* backward of one candidate (gradients.cu:352-420) */
        if (!(valid && has)) continue;
        float gx_, gy_, gz_;
        if (MODE == MODE_GRAD) {
            gx_ = 2.0f * (pr - tr); gy_ = 2.0f * (pg - tg); gz_ = 2.0f * (pb - tb);
        } else {
            const float* go = A.grad_out + 3u * pix;
            gx_ = go[0]; gy_ = go[1]; gz_ = go[2];
        }
        float c[NACC];
        c[4] = w * gx_;
        c[5] = w * gy_;
        c[6] = w * gz_;
        float dlogit = c[4] * (t.p2.x - pr) + c[5] * (t.p2.y - pg) + c[6] * (t.p2.z - pb);
        float tau = t.p0.z;
        float l1 = t.p1.z, l2 = t.p2.w;
        float d2s = fmaxf(fmaf(l1, proj2, l2 * perp2) * A.iss, 1e-8f);
        float isd = fast_rsq(d2s);
        float dmix = fast_sqrt(d2s) - t.p0.w;
        float dl12 = l1 - l2;
        float gxd = fmaf(dl12 * t.p1.x, proj, l2 * dx);
        float gyd = fmaf(dl12 * t.p1.y, proj, l2 * dy);
        float cpos = dlogit * (tau * A.iss * isd);
        c[0] = cpos * gxd;
        c[1] = cpos * gyd;
        c[2] = -dlogit * dmix * tau;
        c[3] = dlogit * tau * A.inv_scale;
        float hq = -dlogit * tau * 0.5f * isd;
        c[9] = hq * (fmaf(l1, proj2, -(l2 * perp2)) * A.iss);
        float cd = 2.0f * dl12 * proj * A.iss;
        c[7] = hq * (cd * dx);
        c[8] = hq * (cd * dy);
        if (slot < (uint16_t)H) {
#pragma unroll
            for (int j = 0; j < NACC; ++j) atomicAdd(&L.acc[j][slot], c[j]);
        } else if (MODE == MODE_GRAD) {
#pragma unroll
            for (int j = 0; j < NACC; ++j) unsafeAtomicAdd(A.g[j] + id, c[j]);
        } else {
#pragma unroll
            for (int j = 0; j < NACC; ++j) unsafeAtomicAdd(A.out + 10u * id + (uint32_t)j, c[j]);
        }
    }

    if (MODE == MODE_RENDER) return;
    __syncthreads();

    /* stage 4: flush the tile's sums */
    for (uint32_t s = tid; s < (uint32_t)H; s += 256u) {
        uint32_t id = L.key[s];
        if (id == SAD_EMPTY || !(L.p0[s].x >= 0.0f)) continue;
        if (MODE == MODE_GRAD) {
#pragma unroll
            for (int j = 0; j < NACC; ++j) unsafeAtomicAdd(A.g[j] + id, L.acc[j][s]);
            if (REMOVAL && L.delta[s] != 0.0f) unsafeAtomicAdd(A.removal + id, L.delta[s]);
        } else if (MODE == MODE_BWD) {
#pragma unroll
            for (int j = 0; j < NACC; ++j) unsafeAtomicAdd(A.out + 10u * id + (uint32_t)j, L.acc[j][s]);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) unsafeAtomicAdd(A.g[j] + id, L.acc[j][s]);
        }
    }
}

/* ------------------------------------------------------------ launchers -- */

constexpr int HASH_SLOTS = 256;

template <int MODE, bool REMOVAL>
static int launch_pixel(PixelArgs& A, hipStream_t stream, const char* what)
{
    A.tiles_x = div_up((uint32_t)A.width, 16u);
    A.ntiles = A.tiles_x * div_up((uint32_t)A.height, 16u);
    A.inv_scale = sqrtf(A.iss);
    hipLaunchKernelGGL((k_pixel<MODE, REMOVAL, HASH_SLOTS>), dim3(A.ntiles), dim3(256), 0, stream, A);
    return launch_status(what);
}

int launch_render(const uint32_t* cand0, const uint32_t* cand1, float* out, const SadSite* sites, float iss,
                  uint32_t n, int width, int height, int cw, int ch, hipStream_t stream)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.out = out; A.iss = iss; A.n = n;
    A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    return launch_pixel<MODE_RENDER, false>(A, stream, "render_voronoi");
}

int launch_gradients(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                     float* const g[10], const SadSite* sites, float iss, uint32_t n, float* removal,
                     bool compute_removal, int width, int height, int cw, int ch, hipStream_t stream)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    for (int j = 0; j < 10; ++j) A.g[j] = g[j];
    A.removal = removal; A.iss = iss; A.n = n;
    A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    if (compute_removal) return launch_pixel<MODE_GRAD, true>(A, stream, "compute_gradients_tiled");
    return launch_pixel<MODE_GRAD, false>(A, stream, "compute_gradients_tiled");
}

int launch_site_stats(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                      const SadSite* sites, float iss, uint32_t n, float* const st[8], int width, int height,
                      int cw, int ch, hipStream_t stream)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    for (int j = 0; j < 8; ++j) A.g[j] = st[j];
    A.iss = iss; A.n = n; A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    return launch_pixel<MODE_STATS, false>(A, stream, "compute_site_stats");
}

}  // namespace sadh

using namespace sadh;

extern "C" {

int sadh_render_voronoi(const uint32_t* cand0, const uint32_t* cand1, float* output, const SadSite* sites,
                        float inv_scale_sq, uint32_t site_count, int width, int height, int cand_width,
                        int cand_height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && output && sites, "render_voronoi: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "render_voronoi: bad sizes");
    return launch_render(cand0, cand1, output, sites, inv_scale_sq, site_count, width, height, cand_width,
                         cand_height, to_hip(stream));
}

int sadh_render_backward(const uint32_t* cand0, const uint32_t* cand1, const float* grad_output,
                         float* grad_sites, const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                         int width, int height, int cand_width, int cand_height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && grad_output && grad_sites && sites, "render_backward: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "render_backward: bad sizes");
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.grad_out = grad_output; A.out = grad_sites;
    A.iss = inv_scale_sq; A.n = site_count; A.width = width; A.height = height;
    A.cw = cand_width; A.ch = cand_height;
    return launch_pixel<MODE_BWD, false>(A, to_hip(stream), "render_backward");
}

int sadh_compute_gradients_tiled(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                                 const float* rendered, const float* mask, float* grad_pos_x, float* grad_pos_y,
                                 float* grad_log_tau, float* grad_radius, float* grad_color_r, float* grad_color_g,
                                 float* grad_color_b, float* grad_dir_x, float* grad_dir_y, float* grad_log_aniso,
                                 const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                                 float* removal_delta, uint32_t compute_removal, float ssim_weight, int width,
                                 int height, int cand_width, int cand_height, sadh_stream_t stream)
{
    (void)rendered;
    if (ssim_weight > 0.0f) return fail(SADH_ERR_UNSUPPORTED, "compute_gradients_tiled: SSIM loss is not implemented");
    SADH_REQUIRE(cand0 && cand1 && target && mask && sites, "compute_gradients_tiled: null pointer");
    SADH_REQUIRE(grad_pos_x && grad_pos_y && grad_log_tau && grad_radius && grad_color_r && grad_color_g &&
                     grad_color_b && grad_dir_x && grad_dir_y && grad_log_aniso,
                 "compute_gradients_tiled: null gradient array");
    SADH_REQUIRE(!compute_removal || removal_delta, "compute_gradients_tiled: removal_delta missing");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "compute_gradients_tiled: bad sizes");
    float* const g[10] = {grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r,
                          grad_color_g, grad_color_b, grad_dir_x, grad_dir_y, grad_log_aniso};
    return launch_gradients(cand0, cand1, target, mask, g, sites, inv_scale_sq, site_count, removal_delta,
                            compute_removal != 0, width, height, cand_width, cand_height, to_hip(stream));
}

}  // extern "C"
