/*
 * k_sites.hip - per-site kernels: fused Adam, gradient clears, tau-gradient
 * diffusion, densify/prune scores and site edits, Hilbert keys, metrics.
 *
 * All of these stream N-sized arrays once (HBM-bound); one thread per site,
 * 256-thread blocks.  Sites and Adam state are AoS (40 B / 80 B per site, read
 * and written whole with 8- and 16-byte accesses); gradients are SoA.
 */
#include <algorithm>
#include <type_traits>

#include "fit_common.h"

namespace sadh {

/* ---------------------------------------------------------------- Adam --
 * adamUpdateKernel, adam.cu:5-180.  b1c = 1/(1-beta1^t), b2c = 1/(1-beta2^t)
 * arrive precomputed (host sad_powi_f, once per launch).  CLEAR zeroes the ten
 * gradient entries after reading them (adam_encoder.metal:117-127). */
struct AdamArgs {
    SadSite* sites;
    SadAdamState* st;
    float* g[10];              /* gradient j of site i at g[j][i * gstride] */
    uint32_t gstride;          /* 1: SoA arrays; 16: AoS [N][16] lines */
    long long* gfix;           /* non-null: gradient j of site i is the (lo, hi) pair at gfix + i * FIX_GRAD_STRIDE + 2 j */
    float lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso;
    float beta1, beta2, eps, b1c, b2c, norm, maxx, maxy;
    uint32_t n;
};

/* a site and its Adam moments as read by adam_apply */
struct AdamIn {
    float2 s0, s1, s2, s3, s4;
    float4 a0, a1, a2, a3, a4;
};

__device__ __forceinline__ AdamIn adam_load(const AdamArgs& A, uint32_t i)
{
    const float2* sp = reinterpret_cast<const float2*>(A.sites + i);
    const float4* ap = reinterpret_cast<const float4*>(A.st + i);
    AdamIn in;
    in.s0 = sp[0]; in.s1 = sp[1]; in.s2 = sp[2]; in.s3 = sp[3]; in.s4 = sp[4];
    in.a0 = ap[0]; in.a1 = ap[1]; in.a2 = ap[2]; in.a3 = ap[3]; in.a4 = ap[4];
    return in;
}

/* one site's update from its normalised gradient g[10], given the site and moments read by adam_load */
__device__ __forceinline__ void adam_apply(const AdamArgs& A, uint32_t i, const AdamIn& in, const float g[10])
{
    float2 s0 = in.s0, s1 = in.s1, s2 = in.s2, s3 = in.s3, s4 = in.s4;
    /* s0 = pos, s1 = (log_tau, radius), s2 = (r, g), s3 = (b, dir.x), s4 = (dir.y, log_aniso) */
    if (s0.x < 0.0f) return;
    float4 a0 = in.a0, a1 = in.a1, a2 = in.a2, a3 = in.a3, a4 = in.a4;
    /* a0 = (m_pos.x, m_pos.y, v_pos, m_log_tau), a1 = (v_log_tau, m_radius, v_radius, m_r),
     * a2 = (m_g, m_b, v_r, v_g), a3 = (v_b, m_dir.x, m_dir.y, v_dir), a4 = (m_la, v_la, pad, pad) */
    const float b1 = A.beta1, b2 = A.beta2, ob1 = 1.0f - A.beta1, ob2 = 1.0f - A.beta2;
    const float b1c = A.b1c, b2c = A.b2c, eps = A.eps;

    /* position: one second moment for both axes */
    a0.x = b1 * a0.x + ob1 * g[0];
    a0.y = b1 * a0.y + ob1 * g[1];
    a0.z = b2 * a0.z + ob2 * (g[0] * g[0] + g[1] * g[1]);
    float idp = 1.0f / (sqrtf(a0.z * b2c) + eps);
    s0.x = sad_clampf(s0.x - A.lr_pos * (a0.x * b1c) * idp, 0.0f, A.maxx);
    s0.y = sad_clampf(s0.y - A.lr_pos * (a0.y * b1c) * idp, 0.0f, A.maxy);
    /* log_tau */
    a0.w = b1 * a0.w + ob1 * g[2];
    a1.x = b2 * a1.x + ob2 * (g[2] * g[2]);
    s1.x = s1.x - A.lr_tau * (a0.w * b1c) / (sqrtf(a1.x * b2c) + eps);
    /* radius in [1, 512] */
    a1.y = b1 * a1.y + ob1 * g[3];
    a1.z = b2 * a1.z + ob2 * (g[3] * g[3]);
    s1.y = sad_clampf(s1.y - A.lr_radius * (a1.y * b1c) / (sqrtf(a1.z * b2c) + eps), 1.0f, 512.0f);
    /* colour, per channel, in [0, 1] */
    a1.w = b1 * a1.w + ob1 * g[4];
    a2.x = b1 * a2.x + ob1 * g[5];
    a2.y = b1 * a2.y + ob1 * g[6];
    a2.z = b2 * a2.z + ob2 * (g[4] * g[4]);
    a2.w = b2 * a2.w + ob2 * (g[5] * g[5]);
    a3.x = b2 * a3.x + ob2 * (g[6] * g[6]);
    s2.x = sad_clampf(s2.x - A.lr_color * (a1.w * b1c) / (sqrtf(a2.z * b2c) + eps), 0.0f, 1.0f);
    s2.y = sad_clampf(s2.y - A.lr_color * (a2.x * b1c) / (sqrtf(a2.w * b2c) + eps), 0.0f, 1.0f);
    s3.x = sad_clampf(s3.x - A.lr_color * (a2.y * b1c) / (sqrtf(a3.x * b2c) + eps), 0.0f, 1.0f);
    /* direction: tangent projection, isotropic Adam, renormalisation */
    float dxv = s3.y, dyv = s4.x;
    float gd = g[7] * dxv + g[8] * dyv;
    float tx = g[7] - dxv * gd, ty = g[8] - dyv * gd;
    a3.y = b1 * a3.y + ob1 * tx;
    a3.z = b1 * a3.z + ob1 * ty;
    a3.w = b2 * a3.w + ob2 * (tx * tx + ty * ty);
    float idd = 1.0f / (sqrtf(a3.w * b2c) + eps);
    float nx = dxv - A.lr_dir * (a3.y * b1c) * idd;
    float ny = dyv - A.lr_dir * (a3.z * b1c) * idd;
    float len2 = nx * nx + ny * ny;
    if (len2 < 1e-12f) { nx = 1.0f; ny = 0.0f; }
    else { float il = 1.0f / sqrtf(len2); nx *= il; ny *= il; }
    s3.y = nx;
    s4.x = ny;
    /* log_aniso in [-4, 4] */
    a4.x = b1 * a4.x + ob1 * g[9];
    a4.y = b2 * a4.y + ob2 * (g[9] * g[9]);
    s4.y = sad_clampf(s4.y - A.lr_aniso * (a4.x * b1c) / (sqrtf(a4.y * b2c) + eps), -4.0f, 4.0f);

    float2* sp = reinterpret_cast<float2*>(A.sites + i);
    float4* ap = reinterpret_cast<float4*>(A.st + i);
    sp[0] = s0; sp[1] = s1; sp[2] = s2; sp[3] = s3; sp[4] = s4;
    ap[0] = a0; ap[1] = a1; ap[2] = a2; ap[3] = a3; ap[4] = a4;
}

/* one site's update from its normalised gradient g[10] */
__device__ __forceinline__ void adam_site(const AdamArgs& A, uint32_t i, const float g[10])
{
    adam_apply(A, i, adam_load(A, i), g);
}

/* site i's ten gradients times norm, with log_tau replaced by `tau`, then
 * zeroed at the source.  Fixed-point lines round once to float here: the
 * cross-tile sums that filled them are integer, so the float seen by Adam does
 * not depend on the order of the atomics. */
__device__ __forceinline__ void take_gradients(const AdamArgs& A, uint32_t i, float tau, float g[10])
{
    if (A.gfix) {
        longlong2* line = reinterpret_cast<longlong2*>(A.gfix + (size_t)i * FIX_GRAD_STRIDE);
#pragma unroll
        for (int j = 0; j < 10; ++j) {
            const longlong2 d = line[j];
            g[j] = ((j == 2) ? tau : fix_value(d.x, d.y)) * A.norm;
        }
#pragma unroll
        for (int j = 0; j < 10; ++j) line[j] = make_longlong2(0, 0);
        return;
    }
#pragma unroll
    for (int j = 0; j < 10; ++j) g[j] = ((j == 2) ? tau : A.g[j][i * A.gstride]) * A.norm;
#pragma unroll
    for (int j = 0; j < 10; ++j) A.g[j][i * A.gstride] = 0.0f;
}

/* take_gradients for fixed-point lines in two steps: the line's words are read
 * first (grad_line_load), so the read can be issued early, and converted and
 * cleared later (grad_line_take) */
struct GradLine {
    longlong2 d[10];
};

__device__ __forceinline__ GradLine grad_line_load(const AdamArgs& A, uint32_t i)
{
    const longlong2* line = reinterpret_cast<const longlong2*>(A.gfix + (size_t)i * FIX_GRAD_STRIDE);
    GradLine L;
#pragma unroll
    for (int j = 0; j < 10; ++j) L.d[j] = line[j];
    return L;
}

__device__ __forceinline__ void grad_line_take(const AdamArgs& A, uint32_t i, const GradLine& L, float tau, float g[10])
{
#pragma unroll
    for (int j = 0; j < 10; ++j) g[j] = ((j == 2) ? tau : fix_value(L.d[j].x, L.d[j].y)) * A.norm;
    longlong2* line = reinterpret_cast<longlong2*>(A.gfix + (size_t)i * FIX_GRAD_STRIDE);
#pragma unroll
    for (int j = 0; j < 10; ++j) line[j] = make_longlong2(0, 0);
}

template <bool CLEAR>
__global__ __launch_bounds__(256) void k_adam(AdamArgs A)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= A.n) return;
    float g[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) g[j] = A.g[j][i * A.gstride] * A.norm;
    if (CLEAR) {
#pragma unroll
        for (int j = 0; j < 10; ++j) A.g[j][i * A.gstride] = 0.0f;
    }
    adam_site(A, i, g);
}

/* clearGradientsKernel / clearBufferKernel, adam.cu:183-219 */
struct ClearArgs {
    float* g[10];
    int count_arrays;
    uint32_t n;
};

__global__ __launch_bounds__(256) void k_clear(ClearArgs A)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= A.n) return;
    for (int j = 0; j < A.count_arrays; ++j) A.g[j][i] = 0.0f;
}

/* -------------------------------------------------------- tau diffusion --
 * tau_diffuse_kernel, tau_diffusion.cu:27-89: one Jacobi step over the unique
 * active neighbours found in the candidate list of the site's home cell.
 * raw_copy (nullable) receives grad_raw[i], folding the copy the trainer makes
 * before the first pass (sad_cuda.cu:2335-2337) into that pass. */
__device__ __forceinline__ void tau_site(uint32_t i, const uint4* __restrict__ c0, const uint4* __restrict__ c1,
                                         const SadSite* __restrict__ sites, const float* __restrict__ graw,
                                         const float* __restrict__ gin, float* __restrict__ gout,
                                         float* __restrict__ raw_copy, uint32_t n, float lambda, int width,
                                         int height, float inv_ds, uint32_t in_stride = 1u)
{
    /* graw and gin hold site i at [i * in_stride]; gout and raw_copy are contiguous */
    float raw = graw[i * in_stride];
    if (raw_copy) raw_copy[i] = raw;
    const float2* sp = reinterpret_cast<const float2*>(sites + i);
    float2 pos = sp[0];
    if (lambda <= 0.0f || pos.x < 0.0f) { gout[i] = gin[i * in_stride]; return; }
    int px = sad_clampi((int)(pos.x * inv_ds + 0.5f), 0, width - 1);
    int py = sad_clampi((int)(pos.y * inv_ds + 0.5f), 0, height - 1);
    uint4 a = c0[py * width + px], b = c1[py * width + px];
    uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    float sum = 0.0f;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t nb = ids[k];
        bool ok = nb < n && nb != i;
#pragma unroll
        for (int j = 0; j < k; ++j) ok = ok && ids[j] != nb;
        if (ok && sites[nb].pos_x >= 0.0f) {
            sum += gin[nb * in_stride];
            cnt += 1u;
        }
    }
    float diag = 1.0f + lambda * (float)cnt;
    gout[i] = (raw + lambda * sum) / diag;
}

/* ------------------------------------------------ tau passes from lists --
 * The four passes of one iteration read the same neighbour sets: positions,
 * the candidate field and activity do not change between them.  The first pass
 * writes each site's set as 8 ids in candidate order (SAD_EMPTY for ids that do
 * not count, NBR_COPY in slot 0 when the pass copies gin through), and the
 * later passes read that list instead of chasing position -> home cell ->
 * candidates -> neighbour activity again.  Sums run in the same order as in
 * tau_site, so the results are bit-identical. */

template <typename TR, typename TI>
__device__ __forceinline__ void tau_site_list(uint32_t i, const uint4* __restrict__ c0, const uint4* __restrict__ c1,
                                              const SadSite* __restrict__ sites, TR graw, TI gin,
                                              float* __restrict__ gout,
                                              float* __restrict__ raw_copy, uint32_t n, float lambda, int width,
                                              int height, float inv_ds, uint32_t in_stride, uint4* __restrict__ nbr)
{
    const float raw = (float)graw[i * in_stride];
    if (raw_copy) raw_copy[i] = raw;
    const float2 pos = reinterpret_cast<const float2*>(sites + i)[0];
    uint32_t keep[8];
    if (lambda <= 0.0f || pos.x < 0.0f) {
        gout[i] = (float)gin[i * in_stride];
        nbr[2u * i] = make_uint4(NBR_COPY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY);
        nbr[2u * i + 1u] = make_uint4(SAD_EMPTY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY);
        return;
    }
    const int px = sad_clampi((int)(pos.x * inv_ds + 0.5f), 0, width - 1);
    const int py = sad_clampi((int)(pos.y * inv_ds + 0.5f), 0, height - 1);
    const uint4 a = c0[py * width + px], b = c1[py * width + px];
    const uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    float sum = 0.0f;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t nb = ids[k];
        bool ok = nb < n && nb != i;
#pragma unroll
        for (int j = 0; j < k; ++j) ok = ok && ids[j] != nb;
        ok = ok && sites[nb].pos_x >= 0.0f;
        keep[k] = ok ? nb : SAD_EMPTY;
        if (ok) {
            sum += (float)gin[nb * in_stride];
            cnt += 1u;
        }
    }
    nbr[2u * i] = make_uint4(keep[0], keep[1], keep[2], keep[3]);
    nbr[2u * i + 1u] = make_uint4(keep[4], keep[5], keep[6], keep[7]);
    gout[i] = (raw + lambda * sum) / (1.0f + lambda * (float)cnt);
}

/* one pass from the list: the value tau_site computes (graw and gin hold site i at [i * stride]; float
 * arrays or FixCol columns of fixed-point lines) */
template <typename TR, typename TI>
__device__ __forceinline__ float tau_from_list(uint32_t i, const uint4* __restrict__ nbr, TR graw, TI gin,
                                               float lambda, uint32_t stride = 1u)
{
    const uint4 a = nbr[2u * i], b = nbr[2u * i + 1u];
    if (a.x == NBR_COPY) return (float)gin[i * stride];
    const uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    float sum = 0.0f;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (ids[k] != SAD_EMPTY) {
            sum += (float)gin[ids[k] * stride];
            cnt += 1u;
        }
    }
    return ((float)graw[i * stride] + lambda * sum) / (1.0f + lambda * (float)cnt);
}

template <typename T>
__global__ __launch_bounds__(256) void k_tau_list_first(const uint4* __restrict__ c0, const uint4* __restrict__ c1,
                                                        const SadSite* __restrict__ sites, T g2,
                                                        float* __restrict__ gout, float* __restrict__ raw_copy,
                                                        uint32_t n, float lambda, int width, int height, float inv_ds,
                                                        uint4* __restrict__ nbr, uint32_t in_stride)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tau_site_list(i, c0, c1, sites, g2, g2, gout, raw_copy, n, lambda, width, height, inv_ds, in_stride,
                             nbr);
}

__global__ __launch_bounds__(256) void k_tau_list_pass(const uint4* __restrict__ nbr, const float* __restrict__ graw,
                                                       const float* __restrict__ gin, float* __restrict__ gout,
                                                       uint32_t n, float lambda)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gout[i] = tau_from_list(i, nbr, graw, gin, lambda);
}

__global__ __launch_bounds__(256) void k_tau_list_last_adam(const uint4* __restrict__ nbr,
                                                            const float* __restrict__ graw,
                                                            const float* __restrict__ gin, float lambda, AdamArgs A)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    float g[10];
    take_gradients(A, i, tau_from_list(i, nbr, graw, gin, lambda), g);
    adam_site(A, i, g);
}

/* Two passes in one launch.  Site i's pass-(k+1) value needs its neighbours'
 * pass-k values; each is recomputed here from the neighbour's own list with the
 * same arithmetic as its separate pass (tau_from_list), so the result has the
 * same bits as two launches.  Eight lanes serve one site, lane k recomputing
 * neighbour k's value; the group's first lane sums them in slot order. */
template <typename TR, typename TI>
__device__ __forceinline__ float tau_two_passes8(uint32_t i, uint32_t lane, const uint4* __restrict__ nbr,
                                                 TR graw, TI gin, float lambda, uint32_t stride)
{
    /* site i's own raw value and list are read first, alongside each other */
    const float raw_i = (float)graw[i * stride];
    const uint4 a = nbr[2u * i], b = nbr[2u * i + 1u];
    const uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t mine = ids[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) mine = lane == (uint32_t)k ? ids[k] : mine;
    const bool copy = a.x == NBR_COPY;
    float v = 0.0f;
    if (!copy && mine != SAD_EMPTY) v = tau_from_list(mine, nbr, graw, gin, lambda, stride);
    const int base = (int)(threadIdx.x & ~7u);
    float vals[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) vals[k] = __shfl(v, base + k, 64);
    if (copy) return (float)gin[i * stride];
    float sum = 0.0f;
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (ids[k] != SAD_EMPTY) {
            sum += vals[k];
            cnt += 1u;
        }
    }
    return (raw_i + lambda * sum) / (1.0f + lambda * (float)cnt);
}

/* passes 1 and 2 from lists built during the gradient launch: reads the
 * log_tau column of the gradient lines, saves raw, writes the pass-2 value.
 * 8 threads per site; blockDim is a multiple of 64. */
template <typename T>
__global__ __launch_bounds__(256) void k_tau_list_12_built(const uint4* __restrict__ nbr, T g2,
                                                           float* __restrict__ gout, float* __restrict__ raw_copy,
                                                           uint32_t n, float lambda)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t >> 3, lane = t & 7u;
    if (i >= n) return;                        /* whole groups of 8 leave together */
    const float v = tau_two_passes8(i, lane, nbr, g2, g2, lambda, 1u);
    if (lane == 0u) {
        raw_copy[i] = (float)g2[i];   /* the same word tau_two_passes8 read first: no second wait */
        gout[i] = v;
    }
}

/* passes 3 and 4 from the pass-2 values, then Adam (which reads and clears the
 * site's gradient line); 8 threads per site */
__global__ __launch_bounds__(256) void k_tau_list_34_adam(const uint4* __restrict__ nbr,
                                                          const float* __restrict__ graw,
                                                          const float* __restrict__ gin, float lambda, AdamArgs A)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t >> 3, lane = t & 7u;
    if (i >= A.n) return;
    /* the group's first lane reads the site, its moments and its gradient line before the tau passes, so
     * those reads wait alongside the passes' first reads instead of after the passes */
    AdamIn in = {};
    GradLine gl = {};
    if (lane == 0u) {
        in = adam_load(A, i);
        gl = grad_line_load(A, i);
    }
    const float tau = tau_two_passes8(i, lane, nbr, graw, gin, lambda, 1u);
    if (lane != 0u) return;
    float g[10];
    grad_line_take(A, i, gl, tau, g);
    adam_apply(A, i, in, g);
}

__global__ __launch_bounds__(256) void k_tau_diffuse(const uint4* __restrict__ c0, const uint4* __restrict__ c1,
                                                     const SadSite* __restrict__ sites,
                                                     const float* __restrict__ graw, const float* __restrict__ gin,
                                                     float* __restrict__ gout, float* __restrict__ raw_copy,
                                                     uint32_t n, float lambda, int width, int height, float inv_ds)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) tau_site(i, c0, c1, sites, graw, gin, gout, raw_copy, n, lambda, width, height, inv_ds);
}

/* Batched fit: tau pass `pass` (0-2: log_tau column of the gradient lines ->
 * tmp -> tmp2 -> tmp, raw saved by pass 0) or, for pass 3, the last pass fused
 * with Adam + clear, over the concatenated sites of every job; the candidate
 * field is each job's current one. */
__global__ __launch_bounds__(256) void k_tau_batch(const BatchJob* __restrict__ jobs, uint32_t njobs, uint32_t pass,
                                                   const JobStep* __restrict__ steps, BatchStep st)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t jx = batch_find(jobs, njobs, f, &BatchJob::site_begin);
    const BatchJob J = job_global(jobs, jx);
    const JobStep js = steps[jx];
    const uint32_t i = f - J.site_begin;
    if (i >= J.n) return;
    const uint4* c0 = reinterpret_cast<const uint4*>(J.cur ? J.fb0 : J.fa0);
    const uint4* c1 = reinterpret_cast<const uint4*>(J.cur ? J.fb1 : J.fa1);
    const float inv_ds = 1.0f / fmaxf((float)J.ds, 1.0f);
    uint4* nbr = reinterpret_cast<uint4*>(J.tau_nbr);
    if (pass == 0u) {
        const FixCol g2{J.gacc + 4, FIX_GRAD_STRIDE};
        tau_site_list(i, c0, c1, J.sites, g2, g2, J.tau_tmp, J.tau_raw, J.n, js.lambda, J.cw, J.ch, inv_ds, 1u,
                      nbr);
    } else if (pass == 1u) {
        J.tau_tmp2[i] = tau_from_list(i, nbr, (const float*)J.tau_raw, (const float*)J.tau_tmp, js.lambda);
    } else if (pass == 2u) {
        J.tau_tmp[i] = tau_from_list(i, nbr, (const float*)J.tau_raw, (const float*)J.tau_tmp2, js.lambda);
    } else {
        AdamArgs A;
        A.sites = J.sites; A.st = J.adam;
        A.gfix = J.gacc;
        A.lr_pos = st.lr_pos; A.lr_tau = st.lr_tau; A.lr_radius = st.lr_radius; A.lr_color = st.lr_color;
        A.lr_dir = st.lr_dir; A.lr_aniso = st.lr_aniso; A.beta1 = st.beta1; A.beta2 = st.beta2; A.eps = st.eps;
        A.b1c = js.b1c; A.b2c = js.b2c; A.norm = J.norm; A.maxx = J.maxx; A.maxy = J.maxy; A.n = J.n;
        /* the site, its moments and its gradient line are read before the last pass's gathers */
        const AdamIn in = adam_load(A, i);
        const GradLine gl = grad_line_load(A, i);
        const float tau = tau_from_list(i, nbr, (const float*)J.tau_raw, (const float*)J.tau_tmp, js.lambda);
        float g[10];
        grad_line_take(A, i, gl, tau, g);
        adam_apply(A, i, in, g);
    }
}

/* Batched fit: set n and/or flip the current field of listed jobs, in stream order. */
struct JobUpdate {
    uint32_t count;
    uint32_t job[64];
    uint32_t n[64];            /* 0xFFFFFFFF: keep */
    uint32_t flip[64];         /* 1: swap the current candidate field */
};

__global__ void k_update_jobs(BatchJob* jobs, JobUpdate u)
{
    const uint32_t t = threadIdx.x;
    if (t >= u.count) return;
    BatchJob& J = jobs[u.job[t]];
    if (u.n[t] != 0xFFFFFFFFu) J.n = u.n[t];
    if (u.flip[t]) J.cur ^= 1u;
}

/* ----------------------------------------------------- densify / prune --- */

/* computeDensifyScorePairsKernel, densify_prune.cu:223-254.  pow(m, alpha) is
 * the exact exp(alpha * log m) of sad_math.h so keys match the CPU bit for bit. */
template <typename T>
__global__ __launch_bounds__(256) void k_densify_pairs(const SadSite* __restrict__ sites, T mass, T energy,
                                                       SadPair* __restrict__ pairs,
                                                       uint32_t n, float min_mass, float alpha, uint32_t count)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    uint2 out = make_uint2(SAD_EMPTY, SAD_EMPTY);
    if (i < n) {
        const float m = (float)mass[i], e = (float)energy[i];
        if (sites[i].pos_x >= 0.0f && m > min_mass && isfinite(e)) {
            float den = sad_powf_pos(fmaxf(m, 1e-8f), alpha);
            float score = fmaxf(e, 0.0f) / den;
            out = make_uint2(0xFFFFFFFFu - __float_as_uint(score), i);
        }
    }
    reinterpret_cast<uint2*>(pairs)[i] = out;
}

/* computePruneScorePairsKernel, densify_prune.cu:257-281 */
template <typename T>
__global__ __launch_bounds__(256) void k_prune_pairs(const SadSite* __restrict__ sites, T rdelta,
                                                     SadPair* __restrict__ pairs, uint32_t n, float delta_norm,
                                                     uint32_t count)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    uint2 out = make_uint2(SAD_EMPTY, SAD_EMPTY);
    if (i < n) {
        float d = (float)rdelta[i] * delta_norm;
        if (sites[i].pos_x >= 0.0f && isfinite(d)) out = make_uint2(__float_as_uint(fmaxf(d, 0.0f)), i);
    }
    reinterpret_cast<uint2*>(pairs)[i] = out;
}

/* writeSplitIndicesFromSortedKernel, densify_prune.cu:284-292 */
__global__ __launch_bounds__(256) void k_write_indices(const SadPair* __restrict__ sorted, uint32_t* __restrict__ idx,
                                                       uint32_t count)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < count) idx[i] = sorted[i].val;
}

/* pruneSitesByIndexKernel, densify_prune.cu:295-307 */
__global__ __launch_bounds__(256) void k_prune_sites(SadSite* __restrict__ sites, const uint32_t* __restrict__ idx,
                                                     uint32_t count)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= count) return;
    uint32_t s = idx[i];
    if (s == SAD_EMPTY) return;
    reinterpret_cast<float2*>(sites + s)[0] = make_float2(-1.0f, -1.0f);
}

/* splitSitesKernel, densify_prune.cu:86-220, with parents >= cur skipped */
template <typename T>
struct SplitArgs {
    SadSite* sites;
    SadAdamState* st;
    const uint32_t* idx;
    T mass, ew, ewx, ewy, ewxx, ewxy, ewyy;        /* float arrays or FixCol columns */
    const float* target;
    uint32_t k, cur;
    int width, height;
};

template <typename T>
__global__ __launch_bounds__(256) void k_split(SplitArgs<T> A)
{
    uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    if (gid >= A.k) return;
    uint32_t pi = A.idx[gid];
    if (pi >= A.cur) return;
    SadSite parent = A.sites[pi];
    if (parent.pos_x < 0.0f) return;
    const float w = (float)A.ew[pi];
    float ax = parent.dir_x, ay = parent.dir_y;
    float cx = parent.pos_x, cy = parent.pos_y;
    float la = parent.log_aniso * 0.8f;
    const int W = A.width, Hh = A.height;
    if (w > 1e-3f) {
        float mx = (float)A.ewx[pi] / w, my = (float)A.ewy[pi] / w;
        cx = 0.4f * parent.pos_x + 0.6f * mx;
        cy = 0.4f * parent.pos_y + 0.6f * my;
        float exx = fmaxf((float)A.ewxx[pi] / w - mx * mx, 1e-4f);
        float exy = (float)A.ewxy[pi] / w - mx * my;
        float eyy = fmaxf((float)A.ewyy[pi] / w - my * my, 1e-4f);
        /* the library's own atan2 and sincos (sad_math.h), so children are bit-identical to the oracle's */
        float th = 0.5f * sad_atan2f(2.0f * exy, exx - eyy);
        sad_sincosf(th, &ay, &ax);
        float tr = exx + eyy;
        float disc = sqrtf(fmaxf(0.0f, 0.25f * (exx - eyy) * (exx - eyy) + exy * exy));
        float l1 = fmaxf(1e-4f, 0.5f * tr + disc);
        float l2 = fmaxf(1e-4f, 0.5f * tr - disc);
        la = sad_clampf(-0.5f * sad_logf(l1 / l2), -2.0f, 2.0f);
    } else {
        int ix = sad_clampi((int)parent.pos_x, 0, W - 1);
        int iy = sad_clampi((int)parent.pos_y, 0, Hh - 1);
        float lum[3][3];
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                int sx = sad_clampi(ix + dx, 0, W - 1), sy = sad_clampi(iy + dy, 0, Hh - 1);
                const float* c = A.target + 3u * (uint32_t)(sy * W + sx);
                lum[dy + 1][dx + 1] = 0.299f * c[0] + 0.587f * c[1] + 0.114f * c[2];
            }
        float gx = (-lum[0][0] + lum[0][2] - 2.0f * lum[1][0] + 2.0f * lum[1][2] - lum[2][0] + lum[2][2]) / 4.0f;
        float gy = (-lum[0][0] - 2.0f * lum[0][1] - lum[0][2] + lum[2][0] + 2.0f * lum[2][1] + lum[2][2]) / 4.0f;
        float g2 = gx * gx + gy * gy;
        if (g2 > 1e-8f) {
            float len = sqrtf(g2);
            if (len > 1e-8f) { ax = gx / len; ay = gy / len; } else { ax = 1.0f; ay = 0.0f; }
        }
    }
    float m = fmaxf((float)A.mass[pi], 1.0f);
    float od = sad_clampf(sqrtf(m) * 0.5f, 1.5f, 48.0f);
    float ox = ax * od, oy = ay * od;
    float pax = sad_clampf(cx + ox, 0.0f, (float)(W - 1)), pay = sad_clampf(cy + oy, 0.0f, (float)(Hh - 1));
    float pbx = sad_clampf(cx - ox, 0.0f, (float)(W - 1)), pby = sad_clampf(cy - oy, 0.0f, (float)(Hh - 1));
    const float* ca = A.target + 3u * (uint32_t)((int)pay * W + (int)pax);
    const float* cb = A.target + 3u * (uint32_t)((int)pby * W + (int)pbx);
    uint32_t child = A.cur + gid;
    if (child == pi) return;
    SadSite c0 = parent, c1 = parent;
    c0.pos_x = pax; c0.pos_y = pay; c1.pos_x = pbx; c1.pos_y = pby;
    c0.col_r = ca[0]; c0.col_g = ca[1]; c0.col_b = ca[2];
    c1.col_r = cb[0]; c1.col_g = cb[1]; c1.col_b = cb[2];
    c0.log_tau = c1.log_tau = parent.log_tau - 0.25f;
    c0.radius = c1.radius = parent.radius * 0.85f;
    c0.dir_x = c1.dir_x = ax;
    c0.dir_y = c1.dir_y = ay;
    c0.log_aniso = c1.log_aniso = la;
    A.sites[pi] = c0;
    A.sites[child] = c1;
    float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4* sp = reinterpret_cast<float4*>(A.st + pi);
    float4* sc = reinterpret_cast<float4*>(A.st + child);
#pragma unroll
    for (int j = 0; j < 5; ++j) { sp[j] = z; sc[j] = z; }
}

/* ------------------------------------------------------------- Hilbert --- */

/* computeHilbertPairsKernel, sad_cuda.cu:1547-1571 */
__global__ __launch_bounds__(256) void k_hilbert_pairs(const SadSite* __restrict__ sites, SadPair* __restrict__ pairs,
                                                       uint32_t n, uint32_t padded, int width, int height, int bits)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= padded) return;
    uint2 out = make_uint2(SAD_EMPTY, 0u);
    if (i < n) {
        float2 pos = reinterpret_cast<const float2*>(sites + i)[0];
        int px = (int)pos.x, py = (int)pos.y;
        px = px < 0 ? 0 : (px >= width ? width - 1 : px);
        py = py < 0 ? 0 : (py >= height ? height - 1 : py);
        out = make_uint2(sad_hilbert_index((uint32_t)px, (uint32_t)py, bits), i);
    }
    reinterpret_cast<uint2*>(pairs)[i] = out;
}

/* writeHilbertOrderKernel, sad_cuda.cu:1573-1584 */
__global__ __launch_bounds__(256) void k_hilbert_order(const SadPair* __restrict__ sorted, uint32_t* __restrict__ order,
                                                       uint32_t* __restrict__ pos, uint32_t n)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    uint32_t s = sorted[i].val;
    order[i] = s;
    pos[s] = i;
}

/* Batched fit: the Hilbert pairs of every job with hil_n > 0 in one array.  The
 * key is (job << shift) | Hilbert index, so one stable sort leaves each job's
 * sites in a range of their own, in the order the job's own sort gives them;
 * entries past `total` are empty and sort last. */
__global__ __launch_bounds__(256) void k_hilbert_pairs_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                             const JobStep* __restrict__ steps,
                                                             SadPair* __restrict__ pairs, uint32_t total,
                                                             uint32_t padded, uint32_t shift)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= padded) return;
    uint2 out = make_uint2(SAD_EMPTY, 0u);
    if (i < total) {
        const uint32_t j = hil_find(steps, njobs, i);
        const uint32_t local = i - steps[j].hil_begin;
        const SadSite* sites = rebase(jobs, jobs[j].sites);
        const int width = jobs[j].width, height = jobs[j].height;
        const float2 pos = reinterpret_cast<const float2*>(sites + local)[0];
        int px = (int)pos.x, py = (int)pos.y;
        px = px < 0 ? 0 : (px >= width ? width - 1 : px);
        py = py < 0 ? 0 : (py >= height ? height - 1 : py);
        const uint32_t h = sad_hilbert_index((uint32_t)px, (uint32_t)py, (int)steps[j].hil_bits);
        out = make_uint2((j << shift) | h, local);
    }
    reinterpret_cast<uint2*>(pairs)[i] = out;
}

/* Batched fit, last iteration: the sites of every job whose step flag `retire`
 * is set, at the job's offset of the flattened site range (site_begin). */
__global__ __launch_bounds__(256) void k_gather_sites_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                            const JobStep* __restrict__ steps, uint32_t total,
                                                            SadSite* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const uint32_t j = batch_find(jobs, njobs, i, &BatchJob::site_begin);
    if (!steps[j].retire) return;
    const float2* src = reinterpret_cast<const float2*>(rebase(jobs, jobs[j].sites) + (i - jobs[j].site_begin));
    float2* dst = reinterpret_cast<float2*>(out + i);
#pragma unroll
    for (int k = 0; k < 5; ++k) dst[k] = src[k];
}

/* each job's range of the sorted pairs into its order / position tables */
__global__ __launch_bounds__(256) void k_hilbert_order_batch(const BatchJob* __restrict__ jobs,
                                                             const JobStep* __restrict__ steps,
                                                             const SadPair* __restrict__ sorted, uint32_t total,
                                                             uint32_t shift)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= total) return;
    const uint2 p = reinterpret_cast<const uint2*>(sorted)[i];
    const uint32_t j = p.x >> shift;
    const uint32_t r = i - steps[j].hil_begin;
    /* the descriptor holds the tables read-only for the VPT pass; this is their writer */
    const_cast<uint32_t*>(rebase(jobs, jobs[j].h_order))[r] = p.y;
    const_cast<uint32_t*>(rebase(jobs, jobs[j].h_pos))[p.y] = r;
}

/* ------------------------------------------------------------- metrics --- */

__global__ void k_zero_word(uint32_t* p) { *p = 0u; }

/* computePSNRKernel, utils.cu:5-26: block sum of squared error, one atomic per block.
 * Each block's partial is a fixed tree (a shuffle-down tree per wave, then
 * (w0 + w1) + (w2 + w3)); T = float is the reference's float atomic, and the
 * fixed-point kind (k_sse_fix) adds the partial as an exact (lo, hi) pair, so
 * the fit's PSNR log does not depend on the blocks' order. */
__device__ __forceinline__ float sse_block(const float* __restrict__ r, const float* __restrict__ t,
                                           const float* __restrict__ mask, uint32_t np, float* part, uint32_t block)
{
    uint32_t i = block * 256u + threadIdx.x;
    float v = 0.0f;
    if (i < np && mask[i] > 0.0f) {
        float d0 = r[3 * i] - t[3 * i], d1 = r[3 * i + 1] - t[3 * i + 1], d2 = r[3 * i + 2] - t[3 * i + 2];
        v = d0 * d0 + d1 * d1 + d2 * d2;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return (part[0] + part[1]) + (part[2] + part[3]);
}

__global__ __launch_bounds__(256) void k_sse_fix(const float* __restrict__ r, const float* __restrict__ t,
                                                 const float* __restrict__ mask, long long* __restrict__ acc,
                                                 uint32_t np)
{
    __shared__ float part[4];
    const float v = sse_block(r, t, mask, np, part, blockIdx.x);
    if (threadIdx.x < 2u) {
        long long lo, hi;
        fix_split(v, &lo, &hi);
        fix_atomic_add(acc + threadIdx.x, threadIdx.x ? hi : lo);
    }
}

__global__ void k_zero_pair(long long* p)
{
    if (threadIdx.x < 2u) p[threadIdx.x] = 0;
}

/* Batched fit, log iterations: the first total_pblocks blocks add each logging
 * job's squared error, block by block as k_sse_fix (the same 256-pixel blocks,
 * so the same exact sum), into out[4 j] and out[4 j + 1]; the rest count its
 * active sites into out[4 j + 2].  out is cleared by k_clear before. */
__global__ __launch_bounds__(256) void k_log_sums_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                        const JobStep* __restrict__ steps, uint32_t total_pblocks,
                                                        long long* __restrict__ out)
{
    __shared__ float part[4];
    if (blockIdx.x < total_pblocks) {
        const uint32_t j = batch_find(jobs, njobs, blockIdx.x, &BatchJob::pblock_begin);
        if (!steps[j].log) return;
        const BatchJob J = job_global(jobs, j);
        const uint32_t np = (uint32_t)J.width * (uint32_t)J.height;
        const float v = sse_block(J.render, J.target, J.mask, np, part, blockIdx.x - J.pblock_begin);
        if (threadIdx.x < 2u) {
            long long lo, hi;
            fix_split(v, &lo, &hi);
            fix_atomic_add(out + 4u * j + threadIdx.x, threadIdx.x ? hi : lo);
        }
        return;
    }
    const uint32_t sb = blockIdx.x - total_pblocks;
    const uint32_t j = batch_find(jobs, njobs, sb, &BatchJob::sblock_begin);
    if (!steps[j].log) return;
    const uint32_t n = jobs[j].n;
    const SadSite* sites = rebase(jobs, jobs[j].sites);
    const uint32_t i = (sb - jobs[j].sblock_begin) * 256u + threadIdx.x;
    const bool a = i < n && sites[i].pos_x >= 0.0f;
    const unsigned long long b = __ballot(a);
    if ((threadIdx.x & 63u) == 0u && b) fix_atomic_add(out + 4u * j + 2u, (long long)__popcll(b));
}

template <typename T>
__global__ __launch_bounds__(256) void k_sse(const float* __restrict__ r, const float* __restrict__ t,
                                             const float* __restrict__ mask, T* __restrict__ acc, uint32_t np)
{
    __shared__ float part[4];
    const float v = sse_block(r, t, mask, np, part, blockIdx.x);
    if (threadIdx.x == 0) unsafeAtomicAdd(acc, (T)v);
}

/* computeSSIMKernel, utils.cu:29-87: sum over masked pixels of the SSIM of the
 * pixel's 3x3 luma window; 16x16 tiles so the windows overlap in cache */
template <typename T>
__global__ __launch_bounds__(256) void k_ssim(const float* __restrict__ r, const float* __restrict__ t,
                                              const float* __restrict__ mask, T* __restrict__ acc, int width,
                                              int height, uint32_t tiles_x)
{
    __shared__ float part[4];
    const int x = (int)(blockIdx.x % tiles_x) * 16 + (int)(threadIdx.x & 15u);
    const int y = (int)(blockIdx.x / tiles_x) * 16 + (int)(threadIdx.x >> 4);
    float v = 0.0f;
    if (x < width && y < height && mask[(size_t)y * width + x] > 0.0f) {
        float A, B, C, D;
        v = sad_ssim_terms(sad_ssim_window(r, t, width, height, x, y), &A, &B, &C, &D);
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63u) == 0u) part[threadIdx.x >> 6] = v;
    __syncthreads();
    const float bsum = (part[0] + part[1]) + (part[2] + part[3]);
    if constexpr (std::is_same<T, long long>::value) {
        if (threadIdx.x < 2u) {
            long long lo, hi;
            fix_split(bsum, &lo, &hi);
            fix_atomic_add(acc + threadIdx.x, threadIdx.x ? hi : lo);
        }
    } else {
        if (threadIdx.x == 0) unsafeAtomicAdd(acc, (T)bsum);
    }
}

/* -------------------------------------------- quantised inference sites --
 * makePackedInferenceSites (packed_inference.swift:86-199) on the device: the
 * field ranges are min/max reductions over active sites, kept as order-
 * preserving integer images of the floats in the 10 words of the quant
 * struct (min keys 0-4, max keys 5-9) until one thread turns them into
 * (min, scale) pairs in place; then one thread per site packs. */
__device__ __forceinline__ uint32_t f2ord(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_quant_init(uint32_t* keys)
{
    if (threadIdx.x < 10u) keys[threadIdx.x] = threadIdx.x < 5u ? 0xFFFFFFFFu : 0u;
}

__global__ __launch_bounds__(256) void k_quant_range(const SadSite* __restrict__ sites, uint32_t n, uint32_t* keys)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    uint32_t lo[5], hi[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) { lo[k] = 0xFFFFFFFFu; hi[k] = 0u; }
    if (i < n) {
        const SadSite st = sites[i];
        if (st.pos_x >= 0.0f && st.pos_y >= 0.0f) {
            const float v[5] = {st.log_tau, st.radius, st.col_r, st.col_g, st.col_b};
#pragma unroll
            for (int k = 0; k < 5; ++k)
                if (isfinite(v[k])) { lo[k] = f2ord(v[k]); hi[k] = lo[k]; }
        }
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = min(lo[k], (uint32_t)__shfl_xor((int)lo[k], off));
            hi[k] = max(hi[k], (uint32_t)__shfl_xor((int)hi[k], off));
        }
    }
    if ((threadIdx.x & 63u) == 0u) {
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (lo[k] != 0xFFFFFFFFu) atomicMin(keys + k, lo[k]);
            if (hi[k] != 0u) atomicMax(keys + 5 + k, hi[k]);
        }
    }
}

__global__ void k_quant_finish(SadInferenceQuant* q)
{
    if (threadIdx.x != 0u) return;
    uint32_t* keys = reinterpret_cast<uint32_t*>(q);
    float mn[5], mx[5];
    for (int k = 0; k < 5; ++k) {
        mn[k] = keys[k] == 0xFFFFFFFFu ? INFINITY : ord2f(keys[k]);
        mx[k] = keys[5 + k] == 0u ? -INFINITY : ord2f(keys[5 + k]);
    }
    sad_quant_finish(mn, mx, q);
}

__global__ __launch_bounds__(256) void k_pack_inference(const SadSite* __restrict__ sites, uint32_t n,
                                                        const SadInferenceQuant* __restrict__ q, float maxx, float maxy,
                                                        SadInferenceSite* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] = sad_pack_inference(&sites[i], q, maxx, maxy);
}

__global__ __launch_bounds__(256) void k_unpack_inference(const SadInferenceSite* __restrict__ in, uint32_t n,
                                                          const SadInferenceQuant* __restrict__ q, float maxx,
                                                          float maxy, SadSite* __restrict__ out)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) out[i] = sad_unpack_inference(in[i], q, maxx, maxy);
}

/* countActiveSitesKernel, utils.cu:90-101 */
__global__ __launch_bounds__(256) void k_count_active(const SadSite* __restrict__ sites, uint32_t* count, uint32_t n)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    bool a = i < n && sites[i].pos_x >= 0.0f;
    unsigned long long b = __ballot(a);
    if ((threadIdx.x & 63u) == 0u && b) atomicAdd(count, (uint32_t)__popcll(b));
}

/* ----------------------------------------------- gradient-weighted init --
 * initGradientWeightedKernel, densify_prune.cu:5-83.  There, every site thread
 * draws seeds from one atomic counter until one is accepted or max_attempts
 * are spent, so which site gets which seed depends on thread timing.  Here the
 * draws are serialised in site order: site i draws from the seed after site
 * i-1's last draw.  When no site runs out of attempts the set of sites is the
 * same as the reference's (the first n accepted seeds; every seed the counter
 * handed out was drawn by some site); the order is seed order.
 *
 * Serial walk: from q (one past the previous accepted seed), the next accepted
 * seed a is preceded by (a - q) / M failed sites of M draws each, then taken by
 * the next site; q = a + 1.  One workgroup evaluates IG_ROUND consecutive
 * seeds per round (IG_PER per thread, a bit each), and two block scans place
 * each thread's accepted seeds: an exclusive max-scan gives the end of the
 * previous accepted seed, an exclusive sum-scan the site index.  Failed sites
 * are all equal and are written up front.  The walk ends when n sites are
 * placed, or when the open gap after q holds the remaining sites' failures. */
#define IG_THREADS 1024
#define IG_PER 32
#define IG_ROUND (IG_THREADS * IG_PER)

struct InitGradArgs {
    SadSite* sites;
    uint32_t* seed_counter;    /* in: first seed; out: one past the last seed drawn (null: from 0) */
    const float* target;
    const float* mask;         /* null: all ones */
    float thr, log_tau0, radius0;
    uint32_t n, attempts;
    int w, h;
};

/* exclusive scan over the workgroup (sum, or max of non-negative values) and the block total */
template <bool MAX>
__device__ __forceinline__ uint32_t ig_scan(uint32_t v, uint32_t* lds, uint32_t* total)
{
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (uint32_t o = 1u; o < 64u; o <<= 1) {
        uint32_t t = __shfl_up(x, o, 64);
        if (lane >= o) x = MAX ? (t > x ? t : x) : x + t;
    }
    uint32_t xe = __shfl_up(x, 1u, 64);
    if (lane == 0u) xe = 0u;
    if (lane == 63u) lds[wid] = x;
    __syncthreads();
    uint32_t pre = 0u, tot = 0u;
    for (uint32_t k = 0u; k < IG_THREADS / 64u; ++k) {
        uint32_t y = lds[k];
        if (k < wid) pre = MAX ? (y > pre ? y : pre) : pre + y;
        tot = MAX ? (y > tot ? y : tot) : tot + y;
    }
    __syncthreads();
    *total = tot;
    return MAX ? (xe > pre ? xe : pre) : pre + xe;
}

__global__ __launch_bounds__(IG_THREADS) void k_init_gradient(InitGradArgs A)
{
    __shared__ uint32_t lds[IG_THREADS / 64u];
    const uint32_t tid = threadIdx.x;
    const SadSite failed = sad_init_grad_site(A.target, 0.0f, 0.0f, -1, 0.0f, A.log_tau0, A.radius0);
    for (uint32_t i = tid; i < A.n; i += IG_THREADS) A.sites[i] = failed;
    if (A.attempts == 0u) return;       /* no draws: every site fails, the counter is unchanged */
    __syncthreads();

    const uint64_t M = A.attempts;
    uint64_t p = A.seed_counter ? *A.seed_counter : 0u;   /* first seed of this round */
    uint64_t q = p;                                       /* one past the last accepted seed */
    uint32_t E = 0u;                                      /* sites placed */
    for (;;) {
        const uint64_t s0 = p + (uint64_t)tid * IG_PER;
        uint32_t bits = 0u;
        for (uint32_t j = 0u; j < IG_PER; ++j) {
            float px, py, g;
            int pix;
            if (sad_init_grad_eval(A.target, A.mask, A.w, A.h, (uint32_t)(s0 + j), A.thr, &px, &py, &pix, &g))
                bits |= 1u << j;
        }
        const uint32_t c = (uint32_t)__popc(bits);
        const uint32_t first = tid * IG_PER + (c ? (uint32_t)__ffs(bits) - 1u : 0u);
        const uint32_t end = c ? tid * IG_PER + 32u - (uint32_t)__clz(bits) : 0u;   /* one past the last */
        uint32_t max_end, placed;
        const uint32_t prev = ig_scan<true>(end, lds, &max_end);
        const uint64_t prev_end = prev ? p + prev : q;
        const uint32_t f = c ? (uint32_t)((p + first - prev_end) / M) : 0u;
        const uint32_t base = ig_scan<false>(c + f, lds, &placed);
        if (c) {
            uint32_t i = E + base;
            if (i < A.n && A.n - i <= f && A.seed_counter)   /* the last site is one of these failures */
                *A.seed_counter = (uint32_t)(prev_end + (uint64_t)(A.n - i) * M);
            i += f;
            uint32_t b = bits;
            while (b && i < A.n) {
                const uint32_t j = (uint32_t)__ffs(b) - 1u;
                b &= b - 1u;
                const uint64_t s = s0 + j;
                float px, py, g;
                int pix;
                sad_init_grad_eval(A.target, A.mask, A.w, A.h, (uint32_t)s, A.thr, &px, &py, &pix, &g);
                A.sites[i] = sad_init_grad_site(A.target, px, py, pix, g, A.log_tau0, A.radius0);
                if (i == A.n - 1u && A.seed_counter) *A.seed_counter = (uint32_t)(s + 1u);
                ++i;
            }
        }
        E += placed;
        if (max_end) q = p + max_end;
        p += IG_ROUND;
        if (E >= A.n) return;
        if ((p - q) / M >= (uint64_t)(A.n - E)) {     /* the open gap fails every remaining site */
            if (tid == 0u && A.seed_counter) *A.seed_counter = (uint32_t)(q + (uint64_t)(A.n - E) * M);
            return;
        }
    }
}

/* ------------------------------------------------------------ launchers -- */

static void fill_adam(AdamArgs& A, SadSite* sites, SadAdamState* st, float* const g[10], float lr_pos, float lr_tau,
                      float lr_radius, float lr_color, float lr_dir, float lr_aniso, float beta1, float beta2, float eps,
                      uint32_t t, uint32_t n, int width, int height)
{
    A.sites = sites; A.st = st;
    for (int j = 0; j < 10; ++j) A.g[j] = g[j];
    A.gstride = 1u;
    A.gfix = nullptr;
    A.lr_pos = lr_pos; A.lr_tau = lr_tau; A.lr_radius = lr_radius; A.lr_color = lr_color;
    A.lr_dir = lr_dir; A.lr_aniso = lr_aniso; A.beta1 = beta1; A.beta2 = beta2; A.eps = eps;
    A.b1c = 1.0f / (1.0f - sad_powi_f(beta1, t));
    A.b2c = 1.0f / (1.0f - sad_powi_f(beta2, t));
    A.norm = 1.0f / (float)(width * height);
    A.maxx = (float)(width - 1);
    A.maxy = (float)(height - 1);
    A.n = n;
}

int launch_adam(SadSite* sites, SadAdamState* st, float* const g[10], float lr_pos, float lr_tau, float lr_radius,
                float lr_color, float lr_dir, float lr_aniso, float beta1, float beta2, float eps, uint32_t t,
                uint32_t n, int width, int height, bool clear, hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    AdamArgs A;
    fill_adam(A, sites, st, g, lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso, beta1, beta2, eps, t, n, width,
              height);
    if (clear) hipLaunchKernelGGL(k_adam<true>, dim3(div_up(n, 256u)), dim3(256), 0, stream, A);
    else hipLaunchKernelGGL(k_adam<false>, dim3(div_up(n, 256u)), dim3(256), 0, stream, A);
    return launch_status("adam_update");
}

int launch_tau_diffuse(const uint32_t* c0, const uint32_t* c1, const SadSite* sites, const float* graw,
                       const float* gin, float* gout, float* raw_copy, uint32_t n, float lambda, int width,
                       int height, int ds, hipStream_t stream);

/* The trainer's tail: tau passes 1-3 (g2 -> tmp -> g2 -> tmp, raw saved by the
 * first) and pass 4 fused with Adam + clear. */
/* SoA gradients; nbr: 8 uint32 per site of neighbour-list scratch */
int launch_tau_adam(const uint32_t* c0, const uint32_t* c1, float* raw, float* tmp, uint32_t* nbr_lists, float lambda,
                    int cand_width, int cand_height, int ds, SadSite* sites, SadAdamState* st, float* const g[10],
                    float lr_pos, float lr_tau, float lr_radius, float lr_color, float lr_dir, float lr_aniso,
                    float beta1, float beta2, float eps, uint32_t t, uint32_t n, int width, int height,
                    hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    const dim3 grid(div_up(n, 256u));
    const uint4* nbr = reinterpret_cast<const uint4*>(nbr_lists);
    hipLaunchKernelGGL(k_tau_list_first<const float*>, grid, dim3(256), 0, stream, reinterpret_cast<const uint4*>(c0),
                       reinterpret_cast<const uint4*>(c1), sites, g[2], tmp, raw, n, lambda, cand_width, cand_height,
                       1.0f / fmaxf((float)ds, 1.0f), reinterpret_cast<uint4*>(nbr_lists), 1u);
    hipLaunchKernelGGL(k_tau_list_pass, grid, dim3(256), 0, stream, nbr, (const float*)raw, (const float*)tmp, g[2],
                       n, lambda);
    hipLaunchKernelGGL(k_tau_list_pass, grid, dim3(256), 0, stream, nbr, (const float*)raw, (const float*)g[2], tmp,
                       n, lambda);
    AdamArgs A;
    fill_adam(A, sites, st, g, lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso, beta1, beta2, eps, t, n, width,
              height);
    hipLaunchKernelGGL(k_tau_list_last_adam, grid, dim3(256), 0, stream, nbr, raw, tmp, lambda, A);
    return launch_status("tau_diffuse_adam");
}

/* block: threads per block, 64 or 256 (the same results) */
int launch_tau_adam_batch(const BatchJob* jobs, uint32_t njobs, const JobStep* steps, uint32_t total_sites,
                          const BatchStep& st, hipStream_t stream, uint32_t block)
{
    if (total_sites == 0) return SADH_OK;
    if (block != 64u) block = 256u;
    for (uint32_t pass = 0; pass < 4u; ++pass)
        hipLaunchKernelGGL(k_tau_batch, dim3(div_up(total_sites, block)), dim3(block), 0, stream, jobs, njobs, pass,
                           steps, st);
    return launch_status("tau_diffuse_adam_batch");
}

/* job_index[k] gets n[k] (0xFFFFFFFF keeps it) and flips its field when flip[k] */
int launch_update_jobs(BatchJob* jobs, const uint32_t* job_index, const uint32_t* n, const uint32_t* flip, uint32_t k,
                       hipStream_t stream)
{
    for (uint32_t b = 0; b < k; b += 64u) {
        JobUpdate u = {};
        u.count = std::min(64u, k - b);
        for (uint32_t t = 0; t < u.count; ++t) {
            u.job[t] = job_index[b + t];
            u.n[t] = n[b + t];
            u.flip[t] = flip[b + t];
        }
        hipLaunchKernelGGL(k_update_jobs, dim3(1), dim3(64), 0, stream, jobs, u);
    }
    return launch_status("batch_update_jobs");
}

/* The list-driven tail on fixed-point gradient lines: pass 1 reads the
 * log_tau column (raw saved).  With lists_built (written during the gradient
 * launch) passes 1+2 are one launch into tmp2; otherwise pass 1 writes the
 * neighbour lists into tmp and pass 2 reads them into tmp2.  Passes 3+4 are one
 * launch that feeds Adam, which reads and clears each site's line. */
static void tau_adam_lines(const uint4* u0, const uint4* u1, long long* gacc, float* raw, float* tmp, float* tmp2,
                           float lambda, int cand_width, int cand_height, float inv_ds, SadSite* sites,
                           uint32_t* nbr_lists, bool lists_built, const AdamArgs& A, hipStream_t stream)
{
    /* 64-thread blocks: these passes are dependent gathers over a few 10k sites, and smaller blocks spread the
     * same waves over more CUs and their caches */
    constexpr uint32_t TB = 64u;
    const uint32_t n = A.n;
    const dim3 grid(div_up(n, TB));
    const dim3 grid8(div_up(n, TB / 8u));          /* 8 threads per site */
    uint4* nbr = reinterpret_cast<uint4*>(nbr_lists);
    const FixCol g2{gacc + 4, FIX_GRAD_STRIDE};
    if (lists_built) {
        /* passes 1+2, then 3+4 with Adam: two launches */
        hipLaunchKernelGGL(k_tau_list_12_built<FixCol>, grid8, dim3(TB), 0, stream, (const uint4*)nbr, g2, tmp2, raw,
                           n, lambda);
    } else {
        /* pass 1 writes the lists, so pass 2 cannot share its launch */
        hipLaunchKernelGGL(k_tau_list_first<FixCol>, grid, dim3(TB), 0, stream, u0, u1, (const SadSite*)sites, g2,
                           tmp, raw, n, lambda, cand_width, cand_height, inv_ds, nbr, 1u);
        hipLaunchKernelGGL(k_tau_list_pass, grid, dim3(TB), 0, stream, (const uint4*)nbr, (const float*)raw,
                           (const float*)tmp, tmp2, n, lambda);
    }
    hipLaunchKernelGGL(k_tau_list_34_adam, grid8, dim3(TB), 0, stream, (const uint4*)nbr, (const float*)raw,
                       (const float*)tmp2, lambda, A);
}

/* The same tail on the fit's fixed-point gradient lines, which always runs
 * from neighbour lists (nbr_lists: 8 uint32 per site). */
int launch_tau_adam_fix(const uint32_t* c0, const uint32_t* c1, long long* gacc, float* raw, float* tmp, float* tmp2,
                            float lambda, int cand_width, int cand_height, int ds, SadSite* sites, SadAdamState* st,
                            float lr_pos, float lr_tau, float lr_radius, float lr_color, float lr_dir, float lr_aniso,
                            float beta1, float beta2, float eps, uint32_t t, uint32_t n, int width, int height,
                            hipStream_t stream, uint32_t* nbr_lists, bool lists_built)
{
    if (n == 0) return SADH_OK;
    if (!nbr_lists) return fail(SADH_ERR_INVALID, "tau_diffuse_adam(fix): neighbour lists are required");
    float* g[10] = {};
    AdamArgs A;
    fill_adam(A, sites, st, g, lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso, beta1, beta2, eps, t, n, width,
              height);
    A.gfix = gacc;
    tau_adam_lines(reinterpret_cast<const uint4*>(c0), reinterpret_cast<const uint4*>(c1), gacc, raw, tmp, tmp2,
                   lambda, cand_width, cand_height, 1.0f / fmaxf((float)ds, 1.0f), sites, nbr_lists, lists_built, A,
                   stream);
    return launch_status("tau_diffuse_adam(fix)");
}

int launch_clear(float* const* arrays, int count_arrays, uint32_t n, hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    ClearArgs A;
    for (int j = 0; j < count_arrays; ++j) A.g[j] = arrays[j];
    A.count_arrays = count_arrays;
    A.n = n;
    hipLaunchKernelGGL(k_clear, dim3(div_up(n, 256u)), dim3(256), 0, stream, A);
    return launch_status("clear");
}

/* zeroes `words` int64 words, as 2 * words floats (the same all-zero bits); a kernel, not a memset, so the
 * clear can sit in a captured graph */
int launch_clear_fix(long long* p, size_t words, hipStream_t stream)
{
    if (words == 0) return SADH_OK;
    if (2u * words > 0xFFFFFFFFu) return fail(SADH_ERR_INVALID, "clear_fix: buffer too large");
    float* f[1] = {reinterpret_cast<float*>(p)};
    return launch_clear(f, 1, (uint32_t)(2u * words), stream);
}

int launch_tau_diffuse(const uint32_t* c0, const uint32_t* c1, const SadSite* sites, const float* graw,
                       const float* gin, float* gout, float* raw_copy, uint32_t n, float lambda, int width,
                       int height, int ds, hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    float inv_ds = 1.0f / fmaxf((float)ds, 1.0f);
    hipLaunchKernelGGL(k_tau_diffuse, dim3(div_up(n, 256u)), dim3(256), 0, stream,
                       reinterpret_cast<const uint4*>(c0), reinterpret_cast<const uint4*>(c1), sites, graw, gin,
                       gout, raw_copy, n, lambda, width, height, inv_ds);
    return launch_status("tau_diffuse");
}

template <typename T>
static int split_impl(SadSite* sites, SadAdamState* st, const uint32_t* idx, uint32_t k, const T stats[8],
                      uint32_t cur, const float* target, int width, int height, hipStream_t stream)
{
    if (k == 0) return SADH_OK;
    SplitArgs<T> A;
    A.sites = sites; A.st = st; A.idx = idx;
    A.mass = stats[0]; A.ew = stats[2]; A.ewx = stats[3]; A.ewy = stats[4];
    A.ewxx = stats[5]; A.ewxy = stats[6]; A.ewyy = stats[7];
    A.target = target; A.k = k; A.cur = cur; A.width = width; A.height = height;
    hipLaunchKernelGGL(k_split<T>, dim3(div_up(k, 256u)), dim3(256), 0, stream, A);
    return launch_status("split_sites");
}

int launch_split(SadSite* sites, SadAdamState* st, const uint32_t* idx, uint32_t k, const float* const stats[8],
                 uint32_t cur, const float* target, int width, int height, hipStream_t stream)
{
    const float* s[8];
    for (int j = 0; j < 8; ++j) s[j] = stats[j];
    return split_impl(sites, st, idx, k, s, cur, target, width, height, stream);
}

/* statistics as fixed-point lines (FIX_STAT_STRIDE words per site) */
static void stat_cols(const long long* stats, FixCol cols[8])
{
    for (int j = 0; j < 8; ++j) cols[j] = FixCol{stats + 2 * j, FIX_STAT_STRIDE};
}

int launch_split_fix(SadSite* sites, SadAdamState* st, const uint32_t* idx, uint32_t k, const long long* stats,
                     uint32_t cur, const float* target, int width, int height, hipStream_t stream)
{
    FixCol c[8];
    stat_cols(stats, c);
    return split_impl(sites, st, idx, k, c, cur, target, width, height, stream);
}

__global__ void k_fix_store(const long long* __restrict__ w, float* __restrict__ out)
{
    if (threadIdx.x == 0) *out = fix_value(w[0], w[1]);
}

/* launchComputePSNR's float accumulator (utils.cu:258-276): the block sums meet in a fixed-point pair of a
 * per-stream workspace, then the float is stored, so the value does not depend on the blocks' order */
int launch_sse(const float* rendered, const float* target, const float* mask, float* acc, int width, int height,
               hipStream_t stream)
{
    void* w = nullptr;
    SAD_TRY(stream_workspace(stream, 2 * sizeof(long long), &w));
    SAD_TRY(launch_sse_fix(rendered, target, mask, static_cast<long long*>(w), width, height, stream));
    hipLaunchKernelGGL(k_fix_store, dim3(1), dim3(64), 0, stream, static_cast<const long long*>(w), acc);
    return launch_status("compute_psnr");
}

int launch_sse_fix(const float* rendered, const float* target, const float* mask, long long* acc, int width,
                   int height, hipStream_t stream)
{
    uint32_t np = (uint32_t)width * (uint32_t)height;
    hipLaunchKernelGGL(k_zero_pair, dim3(1), dim3(64), 0, stream, acc);
    hipLaunchKernelGGL(k_sse_fix, dim3(div_up(np, 256u)), dim3(256), 0, stream, rendered, target, mask, acc, np);
    return launch_status("compute_psnr");
}

/* launchComputeSSIM's float accumulator (utils.cu:278-296), through a fixed-point pair as launch_sse */
int launch_ssim(const float* rendered, const float* target, const float* mask, float* acc, int width, int height,
                hipStream_t stream)
{
    void* w = nullptr;
    SAD_TRY(stream_workspace(stream, 2 * sizeof(long long), &w));
    SAD_TRY(launch_ssim_fix(rendered, target, mask, static_cast<long long*>(w), width, height, stream));
    hipLaunchKernelGGL(k_fix_store, dim3(1), dim3(64), 0, stream, static_cast<const long long*>(w), acc);
    return launch_status("compute_ssim");
}

int launch_ssim_fix(const float* rendered, const float* target, const float* mask, long long* acc, int width,
                    int height, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_pair, dim3(1), dim3(64), 0, stream, acc);
    const uint32_t tx = div_up((uint32_t)width, 16u), ty = div_up((uint32_t)height, 16u);
    hipLaunchKernelGGL(k_ssim<long long>, dim3(tx * ty), dim3(256), 0, stream, rendered, target, mask, acc, width,
                       height, tx);
    return launch_status("compute_ssim");
}

int launch_pack_inference(const SadSite* sites, uint32_t n, int width, int height, SadInferenceSite* out,
                          SadInferenceQuant* quant, hipStream_t stream)
{
    uint32_t* keys = reinterpret_cast<uint32_t*>(quant);
    hipLaunchKernelGGL(k_quant_init, dim3(1), dim3(64), 0, stream, keys);
    if (n > 0) hipLaunchKernelGGL(k_quant_range, dim3(div_up(n, 256u)), dim3(256), 0, stream, sites, n, keys);
    hipLaunchKernelGGL(k_quant_finish, dim3(1), dim3(64), 0, stream, quant);
    const float maxx = (float)std::max(width - 1, 0), maxy = (float)std::max(height - 1, 0);
    if (n > 0)
        hipLaunchKernelGGL(k_pack_inference, dim3(div_up(n, 256u)), dim3(256), 0, stream, sites, n, quant, maxx, maxy,
                           out);
    return launch_status("pack_inference_sites");
}

int launch_unpack_inference(const SadInferenceSite* in, const SadInferenceQuant* quant, uint32_t n, int width,
                            int height, SadSite* out, hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    const float maxx = (float)std::max(width - 1, 0), maxy = (float)std::max(height - 1, 0);
    hipLaunchKernelGGL(k_unpack_inference, dim3(div_up(n, 256u)), dim3(256), 0, stream, in, n, quant, maxx, maxy,
                       out);
    return launch_status("unpack_inference_sites");
}

int launch_log_sums_batch(const BatchJob* jobs, uint32_t njobs, const JobStep* steps, uint32_t total_pblocks,
                          uint32_t total_sblocks, long long* out, hipStream_t stream)
{
    SAD_TRY(launch_clear_fix(out, 4u * (size_t)njobs, stream));
    if (total_pblocks + total_sblocks == 0) return SADH_OK;
    hipLaunchKernelGGL(k_log_sums_batch, dim3(total_pblocks + total_sblocks), dim3(256), 0, stream, jobs, njobs, steps,
                       total_pblocks, out);
    return launch_status("log_sums_batch");
}

int launch_count_active(const SadSite* sites, uint32_t* count, uint32_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, stream, count);
    if (n > 0) hipLaunchKernelGGL(k_count_active, dim3(div_up(n, 256u)), dim3(256), 0, stream, sites, count, n);
    return launch_status("count_active_sites");
}

template <typename T>
static int densify_impl(const SadSite* sites, T mass, T energy, SadPair* pairs, uint32_t n,
                        float min_mass, float alpha, uint32_t count, hipStream_t stream)
{
    if (count == 0) return SADH_OK;
    hipLaunchKernelGGL(k_densify_pairs<T>, dim3(div_up(count, 256u)), dim3(256), 0, stream, sites, mass, energy,
                       pairs, n, min_mass, alpha, count);
    return launch_status("compute_densify_score_pairs");
}

int launch_densify_pairs(const SadSite* sites, const float* mass, const float* energy, SadPair* pairs, uint32_t n,
                         float min_mass, float alpha, uint32_t count, hipStream_t stream)
{
    return densify_impl(sites, mass, energy, pairs, n, min_mass, alpha, count, stream);
}

int launch_densify_pairs_fix(const SadSite* sites, const long long* stats, SadPair* pairs, uint32_t n, float min_mass,
                             float alpha, uint32_t count, hipStream_t stream)
{
    FixCol c[8];
    stat_cols(stats, c);
    return densify_impl(sites, c[0], c[1], pairs, n, min_mass, alpha, count, stream);
}

template <typename T>
static int prune_impl(const SadSite* sites, T rdelta, SadPair* pairs, uint32_t n, float delta_norm,
                      uint32_t count, hipStream_t stream)
{
    if (count == 0) return SADH_OK;
    hipLaunchKernelGGL(k_prune_pairs<T>, dim3(div_up(count, 256u)), dim3(256), 0, stream, sites, rdelta, pairs, n,
                       delta_norm, count);
    return launch_status("compute_prune_score_pairs");
}

int launch_prune_pairs(const SadSite* sites, const float* rdelta, SadPair* pairs, uint32_t n, float delta_norm,
                       uint32_t count, hipStream_t stream)
{
    return prune_impl(sites, rdelta, pairs, n, delta_norm, count, stream);
}

/* removal deltas as fixed-point (lo, hi) pairs, one per site */
int launch_prune_pairs_fix(const SadSite* sites, const long long* rdelta, SadPair* pairs, uint32_t n, float delta_norm,
                           uint32_t count, hipStream_t stream)
{
    return prune_impl(sites, FixCol{rdelta, 2u}, pairs, n, delta_norm, count, stream);
}

int launch_write_indices(const SadPair* sorted, uint32_t* idx, uint32_t count, hipStream_t stream)
{
    if (count == 0) return SADH_OK;
    hipLaunchKernelGGL(k_write_indices, dim3(div_up(count, 256u)), dim3(256), 0, stream, sorted, idx, count);
    return launch_status("write_split_indices");
}

int launch_prune_sites(SadSite* sites, const uint32_t* idx, uint32_t count, hipStream_t stream)
{
    if (count == 0) return SADH_OK;
    hipLaunchKernelGGL(k_prune_sites, dim3(div_up(count, 256u)), dim3(256), 0, stream, sites, idx, count);
    return launch_status("prune_sites_by_index");
}

int launch_hilbert_pairs(const SadSite* sites, SadPair* pairs, uint32_t n, uint32_t padded, int width, int height,
                         int bits, hipStream_t stream)
{
    if (padded == 0) return SADH_OK;
    hipLaunchKernelGGL(k_hilbert_pairs, dim3(div_up(padded, 256u)), dim3(256), 0, stream, sites, pairs, n, padded,
                       width, height, bits);
    return launch_status("hilbert_pairs");
}

int launch_hilbert_batch(const BatchJob* jobs, uint32_t njobs, const JobStep* steps, uint32_t total, uint32_t shift,
                         uint32_t max_key_exclusive, SadPair* pairs, SadPair* scratch, uint32_t* hist,
                         hipStream_t stream)
{
    if (total == 0) return SADH_OK;
    const uint32_t padded = div_up(total, 1024u) * 1024u;
    hipLaunchKernelGGL(k_hilbert_pairs_batch, dim3(padded / 256u), dim3(256), 0, stream, jobs, njobs, steps, pairs,
                       total, padded, shift);
    SAD_TRY(launch_status("hilbert_pairs_batch"));
    SAD_TRY(launch_radix_sort(pairs, scratch, hist, padded, max_key_exclusive, stream));
    hipLaunchKernelGGL(k_hilbert_order_batch, dim3(div_up(total, 256u)), dim3(256), 0, stream, jobs, steps, pairs,
                       total, shift);
    return launch_status("hilbert_order_batch");
}

int launch_gather_sites_batch(const BatchJob* jobs, uint32_t njobs, const JobStep* steps, uint32_t total, SadSite* out,
                              hipStream_t stream)
{
    if (total == 0) return SADH_OK;
    hipLaunchKernelGGL(k_gather_sites_batch, dim3(div_up(total, 256u)), dim3(256), 0, stream, jobs, njobs, steps, total,
                       out);
    return launch_status("gather_sites_batch");
}

int launch_hilbert_order(const SadPair* sorted, uint32_t* order, uint32_t* pos, uint32_t n, hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    hipLaunchKernelGGL(k_hilbert_order, dim3(div_up(n, 256u)), dim3(256), 0, stream, sorted, order, pos, n);
    return launch_status("hilbert_order");
}

int launch_init_gradient(SadSite* sites, uint32_t n, uint32_t* seed_counter, const float* target, const float* mask,
                         float thr, uint32_t attempts, float log_tau0, float radius0, int width, int height,
                         hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    InitGradArgs A;
    A.sites = sites; A.seed_counter = seed_counter; A.target = target; A.mask = mask;
    A.thr = thr; A.log_tau0 = log_tau0; A.radius0 = radius0;
    A.n = n; A.attempts = attempts; A.w = width; A.h = height;
    hipLaunchKernelGGL(k_init_gradient, dim3(1), dim3(IG_THREADS), 0, stream, A);
    return launch_status("init_gradient_weighted");
}

}  // namespace sadh

using namespace sadh;

extern "C" {

int sadh_adam_update(SadSite* sites, SadAdamState* adam, const float* grad_pos_x, const float* grad_pos_y,
                     const float* grad_log_tau, const float* grad_radius, const float* grad_color_r,
                     const float* grad_color_g, const float* grad_color_b, const float* grad_dir_x,
                     const float* grad_dir_y, const float* grad_log_aniso, float lr_pos, float lr_tau,
                     float lr_radius, float lr_color, float lr_dir, float lr_aniso, float beta1, float beta2,
                     float eps, uint32_t t, uint32_t site_count, int width, int height, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && adam && grad_pos_x && grad_pos_y && grad_log_tau && grad_radius && grad_color_r &&
                     grad_color_g && grad_color_b && grad_dir_x && grad_dir_y && grad_log_aniso,
                 "adam_update: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && t > 0, "adam_update: bad sizes or step");
    float* const g[10] = {(float*)grad_pos_x, (float*)grad_pos_y, (float*)grad_log_tau, (float*)grad_radius,
                          (float*)grad_color_r, (float*)grad_color_g, (float*)grad_color_b, (float*)grad_dir_x,
                          (float*)grad_dir_y, (float*)grad_log_aniso};
    return launch_adam(sites, adam, g, lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso, beta1, beta2, eps, t,
                       site_count, width, height, false, to_hip(stream));
}

int sadh_adam_update_clear(SadSite* sites, SadAdamState* adam, float* grad_pos_x, float* grad_pos_y,
                           float* grad_log_tau, float* grad_radius, float* grad_color_r, float* grad_color_g,
                           float* grad_color_b, float* grad_dir_x, float* grad_dir_y, float* grad_log_aniso,
                           float lr_pos, float lr_tau, float lr_radius, float lr_color, float lr_dir, float lr_aniso,
                           float beta1, float beta2, float eps, uint32_t t, uint32_t site_count, int width,
                           int height, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && adam && grad_pos_x && grad_pos_y && grad_log_tau && grad_radius && grad_color_r &&
                     grad_color_g && grad_color_b && grad_dir_x && grad_dir_y && grad_log_aniso,
                 "adam_update_clear: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && t > 0, "adam_update_clear: bad sizes or step");
    float* const g[10] = {grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r,
                          grad_color_g, grad_color_b, grad_dir_x, grad_dir_y, grad_log_aniso};
    return launch_adam(sites, adam, g, lr_pos, lr_tau, lr_radius, lr_color, lr_dir, lr_aniso, beta1, beta2, eps, t,
                       site_count, width, height, true, to_hip(stream));
}

int sadh_clear_gradients(float* grad_pos_x, float* grad_pos_y, float* grad_log_tau, float* grad_radius,
                         float* grad_color_r, float* grad_color_g, float* grad_color_b, float* grad_dir_x,
                         float* grad_dir_y, float* grad_log_aniso, uint32_t count, sadh_stream_t stream)
{
    float* const g[10] = {grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r,
                          grad_color_g, grad_color_b, grad_dir_x, grad_dir_y, grad_log_aniso};
    for (int j = 0; j < 10; ++j) SADH_REQUIRE(g[j], "clear_gradients: null pointer");
    return launch_clear(g, 10, count, to_hip(stream));
}

int sadh_clear_buffer(float* buffer, uint32_t count, sadh_stream_t stream)
{
    SADH_REQUIRE(buffer, "clear_buffer: null pointer");
    return launch_clear(&buffer, 1, count, to_hip(stream));
}

int sadh_tau_diffuse_adam(const uint32_t* cand0, const uint32_t* cand1, SadSite* sites, SadAdamState* adam,
                          float* grad_pos_x, float* grad_pos_y, float* grad_log_tau, float* grad_radius,
                          float* grad_color_r, float* grad_color_g, float* grad_color_b, float* grad_dir_x,
                          float* grad_dir_y, float* grad_log_aniso, float* tau_scratch, float lambda, int cand_width,
                          int cand_height, int cand_downscale, float lr_pos, float lr_tau, float lr_radius,
                          float lr_color, float lr_dir, float lr_aniso, float beta1, float beta2, float eps, uint32_t t,
                          uint32_t site_count, int width, int height, sadh_stream_t stream)
{
    float* const g[10] = {grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r,
                          grad_color_g, grad_color_b, grad_dir_x, grad_dir_y, grad_log_aniso};
    for (int j = 0; j < 10; ++j) SADH_REQUIRE(g[j], "tau_diffuse_adam: null gradient array");
    SADH_REQUIRE(cand0 && cand1 && sites && adam && tau_scratch, "tau_diffuse_adam: null pointer");
    SADH_REQUIRE(cand_width > 0 && cand_height > 0 && width > 0 && height > 0, "tau_diffuse_adam: bad sizes");
    /* tau_scratch: raw [n], tmp [n], neighbour lists [n][8] */
    return launch_tau_adam(cand0, cand1, tau_scratch, tau_scratch + site_count,
                           reinterpret_cast<uint32_t*>(tau_scratch + 2 * (size_t)site_count), lambda, cand_width,
                           cand_height, cand_downscale, sites, adam, g, lr_pos, lr_tau, lr_radius, lr_color, lr_dir,
                           lr_aniso, beta1, beta2, eps, t, site_count, width, height, to_hip(stream));
}

int sadh_tau_diffuse(const uint32_t* cand0, const uint32_t* cand1, const SadSite* sites, const float* grad_raw,
                     const float* grad_in, float* grad_out, uint32_t site_count, float lambda, int width, int height,
                     int cand_downscale, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && sites && grad_raw && grad_in && grad_out, "tau_diffuse: null pointer");
    SADH_REQUIRE(grad_in != grad_out, "tau_diffuse: grad_in and grad_out must differ");
    SADH_REQUIRE(width > 0 && height > 0, "tau_diffuse: bad sizes");
    return launch_tau_diffuse(cand0, cand1, sites, grad_raw, grad_in, grad_out, nullptr, site_count, lambda, width,
                              height, cand_downscale, to_hip(stream));
}

int sadh_compute_densify_score_pairs(const SadSite* sites, const float* mass, const float* energy, SadPair* pairs,
                                     uint32_t site_count, float min_mass, float score_alpha, uint32_t pair_count,
                                     sadh_stream_t stream)
{
    SADH_REQUIRE(sites && mass && energy && pairs, "compute_densify_score_pairs: null pointer");
    SADH_REQUIRE(site_count <= pair_count, "compute_densify_score_pairs: site_count > pair_count");
    return launch_densify_pairs(sites, mass, energy, pairs, site_count, min_mass, score_alpha, pair_count,
                                to_hip(stream));
}

int sadh_compute_prune_score_pairs(const SadSite* sites, const float* removal_delta, SadPair* pairs,
                                   uint32_t site_count, float delta_norm, uint32_t pair_count, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && removal_delta && pairs, "compute_prune_score_pairs: null pointer");
    SADH_REQUIRE(site_count <= pair_count, "compute_prune_score_pairs: site_count > pair_count");
    return launch_prune_pairs(sites, removal_delta, pairs, site_count, delta_norm, pair_count, to_hip(stream));
}

int sadh_write_split_indices(const SadPair* sorted_pairs, uint32_t* indices, uint32_t count, sadh_stream_t stream)
{
    SADH_REQUIRE(sorted_pairs && indices, "write_split_indices: null pointer");
    return launch_write_indices(sorted_pairs, indices, count, to_hip(stream));
}

int sadh_split_sites(SadSite* sites, SadAdamState* adam, const uint32_t* split_indices, uint32_t num_to_split,
                     const float* mass, const float* err_w, const float* err_wx, const float* err_wy,
                     const float* err_wxx, const float* err_wxy, const float* err_wyy, uint32_t current_site_count,
                     const float* target, int width, int height, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && adam && split_indices && mass && err_w && err_wx && err_wy && err_wxx && err_wxy &&
                     err_wyy && target,
                 "split_sites: null pointer");
    SADH_REQUIRE(width > 0 && height > 0, "split_sites: bad sizes");
    const float* const st[8] = {mass, nullptr, err_w, err_wx, err_wy, err_wxx, err_wxy, err_wyy};
    return launch_split(sites, adam, split_indices, num_to_split, st, current_site_count, target, width, height,
                        to_hip(stream));
}

int sadh_prune_sites_by_index(SadSite* sites, const uint32_t* indices, uint32_t count, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && indices, "prune_sites_by_index: null pointer");
    return launch_prune_sites(sites, indices, count, to_hip(stream));
}

int sadh_hilbert_pairs(const SadSite* sites, SadPair* pairs, uint32_t site_count, uint32_t padded_count, int width,
                       int height, int bits, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && pairs && width > 0 && height > 0, "hilbert_pairs: bad arguments");
    SADH_REQUIRE(site_count <= padded_count && bits >= 1 && bits <= 16, "hilbert_pairs: bad sizes");
    return launch_hilbert_pairs(sites, pairs, site_count, padded_count, width, height, bits, to_hip(stream));
}

int sadh_hilbert_order(const SadPair* sorted_pairs, uint32_t* order, uint32_t* pos, uint32_t site_count,
                       sadh_stream_t stream)
{
    SADH_REQUIRE(sorted_pairs && order && pos, "hilbert_order: null pointer");
    return launch_hilbert_order(sorted_pairs, order, pos, site_count, to_hip(stream));
}

int sadh_compute_psnr(const float* rendered, const float* target, const float* mask, float* mse_accum, int width,
                      int height, sadh_stream_t stream)
{
    SADH_REQUIRE(rendered && target && mask && mse_accum && width > 0 && height > 0, "compute_psnr: bad arguments");
    return launch_sse(rendered, target, mask, mse_accum, width, height, to_hip(stream));
}

int sadh_compute_ssim(const float* rendered, const float* target, const float* mask, float* ssim_accum, int width,
                      int height, sadh_stream_t stream)
{
    SADH_REQUIRE(rendered && target && mask && ssim_accum && width > 0 && height > 0, "compute_ssim: bad arguments");
    return launch_ssim(rendered, target, mask, ssim_accum, width, height, to_hip(stream));
}

int sadh_pack_inference_sites(const SadSite* sites, uint32_t site_count, int width, int height,
                              SadInferenceSite* packed, SadInferenceQuant* quant, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && packed && quant && width > 0 && height > 0, "pack_inference_sites: bad arguments");
    return launch_pack_inference(sites, site_count, width, height, packed, quant, to_hip(stream));
}

int sadh_unpack_inference_sites(const SadInferenceSite* packed, const SadInferenceQuant* quant, uint32_t site_count,
                                int width, int height, SadSite* sites, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && packed && quant && width > 0 && height > 0, "unpack_inference_sites: bad arguments");
    return launch_unpack_inference(packed, quant, site_count, width, height, sites, to_hip(stream));
}

int sadh_count_active_sites(const SadSite* sites, uint32_t* count, uint32_t site_count, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && count, "count_active_sites: null pointer");
    return launch_count_active(sites, count, site_count, to_hip(stream));
}

int sadh_init_gradient_weighted(SadSite* sites, uint32_t num_sites, uint32_t* seed_counter, const float* target,
                                const float* mask, float grad_threshold, uint32_t max_attempts, float init_log_tau,
                                float init_radius, int width, int height, sadh_stream_t stream)
{
    SADH_REQUIRE(sites && target && width > 0 && height > 0, "init_gradient_weighted: bad arguments");
    SADH_REQUIRE(num_sites <= 0x7FFFFFFFu, "init_gradient_weighted: too many sites");
    return launch_init_gradient(sites, num_sites, seed_counter, target, mask, grad_threshold, max_attempts,
                                init_log_tau, init_radius, width, height, to_hip(stream));
}

}  // extern "C"
