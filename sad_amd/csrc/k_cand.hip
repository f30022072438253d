/*
 * k_cand.hip - candidate-field kernels: init, clear, JFA seed/flood, fp16 site
 * pack and the VPT (jump-neighbour propagation) pass.
 *
 * The candidate lists these kernels produce are compared bit for bit with the
 * CPU oracle, so every key is computed from sad_math.h with the exact
 * primitives (own exp, IEEE division and sqrt, no contraction).
 *
 * Layout: one thread per candidate cell, 16x16 cells per 256-thread block (four
 * wave64 of 16x4 cells).  A cell's two planes are 16-byte rows, so a wave reads
 * four contiguous 256-byte segments per plane.  Site data for the key comes
 * either from the 16-byte fp16 pack (the reference's format, kept for the torch
 * op contract) or from a 32-byte "search site" with the per-site terms of the
 * key (tau, e^a, e^-a, r/S) already evaluated, which saves two exp, one division
 * and the fp16 decodes per candidate and gives identical bits.
 */
#include "fit_common.h"

namespace sadh {

/* ------------------------------------------------------ init / clear ---- */

/* initCandidatesKernel, render.cu:95-136 */
__global__ __launch_bounds__(256) void k_init_candidates(uint4* __restrict__ c0, uint4* __restrict__ c1,
                                                         uint32_t n, uint32_t seed, int per_pixel,
                                                         int width, int height)
{
    int x = blockIdx.x * 16 + (threadIdx.x & 15);
    int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= width || y >= height) return;
    uint32_t p = (uint32_t)(y * width + x);
    uint32_t idx[8];
    if (per_pixel && n == (uint32_t)(width * height)) {
        const int nb[7][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {1, -1}, {-1, 1}};
        idx[0] = p;
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            int nx = sad_clampi(x + nb[i][0], 0, width - 1);
            int ny = sad_clampi(y + nb[i][1], 0, height - 1);
            idx[i + 1] = (uint32_t)(ny * width + nx);
        }
    } else {
        uint32_t state = seed ^ ((uint32_t)x * 1973u) ^ ((uint32_t)y * 9277u);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            state = sad_xorshift32(state + (uint32_t)i);
            idx[i] = state % n;
        }
    }
    c0[p] = make_uint4(idx[0], idx[1], idx[2], idx[3]);
    c1[p] = make_uint4(idx[4], idx[5], idx[6], idx[7]);
}

/* clearCandidatesKernel, render.cu:139-154 */
__global__ __launch_bounds__(256) void k_clear_candidates(uint4* __restrict__ c0, uint4* __restrict__ c1,
                                                          uint32_t cells)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= cells) return;
    uint4 e = make_uint4(SAD_EMPTY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY);
    c0[i] = e;
    c1[i] = e;
}

/* ------------------------------------------------------------ JFA seed --
 * jfaSeedKernel, jfa.cu:66-92, made deterministic: pass 1 sets slot .x of
 * every occupied home cell to 0, pass 2 takes the atomic max of the ids, so
 * the cell ends with the largest id of the sites it holds. */
__device__ __forceinline__ int seed_cell(const SadSite* s, int width, int height, float inv_ds)
{
    int x = sad_clampi((int)(s->pos_x * inv_ds), 0, width - 1);
    int y = sad_clampi((int)(s->pos_y * inv_ds), 0, height - 1);
    return y * width + x;
}

__global__ __launch_bounds__(256) void k_jfa_seed_reset(uint32_t* __restrict__ c0, const SadSite* __restrict__ sites,
                                                        uint32_t n, int width, int height, float inv_ds)
{
    uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= n) return;
    SadSite s = sites[g];
    if (s.pos_x < 0.0f) return;
    c0[4 * seed_cell(&s, width, height, inv_ds)] = 0u;
}

__global__ __launch_bounds__(256) void k_jfa_seed_max(uint32_t* __restrict__ c0, const SadSite* __restrict__ sites,
                                                      uint32_t n, int width, int height, float inv_ds)
{
    uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= n) return;
    SadSite s = sites[g];
    if (s.pos_x < 0.0f) return;
    atomicMax(&c0[4 * seed_cell(&s, width, height, inv_ds)], g);
}

/* ------------------------------------------------------- site packing --- */

/* packCandidateSitesKernel, candidate_pack.cu:3-27 */
__global__ __launch_bounds__(256) void k_pack_sites(const SadSite* __restrict__ sites,
                                                    SadPackedSite* __restrict__ packed, uint32_t n)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    SadSite s = sites[i];
    SadPackedSite p = sad_pack_site(&s);
    uint4 v;
    v.x = (uint32_t)p.h[0] | ((uint32_t)p.h[1] << 16);
    v.y = (uint32_t)p.h[2] | ((uint32_t)p.h[3] << 16);
    v.z = (uint32_t)p.h[4] | ((uint32_t)p.h[5] << 16);
    v.w = (uint32_t)p.h[6] | ((uint32_t)p.h[7] << 16);
    reinterpret_cast<uint4*>(packed)[i] = v;
}

/* Search sites from the fp16 pack's values (VPT keys): exactly the inputs
 * candidatePackedDistance decodes (jfa.cu:155-172). */
__global__ __launch_bounds__(256) void k_search_from_sites_f16(const SadSite* __restrict__ sites,
                                                               SadSearchSite* __restrict__ out,
                                                               uint32_t n, float inv_scale)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    SadSite s = sites[i];
    SadPackedSite p = sad_pack_site(&s);
    SadSearchSite q = sad_search_site(h2f(p.h[0]), h2f(p.h[1]), h2f(p.h[2]), h2f(p.h[3]),
                                      h2f(p.h[4]), h2f(p.h[5]), h2f(p.h[6]), inv_scale);
    float4* o = reinterpret_cast<float4*>(out + i);
    o[0] = make_float4(q.px, q.py, q.dx, q.dy);
    o[1] = make_float4(q.l1, q.l2, q.rn, q.tau);
}

/* Search sites from full-precision sites (JFA flood keys, jfa.cu:33-35). */
__global__ __launch_bounds__(256) void k_search_from_sites_f32(const SadSite* __restrict__ sites,
                                                               SadSearchSite* __restrict__ out,
                                                               uint32_t n, float inv_scale)
{
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    SadSite s = sites[i];
    SadSearchSite q = sad_search_site(s.pos_x, s.pos_y, s.log_tau, s.radius, s.dir_x, s.dir_y,
                                      s.log_aniso, inv_scale);
    if (s.pos_x < 0.0f) q.px = -1.0f;
    float4* o = reinterpret_cast<float4*>(out + i);
    o[0] = make_float4(q.px, q.py, q.dx, q.dy);
    o[1] = make_float4(q.l1, q.l2, q.rn, q.tau);
}

/* ------------------------------------------------------- key fetchers --- */

struct PackedFetch {
    const uint4* ps;
    float inv_scale;
    __device__ __forceinline__ float key(uint32_t id, float ux, float uy, float iss) const
    {
        uint4 v = ps[id];
        float px = h2f(v.x & 0xFFFFu);
        if (px < 0.0f) return INFINITY;
        SadSearchSite s = sad_search_site(px, h2f(v.x >> 16), h2f(v.y & 0xFFFFu), h2f(v.y >> 16),
                                          h2f(v.z & 0xFFFFu), h2f(v.z >> 16), h2f(v.w & 0xFFFFu),
                                          inv_scale);
        return s.tau * sad_dmix2_pre(s.px, s.py, s.dx, s.dy, s.l1, s.l2, s.rn, ux, uy, iss);
    }
};

struct SearchFetch {
    const float4* ss;
    __device__ __forceinline__ float key(uint32_t id, float ux, float uy, float iss) const
    {
        float4 a = ss[2 * id];
        if (a.x < 0.0f) return INFINITY;
        float4 b = ss[2 * id + 1];
        return b.w * sad_dmix2_pre(a.x, a.y, a.z, a.w, b.x, b.y, b.z, ux, uy, iss);
    }
};

/* Sorted top-K insertion with strict '<' (jfa.cu:14-63 for K=4, 176-248 for
 * K=8): the new key lands before the first entry it is smaller than.  Written
 * as a branch-free shift over fully unrolled register arrays. */
template <int K>
__device__ __forceinline__ void insert_sorted(uint32_t (&bi)[K], float (&bd)[K], uint32_t c, float d)
{
    if (!(d < bd[K - 1])) return;
#pragma unroll
    for (int i = K - 1; i >= 1; --i) {
        bool shift = d < bd[i - 1];
        bool here = !shift && d < bd[i];
        float nd = shift ? bd[i - 1] : (here ? d : bd[i]);
        uint32_t ni = shift ? bi[i - 1] : (here ? c : bi[i]);
        bd[i] = nd;
        bi[i] = ni;
    }
    if (d < bd[0]) { bd[0] = d; bi[0] = c; }
}

template <int K>
__device__ __forceinline__ bool contains(const uint32_t (&bi)[K], uint32_t c)
{
    bool r = false;
#pragma unroll
    for (int i = 0; i < K; ++i) r |= (bi[i] == c);
    return r;
}

/* ------------------------------------------------------------ JFA flood -- */

/* jfaFloodKernel, jfa.cu:95-144 (insertClosest4 :14-63) */
template <bool SEARCH>
__global__ __launch_bounds__(256) void k_jfa_flood(const uint4* __restrict__ in0, uint4* __restrict__ out0,
                                                   const SadSite* __restrict__ sites,
                                                   const float4* __restrict__ search, uint32_t n,
                                                   int step, float iss, int width, int height,
                                                   int ds, int tw, int th)
{
    int x = blockIdx.x * 16 + (threadIdx.x & 15);
    int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= width || y >= height) return;
    float ux = fminf((float)x * (float)ds, (float)(tw - 1));
    float uy = fminf((float)y * (float)ds, (float)(th - 1));
    uint32_t bi[4] = {SAD_EMPTY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY};
    float bd[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    SearchFetch sf{search};
    for (int dy = -1; dy <= 1; ++dy) {
        int sy = sad_clampi(y + dy * step, 0, height - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            int sx = sad_clampi(x + dx * step, 0, width - 1);
            uint4 c = in0[sy * width + sx];
            uint32_t cs[4] = {c.x, c.y, c.z, c.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uint32_t id = cs[k];
                if (id >= n || contains<4>(bi, id)) continue;
                float d;
                if (SEARCH) {
                    d = sf.key(id, ux, uy, iss);         /* +inf for inactive: never inserted */
                } else {
                    SadSite s = sites[id];
                    if (s.pos_x < 0.0f) continue;
                    d = sad_site_key(&s, ux, uy, iss);
                }
                insert_sorted<4>(bi, bd, id, d);
            }
        }
    }
    out0[y * width + x] = make_uint4(bi[0], bi[1], bi[2], bi[3]);
}

/* ------------------------------------------------------------- VPT pass --
 * updateCandidatesKernel, jfa.cu:271-386 */
/* One candidate cell (x, y) of a VPT pass. */
template <class Fetch>
__device__ __forceinline__ void vpt_cell(int x, int y, const uint4* __restrict__ in0, const uint4* __restrict__ in1,
                                         uint4* __restrict__ out0, uint4* __restrict__ out1,
                                         const Fetch& fetch, uint32_t n, uint32_t step, float iss,
                                         uint32_t step_high, float radius_scale,
                                         uint32_t radius_probes, uint32_t inject_count,
                                         const uint32_t* __restrict__ hil_order,
                                         const uint32_t* __restrict__ hil_pos,
                                         uint32_t hil_probes, uint32_t hil_window, int ds,
                                         int tw, int th, int width, int height)
{
    uint32_t step_index = step & 0xFFFFu;
    uint32_t jump = step >> 16;
    if (jump == 0u) jump = 1u;
    uint32_t full_index = (step_high << 16) | step_index;
    float ux = fminf((float)x * (float)ds, (float)(tw - 1));
    float uy = fminf((float)y * (float)ds, (float)(th - 1));

    uint32_t bi[8];
    float bd[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) { bi[i] = SAD_EMPTY; bd[i] = INFINITY; }

    auto try_insert = [&](uint32_t c) {
        if (c >= n || contains<8>(bi, c)) return;
        insert_sorted<8>(bi, bd, c, fetch.key(c, ux, uy, iss));
    };
    auto merge_cell = [&](int cell) {
        uint4 a = in0[cell];
        uint4 b = in1[cell];
        try_insert(a.x); try_insert(a.y); try_insert(a.z); try_insert(a.w);
        try_insert(b.x); try_insert(b.y); try_insert(b.z); try_insert(b.w);
    };

    merge_cell(y * width + x);
    /* The four neighbour cells' 32 ids, in the reference's order.  An id that is
     * in the list after the cell's own 8 can add nothing: it is either still in
     * the list when it comes up (skipped as already present) or was pushed out
     * by a smaller key, and then its key is not below the list's last.  So only
     * the other ids are tried, in order, each with the full check (an id that
     * occurs in two neighbour cells is then caught as present).  The result is
     * the reference's list, and a wave evaluates keys only as often as its
     * lanes have such ids, instead of 32 times. */
    const int off[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    int ncell[4];
    uint32_t todo = 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int nx = sad_clampi(x + off[j][0] * (int)jump, 0, width - 1);
        const int ny = sad_clampi(y + off[j][1] * (int)jump, 0, height - 1);
        ncell[j] = ny * width + nx;
        const uint4 a = in0[ncell[j]];
        const uint4 b = in1[ncell[j]];
        const uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (ids[k] < n && !contains<8>(bi, ids[k])) todo |= 1u << (8 * j + k);
    }
    while (todo != 0u) {
        const uint32_t t = (uint32_t)__builtin_ctz(todo);
        todo &= todo - 1u;
        const uint32_t j = t >> 3, k = t & 7u;
        const int cell = j == 0u ? ncell[0] : (j == 1u ? ncell[1] : (j == 2u ? ncell[2] : ncell[3]));
        const uint32_t* plane = reinterpret_cast<const uint32_t*>(k < 4u ? in0 : in1);
        try_insert(plane[4u * (uint32_t)cell + (k & 3u)]);
    }

    float rad = radius_scale * ((float)width / 1024.0f);
    uint32_t state = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((step_index + jump) * 83492791u);
    for (uint32_t r = 0; r < radius_probes; ++r) {
        state = sad_xorshift32(state);
        float a = ((float)state * (1.0f / 4294967296.0f)) * 6.2831853f;
        float sa, ca;
        sad_sincosf(a, &sa, &ca);
        int px = sad_clampi(x + (int)(ca * rad), 0, width - 1);
        int py = sad_clampi(y + (int)(sa * rad), 0, height - 1);
        if (px == x && py == y) continue;
        merge_cell(py * width + px);
    }

    if (hil_probes > 0u && hil_window > 0u) {
        uint32_t best = bi[0];
        if (best < n) {
            uint32_t pos = hil_pos[best];
            uint32_t span = hil_window * 2u + 1u;
            uint32_t hs = ((uint32_t)x * 2654435761u) ^ ((uint32_t)y * 1597334677u) ^ (full_index * 374761393u);
            for (uint32_t i = 0; i < hil_probes; ++i) {
                hs = sad_xorshift32(hs + i);
                int o = (int)(hs % span) - (int)hil_window;
                int idx = sad_clampi((int)pos + o, 0, (int)(n - 1u));
                try_insert(hil_order[idx]);
            }
        }
    }

    uint32_t is = ((uint32_t)x * 1664525u) ^ ((uint32_t)y * 1013904223u) ^ (full_index * 374761393u);
    for (uint32_t i = 0; i < inject_count; ++i) {
        is = sad_xorshift32(is + i);
        try_insert(is % n);
    }
    out0[y * width + x] = make_uint4(bi[0], bi[1], bi[2], bi[3]);
    out1[y * width + x] = make_uint4(bi[4], bi[5], bi[6], bi[7]);
}

template <class Fetch>
__global__ __launch_bounds__(256) void k_vpt(const uint4* __restrict__ in0, const uint4* __restrict__ in1,
                                             uint4* __restrict__ out0, uint4* __restrict__ out1,
                                             Fetch fetch, uint32_t n, uint32_t step, float iss,
                                             uint32_t step_high, float radius_scale,
                                             uint32_t radius_probes, uint32_t inject_count,
                                             const uint32_t* __restrict__ hil_order,
                                             const uint32_t* __restrict__ hil_pos,
                                             uint32_t hil_probes, uint32_t hil_window, int ds,
                                             int tw, int th, int width, int height)
{
    int x = blockIdx.x * 16 + (threadIdx.x & 15);
    int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= width || y >= height) return;
    vpt_cell(x, y, in0, in1, out0, out1, fetch, n, step, iss, step_high, radius_scale, radius_probes, inject_count,
             hil_order, hil_pos, hil_probes, hil_window, ds, tw, th, width, height);
}

/* Batched fit: one VPT pass over the concatenated candidate tiles of every job
 * (each job's current field -> its other field); the jump step comes from each
 * job's grid (packJumpStep, sad_cuda.cu:1157-1183). */
__global__ __launch_bounds__(256) void k_vpt_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                   const JobStep* __restrict__ steps, BatchStep st)
{
    const uint32_t f = blockIdx.x;
    const uint32_t j = batch_find(jobs, njobs, f, &BatchJob::ctile_begin);
    if (!steps[j].vpt) return;
    const uint32_t jump_index = steps[j].jump_index + st.vpt_pass;
    const BatchJob J = job_global(jobs, j);
    const uint32_t t = f - J.ctile_begin;
    const int x = (int)(t % J.ctiles_x) * 16 + (int)(threadIdx.x & 15u);
    const int y = (int)(t / J.ctiles_x) * 16 + (int)(threadIdx.x >> 4);
    if (x >= J.cw || y >= J.ch) return;
    const uint32_t src = J.cur ^ (st.vpt_pass & 1u);
    const uint4* in0 = reinterpret_cast<const uint4*>(src ? J.fb0 : J.fa0);
    const uint4* in1 = reinterpret_cast<const uint4*>(src ? J.fb1 : J.fa1);
    uint4* out0 = reinterpret_cast<uint4*>(src ? J.fa0 : J.fb0);
    uint4* out1 = reinterpret_cast<uint4*>(src ? J.fa1 : J.fb1);
    SearchFetch fetch{reinterpret_cast<const float4*>(J.search)};
    const bool hil = J.h_order != nullptr && st.hil_probes > 0u && st.hil_window > 0u;
    vpt_cell(x, y, in0, in1, out0, out1, fetch, J.n, sad_pack_jump_step(jump_index, J.cw, J.ch), J.iss,
             jump_index >> 16, st.radius_scale, st.radius_probes, st.inject_count, J.h_order, J.h_pos,
             hil ? st.hil_probes : 0u, hil ? st.hil_window : 0u, J.ds, J.width, J.height, J.cw, J.ch);
}

/* Batched fit: search sites (fp16-rounded key terms) of every job's sites. */
__global__ __launch_bounds__(256) void k_search_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                      const JobStep* __restrict__ steps)
{
    const uint32_t f = blockIdx.x * 256u + threadIdx.x;
    const uint32_t j = batch_find(jobs, njobs, f, &BatchJob::site_begin);
    if (!steps[j].vpt) return;
    const BatchJob J = job_global(jobs, j);
    const uint32_t i = f - J.site_begin;
    if (i >= J.n) return;
    SadSite s = J.sites[i];
    SadPackedSite p = sad_pack_site(&s);
    SadSearchSite q = sad_search_site(h2f(p.h[0]), h2f(p.h[1]), h2f(p.h[2]), h2f(p.h[3]),
                                      h2f(p.h[4]), h2f(p.h[5]), h2f(p.h[6]), J.inv_scale);
    float4* o = reinterpret_cast<float4*>(J.search + i);
    o[0] = make_float4(q.px, q.py, q.dx, q.dy);
    o[1] = make_float4(q.l1, q.l2, q.rn, q.tau);
}

/* ------------------------------------------------------------- launchers -- */

static inline dim3 grid16(int w, int h) { return dim3(div_up((uint32_t)w, 16u), div_up((uint32_t)h, 16u)); }

int launch_vpt_search(const uint32_t* in0, const uint32_t* in1, uint32_t* out0, uint32_t* out1,
                      const SadSearchSite* search, uint32_t n, uint32_t step, float iss,
                      uint32_t step_high, float radius_scale, uint32_t radius_probes,
                      uint32_t inject_count, const uint32_t* hil_order, const uint32_t* hil_pos,
                      uint32_t hil_probes, uint32_t hil_window, int ds, int tw, int th, int width,
                      int height, hipStream_t stream)
{
    SearchFetch f{reinterpret_cast<const float4*>(search)};
    hipLaunchKernelGGL(k_vpt<SearchFetch>, grid16(width, height), dim3(256), 0, stream,
                       reinterpret_cast<const uint4*>(in0), reinterpret_cast<const uint4*>(in1),
                       reinterpret_cast<uint4*>(out0), reinterpret_cast<uint4*>(out1), f, n, step, iss,
                       step_high, radius_scale, radius_probes, inject_count, hil_order, hil_pos,
                       hil_probes, hil_window, ds, tw, th, width, height);
    return launch_status("update_candidates(search)");
}

/* VPT pass `st.vpt_pass` of this tick's update for the jobs with steps[j].vpt
 * (the first pass also prepares their search sites) */
int launch_vpt_batch(const BatchJob* jobs, uint32_t njobs, const JobStep* steps, uint32_t total_ctiles,
                     uint32_t total_sites, const BatchStep& st, hipStream_t stream)
{
    if (total_sites > 0 && st.vpt_pass == 0u)
        hipLaunchKernelGGL(k_search_batch, dim3(div_up(total_sites, 256u)), dim3(256), 0, stream, jobs, njobs, steps);
    if (total_ctiles > 0)
        hipLaunchKernelGGL(k_vpt_batch, dim3(total_ctiles), dim3(256), 0, stream, jobs, njobs, steps, st);
    return launch_status("update_candidates_batch");
}

int launch_search_sites(const SadSite* sites, SadSearchSite* out, uint32_t n, float iss, bool from_f16,
                        hipStream_t stream)
{
    if (n == 0) return SADH_OK;
    float inv_scale = sqrtf(iss);
    if (from_f16)
        hipLaunchKernelGGL(k_search_from_sites_f16, dim3(div_up(n, 256u)), dim3(256), 0, stream, sites, out, n, inv_scale);
    else
        hipLaunchKernelGGL(k_search_from_sites_f32, dim3(div_up(n, 256u)), dim3(256), 0, stream, sites, out, n, inv_scale);
    return launch_status("search_sites");
}

int launch_jfa_flood_search(const uint32_t* in0, uint32_t* out0, const SadSearchSite* search, uint32_t n,
                            uint32_t step, float iss, int width, int height, int ds, int tw, int th,
                            hipStream_t stream)
{
    hipLaunchKernelGGL(k_jfa_flood<true>, grid16(width, height), dim3(256), 0, stream,
                       reinterpret_cast<const uint4*>(in0), reinterpret_cast<uint4*>(out0),
                       (const SadSite*)nullptr, reinterpret_cast<const float4*>(search), n, (int)step,
                       iss, width, height, ds, tw, th);
    return launch_status("jfa_flood(search)");
}

}  // namespace sadh

using namespace sadh;

extern "C" {

int sadh_init_candidates(uint32_t* cand0, uint32_t* cand1, uint32_t site_count, uint32_t seed,
                         int per_pixel_mode, int width, int height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && width > 0 && height > 0, "init_candidates: bad arguments");
    SADH_REQUIRE(site_count > 0, "init_candidates: site_count must be positive");
    hipLaunchKernelGGL(k_init_candidates, grid16(width, height), dim3(256), 0, to_hip(stream),
                       reinterpret_cast<uint4*>(cand0), reinterpret_cast<uint4*>(cand1), site_count, seed,
                       per_pixel_mode, width, height);
    return launch_status("init_candidates");
}

int sadh_clear_candidates(uint32_t* cand0, uint32_t* cand1, int width, int height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && width > 0 && height > 0, "clear_candidates: bad arguments");
    uint32_t cells = (uint32_t)width * (uint32_t)height;
    hipLaunchKernelGGL(k_clear_candidates, dim3(div_up(cells, 256u)), dim3(256), 0, to_hip(stream),
                       reinterpret_cast<uint4*>(cand0), reinterpret_cast<uint4*>(cand1), cells);
    return launch_status("clear_candidates");
}

int sadh_jfa_seed(uint32_t* cand0, const SadSite* sites, uint32_t site_count, int width, int height,
                  int cand_downscale, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && sites && width > 0 && height > 0, "jfa_seed: bad arguments");
    if (site_count == 0) return SADH_OK;
    float inv_ds = 1.0f / fmaxf((float)cand_downscale, 1.0f);
    dim3 g(div_up(site_count, 256u));
    hipLaunchKernelGGL(k_jfa_seed_reset, g, dim3(256), 0, to_hip(stream), cand0, sites, site_count, width, height, inv_ds);
    int st = launch_status("jfa_seed");
    if (st) return st;
    hipLaunchKernelGGL(k_jfa_seed_max, g, dim3(256), 0, to_hip(stream), cand0, sites, site_count, width, height, inv_ds);
    return launch_status("jfa_seed");
}

int sadh_jfa_flood(const uint32_t* in_cand0, uint32_t* out_cand0, const SadSite* sites,
                   uint32_t site_count, uint32_t step_size, float inv_scale_sq, int width, int height,
                   int cand_downscale, int target_width, int target_height, sadh_stream_t stream)
{
    SADH_REQUIRE(in_cand0 && out_cand0 && sites && width > 0 && height > 0, "jfa_flood: bad arguments");
    SADH_REQUIRE(in_cand0 != out_cand0, "jfa_flood: input and output must differ");
    /* the key's per-site terms (two exps, a reciprocal, a product) are formed once per site into the library's
     * stream workspace, not once per (cell, candidate); the keys are the same floats */
    if (site_count == 0) {
        hipLaunchKernelGGL(k_jfa_flood<false>, grid16(width, height), dim3(256), 0, to_hip(stream),
                           reinterpret_cast<const uint4*>(in_cand0), reinterpret_cast<uint4*>(out_cand0), sites,
                           (const float4*)nullptr, site_count, (int)step_size, inv_scale_sq, width, height,
                           cand_downscale, target_width, target_height);
        return launch_status("jfa_flood");
    }
    void* ws = nullptr;
    SAD_TRY(stream_workspace(to_hip(stream), (size_t)site_count * sizeof(SadSearchSite), &ws));
    SadSearchSite* search = static_cast<SadSearchSite*>(ws);
    SAD_TRY(launch_search_sites(sites, search, site_count, inv_scale_sq, false, to_hip(stream)));
    return launch_jfa_flood_search(in_cand0, out_cand0, search, site_count, step_size, inv_scale_sq, width, height,
                                   cand_downscale, target_width, target_height, to_hip(stream));
}

int sadh_pack_candidate_sites(const SadSite* sites, SadPackedSite* packed, uint32_t site_count,
                              sadh_stream_t stream)
{
    SADH_REQUIRE(sites && packed, "pack_candidate_sites: bad arguments");
    if (site_count == 0) return SADH_OK;
    hipLaunchKernelGGL(k_pack_sites, dim3(div_up(site_count, 256u)), dim3(256), 0, to_hip(stream), sites, packed,
                       site_count);
    return launch_status("pack_candidate_sites");
}

int sadh_update_candidates(const uint32_t* in_cand0, const uint32_t* in_cand1, uint32_t* out_cand0,
                           uint32_t* out_cand1, const SadPackedSite* sites, uint32_t site_count,
                           uint32_t step, float inv_scale_sq, uint32_t step_high, float radius_scale,
                           uint32_t radius_probes, uint32_t inject_count, const uint32_t* hilbert_order,
                           const uint32_t* hilbert_pos, uint32_t hilbert_probes, uint32_t hilbert_window,
                           int cand_downscale, int target_width, int target_height, int width, int height,
                           sadh_stream_t stream)
{
    SADH_REQUIRE(in_cand0 && in_cand1 && out_cand0 && out_cand1 && sites, "update_candidates: bad arguments");
    SADH_REQUIRE(width > 0 && height > 0 && site_count > 0, "update_candidates: bad sizes");
    SADH_REQUIRE(in_cand0 != out_cand0 && in_cand1 != out_cand1, "update_candidates: in-place update");
    if (hilbert_probes > 0 && hilbert_window > 0)
        SADH_REQUIRE(hilbert_order && hilbert_pos, "update_candidates: hilbert buffers missing");
    /* the keys are formed from the 16-byte packed sites at every use: at a million sites, 32-byte search-site
     * records formed once per site (as the fit uses for its ~20k sites) double the footprint of the random
     * injected-id gathers, and the pass measured slower with them (1856 against 1704 us at 4096^2) */
    PackedFetch f{reinterpret_cast<const uint4*>(sites), sqrtf(inv_scale_sq)};
    hipLaunchKernelGGL(k_vpt<PackedFetch>, grid16(width, height), dim3(256), 0, to_hip(stream),
                       reinterpret_cast<const uint4*>(in_cand0), reinterpret_cast<const uint4*>(in_cand1),
                       reinterpret_cast<uint4*>(out_cand0), reinterpret_cast<uint4*>(out_cand1), f, site_count,
                       step, inv_scale_sq, step_high, radius_scale, radius_probes, inject_count, hilbert_order,
                       hilbert_pos, hilbert_probes, hilbert_window, cand_downscale, target_width,
                       target_height, width, height);
    return launch_status("update_candidates");
}

}  // extern "C"
