/*
 * k_pixel.hip - per-pixel kernels: softmax render forward, tiled gradient
 * backward (forward recomputed), autograd backward into [N][10], and the
 * densify site statistics.
 *
 * One 256-thread workgroup per 16x16-pixel tile; consecutive tiles are mapped
 * onto one XCD so neighbouring tiles, which share most candidate sites, share
 * an L2.  Stages:
 *   1. each thread (one pixel) loads its 8 candidate ids and inserts them into
 *      an LDS hash of H slots (multiplicative hash, linear probing, CAS only on
 *      an empty slot).  Ids that find no slot within 8 probes are "pending".
 *   2. one thread per occupied slot loads the 40-byte site once and stores the
 *      per-site terms of the forward (tau, e^a, 1/e^a, r/S, colour) in LDS; the
 *      occupied slots are compacted into a dense list.
 *   3. each thread computes its pixel's softmax over the 8 candidates from the
 *      staged terms (pending ids read their site directly).  The render writes
 *      the colour here.  The backward kernels store the weights and the pixel's
 *      (pred, dL/dpred) in LDS.
 *   Then, in the backward modes, rounds of:
 *   4. a rank per (slot, pixel, candidate) from an integer LDS atomic on a
 *      per-wave counter, and a scan of the per-slot, per-wave counts, give each
 *      slot an entry list (pixel, candidate rank) in the order (wave, rank,
 *      lane); every slot's list is cut into chunks of CHUNK entries.
 *   5. chunk-parallel accumulation: CLANES (4) lanes per chunk differentiate
 *      its (pixel, candidate) pairs and sum in registers; a DPP reduction over
 *      the lane group leaves one partial per chunk in LDS; the partials of each
 *      slot are then summed in chunk order (one thread per site and value) and
 *      issued as one global atomic per (site, value) and tile: a float add for
 *      the ABI's float outputs, two int64 adds (fix_split) for the fit.
 *   6. if some ids were pending, the hash is cleared, the pending ids are
 *      inserted again and stages 2, 4 and 5 run for them; until none is left.
 *   Every (tile, site) sum therefore runs over the same entries in the same
 *   order whatever the hash placed where: the list order follows (wave, rank,
 *   lane) because the lanes of one LDS atomic instruction are served in lane
 *   order and a wave's atomics in program order, and the fit's cross-tile sums
 *   are integer.  A fit repeats bit for bit, and the CPU oracle, which forms
 *   the same per-tile sums, reproduces it.  Chunking balances the work: the
 *   entry counts of a tile's sites range from a few to all 256 pixels, and
 *   giving whole slots to lane groups left most lanes idle behind the largest.
 * Why not LDS float atomics (the reference's tile hash, gradients.cu:422-459):
 * on gfx950, ds_add_f32 measured ~190 cycles per wave instruction per CU
 * against ~5 for ds_add_u32 and ~8 for a read-add-write, and the hashed
 * accumulation needs 10 of them per pixel-candidate; here only one integer
 * atomic per pixel-candidate remains.
 * Numerics: the forward (per-site terms, logits, softmax weights, blended
 * colour) uses the exact primitives of sad_math.h in the oracle's operation
 * order, so weights and colours are bit-identical to the CPU restatement.  This
 * matters beyond the render: dL/dlogit = w g.(c - pred) cancels when one
 * candidate dominates a pixel, so ulp-level differences in pred would become
 * large relative differences in the gradient.  The per-pair derivative is
 * sad_grad_site and sad_grad_pair, the functions the oracle evaluates (its square roots from the
 * reproducible sad_rsqrt_nr), so per-pair values agree bit for bit as well.
 */
#include <mutex>
#include <string>
#include <type_traits>

#include "fit_common.h"

namespace sadh {

enum PixelMode { MODE_RENDER = 0, MODE_GRAD = 1, MODE_BWD = 2, MODE_STATS = 3 };

constexpr uint16_t SLOT_INVALID = 0xFFFFu;   /* id >= site count, pixel not processed, or already summed */
constexpr uint16_t SLOT_OVERFLOW = 0xFFFEu;  /* valid id without a hash slot this round */
constexpr uint16_t ENTRY_FIRST = 0x8000u;    /* entry flag: first rank of its site in the pixel */

#ifndef SAD_CHUNK
#define SAD_CHUNK 20
#endif
constexpr uint32_t CHUNK = SAD_CHUNK;       /* entries per chunk (5 per lane of a 4-lane group) */
#ifndef SAD_CHUNK_LANES
#define SAD_CHUNK_LANES 4
#endif
constexpr uint32_t CLANES = SAD_CHUNK_LANES;   /* lanes per chunk: 4 or 8 */
static_assert(CLANES == 4u || CLANES == 8u, "chunks take 4 or 8 lanes");

/* sum over the CLANES-lane group of a lane; every lane of the group gets it:
 * (v0 + v1) + (v2 + v3) for 4 lanes (float addition is commutative, so the
 * lanes' mirrored orders give the same bits) */
__device__ __forceinline__ float chunk_sum(float v)
{
    if constexpr (CLANES == 8u) return group8_sum(v);
    v = v + dpp<0xB1>(v);
    return v + dpp<0x4E>(v);
}
constexpr uint8_t SLOT8_NONE = 0xFFu;        /* per-pixel table: candidate not accumulated via a slot */

/* values accumulated per site: 10 gradients (+ removal delta), or 8 statistics */
template <int MODE, bool REMOVAL>
struct NumValues {
    static constexpr int value = (MODE == 3) ? 8 : (REMOVAL ? 11 : 10);
};

/* LDS of every mode: the tile's site hash and staged site terms */
template <int H>
struct TileSites {
    uint32_t key[H];
    float4 p0[H];              /* pos.x, pos.y, tau, r/S      (pos.x < 0: inactive) */
    float4 p1[H];              /* dir.x, dir.y, e^a, 1/e^a */
    float4 p2[H];              /* r, g, b, 1/max(e^a, 1e-8) */
    uint32_t count[H];         /* entries per slot */
    uint16_t occ[H];           /* dense list of occupied slots */
    uint32_t n_occ;
};

/* LDS of the backward modes: per-pixel data, entry lists, chunk partials */
template <int H, bool REMOVAL, int NV>
struct TileLds : TileSites<H> {
    static constexpr uint32_t MAX_CHUNKS = 2048u / CHUNK + (uint32_t)H;
    uint32_t start[H];         /* first entry of each slot in `list` */
    /* entries per slot counted per wave, two 16-bit fields per word (waves 2v and 2v+1 in wcount[v]); after
     * the scan, each field holds the offset of that wave's entries within the slot's list */
    uint32_t wcount[2][H];
    uint16_t cstart[H];        /* first chunk of each slot */
    uint32_t n_chunks;
    uint32_t stot[4][2];       /* stage-4 entry and chunk totals of each scanning wave */
    uint32_t hstate[2];        /* [round & 1] nonzero: some id is pending after this round's inserts */
    /* per pixel: slot of each candidate rank (SLOT8_NONE if unused); read only to sum a site's weight over
     * its ranks in a pixel (removal deltas, statistics) */
    uint8_t slot[(REMOVAL || NV == 8) ? 256 : 1][8];
    float w[256][8];           /* per pixel: softmax weights */
    float4 px[256];            /* per pixel: pred.rgb, err (STATS) */
    float4 pg[256];            /* per pixel: dL/dpred.rgb */
    float4 pt[REMOVAL ? 256 : 1];  /* per pixel: target.rgb (removal deltas) */
    uint16_t list[256 * 8];    /* per-slot entry lists: pixel * 8 + rank (| ENTRY_FIRST) */
    uint8_t chunk_slot[MAX_CHUNKS];
    float part[MAX_CHUNKS][NV];    /* per-chunk partial sums */
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

/* Insert-if-absent with linear probing, 8 probes: the slot, or SLOT_OVERFLOW. */
template <class LDS, int H, int PROBES = 8>
__device__ __forceinline__ uint16_t hash_insert(LDS& L, uint32_t id)
{
    uint32_t h = (id * 2654435761u) & (uint32_t)(H - 1);
#pragma unroll 1
    for (int probe = 0; probe < PROBES; ++probe) {
        const uint32_t k = L.key[h];
        if (k == id) return (uint16_t)h;
        if (k == SAD_EMPTY) {
            const uint32_t old = atomicCAS(&L.key[h], SAD_EMPTY, id);
            if (old == SAD_EMPTY || old == id) return (uint16_t)h;
        }
        h = (h + 1u) & (uint32_t)(H - 1);
    }
    return SLOT_OVERFLOW;
}

/* normalised, clamped d^T G d of the det-1 metric for a given 1/e^a */
__device__ __forceinline__ float d2_safe(const SiteTerms& t, float dx, float dy, float l2, float iss,
                                         float* proj_out, float* perp2_out)
{
    float diff2 = fmaf(dx, dx, dy * dy);
    float proj = fmaf(t.p1.x, dx, t.p1.y * dy);
    float proj2 = proj * proj;
    float perp2 = fmaxf(diff2 - proj2, 0.0f);
    *proj_out = proj;
    *perp2_out = perp2;
    return fmaxf(fmaf(t.p1.z, proj2, l2 * perp2) * iss, 1e-8f);
}

/* The staged terms of a site as sad_grad_site (sad_math.h, gradients.cu:352-420) takes them. */
__device__ __forceinline__ SadGradSite grad_site(const SiteTerms& t, float iss, float inv_scale)
{
    return sad_grad_site(t.p0.x, t.p0.y, t.p0.z, t.p0.w, t.p1.x, t.p1.y, t.p1.z, t.p2.w, t.p2.x, t.p2.y, t.p2.z, iss,
                         inv_scale);
}

struct PixelArgs {
    const uint32_t* cand0;
    const uint32_t* cand1;
    const SadSite* sites;
    const float* target;       /* GRAD, STATS */
    const float* mask;         /* GRAD, STATS */
    const float* rendered;     /* GRAD with ssim_weight > 0: the render of the current sites */
    float ssim_weight;         /* GRAD: weight of the -SSIM loss term */
    const float* grad_out;     /* BWD */
    float* out;                /* RENDER: [H][W][3] */
    uint32_t cells;            /* RENDER: 1 = each site's colour is sad_hash_color of its id (cell picture) */
    /* GRAD: the last list_blocks blocks of the launch write the tau passes' neighbour lists (tau_list_only)
     * instead of a tile; they read only sites and the candidate field, which the gradient does not change.
     * Placed after the tiles, these short blocks fill the slots that the last tiles leave. */
    uint4* nbr;
    float nbr_lambda, nbr_inv_ds;
    uint32_t list_blocks;
    float* g[10];              /* DST_SOA: value j of site i accumulates at g[j] + i * gstride */
    float* gline;              /* DST_LINES: value j of site i accumulates at gline + i * gstride + j */
    long long* gfix;           /* DST_FIX: (lo, hi) of gradient j of site i at gfix + i * FIX_GRAD_STRIDE + 2 j */
    long long* rfix;           /* DST_FIX with removal: (lo, hi) of the removal delta of site i at rfix + 2 i */
    long long* sfix;           /* DST_FIXSTATS: (lo, hi) of statistic j of site i at sfix + i * FIX_STAT_STRIDE + 2 j */
    uint32_t gstride;          /* 1: SoA arrays; 10: AoS [N][10] (BWD); 16: AoS [N][16] lines */
    float* removal;            /* GRAD with removal, float destinations */
    float iss, inv_scale;
    uint32_t n;
    int width, height, cw, ch;
    uint32_t tiles_x, ntiles;
};

template <int MODE, bool REMOVAL, int H>
using PixelLds = typename std::conditional<MODE == MODE_RENDER, TileSites<H>,
                                           TileLds<H, REMOVAL, NumValues<MODE, REMOVAL>::value>>::type;

/* Where a tile's per-site sums go.  The float kinds serve the C ABI's float
 * outputs (the reference's layouts; the order of the float atomics rounds
 * differently run to run, as the reference's do).  The fit accumulates its
 * gradients, removal deltas (DST_FIX) and densify statistics (DST_FIXSTATS)
 * in the exact two-word fixed point of fix_split, so its sums do not depend
 * on the tiles' order.  AoS lines index by arithmetic, so a PixelArgs built
 * in registers is never indexed by a lane-varying value. */
enum PixelDst { DST_SOA = 0, DST_LINES = 1, DST_FIX = 2, DST_FIXSTATS = 3 };

template <int DST>
__device__ __forceinline__ void acc_add(const PixelArgs& A, uint32_t j, uint32_t id, float v)
{
    if constexpr (DST == DST_LINES) unsafeAtomicAdd(A.gline + (size_t)id * A.gstride + j, v);
    else unsafeAtomicAdd(A.g[j] + (size_t)id * A.gstride, v);
}

/* Probes of the later rounds.  Every lane of a round probes the same window, so an id is placed and found
 * consistently whatever its size; a tile that needs later rounds holds hundreds of distinct ids (a random
 * candidate field), where most ids find the table full and a short window fails them sooner. */
#ifndef SAD_PENDING_PROBES
#define SAD_PENDING_PROBES 4
#endif

/* Round r's inserts of the ids still pending (slot SLOT_OVERFLOW); summed ids
 * leave the table (SLOT_INVALID).  Returns whether some id is still pending. */
template <class LDS, int H>
__device__ __forceinline__ bool insert_pending(LDS& L, const uint32_t ids[8], uint16_t slots[8])
{
    bool ovf = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (slots[k] == SLOT_OVERFLOW) {
            slots[k] = hash_insert<LDS, H, SAD_PENDING_PROBES>(L, ids[k]);
            ovf = ovf || slots[k] == SLOT_OVERFLOW;
        } else {
            slots[k] = SLOT_INVALID;
        }
    }
    return ovf;
}

/* stage 2: per-site terms of the occupied slots; dense list of them (backward modes) */
template <int MODE, int H, class LDS>
__device__ __forceinline__ void stage_sites(const PixelArgs& A, LDS& L, uint32_t tid)
{
    for (uint32_t s = tid; s < (uint32_t)H; s += 256u) {
        uint32_t id = L.key[s];
        if (id == SAD_EMPTY) continue;
        SiteTerms t = site_terms(A.sites, id, A.inv_scale);
        L.p0[s] = t.p0;
        L.p1[s] = t.p1;
        L.p2[s] = t.p2;
    }
    if (MODE != MODE_RENDER && tid < 64u) {
        uint32_t base = 0u;
        const unsigned long long lt = (tid == 0u) ? 0ull : (~0ull >> (64u - tid));
        for (uint32_t s0 = 0; s0 < (uint32_t)H; s0 += 64u) {
            bool occ = L.key[s0 + tid] != SAD_EMPTY;
            unsigned long long b = __ballot(occ);
            if (occ) L.occ[base + (uint32_t)__popcll(b & lt)] = (uint16_t)(s0 + tid);
            base += (uint32_t)__popcll(b);
        }
        if (tid == 0u) L.n_occ = base;
    }
}

/* One 16x16 tile (index `tile` of A's image) of any mode; 256 threads, one pixel each. */
template <int MODE, bool REMOVAL, int H, int DST>
__device__ __forceinline__ void pixel_tile(const PixelArgs& A, const uint32_t tile, PixelLds<MODE, REMOVAL, H>& L)
{
    constexpr int NV = NumValues<MODE, REMOVAL>::value;
    static_assert(H <= 128, "per-pixel slot table holds 7-bit slots");
    static_assert(MODE != MODE_STATS || DST == DST_FIXSTATS || DST == DST_SOA,
                  "statistics go to DST_FIXSTATS or float arrays");
    static_assert(DST != DST_FIXSTATS || MODE == MODE_STATS, "DST_FIXSTATS holds statistics");
    const uint32_t tid = threadIdx.x;
    const uint32_t wv = tid >> 6;   /* wave of the block */
    const int tx0 = (int)(tile % A.tiles_x) * 16, ty0 = (int)(tile / A.tiles_x) * 16;
    const int x = tx0 + (int)(tid & 15u);
    const int y = ty0 + (int)(tid >> 4);
    const bool inb = x < A.width && y < A.height;
    const uint32_t pix = inb ? (uint32_t)(y * A.width + x) : 0u;
    const float ux = (float)x, uy = (float)y;

    for (uint32_t s = tid; s < (uint32_t)H; s += 256u) {
        L.key[s] = SAD_EMPTY;
        L.count[s] = 0u;
        if constexpr (MODE != MODE_RENDER) {
            L.wcount[0][s] = 0u;
            L.wcount[1][s] = 0u;
        }
    }
    if constexpr (MODE != MODE_RENDER) {
        if (tid == 0u) {
            L.hstate[0] = 0u;
            L.hstate[1] = 0u;
        }
    }
    __syncthreads();

    /* stage 1: this pixel's candidate ids into the tile hash.  The mask, the
     * candidates and the pixel's target (or upstream gradient, read in stage 3)
     * are loaded together, so their latencies overlap; a masked-out pixel then
     * drops its candidates. */
    bool act = inb;
    uint32_t ids[8];
    uint16_t slots[8];
    float q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
    {
        const uint4 none = make_uint4(SAD_EMPTY, SAD_EMPTY, SAD_EMPTY, SAD_EMPTY);
        uint4 a = none, b = none;
        float mk = 1.0f;
        if (inb) {
            /* a full-resolution field (the usual case) needs none of the index scaling's divisions */
            const int cell = (A.cw == A.width && A.ch == A.height) ? y * A.width + x
                                                                     : sad_cand_index(x, y, A.width, A.height, A.cw, A.ch);
            a = reinterpret_cast<const uint4*>(A.cand0)[cell];
            b = reinterpret_cast<const uint4*>(A.cand1)[cell];
            if constexpr (MODE == MODE_GRAD || MODE == MODE_STATS) mk = A.mask[pix];
            if constexpr (MODE != MODE_RENDER) {
                const float* qp = (MODE == MODE_BWD ? A.grad_out : A.target) + 3u * pix;
                q0 = qp[0]; q1 = qp[1]; q2 = qp[2];
            }
        }
        if (MODE == MODE_GRAD || MODE == MODE_STATS) act = act && mk > 0.0f;
        if (!act) { a = none; b = none; }
        ids[0] = a.x; ids[1] = a.y; ids[2] = a.z; ids[3] = a.w;
        ids[4] = b.x; ids[5] = b.y; ids[6] = b.z; ids[7] = b.w;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            slots[k] = (act && ids[k] < A.n) ? hash_insert<decltype(L), H>(L, ids[k]) : SLOT_INVALID;
    }
    __syncthreads();

    stage_sites<MODE, H>(A, L, tid);
    __syncthreads();

    /* stage 3: softmax of this pixel (render.cu:34-91) */
    float logit[8];
    bool valid[8];
    float mx = -INFINITY;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint16_t s = slots[k];
        SiteTerms t;
        t.p0 = make_float4(-1.0f, -1.0f, 0.0f, 0.0f);
        t.p1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (s < (uint16_t)H) {
            t.p0 = L.p0[s];
            t.p1 = L.p1[s];
        } else if (s == SLOT_OVERFLOW) {
            SiteTerms f = site_terms(A.sites, ids[k], A.inv_scale);
            t.p0 = f.p0;
            t.p1 = f.p1;
        }
        valid[k] = s != SLOT_INVALID && t.p0.x >= 0.0f;
        float proj, perp2;
        float d2s = d2_safe(t, ux - t.p0.x, uy - t.p0.y, t.p1.w, A.iss, &proj, &perp2);
        logit[k] = valid[k] ? -(t.p0.z * (sad_sqrt_pos(d2s) - t.p0.w)) : -INFINITY;
        mx = fmaxf(mx, logit[k]);
    }
    bool has = mx > -INFINITY;
    if (MODE == MODE_BWD) has = has && mx < INFINITY;   /* sad_bindings.cu:72-74 */
    float w[8];
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        /* evaluated for every rank and then selected, so the eight exps share one basic block (an invalid rank
         * gives 0 or NaN here, which the select drops) */
        const float e = sad_expf_nonpos(logit[k] - mx);   /* logit - max <= 0 */
        w[k] = (valid[k] && has) ? e : 0.0f;
        sum += w[k];
    }
    const float inv = 1.0f / fmaxf(sum, 1e-8f);
    /* the candidates' colours: the staged ones are read from LDS for every rank first (an index in range for
     * every slot value), and a pending id's colour replaces its entry in a separate pass, so the compiler keeps
     * LDS loads here instead of merging the two sources into generic loads */
    float cr[8], cg[8], cb[8];
    bool any_pending = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint16_t s = slots[k];
        const float4 q = L.p2[s < (uint16_t)H ? s : 0u];
        cr[k] = q.x; cg[k] = q.y; cb[k] = q.z;
        any_pending = any_pending || s == SLOT_OVERFLOW;
    }
    if (any_pending) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (slots[k] == SLOT_OVERFLOW) {
                const SadSite* sp = A.sites + ids[k];
                cr[k] = sp->col_r; cg[k] = sp->col_g; cb[k] = sp->col_b;
            }
        }
    }
    float pr = 0.0f, pgr = 0.0f, pb = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        w[k] = w[k] * inv;
        float wc = w[k];
        if (MODE == MODE_RENDER && (isnan(wc) || isinf(wc))) wc = 0.0f;
        if (valid[k]) {
            float4 c = make_float4(cr[k], cg[k], cb[k], 0.0f);
            if (MODE == MODE_RENDER && A.cells) {
                float hc[3];
                sad_hash_color(ids[k], hc);
                c = make_float4(hc[0], hc[1], hc[2], 0.0f);
            }
            pr += wc * c.x;
            pgr += wc * c.y;
            pb += wc * c.z;
        }
    }

    if constexpr (MODE == MODE_RENDER) {
        if (inb) {
            float* o = A.out + 3u * pix;
            o[0] = has ? pr : 0.0f;
            o[1] = has ? pgr : 0.0f;
            o[2] = has ? pb : 0.0f;
        }
    } else {

    /* dL/dpred and the per-pixel data the site-centric pass reads */
    const bool contrib = act && has;
    float gr = 0.0f, gg = 0.0f, gb = 0.0f, err = 0.0f;
    float tr = 0.0f, tg = 0.0f, tb = 0.0f;
    if (contrib) {
        if (MODE == MODE_BWD) {
            gr = q0; gg = q1; gb = q2;
        } else {
            const float* tp = A.target + 3u * pix;
            tr = q0; tg = q1; tb = q2;
            float dr = pr - tr, dg = pgr - tg, db = pb - tb;
            gr = 2.0f * dr; gg = 2.0f * dg; gb = 2.0f * db;
            err = dr * dr + dg * dg + db * db;
            if (MODE == MODE_GRAD && A.ssim_weight > 0.0f) {
                /* -w SSIM of the pixel's 3x3 window, through its own luma (gradients.cu:236-266) */
                SadSsimStats st = sad_ssim_window(A.rendered, A.target, A.width, A.height, x, y);
                float f = -A.ssim_weight * sad_ssim_dcentre(st, sad_luma3(A.rendered + 3u * pix), sad_luma3(tp));
                gr += f * 0.299f;
                gg += f * 0.587f;
                gb += f * 0.114f;
            }
        }
    }
    /* the statistics read the pixel's error from px.w; the gradients read the pixel's coordinates there and in
     * pg.w, which saves decoding them from the entry */
    L.px[tid] = make_float4(pr, pgr, pb, MODE == MODE_STATS ? err : ux);
    L.pg[tid] = make_float4(gr, gg, gb, uy);
    if (REMOVAL) L.pt[tid] = make_float4(tr, tg, tb, 0.0f);
    /* the pairs that are summed keep their slot; a valid id without a slot waits for a later round */
    bool pending = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        L.w[tid][k] = w[k];
        if (!(contrib && valid[k])) slots[k] = SLOT_INVALID;
        pending = pending || slots[k] == SLOT_OVERFLOW;
    }
    if (pending) L.hstate[0] = 1u;

    const uint32_t wshift = 16u * (wv & 1u);
    for (uint32_t round = 0;; ++round) {
    /* stage 4: an entry's place in its slot's list is (entries of earlier waves) + (its rank among this
     * wave's).  The per-wave counters make the list order, and with it every float sum over the list,
     * independent of the order in which the tile's waves reach this point: within one wave the atomics run
     * in program order, and the lanes of one LDS atomic instruction are served in lane order. */
    uint32_t rank[8];
    uint8_t sl[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint16_t s = slots[k];
        sl[k] = (s < (uint16_t)H) ? (uint8_t)s : SLOT8_NONE;
        rank[k] = (sl[k] != SLOT8_NONE) ? ((atomicAdd(&L.wcount[wv >> 1][s], 1u << wshift) >> wshift) & 0xFFFFu)
                                        : 0u;
        if constexpr (MODE == MODE_STATS || REMOVAL) L.slot[tid][k] = sl[k];
    }
    __syncthreads();

    /* (waves 0..H/64-1, 64 slots each): entry offsets and chunk offsets =
     * exclusive scans of the per-slot entry and chunk counts; chunk -> slot table.
     * Each wave scans its slots, posts its totals, and after a barrier adds the
     * totals of the waves before it. */
    static_assert(H % 64 == 0 && H <= 256, "stage 4 scans 64 slots per wave");
    {
        const uint32_t lane = tid & 63u;
        const bool scan = tid < (uint32_t)H;
        uint32_t c = 0u, nc = 0u, v = 0u, vc = 0u, c0 = 0u, c01 = 0u, c012 = 0u;
        if (scan) {
            /* per-wave counts (at most 512 each) -> per-wave offsets within the slot, and the total */
            const uint32_t w01 = L.wcount[0][tid], w23 = L.wcount[1][tid];
            c0 = w01 & 0xFFFFu;
            c01 = c0 + (w01 >> 16);
            c012 = c01 + (w23 & 0xFFFFu);
            c = c012 + (w23 >> 16);
            nc = (c + CHUNK - 1u) / CHUNK;
            v = c;
            vc = nc;
#pragma unroll
            for (uint32_t off = 1; off < 64u; off <<= 1) {
                uint32_t o = (uint32_t)__shfl_up((int)v, off);
                uint32_t oc = (uint32_t)__shfl_up((int)vc, off);
                if (lane >= off) { v += o; vc += oc; }
            }
            if (lane == 63u) {
                L.stot[wv][0] = v;
                L.stot[wv][1] = vc;
            }
        }
        __syncthreads();
        if (scan) {
            uint32_t run = 0u, crun = 0u;
            for (uint32_t j = 0; j < wv; ++j) {
                run += L.stot[j][0];
                crun += L.stot[j][1];
            }
            const uint32_t s = tid;
            L.wcount[0][s] = c0 << 16;
            L.wcount[1][s] = c01 | (c012 << 16);
            L.count[s] = c;
            const uint32_t cs = crun + vc - nc;
            L.start[s] = run + v - c;
            L.cstart[s] = (uint16_t)cs;
            for (uint32_t j = 0; j < nc; ++j) L.chunk_slot[cs + j] = (uint8_t)s;
            if (tid == 0u) {
                uint32_t all = 0u;
                for (uint32_t j = 0; j < (uint32_t)H / 64u; ++j) all += L.stot[j][1];
                L.n_chunks = all;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint8_t s = sl[k];
        if (s != SLOT8_NONE) {
            /* the first rank of a site in the pixel: only the statistics and removal deltas read it */
            bool first = MODE == MODE_STATS || REMOVAL;
            if (MODE == MODE_STATS || REMOVAL) {
#pragma unroll
                for (int j = 0; j < k; ++j) first = first && (sl[j] != s);
            }
            const uint32_t at = L.start[s] + ((L.wcount[wv >> 1][s] >> wshift) & 0xFFFFu) + rank[k];
            L.list[at] = (uint16_t)((tid << 3) | (uint32_t)k | (first ? ENTRY_FIRST : 0u));
        }
    }
    __syncthreads();

    /* stage 5: CLANES lanes per chunk of CHUNK entries; one partial per chunk */
    const uint32_t n_chunks = L.n_chunks;
    const uint32_t part = tid & (CLANES - 1u);
    for (uint32_t base = 0; base < n_chunks; base += 256u / CLANES) {
        const uint32_t q = base + tid / CLANES;
        const bool live = q < n_chunks;
        const uint32_t s = live ? (uint32_t)L.chunk_slot[q] : 0u;
        const uint32_t sbeg = L.start[s];
        const uint32_t beg = sbeg + (q - (uint32_t)L.cstart[s]) * CHUNK;
        const uint32_t end = live ? min(beg + CHUNK, sbeg + L.count[s]) : 0u;
        SiteTerms t;
        t.p0 = L.p0[s];
        t.p1 = L.p1[s];
        t.p2 = L.p2[s];
        const SadGradSite gs = grad_site(t, A.iss, A.inv_scale);
        float acc[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) acc[j] = 0.0f;
        for (uint32_t e = beg + part; e < end; e += CLANES) {
            const uint32_t ent = L.list[e];
            const uint32_t p = (ent & 0x7FFFu) >> 3;
            const float4 pp = L.px[p];
            const float4 gq = L.pg[p];
            const float ex = MODE == MODE_STATS ? (float)(tx0 + (int)(p & 15u)) : pp.w;
            const float ey = MODE == MODE_STATS ? (float)(ty0 + (int)(p >> 4)) : gq.w;
            if (MODE == MODE_STATS || (MODE == MODE_GRAD && REMOVAL)) {
                if (ent & ENTRY_FIRST) {
                    float wt = 0.0f;
#pragma unroll
                    for (int j = 0; j < 8; ++j) wt += (L.slot[p][j] == (uint8_t)s) ? L.w[p][j] : 0.0f;
                    if (wt > 0.0f) {
                        if (MODE == MODE_STATS) {
                            const float er = pp.w;
                            const float we = wt * fminf(er, 1.0f);
                            acc[0] += wt; acc[1] += wt * er; acc[2] += we;
                            acc[3] += we * ex; acc[4] += we * ey;
                            acc[5] += we * ex * ex; acc[6] += we * ex * ey; acc[7] += we * ey * ey;
                        } else {
                            const float4 tq = L.pt[p];
                            acc[NV - 1] += sad_removal_delta(pp.x, pp.y, pp.z, tq.x, tq.y, tq.z, t.p2.x, t.p2.y,
                                                             t.p2.z, wt);
                        }
                    }
                }
            }
            if (MODE != MODE_STATS) {
                float c[10];
                sad_grad_pair(&gs, ex, ey, (&L.w[0][0])[ent & 0x7FFFu], gq.x, gq.y, gq.z, pp.x, pp.y, pp.z, c);
#pragma unroll
                for (int j = 0; j < 10; ++j) acc[j] += c[j];
            }
        }
        /* every lane of the group ends with the same totals; the group's first lane stores them */
#pragma unroll
        for (int j = 0; j < NV; ++j) acc[j] = chunk_sum(acc[j]);
        if (live && part == 0u) {
#pragma unroll
            for (int j = 0; j < NV; ++j) L.part[q][j] = acc[j];
        }
    }
    __syncthreads();

    /* per (site, value): sum the slot's chunk partials in chunk order, then the global atomics */
    const uint32_t n_occ = L.n_occ;
    if constexpr (DST == DST_FIX || DST == DST_FIXSTATS) {
        /* one lane per (site, value, word): a site's (lo, hi) pairs are consecutive lanes and addresses */
        for (uint32_t i = tid; i < n_occ * (uint32_t)(2 * NV); i += 256u) {
            const uint32_t u = i / (uint32_t)(2 * NV), r = i - u * (uint32_t)(2 * NV);
            const uint32_t j = r >> 1, hiw = r & 1u;
            const uint32_t s = L.occ[u];
            /* the slot's words are read together, and the chunk partials four at a time (summed in chunk order):
             * the flush is a chain of LDS round trips, not arithmetic */
            const uint32_t c = L.count[s];
            const float sx = L.p0[s].x;
            const uint32_t qa = L.cstart[s];
            const size_t id = L.key[s];
            if (c == 0u || !(sx >= 0.0f)) continue;
            const uint32_t qb = qa + (c + CHUNK - 1u) / CHUNK;
            float v = 0.0f;
            uint32_t q = qa;
            for (; q + 4u <= qb; q += 4u) {
                const float a0 = L.part[q][j], a1 = L.part[q + 1u][j], a2 = L.part[q + 2u][j], a3 = L.part[q + 3u][j];
                v += a0;
                v += a1;
                v += a2;
                v += a3;
            }
            for (; q < qb; ++q) v += L.part[q][j];
            const long long word = fix_word(v, hiw);
            long long* dst;
            if (DST == DST_FIXSTATS) dst = A.sfix + id * FIX_STAT_STRIDE + 2u * j;
            else if (REMOVAL && j == 10u) dst = A.rfix + 2u * id;
            else dst = A.gfix + id * FIX_GRAD_STRIDE + 2u * j;
            fix_atomic_add(dst + hiw, word);
        }
    } else {
        for (uint32_t i = tid; i < n_occ * (uint32_t)NV; i += 256u) {
            const uint32_t u = i / (uint32_t)NV, j = i - u * (uint32_t)NV;
            const uint32_t s = L.occ[u];
            const uint32_t c = L.count[s];
            if (c == 0u || !(L.p0[s].x >= 0.0f)) continue;
            const uint32_t qa = L.cstart[s], qb = qa + (c + CHUNK - 1u) / CHUNK;
            float v = 0.0f;
            for (uint32_t q = qa; q < qb; ++q) v += L.part[q][j];
            const uint32_t id = L.key[s];
            if (MODE == MODE_GRAD && REMOVAL && j == 10u) {
                if (v != 0.0f) unsafeAtomicAdd(A.removal + id, v);
            } else {
                acc_add<DST>(A, j, id, v);
            }
        }
    }

    /* stage 6: the next round, for the ids that found no slot */
    const uint32_t more = L.hstate[round & 1u];
    if (!more) break;
    __syncthreads();
    for (uint32_t s = tid; s < (uint32_t)H; s += 256u) {
        L.key[s] = SAD_EMPTY;
        L.wcount[0][s] = 0u;
        L.wcount[1][s] = 0u;
    }
    if (tid == 0u) L.hstate[(round + 1u) & 1u] = 0u;
    __syncthreads();
    if (insert_pending<decltype(L), H>(L, ids, slots)) L.hstate[(round + 1u) & 1u] = 1u;
    __syncthreads();
    stage_sites<MODE, H>(A, L, tid);
    __syncthreads();
    }  /* rounds */
    }  /* backward modes */
}


template <int MODE, bool REMOVAL, int H, int DST>
__global__ __launch_bounds__(256) void k_pixel(PixelArgs A)
{
    __shared__ PixelLds<MODE, REMOVAL, H> L;
    if constexpr (MODE == MODE_GRAD) {
        if (blockIdx.x >= A.ntiles) {
            const uint32_t i = (blockIdx.x - A.ntiles) * 256u + threadIdx.x;
            if (i < A.n)
                tau_list_only(i, reinterpret_cast<const uint4*>(A.cand0), reinterpret_cast<const uint4*>(A.cand1),
                              A.sites, A.n, A.nbr_lambda, A.cw, A.ch, A.nbr_inv_ds, A.nbr);
            return;
        }
    }
    pixel_tile<MODE, REMOVAL, H, DST>(A, xcd_tile(blockIdx.x, A.ntiles), L);
}

/* Batched fit: one launch over the concatenated tiles of every job. */
template <bool REMOVAL, int H>
__global__ __launch_bounds__(256) void k_pixel_grad_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                          uint32_t total_tiles)
{
    __shared__ PixelLds<MODE_GRAD, REMOVAL, H> L;
    const uint32_t f = xcd_tile(blockIdx.x, total_tiles);
    const uint32_t j = batch_find(jobs, njobs, f, &BatchJob::tile_begin);
    const BatchJob J = job_global(jobs, j);
    PixelArgs A = {};
    A.cand0 = J.cur ? J.fb0 : J.fa0;
    A.cand1 = J.cur ? J.fb1 : J.fa1;
    A.sites = J.sites;
    A.target = J.target;
    A.mask = J.mask;
    A.gfix = J.gacc;
    A.rfix = J.removal;
    A.rendered = J.rendered;
    A.ssim_weight = J.ssim_weight;
    A.gstride = 16u;
    A.iss = J.iss;
    A.inv_scale = J.inv_scale;
    A.n = J.n;
    A.width = J.width;
    A.height = J.height;
    A.cw = J.cw;
    A.ch = J.ch;
    A.tiles_x = J.tiles_x;
    A.ntiles = J.ntiles;
    pixel_tile<MODE_GRAD, REMOVAL, H, DST_FIX>(A, f - J.tile_begin, L);
}

/* Batched fit, log iterations: the render of every job whose step flag `log` is
 * set, into its render buffer; the same tiles as the single-image render. */
template <int H>
__global__ __launch_bounds__(256) void k_pixel_render_batch(const BatchJob* __restrict__ jobs, uint32_t njobs,
                                                            const JobStep* __restrict__ steps, uint32_t total_tiles)
{
    __shared__ PixelLds<MODE_RENDER, false, H> L;
    const uint32_t f = xcd_tile(blockIdx.x, total_tiles);
    const uint32_t j = batch_find(jobs, njobs, f, &BatchJob::tile_begin);
    if (!steps[j].log) return;
    const BatchJob J = job_global(jobs, j);
    PixelArgs A = {};
    A.cand0 = J.cur ? J.fb0 : J.fa0;
    A.cand1 = J.cur ? J.fb1 : J.fa1;
    A.sites = J.sites;
    A.out = J.render;
    A.iss = J.iss;
    A.inv_scale = J.inv_scale;
    A.n = J.n;
    A.width = J.width;
    A.height = J.height;
    A.cw = J.cw;
    A.ch = J.ch;
    A.tiles_x = J.tiles_x;
    A.ntiles = J.ntiles;
    pixel_tile<MODE_RENDER, false, H, DST_SOA>(A, f - J.tile_begin, L);
}

/* ------------------------------------------------------------ launchers -- */

#ifndef SAD_HASH_SLOTS
#define SAD_HASH_SLOTS 128
#endif
constexpr int HASH_SLOTS = SAD_HASH_SLOTS;

template <int MODE, bool REMOVAL, int DST = DST_SOA>
static int launch_pixel(PixelArgs& A, hipStream_t stream, const char* what)
{
    A.tiles_x = div_up((uint32_t)A.width, 16u);
    A.ntiles = A.tiles_x * div_up((uint32_t)A.height, 16u);
    A.inv_scale = sqrtf(A.iss);
    hipLaunchKernelGGL((k_pixel<MODE, REMOVAL, HASH_SLOTS, DST>), dim3(A.ntiles + A.list_blocks), dim3(256), 0,
                       stream, A);
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

/* renderVoronoiColoring, shaders/encoders/render_encoder.metal:73-134: the render with hashed id colours */
int launch_render_cells(const uint32_t* cand0, const uint32_t* cand1, float* out, const SadSite* sites, float iss,
                        uint32_t n, int width, int height, int cw, int ch, hipStream_t stream)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.out = out; A.iss = iss; A.n = n; A.cells = 1u;
    A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    return launch_pixel<MODE_RENDER, false>(A, stream, "render_cells");
}

/* renderCentroidsTauHeatmap, shaders/encoders/render_encoder.metal:136-190: one thread per pixel, tau dots of
 * the pixel's candidate sites composited front to back (sad_dot_alpha, sad_heat_color) */
__global__ __launch_bounds__(256) void k_tau_heatmap(const uint4* __restrict__ c0, const uint4* __restrict__ c1,
                                                     float* __restrict__ out, const SadSite* __restrict__ sites,
                                                     uint32_t n, float tmin, float tmean, float tmax, float dot,
                                                     int width, int height, int cw, int ch)
{
    const int x = (int)(blockIdx.x * 16u + (threadIdx.x & 15u));
    const int y = (int)(blockIdx.y * 16u + (threadIdx.x >> 4));
    if (x >= width || y >= height) return;
    const int cell = sad_cand_index(x, y, width, height, cw, ch);
    const uint4 a = c0[cell], b = c1[cell];
    const uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    float acc[3] = {0.0f, 0.0f, 0.0f}, acc_a = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t id = ids[k];
        if (id >= n) continue;
        const SadSite s = sites[id];
        if (s.pos_x < 0.0f) continue;
        const float al = sad_dot_alpha((float)x - s.pos_x, (float)y - s.pos_y, dot);
        if (al <= 0.0f) continue;
        float col[3];
        sad_heat_color(sad_expf(s.log_tau), tmin, tmean, tmax, col);
        const float blend = al * (1.0f - acc_a);
        acc[0] += col[0] * blend;
        acc[1] += col[1] * blend;
        acc[2] += col[2] * blend;
        acc_a += blend;
    }
    float* o = out + 3u * (uint32_t)(y * width + x);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = acc[c] + 0.05f * (1.0f - acc_a);
}

int launch_tau_heatmap(const uint32_t* cand0, const uint32_t* cand1, float* out, const SadSite* sites, uint32_t n,
                       float tmin, float tmean, float tmax, float dot, int width, int height, int cw, int ch,
                       hipStream_t stream)
{
    const dim3 grid(div_up((uint32_t)width, 16u), div_up((uint32_t)height, 16u));
    hipLaunchKernelGGL(k_tau_heatmap, grid, dim3(256), 0, stream, reinterpret_cast<const uint4*>(cand0),
                       reinterpret_cast<const uint4*>(cand1), out, sites, n, tmin, tmean, tmax, dot, width, height,
                       cw, ch);
    return launch_status("render_tau_heatmap");
}

/* ----------------------------------------- float outputs via exact lines --
 * The reference-interface entry points write float arrays.  They accumulate
 * into fixed-point lines in a per-stream workspace (stream_workspace), cleared
 * by a kernel first, then k_fix_finalize adds (or, for the statistics, stores)
 * each value into the caller's floats: the same bits whatever the order of
 * the tiles, and one line-shaped atomic target per (tile, site) instead of
 * ten scattered float arrays. */
struct FixOut {
    float* g[10];              /* value j of site i at g[j] + i * gstride */
    uint32_t gstride, nv, lstride, assign;
    float* removal;            /* nullable: removal delta of site i at removal[i] */
    const long long* lines;    /* (lo, hi) of value j of site i at lines + i * lstride + 2 j */
    const long long* rlines;   /* (lo, hi) of the removal delta of site i at rlines + 2 i */
    uint32_t n;
};

/* One block per 256 sites.  The block first copies its sites' (lo, hi) pairs
 * into LDS with consecutive lanes on consecutive 16-byte pairs (a site's NV
 * pairs are contiguous in its line), so the line reads are coalesced; then
 * each thread turns its site's pairs into floats and writes them. */
template <uint32_t NV>
__global__ __launch_bounds__(256) void k_fix_finalize(FixOut F)
{
    __shared__ longlong2 pairs[256 * NV];
    const uint32_t i0 = blockIdx.x * 256u;
    const uint32_t ns = min(256u, F.n - i0);
    for (uint32_t c = threadIdx.x; c < ns * NV; c += 256u) {
        const uint32_t s = c / NV, j = c - s * NV;
        pairs[c] = reinterpret_cast<const longlong2*>(F.lines + (size_t)(i0 + s) * F.lstride)[j];
    }
    __syncthreads();
    const uint32_t t = threadIdx.x;
    if (t >= ns) return;
    const uint32_t i = i0 + t;
#pragma unroll
    for (uint32_t j = 0; j < NV; ++j) {
        const longlong2 w = pairs[t * NV + j];
        const float v = fix_value(w.x, w.y);
        float* d = F.g[j] + (size_t)i * F.gstride;
        *d = F.assign ? v : *d + v;
    }
    if (F.removal) {
        const longlong2 w = reinterpret_cast<const longlong2*>(F.rlines)[i];
        F.removal[i] += fix_value(w.x, w.y);
    }
}

static int launch_fix_finalize(const FixOut& F, hipStream_t stream)
{
    if (F.n == 0) return SADH_OK;
    if (F.nv == 10u)
        hipLaunchKernelGGL(k_fix_finalize<10u>, dim3(div_up(F.n, 256u)), dim3(256), 0, stream, F);
    else if (F.nv == 8u)
        hipLaunchKernelGGL(k_fix_finalize<8u>, dim3(div_up(F.n, 256u)), dim3(256), 0, stream, F);
    else
        return fail(SADH_ERR_INVALID, "finalize_sums: 8 or 10 values per site");
    return launch_status("finalize_sums");
}

/* workspace for n sites of `line_words` words, plus (removal) one pair each; cleared */
static int fix_workspace(hipStream_t stream, uint32_t n, uint32_t line_words, bool removal, long long** lines,
                         long long** rlines)
{
    const size_t words = (size_t)n * line_words + (removal ? 2u * (size_t)n : 0u);
    void* p = nullptr;
    SAD_TRY(stream_workspace(stream, std::max<size_t>(words, 1) * sizeof(long long), &p));
    *lines = static_cast<long long*>(p);
    *rlines = removal ? *lines + (size_t)n * line_words : nullptr;
    return launch_clear_fix(*lines, words, stream);
}

/* the tiled gradient into float destinations: g[j][i * gstride] += value j of site i (and removal[i] += the
 * removal delta) */
static int gradients_to_floats(PixelArgs& A, float* const g[10], uint32_t gstride, float* removal,
                               bool compute_removal, hipStream_t stream)
{
    long long *lines, *rlines;
    SAD_TRY(fix_workspace(stream, A.n, FIX_GRAD_STRIDE, compute_removal, &lines, &rlines));
    A.gfix = lines;
    A.rfix = rlines;
    if (compute_removal) SAD_TRY((launch_pixel<MODE_GRAD, true, DST_FIX>(A, stream, "compute_gradients_tiled")));
    else SAD_TRY((launch_pixel<MODE_GRAD, false, DST_FIX>(A, stream, "compute_gradients_tiled")));
    FixOut F = {};
    for (int j = 0; j < 10; ++j) F.g[j] = g[j];
    F.gstride = gstride; F.nv = 10; F.lstride = FIX_GRAD_STRIDE; F.assign = 0;
    F.removal = compute_removal ? removal : nullptr;
    F.lines = lines; F.rlines = rlines; F.n = A.n;
    return launch_fix_finalize(F, stream);
}

int launch_gradients(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                     float* const g[10], const SadSite* sites, float iss, uint32_t n, float* removal,
                     bool compute_removal, int width, int height, int cw, int ch, hipStream_t stream,
                     const float* rendered, float ssim_weight)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    A.rendered = rendered; A.ssim_weight = ssim_weight;
    A.iss = iss; A.n = n;
    A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    return gradients_to_floats(A, g, 1u, removal, compute_removal, stream);
}

/* gradients into AoS [N][16] float lines (values 0-9 in Site field order) */
int launch_gradients_aos(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                         float* gacc, const SadSite* sites, float iss, uint32_t n, float* removal, bool compute_removal,
                         int width, int height, int cw, int ch, hipStream_t stream, const float* rendered,
                         float ssim_weight)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    A.rendered = rendered; A.ssim_weight = ssim_weight;
    A.iss = iss; A.n = n;
    A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    float* g[10];
    for (int j = 0; j < 10; ++j) g[j] = gacc + j;
    return gradients_to_floats(A, g, 16u, removal, compute_removal, stream);
}

/* the fit's gradients: fixed-point (lo, hi) lines and removal pairs */
int launch_gradients_fix(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                         long long* gacc, const SadSite* sites, float iss, uint32_t n, long long* removal,
                         bool compute_removal, int width, int height, int cw, int ch, hipStream_t stream,
                         const float* rendered, float ssim_weight, uint32_t* nbr_lists, float nbr_lambda, int ds)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    A.rendered = rendered; A.ssim_weight = ssim_weight;
    A.gfix = gacc;
    A.rfix = removal;
    A.iss = iss; A.n = n;
    A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    if (nbr_lists && n > 0u) {
        A.list_blocks = div_up(n, 256u);
        A.nbr = reinterpret_cast<uint4*>(nbr_lists);
        A.nbr_lambda = nbr_lambda;
        A.nbr_inv_ds = 1.0f / fmaxf((float)ds, 1.0f);
    }
    if (compute_removal) return launch_pixel<MODE_GRAD, true, DST_FIX>(A, stream, "compute_gradients(fix)");
    return launch_pixel<MODE_GRAD, false, DST_FIX>(A, stream, "compute_gradients(fix)");
}

template <int H>
static void launch_grad_batch_h(const BatchJob* jobs, uint32_t njobs, uint32_t total_tiles, bool removal,
                                hipStream_t stream)
{
    if (removal)
        hipLaunchKernelGGL((k_pixel_grad_batch<true, H>), dim3(total_tiles), dim3(256), 0, stream, jobs, njobs,
                           total_tiles);
    else
        hipLaunchKernelGGL((k_pixel_grad_batch<false, H>), dim3(total_tiles), dim3(256), 0, stream, jobs, njobs,
                           total_tiles);
}

/* hash_slots: the tile hash size, 64 or the build's (HASH_SLOTS); every result is the same for either */
int launch_gradients_batch(const BatchJob* jobs, uint32_t njobs, uint32_t total_tiles, bool removal,
                           hipStream_t stream, int hash_slots)
{
    if (total_tiles == 0) return SADH_OK;
    if (hash_slots == 64 && HASH_SLOTS != 64) launch_grad_batch_h<64>(jobs, njobs, total_tiles, removal, stream);
    else launch_grad_batch_h<HASH_SLOTS>(jobs, njobs, total_tiles, removal, stream);
    return launch_status("compute_gradients_batch");
}

int launch_render_batch(const BatchJob* jobs, uint32_t njobs, const JobStep* steps, uint32_t total_tiles,
                        hipStream_t stream)
{
    if (total_tiles == 0) return SADH_OK;
    hipLaunchKernelGGL((k_pixel_render_batch<HASH_SLOTS>), dim3(total_tiles), dim3(256), 0, stream, jobs, njobs, steps,
                       total_tiles);
    return launch_status("render_batch");
}

/* computeSiteStatsSimpleKernel's float outputs (utils.cu:106-231): the statistics of site i stored into
 * st[j][i] (the reference clears them first) */
int launch_site_stats(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                      const SadSite* sites, float iss, uint32_t n, float* const st[8], int width, int height,
                      int cw, int ch, hipStream_t stream)
{
    long long *lines, *unused;
    SAD_TRY(fix_workspace(stream, n, FIX_STAT_STRIDE, false, &lines, &unused));
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    A.sfix = lines;
    A.iss = iss; A.n = n; A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    SAD_TRY((launch_pixel<MODE_STATS, false, DST_FIXSTATS>(A, stream, "compute_site_stats")));
    FixOut F = {};
    for (int j = 0; j < 8; ++j) F.g[j] = st[j];
    F.gstride = 1u; F.nv = 8; F.lstride = FIX_STAT_STRIDE; F.assign = 1;
    F.lines = lines; F.n = n;
    return launch_fix_finalize(F, stream);
}

/* the fit's densify statistics, in fixed-point (lo, hi) lines of FIX_STAT_STRIDE words */
int launch_site_stats_fix(const uint32_t* cand0, const uint32_t* cand1, const float* target, const float* mask,
                          const SadSite* sites, float iss, uint32_t n, long long* stats, int width, int height,
                          int cw, int ch, hipStream_t stream)
{
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    A.sfix = stats;
    A.iss = iss; A.n = n; A.width = width; A.height = height; A.cw = cw; A.ch = ch;
    return launch_pixel<MODE_STATS, false, DST_FIXSTATS>(A, stream, "compute_site_stats(fix)");
}


/* ------------------------------------------ the allocation-free gradient --
 * sadh_compute_gradients_tiled_noalloc runs the kernels of
 * sadh_compute_gradients_tiled with their fixed-point lines in `scratch`, a
 * device array of this library instead of the per-stream workspace, so it
 * calls no allocator (for callers under graph capture or in loops that must
 * not allocate).  Every call on a device shares that device's array: calls
 * that overlap on one device race.  The array is cleared by a kernel at the
 * start of each call. */
constexpr uint32_t NOALLOC_MAX_SITES = 131072u;   /* the reference's default site budget is 128,000 */
constexpr size_t NOALLOC_WORDS = (size_t)NOALLOC_MAX_SITES * (FIX_GRAD_STRIDE + 2u);   /* 27.3 MB */

static __device__ __align__(256) long long scratch[NOALLOC_WORDS];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
long long* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

/* the device address of `scratch` on the current device, looked up once per device */
static int noalloc_scratch(long long** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail((int)e, std::string("gradients_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(SADH_ERR_UNSUPPORTED, "gradients_noalloc: device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(e != hipSuccess ? (int)e : SADH_ERR_NO_DEVICE,
                        std::string("gradients_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<long long*>(p);
    }
    *out = g_scratch_ptr[dev];
    return SADH_OK;
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

int sadh_render_cells(const uint32_t* cand0, const uint32_t* cand1, float* output, const SadSite* sites,
                      float inv_scale_sq, uint32_t site_count, int width, int height, int cand_width,
                      int cand_height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && output && sites, "render_cells: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "render_cells: bad sizes");
    return launch_render_cells(cand0, cand1, output, sites, inv_scale_sq, site_count, width, height, cand_width,
                               cand_height, to_hip(stream));
}

int sadh_render_tau_heatmap(const uint32_t* cand0, const uint32_t* cand1, float* output, const SadSite* sites,
                            uint32_t site_count, float min_tau, float mean_tau, float max_tau, float dot_radius,
                            int width, int height, int cand_width, int cand_height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && output && sites, "render_tau_heatmap: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "render_tau_heatmap: bad sizes");
    return launch_tau_heatmap(cand0, cand1, output, sites, site_count, min_tau, mean_tau, max_tau, dot_radius, width,
                              height, cand_width, cand_height, to_hip(stream));
}

int sadh_render_backward(const uint32_t* cand0, const uint32_t* cand1, const float* grad_output,
                         float* grad_sites, const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                         int width, int height, int cand_width, int cand_height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && grad_output && grad_sites && sites, "render_backward: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "render_backward: bad sizes");
    const hipStream_t S = to_hip(stream);
    long long *lines, *unused;
    SAD_TRY(fix_workspace(S, site_count, FIX_GRAD_STRIDE, false, &lines, &unused));
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.grad_out = grad_output;
    A.gfix = lines;
    A.iss = inv_scale_sq; A.n = site_count; A.width = width; A.height = height;
    A.cw = cand_width; A.ch = cand_height;
    SAD_TRY((launch_pixel<MODE_BWD, false, DST_FIX>(A, S, "render_backward")));
    /* grad_sites [N][10] in Site field order, added to */
    FixOut F = {};
    for (int j = 0; j < 10; ++j) F.g[j] = grad_sites + j;
    F.gstride = 10u; F.nv = 10; F.lstride = FIX_GRAD_STRIDE; F.assign = 0;
    F.lines = lines; F.n = site_count;
    return launch_fix_finalize(F, S);
}

int sadh_compute_gradients_tiled(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                                 const float* rendered, const float* mask, float* grad_pos_x, float* grad_pos_y,
                                 float* grad_log_tau, float* grad_radius, float* grad_color_r, float* grad_color_g,
                                 float* grad_color_b, float* grad_dir_x, float* grad_dir_y, float* grad_log_aniso,
                                 const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                                 float* removal_delta, uint32_t compute_removal, float ssim_weight, int width,
                                 int height, int cand_width, int cand_height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && target && mask && sites, "compute_gradients_tiled: null pointer");
    SADH_REQUIRE(!(ssim_weight > 0.0f) || rendered, "compute_gradients_tiled: the SSIM term needs `rendered`");
    SADH_REQUIRE(grad_pos_x && grad_pos_y && grad_log_tau && grad_radius && grad_color_r && grad_color_g &&
                     grad_color_b && grad_dir_x && grad_dir_y && grad_log_aniso,
                 "compute_gradients_tiled: null gradient array");
    SADH_REQUIRE(!compute_removal || removal_delta, "compute_gradients_tiled: removal_delta missing");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "compute_gradients_tiled: bad sizes");
    float* const g[10] = {grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r,
                          grad_color_g, grad_color_b, grad_dir_x, grad_dir_y, grad_log_aniso};
    return launch_gradients(cand0, cand1, target, mask, g, sites, inv_scale_sq, site_count, removal_delta,
                            compute_removal != 0, width, height, cand_width, cand_height, to_hip(stream), rendered,
                            ssim_weight > 0.0f ? ssim_weight : 0.0f);
}

int sadh_compute_gradients_tiled_noalloc(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                                         const float* rendered, const float* mask, float* grad_pos_x,
                                         float* grad_pos_y, float* grad_log_tau, float* grad_radius,
                                         float* grad_color_r, float* grad_color_g, float* grad_color_b,
                                         float* grad_dir_x, float* grad_dir_y, float* grad_log_aniso,
                                         const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                                         float* removal_delta, uint32_t compute_removal, float ssim_weight,
                                         int width, int height, int cand_width, int cand_height,
                                         sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && target && mask && sites, "compute_gradients_tiled_noalloc: null pointer");
    SADH_REQUIRE(!(ssim_weight > 0.0f) || rendered, "compute_gradients_tiled_noalloc: the SSIM term needs `rendered`");
    SADH_REQUIRE(grad_pos_x && grad_pos_y && grad_log_tau && grad_radius && grad_color_r && grad_color_g &&
                     grad_color_b && grad_dir_x && grad_dir_y && grad_log_aniso,
                 "compute_gradients_tiled_noalloc: null gradient array");
    SADH_REQUIRE(!compute_removal || removal_delta, "compute_gradients_tiled_noalloc: removal_delta missing");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0,
                 "compute_gradients_tiled_noalloc: bad sizes");
    if (site_count > NOALLOC_MAX_SITES)
        return fail(SADH_ERR_CAPACITY, "compute_gradients_tiled_noalloc: " + std::to_string(site_count) +
                                           " sites exceed the scratch of " + std::to_string(NOALLOC_MAX_SITES));
    long long* lines = nullptr;
    SAD_TRY(noalloc_scratch(&lines));
    const hipStream_t S = to_hip(stream);
    long long* rlines = lines + (size_t)site_count * FIX_GRAD_STRIDE;
    const size_t words = (size_t)site_count * (FIX_GRAD_STRIDE + (compute_removal ? 2u : 0u));
    SAD_TRY(launch_clear_fix(lines, words, S));
    PixelArgs A = {};
    A.cand0 = cand0; A.cand1 = cand1; A.sites = sites; A.target = target; A.mask = mask;
    A.rendered = rendered; A.ssim_weight = ssim_weight > 0.0f ? ssim_weight : 0.0f;
    A.iss = inv_scale_sq; A.n = site_count;
    A.width = width; A.height = height; A.cw = cand_width; A.ch = cand_height;
    A.gfix = lines;
    A.rfix = compute_removal ? rlines : nullptr;
    if (compute_removal) SAD_TRY((launch_pixel<MODE_GRAD, true, DST_FIX>(A, S, "compute_gradients_tiled_noalloc")));
    else SAD_TRY((launch_pixel<MODE_GRAD, false, DST_FIX>(A, S, "compute_gradients_tiled_noalloc")));
    float* const g[10] = {grad_pos_x, grad_pos_y, grad_log_tau, grad_radius, grad_color_r,
                          grad_color_g, grad_color_b, grad_dir_x, grad_dir_y, grad_log_aniso};
    FixOut F = {};
    for (int j = 0; j < 10; ++j) F.g[j] = g[j];
    F.gstride = 1u; F.nv = 10; F.lstride = FIX_GRAD_STRIDE; F.assign = 0;
    F.removal = compute_removal ? removal_delta : nullptr;
    F.lines = lines; F.rlines = rlines; F.n = site_count;
    return launch_fix_finalize(F, S);
}

int sadh_compute_gradients_tiled_aos(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                                     const float* mask, float* grad_lines, const SadSite* sites, float inv_scale_sq,
                                     uint32_t site_count, float* removal_delta, uint32_t compute_removal, int width,
                                     int height, int cand_width, int cand_height, sadh_stream_t stream)
{
    SADH_REQUIRE(cand0 && cand1 && target && mask && sites && grad_lines, "compute_gradients_tiled_aos: null pointer");
    SADH_REQUIRE(!compute_removal || removal_delta, "compute_gradients_tiled_aos: removal_delta missing");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0,
                 "compute_gradients_tiled_aos: bad sizes");
    return launch_gradients_aos(cand0, cand1, target, mask, grad_lines, sites, inv_scale_sq, site_count, removal_delta,
                                compute_removal != 0, width, height, cand_width, cand_height, to_hip(stream), nullptr,
                                0.0f);
}

int sadh_compute_site_stats_simple(const uint32_t* cand0, const uint32_t* cand1, const float* target,
                                   const float* mask, const SadSite* sites, float inv_scale_sq, uint32_t site_count,
                                   float* mass, float* energy, float* err_w, float* err_wx, float* err_wy,
                                   float* err_wxx, float* err_wxy, float* err_wyy, int width, int height,
                                   int cand_width, int cand_height, sadh_stream_t stream)
{
    float* const st[8] = {mass, energy, err_w, err_wx, err_wy, err_wxx, err_wxy, err_wyy};
    for (int j = 0; j < 8; ++j) SADH_REQUIRE(st[j], "compute_site_stats_simple: null stats array");
    SADH_REQUIRE(cand0 && cand1 && target && mask && sites, "compute_site_stats_simple: null pointer");
    SADH_REQUIRE(width > 0 && height > 0 && cand_width > 0 && cand_height > 0, "compute_site_stats_simple: bad sizes");
    return launch_site_stats(cand0, cand1, target, mask, sites, inv_scale_sq, site_count, st, width, height,
                             cand_width, cand_height, to_hip(stream));
}

}  // extern "C"
